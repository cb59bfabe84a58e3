"""GPU against the committed known answers: keys from the seed rule, the
expanded keys, signatures over ragged messages (empty, one byte, one SHAKE
block and one past it), and the forced-failure hook (retry at ctr 1 and 3,
exhaustion after 255 attempts)."""
import hashlib
import json
import os

import numpy as np
import pytest

from tests.golden.make_golden import LENS, message

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                   "mayo_golden.json")))


def sha(b) -> str:
    return hashlib.sha256(bytes(b)).hexdigest()


@pytest.mark.parametrize("v", [1, 2, 3, 5], ids=lambda v: f"MAYO_{v}")
def test_gpu_matches_golden(engine_factory, v):
    import torch
    eng = engine_factory(v)
    z = eng.sizes
    gold = GOLD[f"MAYO_{v}"]
    print(f"[golden] MAYO_{v}", flush=True)
    pk, sk = eng.keypair(3, rand_index=0)
    for k in gold["keys"]:
        i = k["item"]
        assert sk[i].tobytes().hex() == k["sk"]
        assert sha(pk[i]) == k["pk_sha256"]
    esk, epk = eng.expand_tables(sk, pk)
    esk_h, epk_h = esk.cpu().numpy(), epk.cpu().numpy()
    for k in gold["keys"]:
        assert sha(esk_h[k["item"]]) == k["esk_sha256"]
        assert sha(epk_h[k["item"]]) == k["epk_sha256"]
    # signatures, table form over prepared keys
    pesk, pepk = eng.prepare_tables(esk, epk)
    msgs = [message(v, s["j"], s["msg_len"]) for s in gold["signatures"]]
    lens = np.array([len(m) for m in msgs], np.int32)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = np.frombuffer(b"".join(msgs) + b"\0", np.uint8)
    kidx = np.array([s["key"] for s in gold["signatures"]], np.int32)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.array(a)).to(dev)
    n = len(msgs)
    sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    ws = eng.sign_workspace(n)
    eng.sign_prepared(pesk, t(kidx), t(flat), t(offs), t(lens), sigs, ws)
    res = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    eng.verify_prepared(pepk, t(kidx), t(flat), t(offs), t(lens), sigs, res, eng.verify_workspace(n))
    torch.cuda.synchronize()
    g = sigs.cpu().numpy()
    att = eng.sign_attempts(ws, n)
    for i, s in enumerate(gold["signatures"]):
        assert sha(g[i]) == s["sig_sha256"], f"signature {i} (len {s['msg_len']})"
        if s["sig"]:
            assert g[i].tobytes().hex() == s["sig"]
        assert int(att[i]) == s["attempts"]
    assert res.cpu().numpy().tolist() == [0] * n
    # forced failures: the first f attempts report a singular system
    m = message(v, 100, 64)
    for f in gold["forced_failures"]:
        fl = np.frombuffer(m + b"\0", np.uint8)
        one = torch.zeros((1, z.sig_bytes), dtype=torch.uint8, device=dev)
        w1 = eng.sign_workspace(1)
        eng.sign_prepared(pesk, t(np.zeros(1, np.int32)), t(fl), t(np.zeros(1, np.int64)),
                          t(np.array([64], np.int32)), one, w1, stages=3 | (f["fail_first"] << 8))
        torch.cuda.synchronize()
        assert sha(one.cpu().numpy()[0]) == f["sig_sha256"], f"fail_first {f['fail_first']}"
        a = int(eng.sign_attempts(w1, 1)[0])
        assert a == (0 if f["attempts"] == 256 else f["attempts"])
