"""The workload configurations beside the headline one, at full size on the
GPU: every signature verifies (size-independent), and a sample of each is
bit-exact against the oracle.

* reference flow: MAYO_1, 1,024 key pairs from one keypair launch, one
  32-byte message each, keypair -> signWithesk -> verifyWithepk (the flow
  of the reference's bench_stream.cu);
* distinct keys: MAYO_2, 65,536 signatures under 65,536 distinct compact
  keys, 64-byte messages, key expansion fused into signing (signWithesk) and
  verification (verifyWithepk);
* long messages: MAYO_3 and MAYO_5, 32,768 signatures each under 128
  resident keys (key j / 256), message lengths uniform in [32, 4096] bytes.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)


def fixed_messages(n, mlen, tag):
    import bench
    flat, offs, lens, _ = bench.make_workload(n, 1, mlen, mlen, tag=tag)
    return np.array(flat).reshape(n, mlen)


def test_reference_flow_1024_keypairs(engine_factory):
    v, n = 1, 1024
    eng = engine_factory(v)
    print(f"[config] MAYO_1 reference flow n={n}", flush=True)
    pk, sk = eng.keypair(n)
    msgs = fixed_messages(n, 32, "cfg1")
    sig = eng.sign_with_sk(sk, msgs)
    res = eng.verify_with_pk(pk, msgs, sig)
    assert res.tolist() == [0] * n
    for i in list(range(0, n, 97)) + [n - 1]:
        opk, osk = O.keypair(v, 0, i)
        assert pk[i].tobytes() == opk and sk[i].tobytes() == osk, f"keypair {i}"
        osig, _ = O.sign(v, O.expandsk(v, osk), msgs[i].tobytes())
        assert sig[i].tobytes() == osig, f"sign {i}"


def test_distinct_keys_fused_expansion_mayo2(engine_factory):
    v, n = 2, 65536
    eng = engine_factory(v)
    print(f"[config] MAYO_2 distinct keys n={n}", flush=True)
    pk, sk = eng.keypair(n)
    msgs = fixed_messages(n, 64, "cfg3")
    sig = eng.sign_with_sk(sk, msgs)
    res = eng.verify_with_pk(pk, msgs, sig)
    assert int((res != 0).sum()) == 0
    # a flipped message bit is rejected for a sample of items
    idx = np.arange(0, n, 4099)
    bad = msgs[idx].copy()
    bad[:, 5] ^= 0x04
    assert eng.verify_with_pk(pk[idx], bad, sig[idx]).tolist() == [1] * idx.size
    for i in list(idx) + [n - 1]:
        opk, osk = O.keypair(v, 0, int(i))
        assert sk[i].tobytes() == osk and pk[i].tobytes() == opk, f"keypair {i}"
        osig, _ = O.sign(v, O.expandsk(v, osk), msgs[i].tobytes())
        assert sig[i].tobytes() == osig, f"sign {i}"


@pytest.mark.parametrize("v", [3, 5])
def test_long_messages_resident_keys(engine_factory, v):
    import torch
    import bench
    n, nkeys = 32768, 128
    eng = engine_factory(v)
    z = eng.sizes
    print(f"[config] MAYO_{v} long messages n={n}", flush=True)
    d_esk, d_epk = eng.keypair_tables(nkeys, rand_index=0)
    flat, offs, lens, _ = bench.make_workload(n, nkeys, 32, 4096, tag="cfg4")
    kidx = (np.arange(n) // 256).astype(np.int32)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.array(a)).to(dev)
    d_kidx, d_msgs, d_off, d_len = t(kidx), t(flat), t(offs), t(lens)
    sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    ws = eng.sign_workspace(n)
    eng.sign_table(d_esk, d_kidx, d_msgs, d_off, d_len, sigs, ws)
    res = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    eng.verify_table(d_epk, d_kidx, d_msgs, d_off, d_len, sigs, res, eng.verify_workspace(n))
    torch.cuda.synchronize()
    assert int((res != 0).sum().item()) == 0
    att = eng.sign_attempts(ws, n)
    assert (att >= 1).all()
    # oracle on a strided sample (every key block, the longest and shortest messages)
    sample = np.unique(np.concatenate([np.arange(0, n, 97), [int(lens.argmax()), int(lens.argmin())]]))
    g = sigs.cpu().numpy()[sample]
    osigs = O.sign_batch(v, d_esk.cpu().numpy(), kidx[sample], flat, offs[sample], lens[sample],
                         threads=THREADS)
    bad = np.where((g != osigs).any(axis=1))[0]
    assert bad.size == 0, f"mismatching items {sample[bad[:8]].tolist()}"


def test_headline_batch_through_captured_graphs(engine_factory):
    """The headline workload at full size: MAYO_1, 65,536 signatures under
    64 resident keys (key j mod 64, prepared tables), message lengths
    uniform in [32, 1024] bytes (the bench's generator), signed and verified
    through the captured sign and verify graphs, replayed twice with the
    outputs cleared in between: both replays give the eager bytes, every
    signature verifies, a flipped signature byte is rejected for a strided
    sample, and all 65,536 signatures equal the oracle's."""
    import torch
    import bench
    v, n, nkeys = 1, 65536, 64
    eng = engine_factory(v)
    z = eng.sizes
    print(f"[config] MAYO_1 headline batch through graphs n={n}", flush=True)
    d_esk, d_epk = eng.keypair_tables(nkeys, rand_index=0)
    p_esk, p_epk = eng.prepare_tables(d_esk, d_epk)
    flat, offs, lens, kidx = bench.make_workload(n, nkeys, 32, 1024)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.array(a)).to(dev)
    d_kidx, d_msgs, d_off, d_len = t(kidx), t(flat), t(offs), t(lens)
    ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    eng.sign_prepared(p_esk, d_kidx, d_msgs, d_off, d_len, ref, eng.sign_workspace(n))
    torch.cuda.synchronize()
    sigs = torch.zeros_like(ref)
    res = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    ws_s, ws_v = eng.sign_workspace(n), eng.verify_workspace(n)
    gs = eng.sign_graph(p_esk, d_kidx, d_msgs, d_off, d_len, sigs, ws_s)
    gv = eng.verify_graph(p_epk, d_kidx, d_msgs, d_off, d_len, sigs, res, ws_v)
    try:
        for rep in range(2):
            sigs.zero_()
            res.fill_(7)
            torch.cuda.synchronize()
            gs.launch()
            gv.launch()
            torch.cuda.synchronize()
            assert torch.equal(sigs, ref), f"sign graph replay {rep}"
            assert int((res != 0).sum().item()) == 0, f"verify graph replay {rep}"
        # rejections through the verification graph
        idx = torch.arange(0, n, 257, device=dev)
        sigs[idx, 100] ^= 0x01
        gv.launch()
        torch.cuda.synchronize()
        r = res.cpu().numpy()
        assert r[idx.cpu().numpy()].tolist() == [1] * idx.numel()
        assert int(r.sum()) == idx.numel()
    finally:
        gs.close()
        gv.close()
    att = eng.sign_attempts(ws_s, n)
    assert (att >= 1).all()
    osigs = O.sign_batch(v, d_esk.cpu().numpy(), kidx, flat, offs, lens, threads=THREADS)
    g = ref.cpu().numpy()
    bad = np.where((g != osigs).any(axis=1))[0]
    assert bad.size == 0, f"mismatching items {bad[:8].tolist()}"
