"""GPU parity: every drop-in entry point and the batched key-table form
against the CPU oracle, bit for bit, for all four parameter sets.

Sizes are small so the oracle finishes in seconds. Inputs are seeded.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

VERSIONS = (1, 2, 3, 5)


def msgs_for(version, batch, mlen, tag="parity"):
    out = np.zeros((batch, mlen), np.uint8)
    for j in range(batch):
        out[j] = np.frombuffer(hashlib.shake_256(f"{tag}/{version}/{j}".encode()).digest(mlen), np.uint8)
    return out


@pytest.fixture(scope="module", params=VERSIONS, ids=lambda v: f"MAYO_{v}")
def keys(request, engine_factory):
    v = request.param
    eng = engine_factory(v)
    batch = 3
    print(f"[parity] MAYO_{v} keypair batch={batch}", flush=True)
    pk, sk = eng.keypair(batch, rand_index=0)
    return v, eng, pk, sk


def test_keypair_matches_oracle(keys):
    v, eng, pk, sk = keys
    for i in range(pk.shape[0]):
        opk, osk = O.keypair(v, 0, i)
        assert sk[i].tobytes() == osk, f"sk mismatch item {i}"
        assert pk[i].tobytes() == opk, f"pk mismatch item {i}"


def test_keypair_rand_index(engine_factory):
    eng = engine_factory(1)
    pk, sk = eng.keypair(2, rand_index=1000)
    for i in range(2):
        opk, osk = O.keypair(1, 1000, i)
        assert sk[i].tobytes() == osk and pk[i].tobytes() == opk


def test_expandsk_expandpk_match_oracle(keys):
    v, eng, pk, sk = keys
    esk = eng.expandsk(sk)
    epk = eng.expandpk(pk)
    for i in range(sk.shape[0]):
        assert esk[i].tobytes() == O.expandsk(v, sk[i].tobytes()), f"esk mismatch item {i}"
        assert epk[i].tobytes() == O.expandpk(v, pk[i].tobytes()), f"epk mismatch item {i}"


@pytest.mark.parametrize("mlen", [32, 200])
def test_sign_and_verify_match_oracle(keys, mlen):
    v, eng, pk, sk = keys
    batch = sk.shape[0]
    esk = np.stack([np.frombuffer(O.expandsk(v, sk[i].tobytes()), np.uint8) for i in range(batch)])
    epk = np.stack([np.frombuffer(O.expandpk(v, pk[i].tobytes()), np.uint8) for i in range(batch)])
    msgs = msgs_for(v, batch, mlen)
    print(f"[parity] MAYO_{v} sign mlen={mlen}", flush=True)
    sig = eng.sign(esk, msgs)
    for i in range(batch):
        osig, _ = O.sign(v, esk[i].tobytes(), msgs[i].tobytes())
        assert sig[i].tobytes() == osig, f"signature mismatch item {i}"
    sig2 = eng.sign_with_sk(sk, msgs)
    assert np.array_equal(sig2, sig)
    print(f"[parity] MAYO_{v} verify mlen={mlen}", flush=True)
    assert eng.verify(epk, msgs, sig).tolist() == [0] * batch
    assert eng.verify_with_pk(pk, msgs, sig).tolist() == [0] * batch
    bad = sig.copy()
    bad[:, 0] ^= 1
    assert eng.verify(epk, msgs, bad).tolist() == [1] * batch
    badm = msgs.copy()
    badm[:, -1] ^= 0x80
    assert eng.verify_with_pk(pk, badm, sig).tolist() == [1] * batch


def test_table_form_shared_keys_ragged(engine_factory):
    import torch
    v = 1
    eng = engine_factory(v)
    z = eng.sizes
    nkeys, n = 3, 12
    pk, sk = eng.keypair(nkeys)
    esk = eng.expandsk(sk)
    epk = eng.expandpk(pk)
    rng = np.random.default_rng(7)
    lens = rng.integers(0, 300, n).astype(np.int32)
    lens[0], lens[1] = 0, 136  # empty message and exactly one SHAKE256 block
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = (np.arange(n) % nkeys).astype(np.int32)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    eng.sign_table(t(esk), t(kidx), t(flat), t(offs), t(lens), sigs, eng.sign_workspace(n))
    res = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    eng.verify_table(t(epk), t(kidx), t(flat), t(offs), t(lens), sigs, res, eng.verify_workspace(n))
    torch.cuda.synchronize()
    sigs = sigs.cpu().numpy()
    for i in range(n):
        m = flat[offs[i]:offs[i] + lens[i]].tobytes()
        osig, _ = O.sign(v, esk[kidx[i]].tobytes(), m)
        assert sigs[i].tobytes() == osig, f"item {i} len {lens[i]}"
    assert res.cpu().numpy().tolist() == [0] * n


def test_table_form_large_batch_includes_retries(engine_factory):
    """8,192 signatures over 16 keys: every one bit-exact vs the oracle,
    including the items whose first attempt hit a rank failure (about 1 in
    4,000 for MAYO_1), and every one verifies."""
    import os
    import torch
    import bench
    v = 1
    eng = engine_factory(v)
    z = eng.sizes
    n, nkeys = 8192, 16
    pk, sk = eng.keypair(nkeys)
    d_esk, d_epk = eng.expand_tables(sk, pk)
    flat, offs, lens, kidx = bench.make_workload(n, nkeys, 32, 1024, tag="retry-test")
    dev = eng.device
    t = lambda a: torch.from_numpy(np.array(a)).to(dev)
    sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    ws = eng.sign_workspace(n)
    print(f"[parity] table sign n={n}", flush=True)
    eng.sign_table(d_esk, t(kidx), t(flat), t(offs), t(lens), sigs, ws)
    res = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    eng.verify_table(d_epk, t(kidx), t(flat), t(offs), t(lens), sigs, res, eng.verify_workspace(n))
    torch.cuda.synchronize()
    att = eng.sign_attempts(ws, n)
    print(f"[parity] items with retries: {int((att > 1).sum())}", flush=True)
    assert (att >= 1).all()
    assert res.cpu().numpy().tolist() == [0] * n
    esk = d_esk.cpu().numpy()
    osigs = O.sign_batch(v, esk, kidx, flat, offs, lens, threads=min(os.cpu_count() or 1, 16))
    g = sigs.cpu().numpy()
    bad = np.where((g != osigs).any(axis=1))[0]
    assert bad.size == 0, f"mismatching items {bad[:10].tolist()} attempts {att[bad[:10]].tolist()}"


def test_device_key_tables_match_oracle(engine_factory):
    """keypair_tables: keys generated and expanded without leaving the device."""
    for v in (1, 2):
        eng = engine_factory(v)
        esk, epk = eng.keypair_tables(3, rand_index=5)
        esk, epk = esk.cpu().numpy(), epk.cpu().numpy()
        for i in range(3):
            opk, osk = O.keypair(v, 5, i)
            assert esk[i].tobytes() == O.expandsk(v, osk)
            assert epk[i].tobytes() == O.expandpk(v, opk)
