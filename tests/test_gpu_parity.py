"""GPU parity: every drop-in entry point and the batched key-table form
against the CPU oracle, bit for bit, for all four parameter sets.

Sizes are small so the oracle finishes in seconds. Inputs are seeded.
"""
import hashlib

import numpy as np
import pytest

from cumayo_amd.engine import a128
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
    4,000 for MAYO_1), and every one verifies. The oracle finds exactly two
    natural retries in this workload (items 2768 and 6420, second attempt
    each); the test pins them so that the natural-retry path stays covered."""
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
    assert np.flatnonzero(att > 1).tolist() == [2768, 6420] and int(att.max()) == 2
    assert res.cpu().numpy().tolist() == [0] * n
    # the same batch over prepared key tables (the retry pass included)
    p_esk, _ = eng.prepare_tables(d_esk, None)
    psigs = torch.zeros_like(sigs)
    pws = eng.sign_workspace(n)
    eng.sign_prepared(p_esk, t(kidx), t(flat), t(offs), t(lens), psigs, pws)
    torch.cuda.synchronize()
    assert np.array_equal(eng.sign_attempts(pws, n), att)
    esk = d_esk.cpu().numpy()
    osigs = O.sign_batch(v, esk, kidx, flat, offs, lens, threads=min(os.cpu_count() or 1, 16))
    g = sigs.cpu().numpy()
    bad = np.where((g != osigs).any(axis=1))[0]
    assert bad.size == 0, f"mismatching items {bad[:10].tolist()} attempts {att[bad[:10]].tolist()}"
    assert np.array_equal(psigs.cpu().numpy(), g)


@pytest.mark.parametrize("v", (1, 3, 5), ids=lambda v: f"MAYO_{v}")
def test_retry_pass_many_items(engine_factory, v):
    """Forced rank failures of attempts 0 and 1 for 300 items: every item
    goes through the retry pass (more items than its workgroups, so they
    take several each) and comes out as the oracle's signature under the
    same forced failures, with attempts >= 3."""
    import torch
    eng = engine_factory(v)
    z = eng.sizes
    n, nkeys, ff = 300, 3, 2
    esk, _ = eng.keypair_tables(nkeys, rand_index=7)
    pesk, _ = eng.prepare_tables(esk, None)
    rng = np.random.default_rng(11 + v)
    lens = rng.integers(0, 200, n).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = (np.arange(n) % nkeys).astype(np.int32)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    ws = eng.sign_workspace(n)
    print(f"[parity] forced retries MAYO_{v} n={n}", flush=True)
    eng.sign_prepared(pesk, t(kidx), t(flat), t(offs), t(lens), sigs, ws, stages=3 | (ff << 8))
    torch.cuda.synchronize()
    att = eng.sign_attempts(ws, n)
    assert (att >= ff + 1).all(), att.min()
    g = sigs.cpu().numpy()
    esk_h = esk.cpu().numpy()
    for i in list(range(0, n, 23)) + [n - 1]:
        osig, oatt = O.sign(v, esk_h[kidx[i]].tobytes(), flat[offs[i]:offs[i] + lens[i]].tobytes(), fail_first=ff)
        assert g[i].tobytes() == osig, f"item {i}"
        assert int(att[i]) == oatt, f"item {i}: attempts {att[i]} vs {oatt}"


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


@pytest.mark.parametrize("n", (1024, 2048, 4096, 8192))
def test_device_key_tables_every_aes_launch_shape(engine_factory, n):
    """The AES-CTR expansion gives each key 8, 4, 2 and 1 waves at these
    batch sizes (n times the waves per key is about 8,192) and so takes 2,
    4, 8 and 16 keys per group: rows on both sides of every group size's
    edges and the last row equal the oracle's expanded keys."""
    v = 2
    eng = engine_factory(v)
    print(f"[parity] MAYO_{v} device key tables n={n}", flush=True)
    esk, epk = eng.keypair_tables(n, rand_index=21)
    rows = sorted({0, 1, 3, 4, 7, 8, 15, 16, n // 2 - 1, n // 2, n - 1})
    esk, epk = esk[rows].cpu().numpy(), epk[rows].cpu().numpy()
    for k, i in enumerate(rows):
        opk, osk = O.keypair(v, 21, i)
        assert esk[k].tobytes() == O.expandsk(v, osk), (n, i)
        assert epk[k].tobytes() == O.expandpk(v, opk), (n, i)


def _planar_words(entries):
    """Each little-endian dword of the entries (last axis, bytes) with bit
    4e + d moved to bit 8d + e: byte d then holds bit d of the 8 nibbles."""
    w = np.ascontiguousarray(entries).view("<u4")
    out = np.zeros_like(w)
    for e in range(8):
        for d in range(4):
            out |= ((w >> np.uint32(4 * e + d)) & np.uint32(1)) << np.uint32(8 * d + e)
    return out.view(np.uint8)


def _prepared_layout(z, table, kind):
    """numpy restatement of the prepared layout: entries padded to a
    multiple of 4 bytes and stored planar, blocks 16-byte aligned, rows a
    multiple of 128."""
    mb = z.m // 2
    mbp = (mb + 3) // 4 * 4
    v, o = z.n - z.o, z.o
    tri_v, tri_o = v * (v + 1) // 2, o * (o + 1) // 2
    if kind == "esk":
        hdr = z.sk_bytes + z.o_bytes
        blocks = [(hdr, tri_v), (hdr + tri_v * mb, v * o)]
        dst0 = (hdr + 15) // 16 * 16
    else:
        hdr = 0
        blocks = [(0, tri_v), (tri_v * mb, v * o), (tri_v * mb + v * o * mb, tri_o)]
        dst0 = 0
    size = dst0 + sum(c for _, c in blocks) * mbp
    size = (size + 127) // 128 * 128
    out = np.zeros((table.shape[0], size), np.uint8)
    out[:, :hdr] = table[:, :hdr]
    d = dst0
    for s0, cnt in blocks:
        src = table[:, s0:s0 + cnt * mb].reshape(-1, cnt, mb)
        padded = np.zeros((table.shape[0], cnt, mbp), np.uint8)
        padded[:, :, :mb] = src
        out[:, d:d + cnt * mbp] = _planar_words(padded).reshape(table.shape[0], cnt * mbp)
        d += cnt * mbp
    return out


@pytest.mark.parametrize("v", VERSIONS, ids=lambda v: f"MAYO_{v}")
def test_prepared_tables_match_reference_layout(engine_factory, v):
    """Prepared key tables: the layout matches its numpy restatement, and
    signing / verification over them give the reference-layout results and
    the oracle's signatures, for ragged messages (empty, one SHAKE block)."""
    import torch
    eng = engine_factory(v)
    z = eng.sizes
    nkeys, n = 5, 300
    esk, epk = eng.keypair_tables(nkeys, rand_index=21)
    pesk, pepk = eng.prepare_tables(esk, epk)
    torch.cuda.synchronize()
    assert np.array_equal(pesk.cpu().numpy(), _prepared_layout(z, esk.cpu().numpy(), "esk"))
    assert np.array_equal(pepk.cpu().numpy(), _prepared_layout(z, epk.cpu().numpy(), "epk"))
    rng = np.random.default_rng(100 + v)
    lens = rng.integers(0, 700, n).astype(np.int32)
    lens[:2] = [0, 136]
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = rng.integers(0, nkeys, n).astype(np.int32)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_k, d_m, d_o, d_l = t(kidx), t(flat), t(offs), t(lens)
    s_ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    s_prep = torch.zeros_like(s_ref)
    print(f"[parity] prepared MAYO_{v} n={n}", flush=True)
    eng.sign_table(esk, d_k, d_m, d_o, d_l, s_ref, eng.sign_workspace(n))
    eng.sign_prepared(pesk, d_k, d_m, d_o, d_l, s_prep, eng.sign_workspace(n))
    r_ref = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    r_prep = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    eng.verify_table(epk, d_k, d_m, d_o, d_l, s_prep, r_ref, eng.verify_workspace(n))
    eng.verify_prepared(pepk, d_k, d_m, d_o, d_l, s_prep, r_prep, eng.verify_workspace(n))
    bad = s_prep.clone()
    bad[::3, 7] ^= 0x20
    r_bad = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    eng.verify_prepared(pepk, d_k, d_m, d_o, d_l, bad, r_bad, eng.verify_workspace(n))
    torch.cuda.synchronize()
    assert torch.equal(s_ref, s_prep)
    assert r_ref.cpu().numpy().tolist() == [0] * n
    assert r_prep.cpu().numpy().tolist() == [0] * n
    expect_bad = np.zeros(n, np.uint8)
    expect_bad[::3] = 1
    assert np.array_equal(r_bad.cpu().numpy(), expect_bad)
    esk_h = esk.cpu().numpy()
    sp = s_prep.cpu().numpy()
    for i in list(range(0, n, 37)) + [0, 1]:
        osig, _ = O.sign(v, esk_h[kidx[i]].tobytes(), flat[offs[i]:offs[i] + lens[i]].tobytes())
        assert sp[i].tobytes() == osig, f"item {i}"


def test_pitched_copies_pinned_and_pageable(engine_factory):
    """The HtD / DtH entry points give the same rows from pinned host memory
    (copy kernel) and from pageable host memory (hipMemcpy2DAsync)."""
    import torch
    v = 1
    eng = engine_factory(v)
    z = eng.sizes
    batch = 37
    row = a128(z.sk_bytes) + a128(z.pk_bytes) + a128(z.p_bytes)
    pool, pitch = eng.pool(batch, row)
    st = eng.stream_handle()
    eng.lib.crypto_sign_keypair(pool.data_ptr(), pitch, batch, st, 3)
    pk_pin, sk_pin = eng.pinned((batch, z.pk_bytes)), eng.pinned((batch, z.sk_bytes))
    pk_pag = np.zeros((batch, z.pk_bytes), np.uint8)
    sk_pag = np.zeros((batch, z.sk_bytes), np.uint8)
    eng.lib.crypto_sign_keypair_DtH(pk_pin.data_ptr(), sk_pin.data_ptr(), pool.data_ptr(), pitch, batch, st)
    eng.lib.crypto_sign_keypair_DtH(pk_pag.ctypes.data, sk_pag.ctypes.data, pool.data_ptr(), pitch, batch, st)
    torch.cuda.synchronize()
    eng.check()
    assert np.array_equal(pk_pin.numpy(), pk_pag) and np.array_equal(sk_pin.numpy(), sk_pag)
    opk, osk = O.keypair(v, 3, batch - 1)
    assert pk_pag[-1].tobytes() == opk and sk_pag[-1].tobytes() == osk
    # upload the compact keys both ways and expand: identical esk rows
    outs = []
    for src in (eng.pinned(sk_pag.shape, sk_pag), sk_pag):
        ep, epitch = eng.pool(batch, z.esk_bytes)
        ptr = src.data_ptr() if hasattr(src, "data_ptr") else src.ctypes.data
        eng.lib.crypto_sign_expandsk_HtD(ptr, ep.data_ptr(), epitch, batch, st)
        eng.lib.crypto_sign_expandsk(ep.data_ptr(), epitch, batch, st)
        torch.cuda.synchronize()
        outs.append(ep[:, :z.esk_bytes].cpu().numpy())
    eng.check()
    assert np.array_equal(outs[0], outs[1])
    assert outs[0][-1].tobytes() == O.expandsk(v, osk)
