"""NIST wire format: cumayo_crypto_sign writes sm = signature || message,
cumayo_crypto_sign_open recovers and verifies it. The signature part equals
the detached signature (and the oracle's); tampered or truncated sm are
rejected with mlen -1."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("v", [1, 3])
def test_crypto_sign_and_open_roundtrip(engine_factory, v):
    import torch
    eng = engine_factory(v)
    z = eng.sizes
    nkeys, n = 4, 120
    esk, epk = eng.keypair_tables(nkeys, rand_index=31)
    pesk, pepk = eng.prepare_tables(esk, epk)
    rng = np.random.default_rng(5 + v)
    lens = rng.integers(0, 400, n).astype(np.int32)
    lens[:2] = [0, 1]
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = rng.integers(0, nkeys, n).astype(np.int32)
    sm_len = (lens + z.sig_bytes).astype(np.int32)
    sm_off = (offs + np.arange(n, dtype=np.int64) * z.sig_bytes).astype(np.int64)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_k, d_m, d_o, d_l = t(kidx), t(flat), t(offs), t(lens)
    sm = torch.zeros(int(sm_len.sum()) + 1, dtype=torch.uint8, device=dev)
    print(f"[wire] MAYO_{v} n={n}", flush=True)
    eng.crypto_sign(pesk, d_k, d_m, d_o, d_l, sm, t(sm_off))
    det = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    eng.sign_prepared(pesk, d_k, d_m, d_o, d_l, det, eng.sign_workspace(n))
    torch.cuda.synchronize()
    sm_h, det_h = sm.cpu().numpy(), det.cpu().numpy()
    for i in range(n):
        row = sm_h[sm_off[i]:sm_off[i] + sm_len[i]]
        assert row[:z.sig_bytes].tobytes() == det_h[i].tobytes(), f"sig part {i}"
        assert row[z.sig_bytes:].tobytes() == flat[offs[i]:offs[i] + lens[i]].tobytes(), f"msg part {i}"
    esk_h = esk.cpu().numpy()
    for i in (0, 1, n - 1):
        osig, _ = O.sign(v, esk_h[kidx[i]].tobytes(), flat[offs[i]:offs[i] + lens[i]].tobytes())
        assert det_h[i].tobytes() == osig

    def open_(sm_dev, lens_dev):
        out = torch.zeros(int(lens.sum()) + 1, dtype=torch.uint8, device=dev)
        mlen = torch.full((n,), -7, dtype=torch.int32, device=dev)
        res = torch.full((n,), 9, dtype=torch.uint8, device=dev)
        eng.crypto_sign_open(pepk, d_k, sm_dev, t(sm_off), lens_dev, out, d_o, mlen, res)
        torch.cuda.synchronize()
        return out.cpu().numpy(), mlen.cpu().numpy(), res.cpu().numpy()

    out, mlen, res = open_(sm, t(sm_len))
    assert res.tolist() == [0] * n
    assert np.array_equal(mlen, lens)
    assert np.array_equal(out[:int(lens.sum())], flat[:int(lens.sum())])

    # tamper: one signature byte (items 0 mod 3), one message byte (1 mod 3,
    # non-empty), and a truncated sm (item n-1: shorter than a signature)
    bad = sm_h.copy()
    expect = np.zeros(n, np.uint8)
    for i in range(0, n, 3):
        bad[sm_off[i] + 11] ^= 0x40
        expect[i] = 1
    for i in range(1, n, 3):
        if lens[i]:
            bad[sm_off[i] + z.sig_bytes + lens[i] - 1] ^= 0x01
            expect[i] = 1
    short = sm_len.copy()
    short[n - 1] = z.sig_bytes - 1
    expect[n - 1] = 1
    out, mlen, res = open_(t(bad), t(short))
    assert np.array_equal(res, expect)
    assert np.array_equal(mlen, np.where(expect == 0, lens, -1))
