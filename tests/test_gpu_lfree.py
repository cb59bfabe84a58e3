"""crypto_sign_signWithesk for MAYO_2 signs without L: the pool rows keep P2
where the reference's expansion writes L = (P1 + P1^T) O + P2
(expandsk.cu:37-81), and the signing kernel forms M_j = V_j^T P2 + (W_j)^T O
from W_j = (P1 + P1^T) V_j instead of V_j^T L (sign.cu:256-330). The other
sets expand L and sign with it (faster there).

* every parameter set: distinct compact keys, one message each, the
  signWithesk signatures equal, item by item, those of crypto_sign_expandsk
  followed by crypto_sign_sign (which expand and use L) over the same keys
  and messages, and a sample equals the oracle's;
* forced rank failures of the first two attempts (test hook) through the
  L-free kernel and the L kernel: third-attempt signatures equal to each
  other and to the oracle's under the same forced failures;
* the two natural second attempts of the MAYO_1 retry workload through the
  drop-in path, against the oracle;
* the L-free pool after the call holds P1 || P2 in the esk's P1 || L place.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = {1: 2048, 2: 8192, 3: 1024, 5: 512}


def messages(n, mlen, tag):
    import hashlib
    return np.frombuffer(hashlib.shake_256(f"lfree/{tag}".encode()).digest(n * mlen),
                         np.uint8).reshape(n, mlen).copy()


@pytest.mark.parametrize("v", [1, 2, 3, 5], ids=lambda v: f"MAYO_{v}")
def test_lfree_equals_L_path(engine_factory, v):
    n = SIZES[v]
    eng = engine_factory(v)
    print(f"[lfree] MAYO_{v} n={n}", flush=True)
    pk, sk = eng.keypair(n, rand_index=3)
    msgs = messages(n, 48, f"eq{v}")
    sig = eng.sign_with_sk(sk, msgs)
    esk = eng.expandsk(sk)
    sig_L = eng.sign(esk, msgs)
    bad = np.where((sig != sig_L).any(axis=1))[0]
    assert bad.size == 0, f"MAYO_{v}: items {bad[:10].tolist()} differ from the L path"
    for i in list(range(0, n, n // 8)) + [n - 1]:
        osk = sk[i].tobytes()
        oesk = O.expandsk(v, osk)
        assert esk[i].tobytes() == oesk, f"expandsk {i}"
        osig, _ = O.sign(v, oesk, msgs[i].tobytes())
        assert sig[i].tobytes() == osig, f"sign {i}"
    res = eng.verify_with_pk(pk, msgs, sig)
    assert int((res != 0).sum()) == 0


def test_lfree_forced_retries(engine_factory):
    v, n, ff = 2, 300, 2
    eng = engine_factory(v)
    assert eng.lib.cumayo_signwithesk_lfree() == 1
    _, sk = eng.keypair(n, rand_index=11)
    msgs = messages(n, 33, "ff")
    print(f"[lfree] forced retries MAYO_{v} n={n} fail_first={ff}", flush=True)
    eng.lib.cumayo_test_pitched_fail_first(ff)
    try:
        sig = eng.sign_with_sk(sk, msgs)
        esk = eng.expandsk(sk)
        sig_L = eng.sign(esk, msgs)
    finally:
        eng.lib.cumayo_test_pitched_fail_first(0)
    assert np.array_equal(sig, sig_L)
    for i in list(range(0, n, 37)) + [n - 1]:
        osig, oatt = O.sign(v, esk[i].tobytes(), msgs[i].tobytes(), fail_first=ff)
        assert oatt == ff + 1
        assert sig[i].tobytes() == osig, f"item {i}"
    # the hook is off again: first-attempt signatures
    assert eng.sign_with_sk(sk[:1], msgs[:1])[0].tobytes() == O.sign(v, esk[0].tobytes(), msgs[0].tobytes())[0]


def test_lfree_natural_retries_mayo1(engine_factory):
    """Items 2768 and 6420 of the bench generator's "retry-test" workload
    (16 keys, key j mod 16, lengths in [32, 1024]) need a second attempt;
    each is signed here with its compact key through crypto_sign_signWithesk."""
    import bench
    v = 1
    eng = engine_factory(v)
    pk, sk = eng.keypair(16)
    flat, offs, lens, kidx = bench.make_workload(8192, 16, 32, 1024, tag="retry-test")
    flat = np.asarray(flat)
    for i in (2768, 6420):
        k = int(kidx[i])
        m = flat[offs[i]:offs[i] + lens[i]].copy()
        print(f"[lfree] retry item {i} key {k} len {lens[i]}", flush=True)
        sig = eng.sign_with_sk(sk[k:k + 1], m.reshape(1, -1))
        osig, oatt = O.sign(v, O.expandsk(v, sk[k].tobytes()), m.tobytes())
        assert oatt == 2, f"item {i}: the oracle takes {oatt} attempts"
        assert sig[0].tobytes() == osig, f"item {i}"


def test_lfree_pool_keeps_P2(engine_factory):
    """After crypto_sign_signWithesk a row holds sk || O || P1 || P2 (the
    expanded key with P2 in L's place), then the message and signature."""
    from cumayo_amd.engine import a128
    v, n, mlen = 2, 4, 40
    eng = engine_factory(v)
    assert eng.lib.cumayo_signwithesk_lfree() == 1
    z = eng.sizes
    pk, sk = eng.keypair(n, rand_index=9)
    msgs = messages(n, mlen, "pool")
    row = a128(z.esk_bytes) + a128(mlen) + a128(z.sig_bytes) + a128(z.scratch_bytes)
    pool, pitch = eng.pool(n, row)
    h_sk, h_msg = eng.pinned(sk.shape, sk), eng.pinned(msgs.shape, msgs)
    st = eng.stream_handle()
    eng.lib.crypto_sign_signWithesk_HtD(h_sk.data_ptr(), h_msg.data_ptr(), mlen, pool.data_ptr(), pitch, n, st)
    eng.lib.crypto_sign_signWithesk(pool.data_ptr(), mlen, pitch, n, st)
    eng.torch.cuda.synchronize(eng.device)
    eng.check()
    rows = pool.cpu().numpy()
    p1_off = z.sk_bytes + z.o_bytes
    for i in range(n):
        oesk = O.expandsk(v, sk[i].tobytes())
        # sk || O || P1 as expanded, then P2 (that of the expanded public key)
        assert rows[i, :p1_off + z.p1_bytes].tobytes() == oesk[:p1_off + z.p1_bytes], f"row {i}"
        p2 = O.expandpk(v, pk[i].tobytes())[z.p1_bytes:z.p_bytes]
        assert rows[i, p1_off + z.p1_bytes:p1_off + z.p_bytes].tobytes() == p2, f"row {i} P2"
        assert oesk[p1_off + z.p1_bytes:p1_off + z.p_bytes] != p2  # L differs from P2
        assert rows[i, a128(z.esk_bytes):a128(z.esk_bytes) + mlen].tobytes() == msgs[i].tobytes()


def test_lfree_unaligned_pool(engine_factory):
    """A pool whose pitch is not a multiple of 4 takes the unaligned-load
    kernels (the aligned ones need 4-byte rows): MAYO_2 signWithesk and
    verifyWithepk over such a pool give the same bytes as over an aligned one."""
    import torch
    from cumayo_amd.engine import a128
    v, n, mlen = 2, 96, 45
    eng = engine_factory(v)
    z = eng.sizes
    pk, sk = eng.keypair(n, rand_index=13)
    msgs = messages(n, mlen, "odd")
    ref = eng.sign_with_sk(sk, msgs)
    row = a128(z.esk_bytes) + a128(mlen) + a128(z.sig_bytes) + a128(z.scratch_bytes)
    pitch = row + 3
    buf = torch.zeros(n * pitch + 64, dtype=torch.uint8, device=eng.device)
    base = buf.data_ptr() + 1  # odd base as well
    h_sk, h_msg = eng.pinned(sk.shape, sk), eng.pinned(msgs.shape, msgs)
    sig = eng.pinned((n, z.sig_bytes))
    st = eng.stream_handle()
    print(f"[lfree] unaligned pool pitch={pitch}", flush=True)
    eng.lib.crypto_sign_signWithesk_HtD(h_sk.data_ptr(), h_msg.data_ptr(), mlen, base, pitch, n, st)
    eng.lib.crypto_sign_signWithesk(base, mlen, pitch, n, st)
    eng.lib.crypto_sign_signWithesk_DtH(sig.data_ptr(), base, mlen, pitch, n, st)
    eng.torch.cuda.synchronize(eng.device)
    eng.check()
    assert np.array_equal(sig.numpy(), ref)
    # verification over an unaligned pool (compact pk expanded in the call)
    epk_len = z.p_bytes + z.pk_bytes - 16
    vrow = a128(epk_len) + a128(mlen) + a128(z.sig_bytes) + 1
    vpitch = vrow + 1
    vbuf = torch.zeros(n * vpitch + 64, dtype=torch.uint8, device=eng.device)
    vbase = vbuf.data_ptr() + 2
    h_pk = eng.pinned(pk.shape, pk)
    h_sig = eng.pinned((n, z.sig_bytes), ref)
    res = eng.pinned((n,))
    eng.lib.crypto_sign_verifyWithepk_HtD(h_pk.data_ptr(), h_msg.data_ptr(), h_sig.data_ptr(), mlen, vbase, vpitch, n, st)
    eng.lib.crypto_sign_verifyWithepk(vbase, mlen, vpitch, n, st)
    eng.lib.crypto_sign_verifyWithepk_DtH(res.data_ptr(), mlen, vbase, vpitch, n, st)
    eng.torch.cuda.synchronize(eng.device)
    eng.check()
    assert res.numpy().tolist() == [0] * n
