"""Pins the oracle's MAYO stages against the independent per-equation
restatement in gf16_spec (matrix products, repeated E) and against the
algebraic facts the stages must satisfy.
"""
import hashlib
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests import gf16_spec as G

VERSIONS = O.VERSIONS


def key_material(version, item=0):
    p = O.params(version)
    pk, sk = O.keypair(version, 0, item)
    esk = O.expandsk(version, sk)
    epk = O.expandpk(version, pk)
    return p, pk, sk, esk, epk


def split_esk(p, esk):
    lay = G.Layout(p)
    o_off = p.sk_bytes
    p1_off = o_off + p.o_bytes
    l_off = p1_off + p.p1_bytes
    Om = lay.O(esk[o_off:p1_off])
    P1m = lay.upper_matrices(esk[p1_off:l_off], p.v)
    Lm = lay.rect_matrices(esk[l_off:], p.v, p.o)
    return Om, P1m, Lm


def split_epk(p, epk):
    lay = G.Layout(p)
    P1m = lay.upper_matrices(epk[:p.p1_bytes], p.v)
    P2m = lay.rect_matrices(epk[p.p1_bytes:p.p_bytes], p.v, p.o)
    P3m = lay.upper_matrices(epk[p.p_bytes:], p.o)
    return P1m, P2m, P3m


def test_keypair_seed_rule():
    # keypair.cu:32-45: sk = SHAKE128(LE64(rand_index + 2 i + 1))[0:sk]
    for v in VERSIONS:
        p = O.params(v)
        for ri, i in [(0, 0), (0, 5), (77, 3)]:
            want = hashlib.shake_128((ri + 2 * i + 1).to_bytes(8, "little")).digest(p.sk_bytes)
            assert O.keypair_seed(v, ri, i) == want


@pytest.mark.parametrize("version", VERSIONS)
def test_expanded_keys_match_matrix_form(version):
    p, pk, sk, esk, epk = key_material(version)
    S = hashlib.shake_256(sk).digest(16 + p.o_bytes)
    # esk = sk || O || P1 || L, epk = P1 || P2 || P3
    assert esk[:p.sk_bytes] == sk
    assert esk[p.sk_bytes:p.sk_bytes + p.o_bytes] == S[16:]
    assert pk[:16] == S[:16]
    P = O.aes128_ctr(S[:16], p.p_bytes)
    assert esk[p.sk_bytes + p.o_bytes:p.sk_bytes + p.o_bytes + p.p1_bytes] == P[:p.p1_bytes]
    assert epk[:p.p_bytes] == P and epk[p.p_bytes:] == pk[16:]
    Om, P1m, Lm = split_esk(p, esk)
    _, P2m, P3m = split_epk(p, epk)
    for a in range(p.m):
        # L_a = (P1_a + P1_a^T) O + P2_a  (expandsk.cu:37-80)
        want_L = G.matmul(P1m[a] ^ P1m[a].T, Om) ^ P2m[a]
        assert np.array_equal(Lm[a], want_L), f"L eq {a}"
        # P3_a = Upper(O^T (P1_a O + P2_a))  (keypair.cu:55-119)
        want_P3 = G.upper(G.matmul(Om.T, G.matmul(P1m[a], Om) ^ P2m[a]))
        assert np.array_equal(P3m[a], want_P3), f"P3 eq {a}"


def test_key_expansion_invariant_public_map_vanishes_on_oil_space():
    # For any x, the public map at s = (O x, x) is zero: P*(O x, x) = 0 is
    # what makes O the secret oil space. Checked per equation with k=1 form.
    p, pk, sk, esk, epk = key_material(1)
    Om, _, _ = split_esk(p, esk)
    P1m, P2m, P3m = split_epk(p, epk)
    rng = np.random.default_rng(3)
    for _ in range(3):
        x = rng.integers(0, 16, p.o, dtype=np.uint8)
        s = np.concatenate([G.matvec(Om, x), x])
        for a in range(0, p.m, 7):
            P = np.zeros((p.n, p.n), np.uint8)
            P[:p.v, :p.v], P[:p.v, p.v:], P[p.v:, p.v:] = P1m[a], P2m[a], P3m[a]
            assert np.bitwise_xor.reduce(G.MUL[s, G.matvec(P, s)]) == 0


@pytest.mark.parametrize("version", VERSIONS)
def test_sign_intermediates_match_matrix_form(version):
    p, pk, sk, esk, epk = key_material(version)
    msg = b"intermediates"
    digest, salt, t = O.sign_digest_salt_t(version, esk, msg)
    assert digest == hashlib.shake_256(msg).digest(p.digest_bytes)
    assert salt == hashlib.shake_256(digest + bytes(p.sk_bytes) + sk).digest(p.sk_bytes)  # R = 0
    assert t == hashlib.shake_256(digest + salt).digest(p.m_bytes)
    vr = O.sign_vr(version, digest, salt, sk, 0)
    assert vr == hashlib.shake_256(digest + salt + sk + bytes([0])).digest(p.k * p.n // 2)
    Om, P1m, Lm = split_esk(p, esk)
    vb = p.v // 2
    V = [G.unpack_nibbles(vr[i * vb:(i + 1) * vb], p.v) for i in range(p.k)]
    # M_i[a, :] = V_i^T L_a (sign.cu:256-330)
    M = O.sign_M(version, esk, vr)
    for i in range(p.k):
        want = np.stack([G.matvec(Lm[a].T, V[i]) for a in range(p.m)])
        assert np.array_equal(M[i], want), f"M_{i}"
    # A block i = sum E^l M_j ..., E^l applied column-wise by repeated E
    A = O.sign_A(version, M)
    want_A = np.zeros((p.m, p.k * p.o), np.uint8)
    for ell, i, j in G.pair_schedule(p.k):
        for c in range(p.o):
            want_A[:, i * p.o + c] ^= G.E_pow(p, M[j][:, c], ell)
            if i != j:
                want_A[:, j * p.o + c] ^= G.E_pow(p, M[i][:, c], ell)
    assert np.array_equal(A, want_A)
    # U pairs (sign.cu:396-437)
    U = O.sign_U(version, esk, vr)
    want_U = G.quad_vectors(P1m, V)
    k = p.k
    for (i, j), u in want_U.items():
        pid = i * k - i * (i - 1) // 2 + (j - i)
        assert np.array_equal(G.unpack_nibbles(U[pid].tobytes(), p.m), u), f"u_{i}{j}"
    # y = t + sum_l E^l u_l (Horner, sign.cu:440-459)
    y = O.sign_y(version, U, t)
    want_y = G.unpack_nibbles(t, p.m).copy()
    for ell, i, j in G.pair_schedule(p.k):
        want_y ^= G.E_pow(p, want_U[(i, j)], ell)
    assert np.array_equal(G.unpack_nibbles(y, p.m), want_y)


@pytest.mark.parametrize("version", VERSIONS)
def test_signature_satisfies_public_map_and_structure(version):
    p, pk, sk, esk, epk = key_material(version, item=1)
    msg = bytes(range(100))
    sig, attempts = O.sign(version, esk, msg)
    assert attempts == 1
    digest, salt, t = O.sign_digest_salt_t(version, esk, msg)
    assert sig[p.k * p.n // 2:] == salt
    s = G.unpack_nibbles(sig, p.k * p.n).reshape(p.k, p.n)
    P1m, P2m, P3m = split_epk(p, epk)
    # the defining equation: P*(s) = t
    assert np.array_equal(G.public_map(p, P1m, P2m, P3m, s), G.unpack_nibbles(t, p.m))
    # s_i = (V_i + O x_i, x_i) with V from attempt 0
    Om, _, _ = split_esk(p, esk)
    vr = O.sign_vr(version, digest, salt, sk, 0)
    vb = p.v // 2
    for i in range(p.k):
        Vi = G.unpack_nibbles(vr[i * vb:(i + 1) * vb], p.v)
        xi = s[i, p.v:]
        assert np.array_equal(s[i, :p.v], Vi ^ G.matvec(Om, xi))
    assert O.verify(version, epk, msg, sig) == 0


@pytest.mark.parametrize("version", VERSIONS)
def test_verify_matches_public_map_on_random_vectors(version):
    p, pk, sk, esk, epk = key_material(version)
    P1m, P2m, P3m = split_epk(p, epk)
    rng = np.random.default_rng(11 + version)
    s = rng.integers(0, 16, (p.k, p.n), dtype=np.uint8)
    got = G.unpack_nibbles(O.eval_public(version, epk, s), p.m)
    assert np.array_equal(got, G.public_map(p, P1m, P2m, P3m, s))


def test_verify_rejects_mutations():
    v = 1
    p, pk, sk, esk, epk = key_material(v)
    msg = b"mutations"
    sig, _ = O.sign(v, esk, msg)
    assert O.verify(v, epk, msg, sig) == 0
    for pos in (0, 100, p.k * p.n // 2 - 1, p.k * p.n // 2, p.sig_bytes - 1):
        bad = bytearray(sig)
        bad[pos] ^= 0x10
        assert O.verify(v, epk, msg, bytes(bad)) == 1, pos
    assert O.verify(v, epk, msg + b"x", sig) == 1
    bad_epk = bytearray(epk)
    bad_epk[5] ^= 1
    assert O.verify(v, bytes(bad_epk), msg, sig) == 1


def _brute_solutions(A, y):
    rows, cols = A.shape
    sols = []
    for x in itertools.product(range(16), repeat=cols):
        x = np.array(x, np.uint8)
        if np.array_equal(G.matvec(A, x), y):
            sols.append(x)
    return sols


def _pivot_columns(A):
    """Greedy leftmost set of independent columns (rank via elimination)."""
    rows, cols = A.shape
    piv, basis = [], np.zeros((0, rows), np.uint8)
    for c in range(cols):
        cand = np.vstack([basis, A[:, c]])
        if _rank(cand) > basis.shape[0]:
            basis, piv = cand, piv + [c]
    return piv


def _rank(M):
    M = M.copy()
    r = 0
    for c in range(M.shape[1]):
        pr = next((i for i in range(r, M.shape[0]) if M[i, c]), None)
        if pr is None:
            continue
        M[[r, pr]] = M[[pr, r]]
        M[r] = G.MUL[M[r], G.INV[M[r, c]]]
        for i in range(M.shape[0]):
            if i != r and M[i, c]:
                M[i] ^= G.MUL[M[r], M[i, c]]
        r += 1
    return r


def test_sample_solution_brute_force_tiny_systems():
    rng = np.random.default_rng(5)
    for trial in range(12):
        rows, cols = (2, 4) if trial % 2 else (3, 4)
        A = rng.integers(0, 16, (rows, cols), dtype=np.uint8)
        if trial == 3:
            A[:, 1] = 0  # a column with no pivot
        r = rng.integers(0, 16, cols, dtype=np.uint8)
        y = rng.integers(0, 16, rows, dtype=np.uint8)
        ok, x = O.sample_solution(A, y, r)
        full = _rank(A) == rows
        assert ok == int(full)
        if not full:
            assert np.array_equal(x, r)  # x stays r on rank failure
            continue
        sols = _brute_solutions(A, y)
        assert any(np.array_equal(x, s) for s in sols)
        piv = _pivot_columns(A)
        free = [c for c in range(cols) if c not in piv]
        # unique solution equal to r on the non-pivot columns
        match = [s for s in sols if all(s[c] == r[c] for c in free)]
        assert len(match) == 1 and np.array_equal(match[0], x)


def test_sample_solution_solves_full_size_systems():
    rng = np.random.default_rng(9)
    for v in VERSIONS:
        p = O.params(v)
        A = rng.integers(0, 16, (p.m, p.k * p.o), dtype=np.uint8)
        y = rng.integers(0, 16, p.m, dtype=np.uint8)
        r = rng.integers(0, 16, p.k * p.o, dtype=np.uint8)
        ok, x = O.sample_solution(A, y, r)
        assert ok == 1
        assert np.array_equal(G.matvec(A, x), y)


def test_retry_path_uses_next_counter():
    v = 1
    p, pk, sk, esk, epk = key_material(v)
    msg = b"retry"
    sig0, a0 = O.sign(v, esk, msg)
    sig1, a1 = O.sign(v, esk, msg, fail_first=1)
    assert (a0, a1) == (1, 2)
    assert sig0 != sig1 and O.verify(v, epk, msg, sig1) == 0
    digest, salt, t = O.sign_digest_salt_t(v, esk, msg)
    vr1 = O.sign_vr(v, digest, salt, sk, 1)
    s = G.unpack_nibbles(sig1, p.k * p.n).reshape(p.k, p.n)
    Om, _, _ = split_esk(p, esk)
    V0 = G.unpack_nibbles(vr1[:p.v // 2], p.v)
    assert np.array_equal(s[0, :p.v], V0 ^ G.matvec(Om, s[0, p.v:]))


def test_exhausted_attempts_write_last_vector_with_x_equal_r():
    # sign.cu:244, 472-477: after 255 failures the last attempt is written
    v = 1
    p, pk, sk, esk, epk = key_material(v)
    msg = b"exhaust"
    sig, att = O.sign(v, esk, msg, fail_first=255)
    assert att == 256
    digest, salt, t = O.sign_digest_salt_t(v, esk, msg)
    vr = O.sign_vr(v, digest, salt, sk, 254)
    Om, _, _ = split_esk(p, esk)
    s = G.unpack_nibbles(sig, p.k * p.n).reshape(p.k, p.n)
    rv = G.unpack_nibbles(vr[p.k * (p.v // 2):], p.k * p.o).reshape(p.k, p.o)
    vb = p.v // 2
    for i in range(p.k):
        Vi = G.unpack_nibbles(vr[i * vb:(i + 1) * vb], p.v)
        assert np.array_equal(s[i, p.v:], rv[i])
        assert np.array_equal(s[i, :p.v], Vi ^ G.matvec(Om, rv[i]))
