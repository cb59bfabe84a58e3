"""Pins the oracle's primitives against library routines and closed forms:
SHAKE vs hashlib, AES-128 vs the FIPS 197 known answers and OpenSSL's
AES-128-CTR, GF(16) vs the carry-less closed form, E vs polynomial long
division.
"""
import ctypes
import ctypes.util
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests import gf16_spec as G


@pytest.mark.parametrize("inlen", [0, 1, 7, 8, 31, 135, 136, 137, 167, 168, 169, 271, 272, 273, 600])
@pytest.mark.parametrize("outlen", [1, 24, 32, 136, 137, 168, 169, 430, 924])
def test_shake_matches_hashlib(inlen, outlen):
    data = bytes((i * 37 + inlen) & 0xFF for i in range(inlen))
    assert O.shake256(data, outlen) == hashlib.shake_256(data).digest(outlen)
    assert O.shake128(data, outlen) == hashlib.shake_128(data).digest(outlen)


def test_keccak_f1600_zero_state_known_answer():
    # Keccak-f[1600] applied once to the all-zero state (Keccak team KAT)
    st = O.keccak_f1600(np.zeros(25, np.uint64))
    assert int(st[0]) == 0xF1258F7940E1DDE7
    assert int(st[1]) == 0x84D5CCF933C0478A


def test_aes128_fips197_known_answers():
    # FIPS 197 Appendix C.1 and Appendix B
    assert O.aes128_block(bytes(range(16)), bytes.fromhex("00112233445566778899aabbccddeeff")).hex() == \
        "69c4e0d86a7b0430d8cdb78070b4c55a"
    assert O.aes128_block(bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c"),
                          bytes.fromhex("3243f6a8885a308d313198a2e0370734")).hex() == \
        "3925841d02dc09fbdc118597196a0b32"


def _openssl_ctr(key: bytes, n: int) -> bytes:
    path = ctypes.util.find_library("crypto")
    if not path:
        pytest.skip("libcrypto not found")
    L = ctypes.CDLL(path)
    L.EVP_CIPHER_CTX_new.restype = ctypes.c_void_p
    L.EVP_aes_128_ctr.restype = ctypes.c_void_p
    L.EVP_EncryptInit_ex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
    L.EVP_EncryptUpdate.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    L.EVP_CIPHER_CTX_free.argtypes = [ctypes.c_void_p]
    ctx = L.EVP_CIPHER_CTX_new()
    assert L.EVP_EncryptInit_ex(ctx, L.EVP_aes_128_ctr(), None, key, bytes(16)) == 1
    out = ctypes.create_string_buffer(n + 16)
    outl = ctypes.c_int(0)
    assert L.EVP_EncryptUpdate(ctx, out, ctypes.byref(outl), bytes(n), n) == 1
    L.EVP_CIPHER_CTX_free(ctx)
    return out.raw[: outl.value]


@pytest.mark.parametrize("n", [1, 15, 16, 17, 4096 + 7, 144495])
def test_aes128_ctr_matches_openssl(n):
    key = hashlib.sha256(str(n).encode()).digest()[:16]
    assert O.aes128_ctr(key, n) == _openssl_ctr(key, n)


def test_gf16_mul_inv_closed_form():
    for a in range(16):
        for b in range(16):
            assert O.gf16_mul(a, b) == G.gf16_mul_closed(a, b)
    assert O.gf16_inv(0) == 0
    for a in range(1, 16):
        assert O.gf16_mul(a, O.gf16_inv(a)) == 1
    # field sanity: the multiplicative group is cyclic of order 15 (x generates)
    x, seen = 1, set()
    for _ in range(15):
        seen.add(x)
        x = G.gf16_mul_closed(x, 2)
    assert x == 1 and len(seen) == 15


@pytest.mark.parametrize("version", O.VERSIONS)
def test_E_matches_polynomial_division(version):
    p = O.params(version)
    rng = np.random.default_rng(version)
    for _ in range(20):
        y = rng.integers(0, 16, p.m, dtype=np.uint8)
        got = G.unpack_nibbles(O.mvec_E(version, G.pack_nibbles(y)), p.m)
        assert np.array_equal(got, G.E_times(p, y))
    # E^m x^0 equals the tail polynomial: z^m = C0 + C1 z + C2 z^2 + C3 z^3
    e0 = np.zeros(p.m, np.uint8)
    e0[0] = 1
    y = G.pack_nibbles(e0)
    for _ in range(p.m):
        y = O.mvec_E(version, y)
    want = np.zeros(p.m, np.uint8)
    want[:4] = p.f_tail
    assert np.array_equal(G.unpack_nibbles(y, p.m), want)


def test_params_match_reference_sizes():
    # params.h:13-96 byte sizes
    want = {1: (24, 1420, 144495, 312, 120159, 24336),
            2: (24, 4912, 101376, 544, 66560, 34816),
            3: (32, 2986, 376164, 540, 317844, 58320),
            5: (40, 5554, 841847, 852, 720863, 120984)}
    for v, (sk, pk, pb, ob, p1, p2) in want.items():
        p = O.params(v)
        assert (p.sk_bytes, p.pk_bytes, p.p_bytes, p.o_bytes, p.p1_bytes, p.p2_bytes) == (sk, pk, pb, ob, p1, p2)
        assert p.sig_bytes == p.k * p.n // 2 + p.sk_bytes
