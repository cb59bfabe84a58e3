"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE).

Loads ``oracle/libmayo_oracle.so`` (built by ``oracle/Makefile`` for the
portable x86-64-v2 level). ``baseline_lib()`` loads the copy built for the
highest x86-64 level the running CPU supports (v4 = AVX-512, v3 = AVX2),
which the CPU-baseline leg times. Only the tests, the smoke check and the
CPU-baseline leg of ``bench.py`` import this module; the product path in
``cumayo_amd`` never does.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmayo_oracle.so")
_lib = None
_baseline = None
# x86-64 micro-architecture levels and the CPU flags each needs beyond the
# previous one (the levels of the psABI supplement)
_ISA_FLAGS = {
    "x86-64-v3": ("avx", "avx2", "bmi1", "bmi2", "f16c", "fma", "abm", "movbe", "xsave"),
    "x86-64-v4": ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"),
}

VERSIONS = (1, 2, 3, 5)


class _Params(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int) for name in (
        "version", "n", "m", "o", "k", "v", "sk_bytes", "pk_bytes", "p1_bytes",
        "p2_bytes", "p3_bytes", "p_bytes", "o_bytes", "digest_bytes", "pairs",
        "m_bytes", "sig_bytes", "esk_bytes", "epk_bytes", "vr_bytes")] + [
        ("f_tail", ctypes.c_uint8 * 4)]


@dataclass(frozen=True)
class Params:
    version: int
    n: int
    m: int
    o: int
    k: int
    v: int
    sk_bytes: int
    pk_bytes: int
    p1_bytes: int
    p2_bytes: int
    p3_bytes: int
    p_bytes: int
    o_bytes: int
    digest_bytes: int
    pairs: int
    m_bytes: int
    sig_bytes: int
    esk_bytes: int
    epk_bytes: int
    vr_bytes: int
    f_tail: tuple


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def cpu_isa_level() -> str:
    """The highest x86-64 level (v2, v3, v4) whose flags /proc/cpuinfo lists."""
    flags = set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        pass
    level = "x86-64-v2"
    for name in ("x86-64-v3", "x86-64-v4"):
        if not all(fl in flags for fl in _ISA_FLAGS[name]):
            break
        level = name
    return level


def baseline_lib():
    """(library, isa): the oracle built for the highest x86-64 level this CPU
    supports (falling back to the portable build)."""
    global _baseline
    if _baseline is None:
        isa = cpu_isa_level()
        path = os.path.join(_HERE, f"libmayo_oracle_{isa}.so")
        if isa == "x86-64-v2" or not os.path.exists(path):
            _baseline = (lib(), "x86-64-v2")
        else:
            _baseline = (_bind(ctypes.CDLL(path)), isa)
    return _baseline


def _bind(L):
    P = ctypes.POINTER
    u8p = P(ctypes.c_uint8)
    L.mayo_oracle_params.restype = P(_Params)
    L.mayo_oracle_params.argtypes = [ctypes.c_int]
    for fn in ("oracle_shake128", "oracle_shake256"):
        getattr(L, fn).argtypes = [u8p, ctypes.c_size_t, u8p, ctypes.c_size_t]
    L.oracle_keccak_f1600.argtypes = [P(ctypes.c_uint64)]
    L.oracle_aes128_ecb_block.argtypes = [u8p, u8p, u8p]
    L.oracle_aes128_ctr.argtypes = [u8p, u8p, ctypes.c_size_t]
    L.oracle_gf16_mul.restype = ctypes.c_uint8
    L.oracle_gf16_mul.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
    L.oracle_gf16_inv.restype = ctypes.c_uint8
    L.oracle_gf16_inv.argtypes = [ctypes.c_uint8]
    L.oracle_mvec_E.argtypes = [ctypes.c_int, u8p]
    L.oracle_keypair_seed.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, u8p]
    L.oracle_keypair.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, u8p, u8p]
    L.oracle_keypair_from_sk.argtypes = [ctypes.c_int, u8p, u8p]
    L.oracle_expandsk.argtypes = [ctypes.c_int, u8p, u8p]
    L.oracle_expandpk.argtypes = [ctypes.c_int, u8p, u8p]
    L.oracle_sign.restype = ctypes.c_int
    L.oracle_sign.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_size_t, u8p, ctypes.c_int]
    L.oracle_verify.restype = ctypes.c_int
    L.oracle_verify.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_size_t, u8p]
    L.oracle_sign_digest_salt_t.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_size_t, u8p, u8p, u8p]
    L.oracle_sign_vr.argtypes = [ctypes.c_int, u8p, u8p, u8p, ctypes.c_int, u8p]
    L.oracle_sign_M.argtypes = [ctypes.c_int, u8p, u8p, u8p]
    L.oracle_sign_A.argtypes = [ctypes.c_int, u8p, u8p]
    L.oracle_sign_U.argtypes = [ctypes.c_int, u8p, u8p, u8p]
    L.oracle_sign_y.argtypes = [ctypes.c_int, u8p, u8p, u8p]
    L.oracle_sample_solution.restype = ctypes.c_int
    L.oracle_sample_solution.argtypes = [ctypes.c_int, ctypes.c_int, u8p, u8p, u8p]
    L.oracle_eval_public.argtypes = [ctypes.c_int, u8p, u8p, u8p]
    i32p, i64p = P(ctypes.c_int32), P(ctypes.c_int64)
    L.oracle_sign_batch.argtypes = [ctypes.c_int, u8p, i32p, u8p, i64p, i32p, u8p,
                                    ctypes.c_int64, ctypes.c_int]
    L.oracle_verify_batch.argtypes = [ctypes.c_int, u8p, i32p, u8p, i64p, i32p, u8p, u8p,
                                      ctypes.c_int64, ctypes.c_int]
    return L


def params(version: int) -> Params:
    ptr = lib().mayo_oracle_params(version)
    if not ptr:
        raise ValueError(f"unknown MAYO version {version}")
    s = ptr.contents
    vals = {f[0]: getattr(s, f[0]) for f in _Params._fields_ if f[0] != "f_tail"}
    return Params(f_tail=tuple(s.f_tail), **vals)


def _u8(a) -> ctypes.POINTER(ctypes.c_uint8):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _buf(data) -> np.ndarray:
    if isinstance(data, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(data), dtype=np.uint8).copy()
    return np.ascontiguousarray(data, dtype=np.uint8)


def shake128(data, outlen: int) -> bytes:
    src = _buf(data)
    out = np.zeros(outlen, np.uint8)
    lib().oracle_shake128(_u8(out), outlen, _u8(src), src.size)
    return out.tobytes()


def shake256(data, outlen: int) -> bytes:
    src = _buf(data)
    out = np.zeros(outlen, np.uint8)
    lib().oracle_shake256(_u8(out), outlen, _u8(src), src.size)
    return out.tobytes()


def keccak_f1600(state):
    st = np.ascontiguousarray(state, dtype=np.uint64).copy()
    lib().oracle_keccak_f1600(st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return st


def aes128_block(key: bytes, block: bytes) -> bytes:
    k, b = _buf(key), _buf(block)
    out = np.zeros(16, np.uint8)
    lib().oracle_aes128_ecb_block(_u8(k), _u8(b), _u8(out))
    return out.tobytes()


def aes128_ctr(key: bytes, outlen: int) -> bytes:
    k = _buf(key)
    out = np.zeros(outlen, np.uint8)
    lib().oracle_aes128_ctr(_u8(k), _u8(out), outlen)
    return out.tobytes()


def gf16_mul(a: int, b: int) -> int:
    return int(lib().oracle_gf16_mul(a, b))


def gf16_inv(a: int) -> int:
    return int(lib().oracle_gf16_inv(a))


def mvec_E(version: int, y: bytes) -> bytes:
    buf = _buf(y).copy()
    lib().oracle_mvec_E(version, _u8(buf))
    return buf.tobytes()


def keypair_seed(version: int, rand_index: int, item: int) -> bytes:
    p = params(version)
    sk = np.zeros(p.sk_bytes, np.uint8)
    lib().oracle_keypair_seed(version, rand_index, item, _u8(sk))
    return sk.tobytes()


def keypair(version: int, rand_index: int, item: int):
    p = params(version)
    pk = np.zeros(p.pk_bytes, np.uint8)
    sk = np.zeros(p.sk_bytes, np.uint8)
    lib().oracle_keypair(version, rand_index, item, _u8(pk), _u8(sk))
    return pk.tobytes(), sk.tobytes()


def keypair_from_sk(version: int, sk: bytes) -> bytes:
    p = params(version)
    s = _buf(sk)
    pk = np.zeros(p.pk_bytes, np.uint8)
    lib().oracle_keypair_from_sk(version, _u8(s), _u8(pk))
    return pk.tobytes()


def expandsk(version: int, sk: bytes) -> bytes:
    p = params(version)
    s = _buf(sk)
    esk = np.zeros(p.esk_bytes, np.uint8)
    lib().oracle_expandsk(version, _u8(s), _u8(esk))
    return esk.tobytes()


def expandpk(version: int, cpk: bytes) -> bytes:
    p = params(version)
    c = _buf(cpk)
    epk = np.zeros(p.epk_bytes, np.uint8)
    lib().oracle_expandpk(version, _u8(c), _u8(epk))
    return epk.tobytes()


def sign(version: int, esk: bytes, msg: bytes, fail_first: int = 0):
    """Returns (signature, attempts)."""
    p = params(version)
    e, mm = _buf(esk), _buf(msg)
    if mm.size == 0:
        mm = np.zeros(1, np.uint8)
    sig = np.zeros(p.sig_bytes, np.uint8)
    att = lib().oracle_sign(version, _u8(e), _u8(mm), len(msg), _u8(sig), fail_first)
    return sig.tobytes(), att


def verify(version: int, epk: bytes, msg: bytes, sig: bytes) -> int:
    e, mm, s = _buf(epk), _buf(msg), _buf(sig)
    if mm.size == 0:
        mm = np.zeros(1, np.uint8)
    return int(lib().oracle_verify(version, _u8(e), _u8(mm), len(msg), _u8(s)))


def sign_digest_salt_t(version: int, esk: bytes, msg: bytes):
    p = params(version)
    e, mm = _buf(esk), _buf(msg)
    if mm.size == 0:
        mm = np.zeros(1, np.uint8)
    d = np.zeros(p.digest_bytes, np.uint8)
    s = np.zeros(p.sk_bytes, np.uint8)
    t = np.zeros(p.m_bytes, np.uint8)
    lib().oracle_sign_digest_salt_t(version, _u8(e), _u8(mm), len(msg), _u8(d), _u8(s), _u8(t))
    return d.tobytes(), s.tobytes(), t.tobytes()


def sign_vr(version: int, digest: bytes, salt: bytes, seed_sk: bytes, ctr: int) -> bytes:
    p = params(version)
    out = np.zeros(p.vr_bytes, np.uint8)
    lib().oracle_sign_vr(version, _u8(_buf(digest)), _u8(_buf(salt)), _u8(_buf(seed_sk)), ctr, _u8(out))
    return out.tobytes()


def sign_M(version: int, esk: bytes, vr: bytes) -> np.ndarray:
    p = params(version)
    out = np.zeros((p.k, p.m, p.o), np.uint8)
    lib().oracle_sign_M(version, _u8(_buf(esk)), _u8(_buf(vr)), _u8(out))
    return out


def sign_A(version: int, M: np.ndarray) -> np.ndarray:
    p = params(version)
    Mc = np.ascontiguousarray(M, dtype=np.uint8)
    out = np.zeros((p.m, p.k * p.o), np.uint8)
    lib().oracle_sign_A(version, _u8(Mc), _u8(out))
    return out


def sign_U(version: int, esk: bytes, vr: bytes) -> np.ndarray:
    p = params(version)
    out = np.zeros((p.pairs, p.m_bytes), np.uint8)
    lib().oracle_sign_U(version, _u8(_buf(esk)), _u8(_buf(vr)), _u8(out))
    return out


def sign_y(version: int, U: np.ndarray, t: bytes) -> bytes:
    p = params(version)
    Uc = np.ascontiguousarray(U, dtype=np.uint8)
    out = np.zeros(p.m_bytes, np.uint8)
    lib().oracle_sign_y(version, _u8(Uc), _u8(_buf(t)), _u8(out))
    return out.tobytes()


def sample_solution(A: np.ndarray, y: np.ndarray, r: np.ndarray):
    """Returns (ok, x) for the unpacked system A (rows x cols), y, r."""
    Ac = np.ascontiguousarray(A, dtype=np.uint8).copy()
    yc = np.ascontiguousarray(y, dtype=np.uint8).copy()
    xc = np.ascontiguousarray(r, dtype=np.uint8).copy()
    rows, cols = Ac.shape
    ok = lib().oracle_sample_solution(rows, cols, _u8(Ac), _u8(yc), _u8(xc))
    return int(ok), xc


def eval_public(version: int, epk: bytes, s_unpacked: np.ndarray) -> bytes:
    p = params(version)
    sc = np.ascontiguousarray(s_unpacked, dtype=np.uint8)
    out = np.zeros(p.m_bytes, np.uint8)
    lib().oracle_eval_public(version, _u8(_buf(epk)), _u8(sc), _u8(out))
    return out.tobytes()


def _i32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _i64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def sign_batch(version: int, esk_table: np.ndarray, key_idx: np.ndarray, msgs: np.ndarray,
               msg_off: np.ndarray, msg_len: np.ndarray, threads: int = 0, library=None) -> np.ndarray:
    p = params(version)
    n = int(key_idx.size)
    sigs = np.zeros((n, p.sig_bytes), np.uint8)
    et = np.ascontiguousarray(esk_table, dtype=np.uint8)
    ki = np.ascontiguousarray(key_idx, dtype=np.int32)
    mm = np.ascontiguousarray(msgs, dtype=np.uint8)
    mo = np.ascontiguousarray(msg_off, dtype=np.int64)
    ml = np.ascontiguousarray(msg_len, dtype=np.int32)
    (library or lib()).oracle_sign_batch(version, _u8(et), _i32(ki), _u8(mm), _i64(mo), _i32(ml), _u8(sigs), n,
                                         threads)
    return sigs


def verify_batch(version: int, epk_table: np.ndarray, key_idx: np.ndarray, msgs: np.ndarray,
                 msg_off: np.ndarray, msg_len: np.ndarray, sigs: np.ndarray,
                 threads: int = 0) -> np.ndarray:
    n = int(key_idx.size)
    res = np.zeros(n, np.uint8)
    et = np.ascontiguousarray(epk_table, dtype=np.uint8)
    ki = np.ascontiguousarray(key_idx, dtype=np.int32)
    mm = np.ascontiguousarray(msgs, dtype=np.uint8)
    mo = np.ascontiguousarray(msg_off, dtype=np.int64)
    ml = np.ascontiguousarray(msg_len, dtype=np.int32)
    sg = np.ascontiguousarray(sigs, dtype=np.uint8)
    lib().oracle_verify_batch(version, _u8(et), _i32(ki), _u8(mm), _i64(mo), _i32(ml), _u8(sg),
                              _u8(res), n, threads)
    return res
