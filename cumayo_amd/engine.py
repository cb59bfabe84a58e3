"""Python front end of the HIP MAYO engine (ctypes over the C ABI).

``Engine(version)`` loads ``cumayo_amd/lib/libcumayo_mayo{version}.so`` and
exposes

* the reference's batched API (``include/api.cuh`` of cuMAYO) one-to-one:
  ``crypto_sign_keypair``, ``crypto_sign_sign_HtD`` ... with device pools
  given as torch uint8 tensors (rows ``pitch`` bytes apart) and host buffers
  as numpy arrays, enqueued on a torch/HIP stream;
* host-convenience wrappers that allocate pools, run a stage and copy the
  result back (used by the tests);
* the batched key-table entry points (``sign_table`` / ``verify_table``).

There is no CPU fallback: constructing an Engine without the built library
or without a GPU raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import build as _build

VERSIONS = _build.VERSIONS


def a128(x: int) -> int:
    return (x + 127) // 128 * 128


class _Sizes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "version", "n", "m", "o", "k", "sk_bytes", "pk_bytes", "sig_bytes", "esk_bytes",
        "epk_bytes", "digest_bytes", "p1_bytes", "p2_bytes", "p3_bytes", "o_bytes",
        "sign_ws_stride", "sign_ws_attempts")]


@dataclass(frozen=True)
class Sizes:
    version: int
    n: int
    m: int
    o: int
    k: int
    sk_bytes: int
    pk_bytes: int
    sig_bytes: int
    esk_bytes: int
    epk_bytes: int
    digest_bytes: int
    p1_bytes: int
    p2_bytes: int
    p3_bytes: int
    o_bytes: int
    sign_ws_stride: int
    sign_ws_attempts: int

    @property
    def p_bytes(self) -> int:
        return self.p1_bytes + self.p2_bytes

    @property
    def scratch_bytes(self) -> int:
        # per-row scratch of the sign pools: k*m*ceil(o/2) (bench_stream.cu:82)
        return self.k * self.m * ((self.o + 1) // 2)


class LaunchConfig(ctypes.Structure):
    _fields_ = [("sign_prep", ctypes.c_int32), ("sign_solve", ctypes.c_int32), ("sign_order", ctypes.c_int32),
                ("verify", ctypes.c_int32), ("verify_order", ctypes.c_int32), ("sign_chunks", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {f[0]: int(getattr(self, f[0])) for f in self._fields_}


class TuneReport(ctypes.Structure):
    _fields_ = [("best", LaunchConfig), ("sign_ms", ctypes.c_double), ("verify_ms", ctypes.c_double),
                ("sign_default_ms", ctypes.c_double), ("verify_default_ms", ctypes.c_double),
                ("configs_timed", ctypes.c_int32)]


LAUNCH_FIELDS = {"sign_prep": 0, "sign_solve": 1, "verify": 2, "order": 3, "sign_chunks": 4}
# keys per call of cumayo_prepare_esk / cumayo_prepare_epk
PREPARE_KEYS_MAX = 65535

# (name, argtypes) of every symbol the header declares
_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t
EXPORTS = {
    "crypto_sign_keypair": [_VP, _SZ, _SZ, _VP, _SZ],
    "crypto_sign_keypair_DtH": [_VP, _VP, _VP, _SZ, _SZ, _VP],
    "crypto_sign_expandsk_HtD": [_VP, _VP, _SZ, _SZ, _VP],
    "crypto_sign_expandsk": [_VP, _SZ, _SZ, _VP],
    "crypto_sign_expandsk_DtH": [_VP, _VP, _SZ, _SZ, _VP],
    "crypto_sign_expandpk_HtD": [_VP, _VP, _SZ, _SZ, _VP],
    "crypto_sign_expandpk": [_VP, _SZ, _SZ, _VP],
    "crypto_sign_expandpk_DtH": [_VP, _VP, _SZ, _SZ, _VP],
    "crypto_sign_sign_HtD": [_VP, _VP, _SZ, _VP, _SZ, _SZ, _VP],
    "crypto_sign_sign": [_VP, _SZ, _SZ, _SZ, _VP],
    "crypto_sign_sign_DtH": [_VP, _VP, _SZ, _SZ, _SZ, _VP],
    "crypto_sign_signWithesk_HtD": [_VP, _VP, _SZ, _VP, _SZ, _SZ, _VP],
    "crypto_sign_signWithesk": [_VP, _SZ, _SZ, _SZ, _VP],
    "crypto_sign_signWithesk_DtH": [_VP, _VP, _SZ, _SZ, _SZ, _VP],
    "crypto_sign_verify_HtD": [_VP, _VP, _VP, _SZ, _VP, _SZ, _SZ, _VP],
    "crypto_sign_verify": [_VP, _SZ, _SZ, _SZ, _VP],
    "crypto_sign_verify_DtH": [_VP, _SZ, _VP, _SZ, _SZ, _VP],
    "crypto_sign_verifyWithepk_HtD": [_VP, _VP, _VP, _SZ, _VP, _SZ, _SZ, _VP],
    "crypto_sign_verifyWithepk": [_VP, _SZ, _SZ, _SZ, _VP],
    "crypto_sign_verifyWithepk_DtH": [_VP, _SZ, _VP, _SZ, _SZ, _VP],
    "aes128_tables_init": [],
    "cumayo_get_sizes": [ctypes.POINTER(_Sizes)],
    "cumayo_last_error": [],
    "cumayo_signwithesk_lfree": [],
    "cumayo_test_pitched_fail_first": [ctypes.c_int],
    "cumayo_sign_workspace_size": [_SZ],
    "cumayo_sign_table": [_VP, _SZ, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP],
    "cumayo_verify_workspace_size": [_SZ],
    "cumayo_verify_table": [_VP, _SZ, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP],
    "cumayo_sign_table_stages": [_VP, _SZ, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP, ctypes.c_int],
    "cumayo_verify_table_stages": [_VP, _SZ, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP, ctypes.c_int],
    "cumayo_prepared_esk_bytes": [],
    "cumayo_prepared_epk_bytes": [],
    "cumayo_prepare_esk": [_VP, _SZ, _SZ, _VP, _VP],
    "cumayo_prepare_epk": [_VP, _SZ, _SZ, _VP, _VP],
    "cumayo_sign_prepared": [_VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP, ctypes.c_int],
    "cumayo_verify_prepared": [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP, ctypes.c_int],
    "cumayo_sign_prepared_noalloc": [_VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP],
    "cumayo_crypto_sign_workspace_size": [_SZ],
    "cumayo_crypto_sign": [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP],
    "cumayo_crypto_sign_open_workspace_size": [_SZ],
    "cumayo_crypto_sign_open": [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP],
    "cumayo_get_launch_config": [ctypes.POINTER(LaunchConfig)],
    "cumayo_set_launch_config": [ctypes.POINTER(LaunchConfig)],
    "cumayo_launch_variants": [ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.c_int],
    "cumayo_autotune": [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _VP, _VP, ctypes.c_int,
                        ctypes.POINTER(TuneReport)],
    "cumayo_sign_graph_create": [_VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, ctypes.c_int],
    "cumayo_verify_graph_create": [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, ctypes.c_int],
    "cumayo_graph_launch": [_VP, _VP],
    "cumayo_graph_destroy": [_VP],
    "cumayo_graph_compose": [_VP, ctypes.c_int],
}
_RESTYPES = {
    "cumayo_last_error": ctypes.c_int,
    "cumayo_signwithesk_lfree": ctypes.c_int,
    "cumayo_sign_workspace_size": _SZ,
    "cumayo_verify_workspace_size": _SZ,
    "cumayo_sign_table": ctypes.c_int,
    "cumayo_verify_table": ctypes.c_int,
    "cumayo_sign_table_stages": ctypes.c_int,
    "cumayo_verify_table_stages": ctypes.c_int,
    "cumayo_prepared_esk_bytes": _SZ,
    "cumayo_prepared_epk_bytes": _SZ,
    "cumayo_prepare_esk": ctypes.c_int,
    "cumayo_prepare_epk": ctypes.c_int,
    "cumayo_sign_prepared": ctypes.c_int,
    "cumayo_verify_prepared": ctypes.c_int,
    "cumayo_sign_prepared_noalloc": ctypes.c_int,
    "cumayo_crypto_sign_workspace_size": _SZ,
    "cumayo_crypto_sign": ctypes.c_int,
    "cumayo_crypto_sign_open_workspace_size": _SZ,
    "cumayo_crypto_sign_open": ctypes.c_int,
    "cumayo_set_launch_config": ctypes.c_int,
    "cumayo_launch_variants": ctypes.c_int,
    "cumayo_autotune": ctypes.c_int,
    "cumayo_sign_graph_create": _VP,
    "cumayo_verify_graph_create": _VP,
    "cumayo_graph_launch": ctypes.c_int,
    "cumayo_graph_compose": _VP,
}


def load_library(version: int) -> ctypes.CDLL:
    """Loads the per-set library and binds every exported symbol."""
    path = _build.lib_path(version)
    if not os.path.exists(path):
        raise RuntimeError(f"HIP engine library missing: {path} (run python -m cumayo_amd.build)")
    lib = ctypes.CDLL(path)
    for name, args in EXPORTS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, None)
    return lib


def _ptr(x) -> int:
    """Address of a torch tensor, numpy array or raw int."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return x.data_ptr()


class Engine:
    """Host wrappers stage every host buffer through pinned memory: the
    asynchronous 2-D copies of the C ABI are then plain DMA transfers."""

    def __init__(self, version: int, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("cumayo_amd needs a ROCm GPU; none is visible")
        self.torch = torch
        self.version = version
        self.device = torch.device("cuda", device)
        self.lib = load_library(version)
        s = _Sizes()
        self.lib.cumayo_get_sizes(ctypes.byref(s))
        self.sizes = Sizes(**{f[0]: getattr(s, f[0]) for f in _Sizes._fields_})
        if self.sizes.version != version:
            raise RuntimeError("library/parameter-set mismatch")
        with torch.cuda.device(self.device):
            self.lib.aes128_tables_init()
        self.check()

    # -- plumbing ---------------------------------------------------------
    def stream_handle(self, stream=None) -> int:
        if stream is None:
            stream = self.torch.cuda.current_stream(self.device)
        return stream.cuda_stream

    def check(self):
        err = self.lib.cumayo_last_error()
        if err:
            raise RuntimeError(f"HIP error {err} in cumayo_amd (MAYO_{self.version})")

    def pinned(self, shape, src=None):
        t = self.torch.empty(shape, dtype=self.torch.uint8, pin_memory=True)
        if src is not None:
            t.numpy()[...] = np.asarray(src, np.uint8).reshape(shape)
        return t

    def pool(self, rows: int, row_bytes: int):
        pitch = (row_bytes + 255) // 256 * 256
        t = self.torch.zeros((max(rows, 1), pitch), dtype=self.torch.uint8, device=self.device)
        return t, pitch

    # -- host convenience wrappers (allocate, run, copy back, synchronise) --
    def keypair(self, batch: int, rand_index: int = 0):
        z = self.sizes
        row = a128(z.sk_bytes) + a128(z.pk_bytes) + a128(z.p_bytes)
        pool, pitch = self.pool(batch, row)
        pk = self.pinned((batch, z.pk_bytes))
        sk = self.pinned((batch, z.sk_bytes))
        st = self.stream_handle()
        self.lib.crypto_sign_keypair(pool.data_ptr(), pitch, batch, st, rand_index)
        self.lib.crypto_sign_keypair_DtH(pk.data_ptr(), sk.data_ptr(), pool.data_ptr(), pitch, batch, st)
        self.torch.cuda.synchronize(self.device)
        self.check()
        return pk.numpy().copy(), sk.numpy().copy()

    def expandsk(self, sk: np.ndarray) -> np.ndarray:
        z = self.sizes
        sk = np.ascontiguousarray(sk, np.uint8).reshape(-1, z.sk_bytes)
        batch = sk.shape[0]
        pool, pitch = self.pool(batch, z.esk_bytes)
        h_sk = self.pinned(sk.shape, sk)
        esk = self.pinned((batch, z.esk_bytes))
        st = self.stream_handle()
        self.lib.crypto_sign_expandsk_HtD(h_sk.data_ptr(), pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_expandsk(pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_expandsk_DtH(esk.data_ptr(), pool.data_ptr(), pitch, batch, st)
        self.torch.cuda.synchronize(self.device)
        self.check()
        return esk.numpy().copy()

    def expandpk(self, pk: np.ndarray) -> np.ndarray:
        z = self.sizes
        pk = np.ascontiguousarray(pk, np.uint8).reshape(-1, z.pk_bytes)
        batch = pk.shape[0]
        pool, pitch = self.pool(batch, z.epk_bytes)
        h_pk = self.pinned(pk.shape, pk)
        epk = self.pinned((batch, z.epk_bytes))
        st = self.stream_handle()
        self.lib.crypto_sign_expandpk_HtD(h_pk.data_ptr(), pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_expandpk(pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_expandpk_DtH(epk.data_ptr(), pool.data_ptr(), pitch, batch, st)
        self.torch.cuda.synchronize(self.device)
        self.check()
        return epk.numpy().copy()

    def sign(self, esk: np.ndarray, msgs: np.ndarray) -> np.ndarray:
        """crypto_sign_sign over rows esk || msg || sig || scratch."""
        z = self.sizes
        esk = np.ascontiguousarray(esk, np.uint8).reshape(-1, z.esk_bytes)
        msgs = np.ascontiguousarray(msgs, np.uint8).reshape(esk.shape[0], -1)
        batch, mlen = msgs.shape
        pool, pitch = self.pool(batch, z.esk_bytes + mlen + z.sig_bytes + z.scratch_bytes)
        h_esk, h_msg = self.pinned(esk.shape, esk), self.pinned(msgs.shape, msgs)
        sig = self.pinned((batch, z.sig_bytes))
        st = self.stream_handle()
        self.lib.crypto_sign_sign_HtD(h_esk.data_ptr(), h_msg.data_ptr(), mlen, pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_sign(pool.data_ptr(), mlen, pitch, batch, st)
        self.lib.crypto_sign_sign_DtH(sig.data_ptr(), pool.data_ptr(), mlen, pitch, batch, st)
        self.torch.cuda.synchronize(self.device)
        self.check()
        return sig.numpy().copy()

    def sign_with_sk(self, sk: np.ndarray, msgs: np.ndarray) -> np.ndarray:
        """crypto_sign_signWithesk: compact sk in, expanded on the GPU, signed."""
        z = self.sizes
        sk = np.ascontiguousarray(sk, np.uint8).reshape(-1, z.sk_bytes)
        msgs = np.ascontiguousarray(msgs, np.uint8).reshape(sk.shape[0], -1)
        batch, mlen = msgs.shape
        row = a128(z.esk_bytes) + a128(mlen) + a128(z.sig_bytes) + a128(z.scratch_bytes)
        pool, pitch = self.pool(batch, row)
        h_sk, h_msg = self.pinned(sk.shape, sk), self.pinned(msgs.shape, msgs)
        sig = self.pinned((batch, z.sig_bytes))
        st = self.stream_handle()
        self.lib.crypto_sign_signWithesk_HtD(h_sk.data_ptr(), h_msg.data_ptr(), mlen, pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_signWithesk(pool.data_ptr(), mlen, pitch, batch, st)
        self.lib.crypto_sign_signWithesk_DtH(sig.data_ptr(), pool.data_ptr(), mlen, pitch, batch, st)
        self.torch.cuda.synchronize(self.device)
        self.check()
        return sig.numpy().copy()

    def verify(self, epk: np.ndarray, msgs: np.ndarray, sigs: np.ndarray) -> np.ndarray:
        """crypto_sign_verify: host epk and sig rows A128-padded (api.cu:148-155)."""
        z = self.sizes
        epk_len = z.p_bytes + z.pk_bytes - 16
        epk = np.ascontiguousarray(epk, np.uint8).reshape(-1, epk_len)
        batch = epk.shape[0]
        msgs = np.ascontiguousarray(msgs, np.uint8).reshape(batch, -1)
        mlen = msgs.shape[1]
        epk_pad = np.zeros((batch, a128(epk_len)), np.uint8)
        epk_pad[:, :epk_len] = epk
        sig_pad = np.zeros((batch, a128(z.sig_bytes)), np.uint8)
        sig_pad[:, :z.sig_bytes] = np.asarray(sigs, np.uint8).reshape(batch, -1)
        row = a128(epk_len) + a128(mlen) + a128(z.sig_bytes) + 1
        pool, pitch = self.pool(batch, row)
        res = self.pinned((batch,), np.full(batch, 0xEE, np.uint8))
        h_epk, h_msg, h_sig = (self.pinned(epk_pad.shape, epk_pad), self.pinned(msgs.shape, msgs),
                               self.pinned(sig_pad.shape, sig_pad))
        st = self.stream_handle()
        self.lib.crypto_sign_verify_HtD(h_epk.data_ptr(), h_msg.data_ptr(), h_sig.data_ptr(), mlen,
                                        pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_verify(pool.data_ptr(), mlen, pitch, batch, st)
        self.lib.crypto_sign_verify_DtH(res.data_ptr(), mlen, pool.data_ptr(), pitch, batch, st)
        self.torch.cuda.synchronize(self.device)
        self.check()
        return res.numpy().copy()

    def verify_with_pk(self, pk: np.ndarray, msgs: np.ndarray, sigs: np.ndarray) -> np.ndarray:
        """crypto_sign_verifyWithepk: compact pk in, expanded on the GPU, verified."""
        z = self.sizes
        epk_len = z.p_bytes + z.pk_bytes - 16
        pk = np.ascontiguousarray(pk, np.uint8).reshape(-1, z.pk_bytes)
        batch = pk.shape[0]
        msgs = np.ascontiguousarray(msgs, np.uint8).reshape(batch, -1)
        mlen = msgs.shape[1]
        sigs = np.ascontiguousarray(sigs, np.uint8).reshape(batch, z.sig_bytes)
        row = a128(epk_len) + a128(mlen) + a128(z.sig_bytes) + 1
        pool, pitch = self.pool(batch, row)
        res = self.pinned((batch,), np.full(batch, 0xEE, np.uint8))
        h_pk, h_msg, h_sig = (self.pinned(pk.shape, pk), self.pinned(msgs.shape, msgs),
                              self.pinned(sigs.shape, sigs))
        st = self.stream_handle()
        self.lib.crypto_sign_verifyWithepk_HtD(h_pk.data_ptr(), h_msg.data_ptr(), h_sig.data_ptr(), mlen,
                                               pool.data_ptr(), pitch, batch, st)
        self.lib.crypto_sign_verifyWithepk(pool.data_ptr(), mlen, pitch, batch, st)
        self.lib.crypto_sign_verifyWithepk_DtH(res.data_ptr(), mlen, pool.data_ptr(), pitch, batch, st)
        self.torch.cuda.synchronize(self.device)
        self.check()
        return res.numpy().copy()

    def keypair_tables(self, nkeys: int, rand_index: int = 0):
        """Generates nkeys key pairs and expands them into dense device
        tables (esk rows, epk rows) without leaving the device."""
        z = self.sizes
        torch = self.torch
        row = a128(z.sk_bytes) + a128(z.pk_bytes) + a128(z.p_bytes)
        pool, pitch = self.pool(nkeys, row)
        st = self.stream_handle()
        self.lib.crypto_sign_keypair(pool.data_ptr(), pitch, nkeys, st, rand_index)
        esk = torch.zeros((nkeys, z.esk_bytes), dtype=torch.uint8, device=self.device)
        epk = torch.zeros((nkeys, z.epk_bytes), dtype=torch.uint8, device=self.device)
        pk_off = a128(z.sk_bytes)
        esk[:, :z.sk_bytes].copy_(pool[:nkeys, :z.sk_bytes])
        epk[:, z.p_bytes - 16:].copy_(pool[:nkeys, pk_off:pk_off + z.pk_bytes])
        self.lib.crypto_sign_expandsk(esk.data_ptr(), z.esk_bytes, nkeys, st)
        self.lib.crypto_sign_expandpk(epk.data_ptr(), z.epk_bytes, nkeys, st)
        torch.cuda.synchronize(self.device)
        self.check()
        return esk, epk

    def expand_tables(self, sk: np.ndarray, pk: np.ndarray):
        """Expanded key tables on the device: esk rows and epk rows packed
        densely (pool pitch = row length), built by crypto_sign_expandsk /
        crypto_sign_expandpk without a host round trip."""
        z = self.sizes
        torch = self.torch
        sk = np.ascontiguousarray(sk, np.uint8).reshape(-1, z.sk_bytes)
        pk = np.ascontiguousarray(pk, np.uint8).reshape(-1, z.pk_bytes)
        nk = sk.shape[0]
        st = self.stream_handle()
        h_sk, h_pk = self.pinned(sk.shape, sk), self.pinned(pk.shape, pk)
        esk = torch.zeros((nk, z.esk_bytes), dtype=torch.uint8, device=self.device)
        self.lib.crypto_sign_expandsk_HtD(h_sk.data_ptr(), esk.data_ptr(), z.esk_bytes, nk, st)
        self.lib.crypto_sign_expandsk(esk.data_ptr(), z.esk_bytes, nk, st)
        epk = torch.zeros((nk, z.epk_bytes), dtype=torch.uint8, device=self.device)
        self.lib.crypto_sign_expandpk_HtD(h_pk.data_ptr(), epk.data_ptr(), z.epk_bytes, nk, st)
        self.lib.crypto_sign_expandpk(epk.data_ptr(), z.epk_bytes, nk, st)
        torch.cuda.synchronize(self.device)
        self.check()
        return esk, epk

    # -- batched key-table form (device tensors in, enqueue only) ----------
    @property
    def sign_attempts_offset(self) -> int:
        return self.sizes.sign_ws_attempts

    def sign_attempts(self, workspace, n: int) -> np.ndarray:
        """Attempts each item of the last sign_table call used (0 = exhausted)."""
        z = self.sizes
        return workspace[: n * z.sign_ws_stride].view(n, z.sign_ws_stride)[:, z.sign_ws_attempts].cpu().numpy()

    def sign_workspace(self, n: int):
        return self.torch.empty(max(int(self.lib.cumayo_sign_workspace_size(n)), 1),
                                dtype=self.torch.uint8, device=self.device)

    def verify_workspace(self, n: int):
        return self.torch.empty(max(int(self.lib.cumayo_verify_workspace_size(n)), 1),
                                dtype=self.torch.uint8, device=self.device)

    def sign_table(self, esk_table, key_idx, msgs, msg_off, msg_len, sigs, workspace, stream=None,
                   stages=3):
        z = self.sizes
        n = int(key_idx.numel())
        err = self.lib.cumayo_sign_table_stages(
            esk_table.data_ptr(), esk_table.stride(0) if esk_table.dim() > 1 else z.esk_bytes,
            key_idx.data_ptr(), msgs.data_ptr(), msg_off.data_ptr(), msg_len.data_ptr(),
            sigs.data_ptr(), n, workspace.data_ptr(), self.stream_handle(stream), stages)
        if err:
            raise RuntimeError(f"cumayo_sign_table failed: HIP error {err}")

    def verify_table(self, epk_table, key_idx, msgs, msg_off, msg_len, sigs, results, workspace,
                     stream=None, stages=3):
        z = self.sizes
        n = int(key_idx.numel())
        err = self.lib.cumayo_verify_table_stages(
            epk_table.data_ptr(), epk_table.stride(0) if epk_table.dim() > 1 else z.epk_bytes,
            key_idx.data_ptr(), msgs.data_ptr(), msg_off.data_ptr(), msg_len.data_ptr(),
            sigs.data_ptr(), results.data_ptr(), n, workspace.data_ptr(),
            self.stream_handle(stream), stages)
        if err:
            raise RuntimeError(f"cumayo_verify_table failed: HIP error {err}")

    # -- prepared key tables ------------------------------------------------
    def prepare_tables(self, esk_table=None, epk_table=None, stream=None):
        """Prepared (padded, aligned) copies of reference-format key tables
        (device tensors, one key per row); None where no table is given."""
        torch = self.torch
        st = self.stream_handle(stream)
        out = []
        for tab, nbytes, fn in ((esk_table, self.lib.cumayo_prepared_esk_bytes(), self.lib.cumayo_prepare_esk),
                                (epk_table, self.lib.cumayo_prepared_epk_bytes(), self.lib.cumayo_prepare_epk)):
            if tab is None:
                out.append(None)
                continue
            nk = tab.shape[0]
            prep = torch.empty((nk, int(nbytes)), dtype=torch.uint8, device=self.device)
            # one call takes at most PREPARE_KEYS_MAX keys
            for k0 in range(0, nk, PREPARE_KEYS_MAX):
                k = min(PREPARE_KEYS_MAX, nk - k0)
                err = fn(tab[k0].data_ptr(), tab.stride(0), k, prep[k0].data_ptr(), st)
                if err:
                    raise RuntimeError(f"cumayo_prepare_* failed: HIP error {err}")
            out.append(prep)
        return tuple(out)

    def sign_prepared(self, prep_esk, key_idx, msgs, msg_off, msg_len, sigs, workspace, stream=None,
                      stages=3):
        n = int(key_idx.numel())
        err = self.lib.cumayo_sign_prepared(
            prep_esk.data_ptr(), key_idx.data_ptr(), msgs.data_ptr(), msg_off.data_ptr(),
            msg_len.data_ptr(), sigs.data_ptr(), n, workspace.data_ptr(), self.stream_handle(stream), stages)
        if err:
            raise RuntimeError(f"cumayo_sign_prepared failed: HIP error {err}")

    def verify_prepared(self, prep_epk, key_idx, msgs, msg_off, msg_len, sigs, results, workspace,
                        stream=None, stages=3):
        n = int(key_idx.numel())
        err = self.lib.cumayo_verify_prepared(
            prep_epk.data_ptr(), key_idx.data_ptr(), msgs.data_ptr(), msg_off.data_ptr(),
            msg_len.data_ptr(), sigs.data_ptr(), results.data_ptr(), n, workspace.data_ptr(),
            self.stream_handle(stream), stages)
        if err:
            raise RuntimeError(f"cumayo_verify_prepared failed: HIP error {err}")

    # -- NIST wire format (sm = signature || message) -----------------------
    def crypto_sign(self, prep_esk, key_idx, msgs, msg_off, msg_len, sm, sm_off, stream=None):
        n = int(key_idx.numel())
        ws = self.torch.empty(max(int(self.lib.cumayo_crypto_sign_workspace_size(n)), 1),
                              dtype=self.torch.uint8, device=self.device)
        err = self.lib.cumayo_crypto_sign(prep_esk.data_ptr(), key_idx.data_ptr(), msgs.data_ptr(),
                                          msg_off.data_ptr(), msg_len.data_ptr(), sm.data_ptr(),
                                          sm_off.data_ptr(), n, ws.data_ptr(), self.stream_handle(stream))
        if err:
            raise RuntimeError(f"cumayo_crypto_sign failed: HIP error {err}")
        return ws

    def crypto_sign_open(self, prep_epk, key_idx, sm, sm_off, sm_len, msgs_out, out_off, mlen_out,
                         results, stream=None):
        n = int(key_idx.numel())
        ws = self.torch.empty(max(int(self.lib.cumayo_crypto_sign_open_workspace_size(n)), 1),
                              dtype=self.torch.uint8, device=self.device)
        err = self.lib.cumayo_crypto_sign_open(
            prep_epk.data_ptr(), key_idx.data_ptr(), sm.data_ptr(), sm_off.data_ptr(), sm_len.data_ptr(),
            msgs_out.data_ptr(), out_off.data_ptr(), mlen_out.data_ptr() if mlen_out is not None else None,
            results.data_ptr(), n, ws.data_ptr(), self.stream_handle(stream))
        if err:
            raise RuntimeError(f"cumayo_crypto_sign_open failed: HIP error {err}")
        return ws

    # -- launch autotuner and captured pipelines ---------------------------
    def launch_config(self) -> dict:
        c = LaunchConfig()
        self.lib.cumayo_get_launch_config(ctypes.byref(c))
        return c.as_dict()

    def set_launch_config(self, **fields):
        c = LaunchConfig(**{**self.launch_config(), **fields})
        err = self.lib.cumayo_set_launch_config(ctypes.byref(c))
        if err:
            raise ValueError(f"not a compiled launch variant: {c.as_dict()}")

    def launch_variants(self) -> dict:
        out = {}
        for name, f in LAUNCH_FIELDS.items():
            buf = (ctypes.c_int32 * 32)()
            cnt = min(self.lib.cumayo_launch_variants(f, buf, 32), 32)
            out[name] = [int(buf[i]) for i in range(cnt)]
        return out

    def autotune(self, prep_esk, prep_epk, key_idx, msgs, msg_off, msg_len, sigs, results=None,
                 sign_ws=None, verify_ws=None, reps: int = 3, stream=None) -> dict:
        """Times the linear-algebra pass under every compiled launch variant
        on this batch and keeps the fastest (cumayo_autotune)."""
        n = int(key_idx.numel())
        rep = TuneReport()
        err = self.lib.cumayo_autotune(
            _ptr(prep_esk), _ptr(prep_epk), key_idx.data_ptr(), msgs.data_ptr(), msg_off.data_ptr(),
            msg_len.data_ptr(), sigs.data_ptr(), _ptr(results), n, _ptr(sign_ws), _ptr(verify_ws),
            self.stream_handle(stream), reps, ctypes.byref(rep))
        if err:
            raise RuntimeError(f"cumayo_autotune failed: HIP error {err}")
        return {"best": rep.best.as_dict(), "sign_ms": rep.sign_ms, "verify_ms": rep.verify_ms,
                "sign_default_ms": rep.sign_default_ms, "verify_default_ms": rep.verify_default_ms,
                "configs_timed": int(rep.configs_timed)}

    def sign_graph(self, prep_esk, key_idx, msgs, msg_off, msg_len, sigs, workspace, stages=3) -> "Graph":
        """cumayo_sign_prepared over these buffers, captured as one HIP graph."""
        h = self.lib.cumayo_sign_graph_create(prep_esk.data_ptr(), key_idx.data_ptr(), msgs.data_ptr(),
                                              msg_off.data_ptr(), msg_len.data_ptr(), sigs.data_ptr(),
                                              int(key_idx.numel()), workspace.data_ptr(), stages)
        if not h:
            raise RuntimeError(f"cumayo_sign_graph_create failed: HIP error {self.lib.cumayo_last_error()}")
        return Graph(self, h)

    def verify_graph(self, prep_epk, key_idx, msgs, msg_off, msg_len, sigs, results, workspace,
                     stages=3) -> "Graph":
        h = self.lib.cumayo_verify_graph_create(prep_epk.data_ptr(), key_idx.data_ptr(), msgs.data_ptr(),
                                                msg_off.data_ptr(), msg_len.data_ptr(), sigs.data_ptr(),
                                                results.data_ptr(), int(key_idx.numel()), workspace.data_ptr(),
                                                stages)
        if not h:
            raise RuntimeError(f"cumayo_verify_graph_create failed: HIP error {self.lib.cumayo_last_error()}")
        return Graph(self, h)


    def compose_graphs(self, graphs) -> "Graph":
        """One graph running the given pipelines (Graph objects of any
        parameter set's engine) as independent children: one launch."""
        arr = (ctypes.c_void_p * len(graphs))(*[g.handle for g in graphs])
        h = self.lib.cumayo_graph_compose(ctypes.cast(arr, ctypes.c_void_p), len(graphs))
        if not h:
            raise RuntimeError(f"cumayo_graph_compose failed: HIP error {self.lib.cumayo_last_error()}")
        return Graph(self, h)


class Graph:
    """A captured batch pipeline; launch() replays it on a stream."""

    def __init__(self, engine: Engine, handle: int):
        self.engine, self.handle = engine, handle

    def launch(self, stream=None):
        if not self.handle:
            raise RuntimeError("graph destroyed")
        err = self.engine.lib.cumayo_graph_launch(self.handle, self.engine.stream_handle(stream))
        if err:
            raise RuntimeError(f"cumayo_graph_launch failed: HIP error {err}")

    def close(self):
        if self.handle:
            self.engine.lib.cumayo_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
