"""Continuous-batching server (ctypes over ``cumayo_serve_*``).

``Server(engine, esk_table, epk_table)`` starts the native worker thread of
the engine's library: requests are queued on the host and admitted into free
batch slots as earlier batches finish, so batches overlap on the GPU. The
Python side only packs request structs and keeps their buffers alive until
their completion has been polled.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t

OP_SIGN = 0
OP_VERIFY = 1
STATUS_REFUSED = 2
STATUS_CANCELLED = 3
STATUS_ERROR = 4


class ServeError(RuntimeError):
    """A batch failed on the device: its requests completed with
    STATUS_ERROR and wrote no output. ``completions`` holds everything the
    poll returned (the failed requests included), ``hip_error`` the server's
    last HIP error."""

    def __init__(self, hip_error: int, completions):
        super().__init__(f"server batch failed: HIP error {hip_error}; "
                         f"{sum(1 for c in completions if c[1] == STATUS_ERROR)} requests completed with STATUS_ERROR")
        self.hip_error = hip_error
        self.completions = completions


class ServeConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("esk_table", _VP), ("esk_stride", _SZ),
                ("epk_table", _VP), ("epk_stride", _SZ), ("n_keys", ctypes.c_int32),
                ("max_batch", ctypes.c_int32), ("n_slots", ctypes.c_int32),
                ("slot_msg_bytes", _SZ), ("max_queue", ctypes.c_int64),
                ("host_threads", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint8), ("key", ctypes.c_int32), ("msg", _VP),
                ("len", ctypes.c_uint32), ("sig_in", _VP), ("sig_out", _VP),
                ("result_out", _VP), ("tag", ctypes.c_uint64), ("t_arrival_ns", ctypes.c_int64)]


class Completion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("latency_ns", ctypes.c_int64), ("status", ctypes.c_uint8)]


class ServeStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("requests", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("max_in_flight", ctypes.c_uint64),
                ("overlapped_launches", ctypes.c_uint64), ("device_bytes", _SZ),
                ("cancelled", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("last_error", ctypes.c_int32)]


class ServeReport(ctypes.Structure):
    _fields_ = [("offered_per_s", ctypes.c_double), ("achieved_per_s", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p90_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("submitted", ctypes.c_uint64),
                ("completed", ctypes.c_uint64), ("refused", ctypes.c_uint64),
                ("duration_s", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


EXPORTS = {
    "cumayo_serve_create": ([ctypes.POINTER(ServeConfig)], _VP),
    "cumayo_serve_destroy": ([_VP], None),
    "cumayo_serve_submit": ([_VP, ctypes.POINTER(Request), _SZ], _SZ),
    "cumayo_serve_poll": ([_VP, ctypes.POINTER(Completion), _SZ], _SZ),
    "cumayo_serve_get_stats": ([_VP, ctypes.POINTER(ServeStats)], None),
    "cumayo_serve_clock_ns": ([], ctypes.c_int64),
    "cumayo_serve_inject_failures": ([_VP, ctypes.c_int], None),
    "cumayo_serve_cancel": ([_VP, ctypes.c_uint64], ctypes.c_int),
    "cumayo_serve_run_poisson": ([_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_uint64, _VP, _VP, _VP, _VP, _VP, _SZ,
                                  ctypes.POINTER(ServeReport), _VP, _SZ], ctypes.c_int),
}


def bind(lib: ctypes.CDLL) -> None:
    for name, (args, res) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


class Server:
    def __init__(self, engine, esk_table=None, epk_table=None, max_batch: int = 4096,
                 n_slots: int = 3, slot_msg_bytes: int | None = None, max_queue: int = 0,
                 host_threads: int = 0):
        self.eng = engine
        self.lib = engine.lib
        bind(self.lib)
        z = engine.sizes
        self.sig_bytes = z.sig_bytes
        n_keys = 0
        if esk_table is not None:
            n_keys = esk_table.shape[0]
        if epk_table is not None:
            n_keys = epk_table.shape[0] if esk_table is None else min(n_keys, epk_table.shape[0])
        self._tables = (esk_table, epk_table)  # keep alive
        cfg = ServeConfig(
            device=engine.device.index or 0,
            esk_table=esk_table.data_ptr() if esk_table is not None else None,
            esk_stride=esk_table.stride(0) if esk_table is not None else 0,
            epk_table=epk_table.data_ptr() if epk_table is not None else None,
            epk_stride=epk_table.stride(0) if epk_table is not None else 0,
            n_keys=n_keys, max_batch=max_batch, n_slots=n_slots,
            slot_msg_bytes=slot_msg_bytes or max_batch * 1024, max_queue=max_queue,
            host_threads=host_threads)
        engine.torch.cuda.synchronize(engine.device)
        self.handle = self.lib.cumayo_serve_create(ctypes.byref(cfg))
        if not self.handle:
            raise RuntimeError(f"cumayo_serve_create failed: HIP error {self.lib.cumayo_last_error()}")
        self._live = {}  # tag -> buffers referenced by an in-flight request

    def close(self):
        if self.handle:
            self.lib.cumayo_serve_destroy(self.handle)
            self.handle = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        s = ServeStats()
        self.lib.cumayo_serve_get_stats(self.handle, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def submit(self, reqs) -> int:
        """reqs: iterable of (tag, op, key, msg_bytes, sig_bytes_or_None).
        Output buffers are allocated here; read them with ``result(tag)``."""
        arr = []
        for tag, op, key, msg, sig in reqs:
            m = np.frombuffer(bytes(msg) + b"\0", np.uint8)
            out_sig = np.zeros(self.sig_bytes, np.uint8)
            out_res = np.full(1, 0xEE, np.uint8)
            s_in = None if sig is None else np.frombuffer(bytes(sig), np.uint8)
            if op == OP_VERIFY and (s_in is None or s_in.size != self.sig_bytes):
                raise ValueError("verify request needs a signature of sig_bytes")
            self._live[tag] = (m, out_sig, out_res, s_in)
            arr.append(Request(op=op, key=key, msg=m.ctypes.data, len=len(msg),
                               sig_in=s_in.ctypes.data if s_in is not None else None,
                               sig_out=out_sig.ctypes.data, result_out=out_res.ctypes.data,
                               tag=tag, t_arrival_ns=0))
        if not arr:
            return 0
        ra = (Request * len(arr))(*arr)
        acc = int(self.lib.cumayo_serve_submit(self.handle, ra, len(arr)))
        for r in arr[acc:]:
            self._live.pop(r.tag, None)
        return acc

    def cancel(self, tag: int) -> bool:
        """Withdraws a queued request; it then completes with status 3."""
        return bool(self.lib.cumayo_serve_cancel(self.handle, tag))

    def inject_failures(self, n: int) -> None:
        """Test hook: the next n batches fail (their requests complete with
        STATUS_ERROR, nothing is launched for them)."""
        self.lib.cumayo_serve_inject_failures(self.handle, n)

    def poll(self, max_n: int = 65536):
        """Completed requests as (tag, status, latency_ns, sig, result).
        Raises ServeError (with every polled completion attached) when one of
        them belongs to a batch that failed on the device."""
        buf = (Completion * max_n)()
        k = int(self.lib.cumayo_serve_poll(self.handle, buf, max_n))
        out = []
        for i in range(k):
            c = buf[i]
            m, out_sig, out_res, _ = self._live.pop(c.tag)
            out.append((c.tag, c.status, c.latency_ns, out_sig, int(out_res[0])))
        if any(c[1] == STATUS_ERROR for c in out):
            raise ServeError(self.stats()["last_error"], out)
        return out

    def wait(self, tags, timeout_s: float = 60.0):
        """Polls until every tag has completed; returns {tag: (status, latency_ns, sig, result)}.
        A failed batch raises ServeError from poll()."""
        want = set(tags)
        got = {}
        t_end = time.monotonic() + timeout_s
        while want - got.keys():
            for tag, st, lat, sig, res in self.poll():
                got[tag] = (st, lat, sig, res)
            if time.monotonic() > t_end:
                raise TimeoutError(f"{len(want - got.keys())} requests still pending")
            time.sleep(0.0005)
        return got

    def run_poisson(self, rate: float, duration: float, verify_fraction: float, seed: int,
                    flat: np.ndarray, offs: np.ndarray, lens: np.ndarray, keys: np.ndarray,
                    sigs: np.ndarray | None = None, latencies: bool = False) -> dict:
        """Open-loop Poisson load (native generator thread: the GIL is
        released for the whole run). With latencies=True the report carries
        the raw latencies in ns under "latencies_ns"."""
        rep = ServeReport()
        cap = int(rate * duration * 1.5) + 1024 if latencies else 0
        lat = np.zeros(cap, np.int64)
        flat = np.ascontiguousarray(flat, np.uint8)
        offs = np.ascontiguousarray(offs, np.int64)
        lens = np.ascontiguousarray(lens, np.int32)
        keys = np.ascontiguousarray(keys, np.int32)
        sp = None
        if sigs is not None:
            sigs = np.ascontiguousarray(sigs, np.uint8)
            sp = sigs.ctypes.data
        err = self.lib.cumayo_serve_run_poisson(
            self.handle, rate, duration, verify_fraction, seed, flat.ctypes.data, offs.ctypes.data,
            lens.ctypes.data, keys.ctypes.data, sp, int(keys.size), ctypes.byref(rep),
            lat.ctypes.data if cap else None, cap)
        if err:
            raise RuntimeError(f"cumayo_serve_run_poisson: HIP error {err}")
        out = rep.as_dict()
        if latencies:
            out["latencies_ns"] = lat[:min(cap, int(rep.completed))].copy()
        return out
