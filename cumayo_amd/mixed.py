"""Mixed parameter sets on one GPU.

Each MAYO parameter set is its own library (one build per set, as in the
reference), so a mixed stream is served by one continuous-batching server per
set, all sharing the device: a request is routed to its set's server, the
servers' batches overlap on the GPU, and latencies are merged. Under load,
each set gets its own open-loop Poisson stream (rate x weight); the
superposition of independent Poisson streams is the Poisson stream of the
whole mix.
"""
from __future__ import annotations

import threading

import numpy as np

from .serve import Server


def percentiles(lat_ns: np.ndarray) -> dict:
    if lat_ns.size == 0:
        return {"p50_ms": 0.0, "p90_ms": 0.0, "p99_ms": 0.0, "max_ms": 0.0}
    q = np.percentile(lat_ns, [50, 90, 99]) / 1e6
    return {"p50_ms": float(q[0]), "p90_ms": float(q[1]), "p99_ms": float(q[2]),
            "max_ms": float(lat_ns.max() / 1e6)}


class MixedServer:
    """servers[v] = Server(engines[v], esk_v, epk_v); requests carry their set."""

    def __init__(self, engines: dict, tables: dict, max_batch: int = 4096, n_slots: int = 3,
                 slot_msg_bytes: int | None = None):
        self.servers = {v: Server(engines[v], esk, epk, max_batch=max_batch, n_slots=n_slots,
                                  slot_msg_bytes=slot_msg_bytes)
                        for v, (esk, epk) in tables.items()}

    def close(self):
        for s in self.servers.values():
            s.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, reqs) -> int:
        """reqs: iterable of (tag, version, op, key, msg, sig_or_None); tags
        must be unique across the whole mix."""
        by_set = {}
        for tag, v, op, key, msg, sig in reqs:
            by_set.setdefault(v, []).append((tag, op, key, msg, sig))
        return sum(self.servers[v].submit(r) for v, r in by_set.items())

    def wait(self, tags_by_set: dict, timeout_s: float = 60.0) -> dict:
        out = {}
        for v, tags in tags_by_set.items():
            for tag, val in self.servers[v].wait(tags, timeout_s).items():
                out[tag] = (v,) + tuple(val)
        return out

    def run_poisson(self, rate: float, duration: float, weights: dict, verify_fraction: float,
                    pools: dict, seed: int = 1) -> dict:
        """Open-loop load over the mix: set v receives rate * weights[v]
        arrivals per second from its pool (flat, offs, lens, keys, sigs).
        Returns the merged report plus one report per set."""
        reports, errors = {}, []

        def run(v):
            try:
                flat, offs, lens, keys, sigs = pools[v]
                reports[v] = self.servers[v].run_poisson(rate * weights[v], duration, verify_fraction,
                                                         seed + v, flat, offs, lens, keys, sigs,
                                                         latencies=True)
            except Exception as e:  # reported after the join
                errors.append(e)

        threads = [threading.Thread(target=run, args=(v,)) for v in weights]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        lat = np.concatenate([r.pop("latencies_ns") for r in reports.values()])
        dur = max(r["duration_s"] for r in reports.values())
        merged = {"offered_per_s": sum(r["offered_per_s"] for r in reports.values()),
                  "achieved_per_s": float(lat.size / dur) if dur else 0.0,
                  "completed": int(sum(r["completed"] for r in reports.values())),
                  "refused": int(sum(r["refused"] for r in reports.values())),
                  **percentiles(lat)}
        merged["per_set"] = {f"MAYO_{v}": {k: (round(x, 3) if isinstance(x, float) else x)
                                           for k, x in r.items()} for v, r in reports.items()}
        return merged
