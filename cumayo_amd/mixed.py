"""Mixed parameter sets on one GPU.

Each MAYO parameter set is its own library (one build per set, as in the
reference), so its kernels are compiled for that set's sizes. A batch that
mixes sets (MixedBatch) is split by set on the host and each set's part runs
through its library on a HIP stream of its own, so the sets' kernels overlap
on the device; MixedGraph makes such a pass one HIP graph (each set's
pipeline, captured by its own library, a child of one graph), so that one
graph launch runs it. A mixed request stream is served by one continuous-batching
server per set (MixedServer), all sharing the device: a request is routed to
its set's server, the servers' batches overlap on the GPU, and latencies are
merged. Under load, each set gets its own open-loop Poisson stream (rate x
weight); the superposition of independent Poisson streams is the Poisson
stream of the whole mix.
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


def split_by_set(set_of, sets) -> dict:
    """{v: positions of the items of set v, in batch order} for the sets in
    `sets` that occur in set_of."""
    set_of = np.asarray(set_of)
    return {v: np.flatnonzero(set_of == v) for v in sets if (set_of == v).any()}


class MixedBatch:
    """One batch whose items use different parameter sets.

    engines[v] is the Engine of set v and prepared[v] its prepared key tables
    (prep_esk, prep_epk), either of which may be None. Item i uses set
    set_of[i], key key_idx[i] of that set's tables and message bytes
    msgs[msg_off[i] : msg_off[i] + msg_len[i]] of one flat device buffer.
    sign() and verify() split the items by set (host index arrays), gather
    each set's key indices, offsets and lengths to the device, and launch
    every set's hashing and linear-algebra passes on its own stream after
    the caller's stream; the caller's stream then waits for all of them.
    Results equal those of one single-set batch per set."""

    def __init__(self, engines: dict, prepared: dict):
        self.engines, self.prepared = engines, prepared
        eng = next(iter(engines.values()))
        self.torch = eng.torch
        # one stream per set, default priority
        self.streams = {v: self.torch.cuda.Stream(device=e.device) for v, e in engines.items()}

    def split(self, set_of) -> dict:
        """{v: positions of the items of set v, in batch order}."""
        return split_by_set(set_of, self.engines)

    def _parts(self, set_of, key_idx, msg_off, msg_len):
        torch = self.torch
        parts = {}
        for v, pos in self.split(set_of).items():
            dev = self.engines[v].device
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev, non_blocking=False)
            parts[v] = (pos, t(np.asarray(key_idx, np.int32)[pos]), t(np.asarray(msg_off, np.int64)[pos]),
                        t(np.asarray(msg_len, np.int32)[pos]))
        unknown = set(np.unique(np.asarray(set_of)).tolist()) - set(self.engines)
        if unknown:
            raise ValueError(f"no engine for parameter sets {sorted(unknown)}")
        return parts

    def _launch(self, parts, body, concurrent=True):
        torch = self.torch
        cur = torch.cuda.current_stream()
        if not concurrent:
            # every set on the caller's stream, one after another
            for v, part in parts.items():
                body(v, part, cur)
            return
        for v, part in parts.items():
            st = self.streams[v]
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                body(v, part, st)
        for v in parts:
            cur.wait_stream(self.streams[v])

    def sign(self, set_of, key_idx, msgs, msg_off, msg_len, concurrent=True) -> dict:
        """Returns {v: (positions, sigs)}: sigs[k] (device, sig_bytes of set v)
        is the signature of item positions[k]. concurrent=False runs the sets
        one after another on the caller's stream (same results)."""
        torch = self.torch
        parts = self._parts(set_of, key_idx, msg_off, msg_len)
        out = {}

        def body(v, part, st):
            eng = self.engines[v]
            pos, dk, do, dl = part
            n = pos.size
            sigs = torch.empty((n, eng.sizes.sig_bytes), dtype=torch.uint8, device=eng.device)
            ws = eng.sign_workspace(n)
            eng.sign_prepared(self.prepared[v][0], dk, msgs, do, dl, sigs, ws, stream=st)
            for x in (dk, do, dl, sigs, ws):
                x.record_stream(st)
            msgs.record_stream(st)
            out[v] = (pos, sigs)

        self._launch(parts, body, concurrent)
        return out

    def verify(self, set_of, key_idx, msgs, msg_off, msg_len, sigs: dict, concurrent=True):
        """sigs[v]: (n_v, sig_bytes) device signatures of set v's items in
        batch order (as sign() returns them). Returns a device uint8 tensor of
        results in batch order (0 valid, 1 invalid)."""
        torch = self.torch
        parts = self._parts(set_of, key_idx, msg_off, msg_len)
        dev = next(iter(self.engines.values())).device
        results = torch.full((len(set_of),), 0xEE, dtype=torch.uint8, device=dev)
        per = {}

        def body(v, part, st):
            eng = self.engines[v]
            pos, dk, do, dl = part
            n = pos.size
            res = torch.empty(n, dtype=torch.uint8, device=eng.device)
            ws = eng.verify_workspace(n)
            eng.verify_prepared(self.prepared[v][1], dk, msgs, do, dl, sigs[v], res, ws, stream=st)
            for x in (dk, do, dl, res, ws):
                x.record_stream(st)
            msgs.record_stream(st)
            sigs[v].record_stream(st)
            per[v] = (pos, res)

        self._launch(parts, body, concurrent)
        for v, (pos, res) in per.items():
            results[torch.from_numpy(pos).to(dev)] = res
        return results


class MixedGraph:
    """The sign or verify pass of one mixed-set batch as one HIP graph: each
    set's pipeline (hashing, linear-algebra and retry kernels) is captured by
    its own library (cumayo_sign_graph_create / cumayo_verify_graph_create)
    over buffers this object owns, and the pipelines become the independent
    children of one graph (cumayo_graph_compose), so one graph launch runs
    every set and the sets' kernels may overlap on the device. The graphs
    hold kernel nodes only (no memory-set or copy nodes). The per-item arrays
    (key indices, offsets, lengths and, for verification, the signatures) are
    copied into the owned buffers at construction; messages are read from
    `msgs` at every launch, so the caller may rewrite their bytes in place
    between launches (same offsets). Launches must not overlap each other."""

    def __init__(self, mb: MixedBatch, op: str, set_of, key_idx, msgs, msg_off, msg_len, sigs=None):
        if op not in ("sign", "verify"):
            raise ValueError(f"op must be 'sign' or 'verify', not {op!r}")
        if op == "verify" and sigs is None:
            raise ValueError("verify needs the signatures, {set: tensor} as MixedBatch.sign gives them")
        torch = mb.torch
        self.mb, self.op, self.msgs = mb, op, msgs
        self.parts = mb._parts(set_of, key_idx, msg_off, msg_len)
        self.out, self.ws, self.sigs, self.graphs = {}, {}, {}, {}
        for v, (pos, dk, do, dl) in self.parts.items():
            eng = mb.engines[v]
            n = pos.size
            if op == "sign":
                self.out[v] = torch.zeros((n, eng.sizes.sig_bytes), dtype=torch.uint8, device=eng.device)
                self.ws[v] = eng.sign_workspace(n)
                self.graphs[v] = eng.sign_graph(mb.prepared[v][0], dk, msgs, do, dl, self.out[v], self.ws[v])
            else:
                self.sigs[v] = sigs[v].clone()
                self.out[v] = torch.full((n,), 0xEE, dtype=torch.uint8, device=eng.device)
                self.ws[v] = eng.verify_workspace(n)
                self.graphs[v] = eng.verify_graph(mb.prepared[v][1], dk, msgs, do, dl, self.sigs[v], self.out[v],
                                                  self.ws[v])
        first = mb.engines[next(iter(self.graphs))]
        self.graph = first.compose_graphs(list(self.graphs.values()))

    def launch(self, stream=None):
        """One graph launch on `stream` (the current stream by default)."""
        self.graph.launch(stream)

    def results(self) -> dict:
        """{v: (positions, out)}: signatures (sign) or results (verify, 0
        valid, 1 invalid) of set v's items in batch order, as of the last
        launch."""
        return {v: (part[0], self.out[v]) for v, part in self.parts.items()}

    def close(self):
        self.graph.close()
        for g in self.graphs.values():
            g.close()


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
