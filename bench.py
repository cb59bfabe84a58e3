"""Benchmark: MAYO_1 batched sign and verify on one MI355X.

Workload (the headline configuration): 65,536 signatures under 64 shared
keys (key j mod 64) whose expanded secret and public keys stay resident in
HBM, messages of uniform random length in [32, 1024] bytes. One step is one
pass of the hot path over the whole batch: ``sign`` signs all 65,536
messages, ``verify`` verifies all 65,536 signatures. Inputs are generated
from SHAKE streams and are in HBM before the timed region.

Prints one JSON line. ``value`` is signatures per second; verification
throughput is reported beside it. ``roofline`` covers the signing
linear-algebra pass (the dominant kernels), timed with HIP events on the
stream it runs on. ``serve`` drives the continuous-batching server with
open-loop Poisson arrivals at a fraction of the batch throughput and reports
latency percentiles. ``cpu_baseline`` times the CPU oracle (the restatement
of the reference algorithm) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues per process
# (default 4); the serve runs keep up to 4 servers x 3 slot streams busy at
# once, which on 4 queues serialise behind each other. Raised to 16 (also
# over a smaller value already in the environment) before HIP starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

# Work per operation (GF(16) multiply-accumulates), from the reference loop
# bounds: sign counts one attempt (sign.cu:256-437 plus the elimination),
# verify counts verify.cu:72-182.
SIGN_MACS = {1: 3_677_947, 2: 982_688, 3: 10_147_914, 5: 24_133_255}
VERIFY_MACS = {1: 4_197_180, 2: 998_656, 3: 11_294_316, 5: 26_389_848}
# Nominal INT roofline: 256 CU x 64 lanes x 2.4 GHz 32-bit ops, 32 GF(16)
# MACs per 16 bitsliced AND-XOR ops -> 2 MACs per op.
PEAK_GF16_TMACS = 256 * 64 * 2.4e9 * 2 / 1e12
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the signing linear-algebra pass, measured with
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (scaled x2.00 and x1.00 by the
# same counters over a 256 MiB copy) on the default workload, per launch form
# (wave: one wavefront per item through the whole quadratic form; shared:
# the shared-key table and fold kernels); the runs are summarised in
# profiles/sign_verify_mayo1_65536x64keys.md
STAGE2_TRAFFIC = {(1, 65536, 64, "wave"): 578_434_540, (1, 65536, 64, "shared"): 5_310_404_268}
STAGE2_TRAFFIC_AT = {"wave": "75e698f", "shared": "75e698f"}  # the commit each counter run was made at


def make_workload(n: int, nkeys: int, lo: int, hi: int, tag: str = "cfg2"):
    """Messages j: SHAKE256(tag/msg || LE64(j)); lengths from SHAKE256(tag/len || LE64(j))."""
    lens = np.empty(n, np.int32)
    for j in range(n):
        r = int.from_bytes(hashlib.shake_256(f"mayo-bench/{tag}/len".encode() + j.to_bytes(8, "little")).digest(8), "little")
        lens[j] = lo + r % (hi - lo + 1)
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    flat = bytearray(int(lens.sum()))
    for j in range(n):
        o = int(offs[j])
        flat[o:o + int(lens[j])] = hashlib.shake_256(f"mayo-bench/{tag}/msg".encode() + j.to_bytes(8, "little")).digest(int(lens[j]))
    key_idx = (np.arange(n) % nkeys).astype(np.int32)
    return np.frombuffer(bytes(flat), np.uint8), offs, lens, key_idx


def log_uniform_workload(n: int, nkeys: int, lo: int, hi: int, tag: str):
    """Like make_workload, lengths log-uniform in [lo, hi]: exp of a uniform
    draw on [ln lo, ln hi] from the same SHAKE length stream."""
    import math
    lens = np.empty(n, np.int32)
    for j in range(n):
        r = int.from_bytes(hashlib.shake_256(f"mayo-bench/{tag}/len".encode() + j.to_bytes(8, "little")).digest(8), "little")
        u = r / 2.0**64
        lens[j] = min(hi, max(lo, int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))))
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    flat = bytearray(int(lens.sum()))
    for j in range(n):
        o = int(offs[j])
        flat[o:o + int(lens[j])] = hashlib.shake_256(f"mayo-bench/{tag}/msg".encode() + j.to_bytes(8, "little")).digest(int(lens[j]))
    return np.frombuffer(bytes(flat), np.uint8), offs, lens, (np.arange(n) % nkeys).astype(np.int32)


def mixed_sets_run(eng1, args, weights=None, pool=8192, keys=64, loads=(0.25, 0.5, 0.75), seconds=None):
    """Serve a mix of parameter sets on the one GPU: MAYO_1/2/3/5 = 40/20/20/20,
    64 resident keys per set, message lengths log-uniform in [32 B, 16 KB],
    half of the requests verifications, swept over offered loads given as
    fractions of the mix's batch capacity (from each set's measured sign and
    verify rates at a batch of `pool`)."""
    import torch
    from cumayo_amd import Engine
    from cumayo_amd.mixed import MixedServer
    weights = weights or {1: 0.4, 2: 0.2, 3: 0.2, 5: 0.2}
    seconds = args.serve_seconds / 2 if seconds is None else seconds
    engines, tables, pools, cap = {}, {}, {}, {}
    for v in weights:
        e = eng1 if v == eng1.version else Engine(v)
        engines[v] = e
        esk, epk = e.keypair_tables(keys, rand_index=0)
        pesk, pepk = e.prepare_tables(esk, epk)
        flat, offs, lens, kidx = log_uniform_workload(pool, keys, 32, 16384, tag=f"cfg5-{v}")
        t = lambda a: torch.from_numpy(np.array(a)).to(e.device)
        dk, dm, do, dl = t(kidx), t(flat), t(offs), t(lens)
        sig = torch.zeros((pool, e.sizes.sig_bytes), dtype=torch.uint8, device=e.device)
        res = torch.zeros(pool, dtype=torch.uint8, device=e.device)
        ws_s, ws_v = e.sign_workspace(pool), e.verify_workspace(pool)
        rates = []
        for fn in (lambda: e.sign_prepared(pesk, dk, dm, do, dl, sig, ws_s),
                   lambda: e.verify_prepared(pepk, dk, dm, do, dl, sig, res, ws_v)):
            fn()
            torch.cuda.synchronize(e.device)
            t0 = time.perf_counter()
            for _ in range(3):
                fn()
            torch.cuda.synchronize(e.device)
            rates.append(3 * pool / (time.perf_counter() - t0))
        cap[v] = rates
        tables[v] = (esk, epk)
        pools[v] = (flat, offs, lens, kidx, sig.cpu().numpy())
    # mean GPU time per request of the mix (half signs, half verifications)
    per_req = sum(w * (0.5 / cap[v][0] + 0.5 / cap[v][1]) for v, w in weights.items())
    capacity = 1.0 / per_req
    points = []
    with MixedServer(engines, tables, max_batch=pool, n_slots=3, slot_msg_bytes=pool * 4096) as ms:
        ms.run_poisson(0.9 * capacity, 0.3, weights, 0.5, pools)  # unrecorded warm-up, all slots busy
        for load in loads:
            print(f"[bench] serve mixed sets at {load:.0%} of {capacity:.0f}/s", file=sys.stderr, flush=True)
            rep = ms.run_poisson(load * capacity, seconds, weights, 0.5, pools)
            per_set = {k: {q: v[q] for q in ("offered_per_s", "achieved_per_s", "p50_ms", "p99_ms")}
                       for k, v in rep.pop("per_set").items()}
            points.append({"offered_frac_of_capacity": load,
                           **{k: (round(x, 3) if isinstance(x, float) else x) for k, x in rep.items()},
                           "per_set": per_set})
    return {"mix": {f"MAYO_{v}": w for v, w in weights.items()},
            "messages": "log-uniform [32 B, 16 KB], 50% verification, 64 keys per set",
            "capacity_per_s": round(capacity, 1),
            f"batch_rates_at_{pool}": {f"MAYO_{v}": {"sign_per_s": round(c[0], 1),
                                                    "verify_per_s": round(c[1], 1)}
                                       for v, c in cap.items()},
            "vs_load": points}


def mixed_batch_run(eng1, n, steps, weights=None, keys=64):
    """One batch mixing the four parameter sets: item j belongs to MAYO_1/2/3/5
    with probabilities 40/20/20/20 (from its own SHAKE stream), key j mod 64
    of its set, message lengths uniform in [32, 1024] B. Sign and verify
    passes over the whole batch with every set's part on its own stream,
    against the same calls with the sets one after another on one stream."""
    import torch
    from cumayo_amd import Engine
    from cumayo_amd.mixed import MixedBatch
    weights = weights or {1: 0.4, 2: 0.2, 3: 0.2, 5: 0.2}
    engines, prepared = {}, {}
    for v in weights:
        e = eng1 if v == eng1.version else Engine(v)
        esk, epk = e.keypair_tables(keys, rand_index=0)
        engines[v], prepared[v] = e, e.prepare_tables(esk, epk)
    flat, offs, lens, kidx = make_workload(n, keys, 32, 1024, tag="mixed-batch")
    cuts = np.cumsum([weights[v] for v in weights])
    sets = np.array(list(weights))
    u = np.array([int.from_bytes(hashlib.shake_256(b"mayo-bench/mixed-batch/set" + j.to_bytes(8, "little"))
                                 .digest(8), "little") for j in range(n)], dtype=np.float64) / 2.0**64
    set_of = sets[np.minimum(np.searchsorted(cuts, u, side="right"), len(sets) - 1)].astype(np.int32)
    dev = eng1.device
    d_msgs = torch.from_numpy(np.array(flat)).to(dev)
    mb = MixedBatch(engines, prepared)
    rates = {}
    outs = {}
    for conc in (True, False):
        print(f"[bench] mixed-set batch: {'sets side by side' if conc else 'sets in sequence'}", file=sys.stderr,
              flush=True)
        sig = mb.sign(set_of, kidx, d_msgs, offs, lens, concurrent=conc)
        res = mb.verify(set_of, kidx, d_msgs, offs, lens, {v: s for v, (p, s) in sig.items()}, concurrent=conc)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(steps):
            sig = mb.sign(set_of, kidx, d_msgs, offs, lens, concurrent=conc)
        torch.cuda.synchronize(dev)
        ts = (time.perf_counter() - t0) / steps
        s_in = {v: s for v, (p, s) in sig.items()}
        t0 = time.perf_counter()
        for _ in range(steps):
            res = mb.verify(set_of, kidx, d_msgs, offs, lens, s_in, concurrent=conc)
        torch.cuda.synchronize(dev)
        tv = (time.perf_counter() - t0) / steps
        rates[conc] = (n / ts, n / tv)
        outs[conc] = ({v: s.cpu().numpy() for v, (p, s) in sig.items()}, res.cpu().numpy())
    same = all(np.array_equal(outs[True][0][v], outs[False][0][v]) for v in outs[True][0])
    return {"items": n, "mix": {f"MAYO_{v}": int((set_of == v).sum()) for v in weights},
            "messages": "uniform [32, 1024] B, key j mod 64 of the item's set",
            "sign_per_s": round(rates[True][0], 1), "verify_per_s": round(rates[True][1], 1),
            "sets_in_sequence": {"sign_per_s": round(rates[False][0], 1),
                                 "verify_per_s": round(rates[False][1], 1)},
            "identical_outputs": bool(same), "all_valid": bool((outs[True][1] == 0).all())}


def distinct_keys_run(n: int, steps: int, v: int = 2, mlen: int = 64):
    """n distinct keys (the reference seed rule, i = 0 .. n-1) of one set,
    one mlen-byte message each, through the drop-in entry points over
    pitched pools already in HBM: crypto_sign_signWithesk expands every
    item's secret key and signs with it, crypto_sign_verifyWithepk expands
    every item's public key and verifies. Host copies are outside the timed
    region; each timed verification starts from the compact public keys,
    restored by a device copy between runs (the expansion writes over the
    16-byte seed)."""
    import torch
    from cumayo_amd import Engine
    from cumayo_amd.engine import a128
    eng = Engine(v)
    z = eng.sizes
    dev = eng.device
    pk, sk = eng.keypair(n)
    msgs = np.frombuffer(hashlib.shake_256(f"mayo-bench/cfg3-{v}/msg".encode()).digest(n * mlen),
                         np.uint8).reshape(n, mlen)
    st = eng.stream_handle()
    lib = eng.lib
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    # signing pool: sk -> esk in place, message, signature, per-row scratch
    srow = a128(z.esk_bytes) + a128(mlen) + a128(z.sig_bytes) + a128(z.scratch_bytes)
    spool, spitch = eng.pool(n, srow)
    h_sk, h_msg = eng.pinned(sk.shape, sk), eng.pinned(msgs.shape, msgs)
    lib.crypto_sign_signWithesk_HtD(h_sk.data_ptr(), h_msg.data_ptr(), mlen, spool.data_ptr(), spitch, n, st)
    lib.crypto_sign_signWithesk(spool.data_ptr(), mlen, spitch, n, st)  # untimed first run
    torch.cuda.synchronize(dev)
    a.record()
    for _ in range(steps):
        lib.crypto_sign_signWithesk(spool.data_ptr(), mlen, spitch, n, st)
    b.record()
    torch.cuda.synchronize(dev)
    sign_ms = a.elapsed_time(b) / steps
    sig = eng.pinned((n, z.sig_bytes))
    lib.crypto_sign_signWithesk_DtH(sig.data_ptr(), spool.data_ptr(), mlen, spitch, n, st)
    torch.cuda.synchronize(dev)
    del spool
    # the same keys and messages the reference's way: crypto_sign_expandsk
    # (L = (P1 + P1^T) O + P2 in every row) then crypto_sign_sign with L
    lrow = z.esk_bytes + mlen + z.sig_bytes + z.scratch_bytes
    lpool, lpitch = eng.pool(n, lrow)
    lib.crypto_sign_expandsk_HtD(h_sk.data_ptr(), lpool.data_ptr(), lpitch, n, st)
    lpool[:, z.esk_bytes:z.esk_bytes + mlen].copy_(torch.from_numpy(msgs).to(dev))

    def with_L():
        lib.crypto_sign_expandsk(lpool.data_ptr(), lpitch, n, st)
        lib.crypto_sign_sign(lpool.data_ptr(), mlen, lpitch, n, st)

    with_L()
    torch.cuda.synchronize(dev)
    a.record()
    for _ in range(steps):
        with_L()
    b.record()
    torch.cuda.synchronize(dev)
    withL_ms = a.elapsed_time(b) / steps
    lsig = eng.pinned((n, z.sig_bytes))
    lib.crypto_sign_sign_DtH(lsig.data_ptr(), lpool.data_ptr(), mlen, lpitch, n, st)
    torch.cuda.synchronize(dev)
    del lpool
    # verification pool: compact pk at P - 16, message, signature, result byte
    epk_len = z.p_bytes + z.pk_bytes - 16
    vrow = a128(epk_len) + a128(mlen) + a128(z.sig_bytes) + 1
    vpool, vpitch = eng.pool(n, vrow)
    h_pk = eng.pinned(pk.shape, pk)
    lib.crypto_sign_verifyWithepk_HtD(h_pk.data_ptr(), h_msg.data_ptr(), sig.data_ptr(), mlen, vpool.data_ptr(),
                                      vpitch, n, st)
    torch.cuda.synchronize(dev)
    lo, hi = z.p_bytes - 16, z.p_bytes - 16 + z.pk_bytes
    cpk = vpool[:, lo:hi].clone()
    ver_ms = 0.0
    for k in range(steps + 1):
        vpool[:, lo:hi].copy_(cpk)
        a.record()
        lib.crypto_sign_verifyWithepk(vpool.data_ptr(), mlen, vpitch, n, st)
        b.record()
        torch.cuda.synchronize(dev)
        if k:
            ver_ms += a.elapsed_time(b) / steps
    res = eng.pinned((n,))
    lib.crypto_sign_verifyWithepk_DtH(res.data_ptr(), mlen, vpool.data_ptr(), vpitch, n, st)
    torch.cuda.synchronize(dev)
    eng.check()
    return {"param_set": f"MAYO_{v}", "items": n, "distinct_keys": n, "mlen": mlen,
            "path": "crypto_sign_signWithesk / crypto_sign_verifyWithepk (per-item key expansion; "
                    "signing without L)",
            "sign_per_s": round(n / (sign_ms * 1e-3), 1), "verify_per_s": round(n / (ver_ms * 1e-3), 1),
            "all_valid": bool((res.numpy() == 0).all()),
            "with_L": {"path": "crypto_sign_expandsk (expands L) then crypto_sign_sign",
                       "sign_per_s": round(n / (withL_ms * 1e-3), 1),
                       "identical_outputs": bool(np.array_equal(sig.numpy(), lsig.numpy()))}}


def host_cpu() -> dict:
    """nproc, the cores this process may run on, and the CPU model name
    (lscpu's "Model name", else /proc/cpuinfo's)."""
    import subprocess
    model = ""
    try:
        res = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10)
        for line in res.stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if not model:
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "available": avail, "model": model}


def cpu_baseline(v, d_esk, kidx, flat, offs, lens, sigs, sample, runs=5):
    """The oracle (C restatement of the reference algorithm, OpenMP over
    items) signing the first `sample` items of the workload: compute only
    (the key table is downloaded before the clock starts), median of `runs`
    runs, on the host cores this process is granted (OMP_NUM_THREADS when
    the host sets one, else every available core), with the oracle build
    for the highest x86-64 level the CPU supports."""
    from oracle import oracle as O
    cpu = host_cpu()
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(cpu["available"], omp) if omp > 0 else cpu["available"]
    L, isa = O.baseline_lib()
    S = min(sample, len(kidx))
    print(f"[bench] cpu baseline: oracle ({isa}) sign of {S} items on {threads} threads x {runs} runs",
          file=sys.stderr, flush=True)
    esk = d_esk.cpu().numpy()
    times, osigs = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        osigs = O.sign_batch(v, esk, kidx[:S], flat, offs[:S], lens[:S], threads=threads, library=L)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {
        "value": round(S / med, 1), "unit": "sign/s", "cores": threads, "kind": "port",
        "sample": f"first {S} signatures of the same workload, median of {runs} runs, compute only; "
                  f"oracle (C restatement of the reference algorithm, OpenMP over items)",
        "isa": isa, "nproc": cpu["nproc"], "cores_available": cpu["available"],
        "omp_num_threads_env": omp or None, "cpu_model": cpu["model"],
        "run_s": [round(t, 3) for t in times],
        "bit_exact_vs_gpu": bool(np.array_equal(osigs, sigs[:S])),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--keys", type=int, default=64)
    ap.add_argument("--version", type=int, default=1)
    ap.add_argument("--msg-max", type=int, default=1024,
                    help="message lengths are uniform in [32, msg-max] bytes")
    ap.add_argument("--key-block", type=int, default=0,
                    help="item j uses key j / key-block (default: key j mod keys)")
    ap.add_argument("--cpu-sample", type=int, default=4096,
                    help="signatures the CPU baseline signs (bounded sample)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="length of each open-loop serve run (0 skips the serve runs)")
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="offered Poisson rate as a fraction of the measured batch throughput")
    ap.add_argument("--serve-batch", type=int, default=8192, help="requests per serve slot")
    ap.add_argument("--serve-slots", type=int, default=6, help="serve batches in flight")
    ap.add_argument("--no-mixed-sets", action="store_true",
                    help="skip the mixed-parameter-set serve run (MAYO_1/2/3/5 = 40/20/20/20)")
    ap.add_argument("--flow-batch", type=int, default=20000,
                    help="batch of the reference-flow program (keypair, signWithesk, "
                         "verifyWithepk over 10 streams, 32-byte messages); 0 skips it")
    ap.add_argument("--mixed-batch", type=int, default=65536,
                    help="items of the mixed-parameter-set batch run (MAYO_1/2/3/5 = 40/20/20/20); 0 skips it")
    ap.add_argument("--distinct-keys", type=int, default=65536,
                    help="items of the MAYO_2 distinct-key run (per-item key expansion fused with "
                         "signing and verification, 64-byte messages); 0 skips it")
    ap.add_argument("--headline-only", action="store_true",
                    help="time the headline sign and verify passes only (no graph, plain-form, reference-layout, "
                         "serve, mixed-set, distinct-key, reference-flow or CPU legs): for profiler passes")
    ap.add_argument("--launch", default="",
                    help="launch configuration to use instead of autotuning, e.g. sign_prep=23,sign_solve=8,"
                         "sign_chunks=2 (fields of cumayo_launch_config)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the compiled default launch variants (no launch autotuning)")
    ap.add_argument("--autotune-reps", type=int, default=3,
                    help="timed runs per launch configuration in the autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (for scaling "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(args.copy_calibration << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel()}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine is measured on one GPU; multi-GPU runs are not "
                          "part of this build (items shard by key with no collective)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from cumayo_amd import Engine

    v = args.version
    eng = Engine(v)
    z = eng.sizes
    dev = eng.device
    n = args.batch
    print(f"[bench] MAYO_{v} batch={n} keys={args.keys}: generating inputs", file=sys.stderr, flush=True)
    # keys: generated, then expanded straight into dense device tables
    # (pitch = row length), so the resident keys never leave HBM
    d_esk, d_epk = eng.keypair_tables(args.keys, rand_index=0)
    # the kernel layout of the same keys (entries padded to whole dwords),
    # built once per key like the expansion itself
    p_esk, p_epk = eng.prepare_tables(d_esk, d_epk)
    flat, offs, lens, kidx = make_workload(n, args.keys, 32, args.msg_max,
                                           tag="cfg2" if args.msg_max == 1024 else f"cfg-{args.msg_max}")
    if args.key_block:
        kidx = np.minimum(np.arange(n) // args.key_block, args.keys - 1).astype(np.int32)
    t = lambda a: torch.from_numpy(np.array(a)).to(dev)
    d_msgs, d_off, d_len, d_kidx = t(flat), t(offs), t(lens), t(kidx)
    d_sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    d_res = torch.full((n,), 0xEE, dtype=torch.uint8, device=dev)
    ws_s = eng.sign_workspace(n)
    ws_v = eng.verify_workspace(n)
    stream = torch.cuda.current_stream(dev)

    def sign_step(ev=None):
        if ev is None:
            eng.sign_prepared(p_esk, d_kidx, d_msgs, d_off, d_len, d_sigs, ws_s)
        else:
            eng.sign_prepared(p_esk, d_kidx, d_msgs, d_off, d_len, d_sigs, ws_s, stages=1)
            ev[0].record(stream)
            eng.sign_prepared(p_esk, d_kidx, d_msgs, d_off, d_len, d_sigs, ws_s, stages=2)
            ev[1].record(stream)

    def verify_step():
        eng.verify_prepared(p_epk, d_kidx, d_msgs, d_off, d_len, d_sigs, d_res, ws_v)

    # launch autotuner: every compiled variant and item order of the
    # linear-algebra kernels timed on this batch, the fastest kept (untimed,
    # like the key preparation)
    tune = None
    if args.launch:
        eng.set_launch_config(**{k: int(x) for k, x in (kv.split("=") for kv in args.launch.split(","))})
        args.no_autotune = True
    if not args.no_autotune:
        print("[bench] autotuning launch variants", file=sys.stderr, flush=True)
        tune = eng.autotune(p_esk, p_epk, d_kidx, d_msgs, d_off, d_len, d_sigs, d_res, ws_s, ws_v,
                            reps=args.autotune_reps)
        print(f"[bench] launch config {tune['best']}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        sign_step()
        verify_step()
    torch.cuda.synchronize(dev)
    print("[bench] warmup done", file=sys.stderr, flush=True)

    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for k in range(args.steps):
        sign_step(events[k])
    torch.cuda.synchronize(dev)
    sign_s = (time.perf_counter() - t0) / args.steps
    main_ms = float(np.mean([a.elapsed_time(b) for a, b in events]))

    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        verify_step()
    torch.cuda.synchronize(dev)
    verify_s = (time.perf_counter() - t0) / args.steps
    print(f"[bench] timed: sign {sign_s*1e3:.2f} ms/step verify {verify_s*1e3:.2f} ms/step",
          file=sys.stderr, flush=True)

    comparisons = {}
    if not args.headline_only:
        # the same pipelines captured as HIP graphs (one graph launch per
        # pass) over the eager passes' own buffers, so that the comparison is
        # between launch mechanisms only; the eager outputs are kept to check
        # every replay against
        e_sigs, e_res = d_sigs.clone(), d_res.clone()
        g_sign = eng.sign_graph(p_esk, d_kidx, d_msgs, d_off, d_len, d_sigs, ws_s)
        g_verify = eng.verify_graph(p_epk, d_kidx, d_msgs, d_off, d_len, d_sigs, d_res, ws_v)
        d_sigs.fill_(0)
        d_res.fill_(0xEE)
        for _ in range(max(args.warmup, 2)):
            g_sign.launch()
            g_verify.launch()
        torch.cuda.synchronize(dev)
        graph_same = bool(torch.equal(e_sigs, d_sigs)) and bool(torch.equal(e_res, d_res))
        # graph and eager passes alternate, each timed with events on the
        # stream, so that a drift of the clocks over the run affects both
        # alike; medians over the rounds
        rounds = max(args.steps, 5)
        ga, gb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def ev_ms(fn):
            ga.record(stream)
            fn()
            gb.record(stream)
            gb.synchronize()
            return ga.elapsed_time(gb)

        t_gs, t_es, t_gv, t_ev = [], [], [], []
        for _ in range(rounds):
            t_gs.append(ev_ms(lambda: g_sign.launch()))
            t_es.append(ev_ms(sign_step))
            t_gv.append(ev_ms(lambda: g_verify.launch()))
            t_ev.append(ev_ms(verify_step))
        graph_sign_s = float(np.median(t_gs)) / 1e3
        graph_verify_s = float(np.median(t_gv)) / 1e3
        eager_after_s = float(np.median(t_es)) / 1e3
        eager_verify_after_s = float(np.median(t_ev)) / 1e3
        graph_same = graph_same and bool(torch.equal(e_sigs, d_sigs)) and bool(torch.equal(e_res, d_res))
        g_sign.close()
        g_verify.close()
        del e_sigs, e_res

        # the plain form: one workgroup per item running the whole signing
        # (or verification) stage in one kernel, the reference's own structure
        def timed(fn, reps):
            torch.cuda.synchronize(dev)
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize(dev)
            return (time.perf_counter() - t0) / reps

        plain_sigs = torch.zeros_like(d_sigs)
        plain_res = torch.full_like(d_res, 0xEE)
        plain_sign = lambda: eng.sign_table(d_esk, d_kidx, d_msgs, d_off, d_len, plain_sigs, ws_s, stages=7)
        plain_verify = lambda: eng.verify_table(d_epk, d_kidx, d_msgs, d_off, d_len, d_sigs, plain_res, ws_v,
                                                stages=7)
        plain_sign()
        plain_verify()
        plain_sign_s = timed(plain_sign, max(1, args.steps // 2))
        plain_verify_s = timed(plain_verify, max(1, args.steps // 2))
        plain_same = bool(torch.equal(plain_sigs, d_sigs)) and bool(torch.equal(plain_res, d_res))
        # the same wave kernels over the reference-format key tables
        ref_sigs = torch.zeros_like(d_sigs)
        ref_res = torch.full_like(d_res, 0xEE)
        ref_sign = lambda: eng.sign_table(d_esk, d_kidx, d_msgs, d_off, d_len, ref_sigs, ws_s)
        ref_verify = lambda: eng.verify_table(d_epk, d_kidx, d_msgs, d_off, d_len, d_sigs, ref_res, ws_v)
        ref_sign()
        ref_verify()
        ref_sign_s = timed(ref_sign, max(1, args.steps // 2))
        ref_verify_s = timed(ref_verify, max(1, args.steps // 2))
        ref_same = bool(torch.equal(ref_sigs, d_sigs)) and bool(torch.equal(ref_res, d_res))
        comparisons = {
            "plain_form": {"kernels": "one workgroup per item, whole stage in one kernel",
                           "sign_per_s": round(n / plain_sign_s, 1),
                           "verify_per_s": round(n / plain_verify_s, 1),
                           "speedup_sign": round(plain_sign_s / sign_s, 2),
                           "speedup_verify": round(plain_verify_s / verify_s, 2),
                           "identical_outputs": plain_same},
            "reference_layout_tables": {"sign_per_s": round(n / ref_sign_s, 1),
                                        "verify_per_s": round(n / ref_verify_s, 1),
                                        "identical_outputs": ref_same},
            "graph": {"pipelines": "sign and verify passes captured as HIP graphs, one graph launch each",
                      "timing": "graph and eager passes alternating, events on the stream, medians",
                      "sign_per_s": round(n / graph_sign_s, 1), "verify_per_s": round(n / graph_verify_s, 1),
                      "eager_sign_per_s_after": round(n / eager_after_s, 1),
                      "eager_verify_per_s_after": round(n / eager_verify_after_s, 1),
                      "identical_outputs": graph_same},
        }
    # restore the workspace of the default form for the attempt counts below
    sign_step()
    torch.cuda.synchronize(dev)

    res = d_res.cpu().numpy()
    all_verify = bool((res == 0).all())
    bad = np.where(res != 0)[0]
    sigs = d_sigs.cpu().numpy()
    attempts = eng.sign_attempts(ws_s, n)

    sign_per_s = n / sign_s
    verify_per_s = n / verify_s
    achieved_tmacs = n * SIGN_MACS[v] / (main_ms * 1e-3) / 1e12
    # algorithmic HBM bytes of the main kernel per signature: the message-side
    # reads are in the hashing pass; the main pass reads V||r, t (scratch) and
    # the esk blocks it touches once per signature (L, P1, O) and writes the
    # signature. Keys are shared, so this is an upper bound on compulsory bytes.
    algo_bytes = n * (z.sig_bytes + z.k * z.n // 2 + z.m // 2)
    lc = eng.launch_config()
    form = "shared" if lc["sign_prep"] >= 1000 and lc["sign_chunks"] != 0 else "wave"
    if form == "shared":
        kernels = (f"signing linear-algebra pass, shared-key form (launch code {lc['sign_prep']}): the key "
                   "sort (k_key_hist, k_key_scan, k_key_scatter: items grouped by key), k_sign_table_shared "
                   "(M and T_j = P1 V_j: products read from LDS tables of the key's entry multiples, built "
                   "once per workgroup of items of one key; T to the workspace), k_sign_fold (U and y from T, one wavefront per item), k_sign_solve "
                   "(elimination, one wavefront per item), k_sign_retry (failed attempts); achieved and "
                   "kernel_ms cover the launches together")
    else:
        kernels = ("signing linear-algebra pass: k_sign_prep (M, U, y) + k_sign_solve (elimination), one "
                   "wavefront per signature, + k_sign_retry (failed attempts, one workgroup per listed "
                   "item); achieved and kernel_ms cover the three launches together")
    out = {
        "metric": f"MAYO_{v} sign/s & verify/s at batch {n // 1024}K (bit-exact); serve p50/p99 latency",
        "value": round(sign_per_s, 1),
        "unit": "sign/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sign_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf16 (u4 nibbles in u32 words)",
        "data": "synthetic (SHAKE-seeded messages, keys from the reference seed rule)",
        "config": {"workload": f"MAYO_{v} sign+verify, {n} signatures, {args.keys} resident keys "
                               f"(key {'j / ' + str(args.key_block) if args.key_block else 'j mod ' + str(args.keys)}"
                               f", prepared key tables), message lengths uniform [32,{args.msg_max}] B",
                   "global_batch": n, "keys": args.keys, "param_set": f"MAYO_{v}"},
        "verify": {"value": round(verify_per_s, 1), "unit": "verify/s",
                   "ms_per_step": round(verify_s * 1e3, 3), "all_valid": all_verify,
                   "invalid": int(bad.size), "invalid_first": bad[:8].tolist()},
        **comparisons,
        "launch_tuning": tune if tune is not None else {"best": eng.launch_config(), "autotuned": False},
        "sign_retries": {"items_with_retry": int((attempts > 1).sum()),
                         "exhausted": int((attempts == 0).sum()),
                         "max_attempts": int(attempts.max())},
        "roofline": {
            "bound": "valu",
            "kernel": kernels,
            "launch_form": form,
            "achieved": round(achieved_tmacs, 3),
            "peak": round(PEAK_GF16_TMACS, 1),
            "unit": "T GF16-MAC/s",
            "frac": round(achieved_tmacs / PEAK_GF16_TMACS, 4),
            "kernel_ms": round(main_ms, 3),
            "macs_per_sign": SIGN_MACS[v],
            "hbm_algorithmic_GBs": round(algo_bytes / (main_ms * 1e-3) / 1e9, 2),
            "hbm_peak_GBs": HBM_PEAK_GBS,
            "traffic": STAGE2_TRAFFIC.get((v, n, args.keys, form)),
            "traffic_unit": "bytes per launch (PMC FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "traffic_measured_at": STAGE2_TRAFFIC_AT.get(form) if (v, n, args.keys, form) in STAGE2_TRAFFIC else None,
        },
    }

    # the VALU peak measured on this GPU: a stream of independent v_bitop3_b32
    # AND-XORs, the multiply-accumulate step of the kernels (2 GF(16) MACs per
    # lane-op), next to the nominal 1 op per 4 cycles per SIMD at 2.4 GHz
    import subprocess
    from cumayo_amd import build as _build
    if os.path.exists(_build.peak_path()):
        pk = subprocess.run([_build.peak_path()], capture_output=True, text=True, timeout=60)
        if pk.returncode == 0:
            meas = json.loads(pk.stdout.strip().splitlines()[-1])
            out["roofline"]["peak_measured"] = {
                "program": "cumayo_amd/lib/valu_peak (v_bitop3_b32 streams, 8 waves per SIMD)",
                "gf16_tmacs": meas["gf16_tmacs"], "cycles_per_wave_op": meas["cycles_per_wave_op"],
                "clock_ghz": meas["clock_ghz"],
                "frac_of_measured": round(achieved_tmacs / meas["gf16_tmacs"], 4),
                # with a VGPR mask two waves' AND-XORs issue in one quad-cycle
                # (the kernels' masks are SGPRs: at most one of a pair may read one)
                "vgpr_mask_gf16_tmacs": meas.get("vgpr_mask_gf16_tmacs"),
                "vgpr_mask_cycles_per_wave_op": meas.get("vgpr_mask_cycles_per_wave_op")}

    if not args.headline_only and os.path.exists(_build.contraction_path()):
        # the shared-key P1 contraction in its bitsliced VALU form, its
        # int8-MFMA GF(2)-plane form and its LDS-table form, same data, all
        # checked against a CPU restatement (cumayo_amd/csrc/contraction_bench.hip)
        cb = subprocess.run([_build.contraction_path()], capture_output=True, text=True, timeout=120)
        if cb.returncode == 0 and cb.stdout.strip():
            out["mfma_contraction"] = json.loads(cb.stdout.strip().splitlines()[-1])
    if args.headline_only:
        args.flow_batch = args.serve_seconds = args.mixed_batch = args.distinct_keys = 0
        args.no_cpu_baseline = True
    if args.flow_batch > 0:
        # the reference's own benchmark method through the drop-in C ABI
        # (a C++ program linked against the library): per-item key expansion
        # and both PCIe copies inside the timed region. Run before the serve
        # runs, on an otherwise idle device.
        print(f"[bench] reference flow, batch {args.flow_batch}", file=sys.stderr, flush=True)
        res = subprocess.run([_build.flow_path(v), str(args.flow_batch), "10", "10", "32"],
                             capture_output=True, text=True, timeout=300)
        if res.returncode == 0:
            out["reference_flow"] = json.loads(res.stdout.strip().splitlines()[-1])
        else:
            out["reference_flow"] = {"error": (res.stdout + res.stderr)[-300:]}

    if args.serve_seconds > 0:
        # Continuous batching under open-loop Poisson arrivals drawn from the
        # same workload (request i: message i, key i mod keys; verification
        # requests carry the signature just made for message i). Latency is
        # arrival to result in host memory, including both PCIe copies.
        from cumayo_amd.serve import Server
        serve = {}
        mixed_rate = 1.0 / (0.5 / sign_per_s + 0.5 / verify_per_s)
        sweep = sorted(set([0.3, 0.5, args.serve_load, 0.9]))
        with Server(eng, d_esk, d_epk, max_batch=args.serve_batch, n_slots=args.serve_slots,
                    slot_msg_bytes=args.serve_batch * 1024) as srv:
            def point(frac_v, rate, seconds):
                rep = srv.run_poisson(rate, seconds, frac_v, 17, flat, offs, lens, kidx, sigs)
                return {"offered_per_s": round(rep["offered_per_s"], 1),
                        "achieved_per_s": round(rep["achieved_per_s"], 1),
                        "p50_ms": round(rep["p50_ms"], 3), "p90_ms": round(rep["p90_ms"], 3),
                        "p99_ms": round(rep["p99_ms"], 3), "max_ms": round(rep["max_ms"], 3),
                        "completed": rep["completed"], "refused": rep["refused"]}
            # a short unrecorded run first, at a load that keeps every slot
            # busy: the first batches through each slot pay one-time costs
            # (first launches, first-touch of pinned pages)
            srv.run_poisson(0.9 * sign_per_s, 0.3, 0.0, 3, flat, offs, lens, kidx, sigs)
            serve["sign_vs_load"] = []
            for frac in sweep:
                print(f"[bench] serve sign at {frac:.0%} of batch throughput", file=sys.stderr, flush=True)
                secs = args.serve_seconds if frac == args.serve_load else args.serve_seconds / 2
                serve["sign_vs_load"].append({"offered_frac_of_batch": frac,
                                              **point(0.0, frac * sign_per_s, secs)})
            serve["sign"] = next(x for x in serve["sign_vs_load"]
                                 if x["offered_frac_of_batch"] == args.serve_load)
            print("[bench] serve sign/verify 50/50", file=sys.stderr, flush=True)
            serve["mixed_50_50"] = {"offered_frac_of_batch": args.serve_load,
                                    **point(0.5, args.serve_load * mixed_rate, args.serve_seconds)}
            st = srv.stats()
        serve["slots"] = {"max_batch": args.serve_batch, "n_slots": args.serve_slots,
                          "batches": st["batches"], "overlapped_launches": st["overlapped_launches"],
                          "gpu_max_hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"])}
        if not args.no_mixed_sets:
            serve["mixed_sets"] = mixed_sets_run(eng, args)
        out["serve"] = serve
        out["serve_p50_ms"] = serve["sign"]["p50_ms"]
        out["serve_p99_ms"] = serve["sign"]["p99_ms"]

    if args.mixed_batch > 0:
        # after the serve runs: its per-set streams would otherwise share the
        # process's hardware queues with the servers' slot streams
        print(f"[bench] mixed-set batch of {args.mixed_batch}", file=sys.stderr, flush=True)
        out["mixed_set_batch"] = mixed_batch_run(eng, args.mixed_batch, max(1, args.steps // 2))

    if args.distinct_keys > 0:
        print(f"[bench] MAYO_2 distinct keys, {args.distinct_keys} items", file=sys.stderr, flush=True)
        out["distinct_keys"] = distinct_keys_run(args.distinct_keys, max(1, args.steps // 2))

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(v, d_esk, kidx, flat, offs, lens, sigs, args.cpu_sample)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
