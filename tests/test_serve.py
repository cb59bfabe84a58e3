"""Continuous-batching server: a request gives the same bytes when it is
served alone, in a mix of overlapping ragged batches, through the batch
entry points and by the oracle."""
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _requests(eng, esk, epk, n, nkeys, seed):
    """n mixed requests over ragged messages (lengths 0, 136, 1 KiB among
    them); about 40% verifications of batch-made signatures, a quarter of
    which are corrupted (signature byte or message bit)."""
    import torch
    z = eng.sizes
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 1100, n).astype(np.int32)
    lens[:3] = [0, 136, 1024]
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    keys = rng.integers(0, nkeys, n).astype(np.int32)
    dev = eng.device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    eng.sign_table(esk, t(keys), t(flat), t(offs), t(lens), sigs, eng.sign_workspace(n))
    torch.cuda.synchronize()
    batch_sigs = sigs.cpu().numpy()
    ops = (rng.random(n) < 0.4).astype(np.uint8)
    ops[:2] = [0, 1]
    reqs, expect_valid = [], {}
    for i in range(n):
        msg = flat[offs[i]:offs[i] + lens[i]].tobytes()
        if ops[i] == 1:
            sig = bytearray(batch_sigs[i].tobytes())
            bad = rng.random() < 0.25
            if bad:
                if lens[i] and rng.random() < 0.5:
                    msg = bytearray(msg)
                    msg[rng.integers(0, lens[i])] ^= 0x10
                    msg = bytes(msg)
                else:
                    sig[rng.integers(0, z.sig_bytes)] ^= 0x01
            expect_valid[i] = 1 if bad else 0
            reqs.append((i, 1, int(keys[i]), msg, bytes(sig)))
        else:
            reqs.append((i, 0, int(keys[i]), msg, None))
    return reqs, batch_sigs, expect_valid


@pytest.mark.parametrize("version", [1, 2])
def test_serve_alone_equals_mix_equals_batch(engine_factory, version):
    from cumayo_amd.serve import Server
    eng = engine_factory(version)
    nkeys, n = 5, 700
    esk, epk = eng.keypair_tables(nkeys, rand_index=11)
    reqs, batch_sigs, expect = _requests(eng, esk, epk, n, nkeys, seed=version)
    print(f"[serve] MAYO_{version} n={n}", flush=True)

    # in the mix: many small overlapping batches, submitted in bursts
    with Server(eng, esk, epk, max_batch=48, n_slots=4) as srv:
        rng = np.random.default_rng(99)
        i = 0
        while i < n:
            k = int(rng.integers(1, 120))
            assert srv.submit(reqs[i:i + k]) == len(reqs[i:i + k])
            i += k
            time.sleep(float(rng.random()) * 0.002)
        mix = srv.wait(range(n))
        st = srv.stats()
    assert st["requests"] == n and st["batches"] >= n // 48
    assert st["max_in_flight"] >= 2, st

    # alone: one request per batch, nothing else in flight
    alone_ids = list(range(0, n, 23)) + [0, 1, 2]
    with Server(eng, esk, epk, max_batch=8, n_slots=2) as srv:
        alone = {}
        for j in alone_ids:
            srv.submit([reqs[j]])
            alone[j] = srv.wait([j])[j]
        assert srv.stats()["max_in_flight"] == 1

    for j in range(n):
        status, _, sig, res = mix[j]
        if reqs[j][1] == 0:
            assert status == 0 and res >= 1
            assert sig.tobytes() == batch_sigs[j].tobytes(), f"sign {j}"
        else:
            assert status == expect[j] and res == expect[j], f"verify {j}"
    for j in alone_ids:
        a, m = alone[j], mix[j]
        assert a[0] == m[0] and a[3] == m[3]
        if reqs[j][1] == 0:
            assert a[2].tobytes() == m[2].tobytes()

    # oracle on a sample of signatures
    esk_h = esk.cpu().numpy()
    for j in [r[0] for r in reqs if r[1] == 0][:12]:
        osig, _ = O.sign(version, esk_h[reqs[j][2]].tobytes(), reqs[j][3])
        assert mix[j][2].tobytes() == osig, f"oracle {j}"


@pytest.mark.parametrize("host_threads", [1, 4])
def test_serve_large_batches_host_pool(engine_factory, host_threads):
    """Batches of several hundred requests, whose staging copies and result
    copies are split over the host thread pool (chunks of 256 requests) or
    done by the worker alone: the same bytes as the batch entry points."""
    from cumayo_amd.serve import Server
    eng = engine_factory(1)
    nkeys, n = 4, 900
    esk, epk = eng.keypair_tables(nkeys, rand_index=17)
    reqs, batch_sigs, expect = _requests(eng, esk, epk, n, nkeys, seed=41)
    print(f"[serve] large batches, host_threads={host_threads}", flush=True)
    with Server(eng, esk, epk, max_batch=1024, n_slots=2, host_threads=host_threads) as srv:
        assert srv.submit(reqs) == n
        got = srv.wait(range(n))
        st = srv.stats()
    assert st["requests"] == n and st["batches"] <= 3
    for j in range(n):
        status, _, sig, res = got[j]
        if reqs[j][1] == 0:
            assert status == 0 and res >= 1
            assert sig.tobytes() == batch_sigs[j].tobytes(), f"sign {j}"
        else:
            assert status == expect[j] and res == expect[j], f"verify {j}"


def test_serve_refusals(engine_factory):
    from cumayo_amd.serve import STATUS_REFUSED, Server
    eng = engine_factory(1)
    esk, epk = eng.keypair_tables(2, rand_index=3)
    z = eng.sizes
    with Server(eng, esk, None, max_batch=16, n_slots=2, slot_msg_bytes=4096) as srv:
        bad = [(0, 0, 7, b"abc", None),                    # key out of range
               (1, 1, 0, b"abc", bytes(z.sig_bytes)),       # no epk table given
               (2, 0, 0, bytes(5000), None),                # longer than a slot holds
               (3, 0, 1, b"fine", None)]
        assert srv.submit(bad) == 4
        got = srv.wait(range(4))
    assert [got[i][0] for i in range(3)] == [STATUS_REFUSED] * 3
    assert got[3][0] == 0
    osig, _ = O.sign(1, esk[1].cpu().numpy().tobytes(), b"fine")
    assert got[3][2].tobytes() == osig


def test_serve_poisson_load(engine_factory):
    """Open-loop Poisson arrivals at 200k signatures/s (about 3% of the batch
    throughput): everything offered completes, and the tail stays within a
    few batch times (measured p99 about 10 ms)."""
    import bench
    from cumayo_amd.serve import Server
    eng = engine_factory(1)
    nkeys, pool = 8, 4096
    esk, epk = eng.keypair_tables(nkeys, rand_index=1)
    flat, offs, lens, kidx = bench.make_workload(pool, nkeys, 32, 1024, tag="serve-test")
    with Server(eng, esk, epk, max_batch=2048, n_slots=3) as srv:
        rep = srv.run_poisson(200_000.0, 0.5, 0.0, 5, flat, offs, lens, kidx)
    print(f"[serve] poisson {rep}", flush=True)
    assert rep["refused"] == 0
    assert rep["completed"] == rep["submitted"] > 50_000
    assert 0 < rep["p50_ms"] <= rep["p99_ms"] < 100


def test_mixed_parameter_sets(engine_factory):
    """MAYO_1 and MAYO_2 requests interleaved through one MixedServer give
    each set's batch results."""
    from cumayo_amd.mixed import MixedServer
    engines, tables, reqs, expect = {}, {}, [], {}
    tag = 0
    for v in (1, 2):
        eng = engine_factory(v)
        esk, epk = eng.keypair_tables(3, rand_index=40 + v)
        engines[v], tables[v] = eng, (esk, epk)
        r, batch_sigs, exp_valid = _requests(eng, esk, epk, 90, 3, seed=50 + v)
        for (i, op, key, msg, sig) in r:
            reqs.append((tag, v, op, key, msg, sig))
            expect[tag] = (v, op, batch_sigs[i].tobytes(), exp_valid.get(i, 0))
            tag += 1
    rng = np.random.default_rng(3)
    order = rng.permutation(len(reqs))
    with MixedServer(engines, tables, max_batch=32, n_slots=3) as ms:
        assert ms.submit([reqs[k] for k in order]) == len(reqs)
        got = ms.wait({v: [t for (t, vv, *_rest) in reqs if vv == v] for v in (1, 2)})
    for t, (v, op, sig, bad) in expect.items():
        gv, status, _, gsig, res = got[t]
        assert gv == v
        if op == 0:
            assert status == 0 and gsig.tobytes() == sig, f"sign tag {t}"
        else:
            assert status == bad and res == bad, f"verify tag {t}"


def test_serve_backpressure_and_cancel(engine_factory):
    """A full queue refuses further requests (counted as rejected); queued
    requests can be withdrawn and complete with status 3, the others with
    the batch results."""
    from cumayo_amd.serve import Server
    eng = engine_factory(1)
    esk, epk = eng.keypair_tables(2, rand_index=8)
    reqs, batch_sigs, _ = _requests(eng, esk, epk, 400, 2, seed=77)
    signs = [r for r in reqs if r[1] == 0][:200]
    with Server(eng, esk, epk, max_batch=4, n_slots=1, max_queue=60) as srv:
        acc = srv.submit(signs)
        assert 60 <= acc < len(signs)
        assert srv.stats()["rejected"] == len(signs) - acc
        accepted = [r[0] for r in signs[:acc]]
        withdrawn = [tag for tag in accepted[-30:] if srv.cancel(tag)]
        assert len(withdrawn) > 0
        got = srv.wait(accepted)
        assert srv.stats()["cancelled"] == len(withdrawn)
    for tag in accepted:
        status, _, sig, _ = got[tag]
        if tag in withdrawn:
            assert status == 3
        else:
            assert status == 0 and sig.tobytes() == batch_sigs[tag].tobytes()


def test_serve_key_table_beyond_one_preparation_call(engine_factory):
    """A server over 70,000 keys prepares its resident tables in chunks of
    65,535 keys (the most one preparation call takes). Real key pairs sit on
    both sides of the chunk edge (rows 65,534 .. 65,536) and at the ends, the
    other rows are random bytes: every request signs and verifies exactly as
    the batch entry points do on the same table rows."""
    import torch
    from cumayo_amd.serve import Server
    v = 2
    eng = engine_factory(v)
    z = eng.sizes
    dev = eng.device
    nk = 70_000
    rows = [0, 65_534, 65_535, 65_536, nk - 1]
    esk_real, epk_real = eng.keypair_tables(len(rows), rand_index=77)
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    esk = torch.randint(0, 256, (nk, z.esk_bytes), dtype=torch.uint8, device=dev, generator=gen)
    epk = torch.randint(0, 256, (nk, z.epk_bytes), dtype=torch.uint8, device=dev, generator=gen)
    at = torch.tensor(rows, device=dev)
    esk[at] = esk_real
    epk[at] = epk_real
    keys = np.array(rows * 2 + [1, 65_537], np.int32)  # the last two are random rows
    n = keys.size
    rng = np.random.default_rng(8)
    lens = rng.integers(0, 300, n).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    print(f"[serve] MAYO_{v} {nk} keys, n={n}", flush=True)
    sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
    eng.sign_table(esk, t(keys), t(flat), t(offs), t(lens), sigs, eng.sign_workspace(n))
    res = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    eng.verify_table(epk, t(keys), t(flat), t(offs), t(lens), sigs, res, eng.verify_workspace(n))
    torch.cuda.synchronize()
    batch_sigs, batch_res = sigs.cpu().numpy(), res.cpu().numpy()
    assert batch_res[:2 * len(rows)].tolist() == [0] * (2 * len(rows))
    # the Python wrapper prepares the whole table in chunks too: its rows
    # equal those prepared from the real key rows alone
    pesk, pepk = eng.prepare_tables(esk, epk)
    pesk_real, pepk_real = eng.prepare_tables(esk_real, epk_real)
    torch.cuda.synchronize()
    assert torch.equal(pesk[at], pesk_real) and torch.equal(pepk[at], pepk_real)
    del pesk, pepk
    msg = lambda i: flat[offs[i]:offs[i] + lens[i]].tobytes()
    reqs = [(i, 0, int(keys[i]), msg(i), None) for i in range(n)]
    reqs += [(n + i, 1, int(keys[i]), msg(i), batch_sigs[i].tobytes()) for i in range(n)]
    try:
        with Server(eng, esk, epk, max_batch=8, n_slots=2) as srv:
            assert srv.submit(reqs) == len(reqs)
            got = srv.wait(range(2 * n))
    finally:
        del esk, epk
        torch.cuda.empty_cache()
    for i in range(n):
        st, _, sig, _ = got[i]
        assert st == 0 and sig.tobytes() == batch_sigs[i].tobytes(), (i, int(keys[i]))
        assert got[n + i][0] == int(batch_res[i]), (i, int(keys[i]))


def test_serve_failed_batch_fails_closed(engine_factory):
    """A batch whose enqueue fails (forced from the host by the server's
    test hook; no kernel runs for it) completes every request with
    STATUS_ERROR and writes no output: a verification of a valid signature
    does not come back valid. poll() raises ServeError carrying those
    completions, the server reports the error in its stats, and the next
    batch is served normally."""
    from cumayo_amd.serve import STATUS_ERROR, Server, ServeError
    eng = engine_factory(1)
    esk, epk = eng.keypair_tables(2, rand_index=21)
    reqs, batch_sigs, expect = _requests(eng, esk, epk, 40, 2, seed=5)
    good_verifies = [r for r in reqs if r[1] == 1 and expect[r[0]] == 0][:4]
    signs = [r for r in reqs if r[1] == 0][:4]
    assert good_verifies and signs
    first = good_verifies + signs
    with Server(eng, esk, epk, max_batch=64, n_slots=1) as srv:
        srv.inject_failures(1)
        assert srv.submit(first) == len(first)
        got = []
        t_end = time.monotonic() + 60
        while len(got) < len(first):
            try:
                got += srv.poll()
            except ServeError as e:
                got += e.completions
                assert e.hip_error != 0
            assert time.monotonic() < t_end, "failed batch never completed"
            time.sleep(0.001)
        for tag, status, _, sig, res in got:
            assert status == STATUS_ERROR, f"tag {tag} status {status}"
            assert res == 0xEE and not sig.any(), f"tag {tag}: output written for a failed batch"
        st = srv.stats()
        assert st["failed"] == len(first) and st["last_error"] != 0
        # the server keeps serving: the same requests again, now bit-exact
        again = [(tag + 1000, op, key, msg, sig) for tag, op, key, msg, sig in first]
        res2 = srv.wait([r[0] for r in again]) if srv.submit(again) == len(again) else {}
    assert len(res2) == len(again)
    for tag, op, *_ in again:
        status, _, sig, res = res2[tag]
        if op == 1:
            assert status == 0 and res == 0
        else:
            assert status == 0 and sig.tobytes() == batch_sigs[tag - 1000].tobytes()


def test_serve_four_set_poisson_mix(engine_factory):
    """The serve configuration's shape: MAYO_1/2/3/5 requests in a 40/20/20/20
    mix, half of them verifications (a quarter of those of corrupted
    signatures or messages), message lengths log-uniform in [32 B, 16 KB],
    64 resident keys per set, submitted at Poisson arrival times into one
    MixedServer (one server per set, all on the one GPU). Every signature
    equals its set's batch signature, every verification its batch verdict,
    and a strided sample of each set's signatures equals the oracle."""
    import torch
    import bench
    from cumayo_amd.mixed import MixedServer
    weights = {1: 0.4, 2: 0.2, 3: 0.2, 5: 0.2}
    total, nkeys = 1000, 64
    rng = np.random.default_rng(55)
    engines, tables, data = {}, {}, {}
    for v, w in weights.items():
        eng = engine_factory(v)
        z = eng.sizes
        n = int(total * w)
        esk, epk = eng.keypair_tables(nkeys, rand_index=0)
        flat, offs, lens, kidx = bench.log_uniform_workload(n, nkeys, 32, 16384, tag=f"serve-mix-{v}")
        dev = eng.device
        t = lambda a: torch.from_numpy(np.array(a)).to(dev)
        sigs = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=dev)
        eng.sign_table(esk, t(kidx), t(flat), t(offs), t(lens), sigs, eng.sign_workspace(n))
        torch.cuda.synchronize()
        batch_sigs = sigs.cpu().numpy()
        ops = (rng.random(n) < 0.5).astype(np.uint8)
        vmsgs, vsigs = [], []
        for i in range(n):
            msg = bytearray(flat[offs[i]:offs[i] + lens[i]].tobytes())
            sig = bytearray(batch_sigs[i].tobytes())
            if ops[i] and rng.random() < 0.25:
                if rng.random() < 0.5:
                    msg[int(rng.integers(0, len(msg)))] ^= 0x20
                else:
                    sig[int(rng.integers(0, z.sig_bytes))] ^= 0x02
            vmsgs.append(bytes(msg))
            vsigs.append(bytes(sig))
        # the batch verdicts of exactly the (message, signature) pairs served
        vl = np.array([len(m) for m in vmsgs], np.int32)
        vo = np.concatenate([[0], np.cumsum(vl)[:-1]]).astype(np.int64)
        vflat = np.frombuffer(b"".join(vmsgs), np.uint8)
        vres = torch.full((n,), 7, dtype=torch.uint8, device=dev)
        eng.verify_table(epk, t(kidx), t(vflat), t(vo), t(vl),
                         t(np.frombuffer(b"".join(vsigs), np.uint8).reshape(n, z.sig_bytes)), vres,
                         eng.verify_workspace(n))
        torch.cuda.synchronize()
        engines[v], tables[v] = eng, (esk, epk)
        data[v] = dict(n=n, ops=ops, kidx=kidx, vmsgs=vmsgs, vsigs=vsigs, flat=flat, offs=offs,
                       lens=lens, batch_sigs=batch_sigs, vres=vres.cpu().numpy(), esk=esk)
    # one arrival stream over the mix: set by weight, exponential gaps at 4,000/s
    reqs = []
    nxt = {v: 0 for v in weights}
    order = [v for v in weights for _ in range(data[v]["n"])]
    rng.shuffle(order)
    arrivals = np.cumsum(rng.exponential(1 / 4000.0, len(order)))
    for k, v in enumerate(order):
        i = nxt[v]
        nxt[v] += 1
        d = data[v]
        op = int(d["ops"][i])
        tag = v * 100000 + i
        reqs.append((float(arrivals[k]), (tag, v, op, int(d["kidx"][i]), d["vmsgs"][i],
                                                 d["vsigs"][i] if op else None)))
    print(f"[serve] four-set mix, {len(reqs)} Poisson arrivals", flush=True)
    with MixedServer(engines, tables, max_batch=64, n_slots=3, slot_msg_bytes=64 * 16384) as ms:
        t0 = time.perf_counter()
        k = 0
        while k < len(reqs):
            now = time.perf_counter() - t0
            due = []
            while k < len(reqs) and reqs[k][0] <= now:
                due.append(reqs[k][1])
                k += 1
            if due:
                assert ms.submit(due) == len(due)
            else:
                time.sleep(min(0.0005, max(0.0, reqs[k][0] - now)))
        got = ms.wait({v: [v * 100000 + i for i in range(data[v]["n"])] for v in weights})
    for v, d in data.items():
        for i in range(d["n"]):
            gv, status, _, sig, res = got[v * 100000 + i]
            assert gv == v
            if d["ops"][i]:
                assert status == d["vres"][i] and res == d["vres"][i], f"MAYO_{v} verify {i}"
            else:
                assert status == 0 and sig.tobytes() == d["batch_sigs"][i].tobytes(), f"MAYO_{v} sign {i}"
        signs = np.flatnonzero(d["ops"] == 0)[::7]
        osigs = O.sign_batch(v, d["esk"].cpu().numpy(), d["kidx"][signs], d["flat"], d["offs"][signs],
                             d["lens"][signs], threads=16)
        assert np.array_equal(osigs, d["batch_sigs"][signs]), f"MAYO_{v} oracle sample"
        assert d["vres"][d["ops"] == 1].sum() > 0 and (d["vres"][d["ops"] == 1] == 0).sum() > 0
