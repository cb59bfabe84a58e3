"""GPU: the launch autotuner's variants and the HIP-graph captures of the
batch pipelines. Every launch variant and item order must give the bytes of
the default launch (which the parity tests pin to the oracle); every replay
of a captured graph must give the eager result, with eager work and
synchronisations between replays.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _batch(eng, n, nkeys, seed, maxlen=600):
    import torch
    esk, epk = eng.keypair_tables(nkeys, rand_index=31 + seed)
    pesk, pepk = eng.prepare_tables(esk, epk)
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, maxlen, n).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = rng.integers(0, nkeys, n).astype(np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(eng.device)
    return {"esk": esk, "pesk": pesk, "pepk": pepk, "k": t(kidx), "m": t(flat), "o": t(offs), "l": t(lens),
            "kidx": kidx, "flat": flat, "offs": offs, "lens": lens}


@pytest.fixture
def restore_config(engine_factory):
    saved = {}

    def keep(eng):
        saved.setdefault(eng.version, (eng, eng.launch_config()))
        return eng

    yield keep
    for eng, cfg in saved.values():
        eng.set_launch_config(**cfg)


@pytest.mark.parametrize("v", (1, 2, 3, 5), ids=lambda v: f"MAYO_{v}")
def test_every_launch_variant_gives_the_same_bytes(engine_factory, restore_config, v):
    import torch
    eng = restore_config(engine_factory(v))
    z = eng.sizes
    n = 300
    b = _batch(eng, n, 5, 40 + v)
    var = eng.launch_variants()
    print(f"[tuning] MAYO_{v} variants {var}", flush=True)
    ws = eng.sign_workspace(n)
    ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=eng.device)
    eng.set_launch_config(**{k: var[k][0] for k in ("sign_prep", "sign_solve", "verify")},
                          sign_order=0, verify_order=0)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], ref, ws)
    ref_att = eng.sign_attempts(ws, n)
    bad = ref.clone()
    bad[::4, 3] ^= 0x11
    vref = torch.full((n,), 7, dtype=torch.uint8, device=eng.device)
    eng.verify_prepared(b["pepk"], b["k"], b["m"], b["o"], b["l"], bad, vref, eng.verify_workspace(n))
    torch.cuda.synchronize()
    expect = np.zeros(n, np.uint8)
    expect[::4] = 1
    assert np.array_equal(vref.cpu().numpy(), expect)
    for p, s, order in itertools.product(var["sign_prep"], var["sign_solve"], var["order"]):
        eng.set_launch_config(sign_prep=p, sign_solve=s, sign_order=order)
        out = torch.full_like(ref, 0xA5)
        eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], out, ws)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (p, s, order)
        assert np.array_equal(eng.sign_attempts(ws, n), ref_att), (p, s, order)
    # the persistent fused kernel (sign_chunks 0) with every quadratic-form
    # variant, also with every item's first two attempts forced to fail
    for p, order in itertools.product(var["sign_prep"], var["order"]):
        eng.set_launch_config(sign_prep=p, sign_order=order, sign_chunks=0)
        out = torch.full_like(ref, 0xA5)
        eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], out, ws)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), ("fused", p, order)
        assert np.array_equal(eng.sign_attempts(ws, n), ref_att), ("fused", p, order)
    eng.set_launch_config(sign_prep=var["sign_prep"][0], sign_order=0, sign_chunks=1)
    forced = torch.zeros_like(ref)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], forced, ws, stages=3 | (2 << 8))
    torch.cuda.synchronize()
    forced_att = eng.sign_attempts(ws, n)
    assert (forced_att >= 3).all()
    eng.set_launch_config(sign_chunks=0)
    out = torch.full_like(ref, 0xA5)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], out, ws, stages=3 | (2 << 8))
    torch.cuda.synchronize()
    assert torch.equal(out, forced) and np.array_equal(eng.sign_attempts(ws, n), forced_att)
    eng.set_launch_config(sign_chunks=1)
    for vv, order in itertools.product(var["verify"], var["order"]):
        eng.set_launch_config(verify=vv, verify_order=order)
        res = torch.full((n,), 7, dtype=torch.uint8, device=eng.device)
        eng.verify_prepared(b["pepk"], b["k"], b["m"], b["o"], b["l"], bad, res, eng.verify_workspace(n))
        torch.cuda.synchronize()
        assert torch.equal(res, vref), (vv, order)
    # the default launch is the oracle's (spot check)
    esk = b["esk"].cpu().numpy()
    r = ref.cpu().numpy()
    for i in (0, 1, n - 1):
        m = b["flat"][b["offs"][i]:b["offs"][i] + b["lens"][i]].tobytes()
        assert r[i].tobytes() == O.sign(v, esk[b["kidx"][i]].tobytes(), m)[0], i


@pytest.mark.parametrize("nkeys", (40, 5000), ids=("40keys", "5000keys"))
def test_shared_key_kernels_over_several_sort_batches(engine_factory, restore_config, nkeys):
    """The shared-key kernels on 20,000 items (the key sort loads its keys in
    batches of 8,192, so three of them, the last ragged) give the wave
    kernels' bytes: with 40 keys (about 500 items per key) and with 5,000
    keys, past the sort's 4,096 exact bins, where a workgroup holds items of
    several keys and runs its table loop once per key."""
    import torch
    v = 1
    eng = restore_config(engine_factory(v))
    z = eng.sizes
    n = 20000
    b = _batch(eng, n, nkeys, 90 + nkeys % 7, maxlen=200)
    var = eng.launch_variants()
    wave_p = [p for p in var["sign_prep"] if p < 1000][0]
    wave_v = [p for p in var["verify"] if p < 1000][0]
    shared = [p for p in var["sign_prep"] if p >= 1000]
    print(f"[tuning] shared-key kernels, {n} items, {nkeys} keys, codes {shared}", flush=True)
    ws, wv = eng.sign_workspace(n), eng.verify_workspace(n)
    eng.set_launch_config(sign_prep=wave_p, sign_order=0, sign_chunks=1, verify=wave_v, verify_order=0)
    ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=eng.device)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], ref, ws)
    bad = ref.clone()
    bad[::3, 5] ^= 0x21
    vref = torch.full((n,), 7, dtype=torch.uint8, device=eng.device)
    eng.verify_prepared(b["pepk"], b["k"], b["m"], b["o"], b["l"], bad, vref, wv)
    torch.cuda.synchronize()
    expect = np.zeros(n, np.uint8)
    expect[::3] = 1
    assert np.array_equal(vref.cpu().numpy(), expect)
    for p, chunks in itertools.product(shared, (1, 2)):
        eng.set_launch_config(sign_prep=p, sign_chunks=chunks, verify=p)
        out = torch.full_like(ref, 0xA5)
        eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], out, ws)
        res = torch.full((n,), 7, dtype=torch.uint8, device=eng.device)
        eng.verify_prepared(b["pepk"], b["k"], b["m"], b["o"], b["l"], bad, res, wv)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (p, chunks)
        assert torch.equal(res, vref), (p, chunks)


def test_autotune_keeps_a_compiled_variant_and_the_results(engine_factory, restore_config):
    import torch
    v = 1
    eng = restore_config(engine_factory(v))
    z = eng.sizes
    n = 4096
    b = _batch(eng, n, 16, 7)
    ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=eng.device)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], ref, eng.sign_workspace(n))
    sigs = torch.zeros_like(ref)
    res = torch.full((n,), 7, dtype=torch.uint8, device=eng.device)
    print(f"[tuning] autotune MAYO_{v} n={n}", flush=True)
    rep = eng.autotune(b["pesk"], b["pepk"], b["k"], b["m"], b["o"], b["l"], sigs, res,
                       eng.sign_workspace(n), eng.verify_workspace(n), reps=2)
    torch.cuda.synchronize()
    print(f"[tuning] {rep}", flush=True)
    var = eng.launch_variants()
    best = rep["best"]
    assert best == eng.launch_config()
    assert best["sign_prep"] in var["sign_prep"] and best["sign_solve"] in var["sign_solve"]
    assert best["verify"] in var["verify"] and best["sign_order"] in (0, 1) and best["verify_order"] in (0, 1)
    chunk_options = 1 if n < 16384 else 2  # the two-part launch applies from 16,384 items on
    # wave-per-item variants: both item orders, plus the persistent fused
    # kernel, which has no solve variant; shared-key variants (codes from
    # 1000) run in their own key order and are not a fused-kernel option
    wave_p = [p for p in var["sign_prep"] if p < 1000]
    shared_p = [p for p in var["sign_prep"] if p >= 1000]
    wave_v = [p for p in var["verify"] if p < 1000]
    shared_v = [p for p in var["verify"] if p >= 1000]
    assert shared_p and shared_v
    assert rep["configs_timed"] == (2 * len(wave_p) * (1 + chunk_options * len(var["sign_solve"])) +
                                    len(shared_p) * chunk_options * len(var["sign_solve"]) +
                                    2 * len(wave_v) + len(shared_v))
    assert 0 < rep["sign_ms"] <= rep["sign_default_ms"] and 0 < rep["verify_ms"] <= rep["verify_default_ms"]
    assert torch.equal(sigs, ref)
    assert res.cpu().numpy().tolist() == [0] * n
    # the tuned configuration signs the same bytes
    again = torch.zeros_like(ref)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], again, eng.sign_workspace(n))
    torch.cuda.synchronize()
    assert torch.equal(again, ref)


@pytest.mark.parametrize("shared", (False, True), ids=("wave-kernels", "shared-key-kernels"))
@pytest.mark.parametrize("forced", (0, 2), ids=("plain", "forced-retries"))
def test_graph_replays_match_eager(engine_factory, restore_config, forced, shared):
    """Sign and verify pipelines captured as HIP graphs, replayed three
    times with eager work (clearing the outputs) and a synchronisation
    between replays: every replay gives the eager bytes, also when every
    item goes through the retry pass (its failure list is reset inside the
    kernels, not by a memory-set node), with the wave kernels and with the
    shared-key kernels (key sort, table and fold kernels)."""
    import torch
    v = 1
    eng = restore_config(engine_factory(v))
    if shared:
        code = [p for p in eng.launch_variants()["sign_prep"] if p >= 1000][0]
        eng.set_launch_config(sign_prep=code, verify=code, sign_chunks=1)
    z = eng.sizes
    n = 1536
    b = _batch(eng, n, 6, 3 + forced, maxlen=900)
    stages = 3 | (forced << 8)
    ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=eng.device)
    ws_e = eng.sign_workspace(n)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], ref, ws_e, stages=stages)
    vref = torch.full((n,), 7, dtype=torch.uint8, device=eng.device)
    eng.verify_prepared(b["pepk"], b["k"], b["m"], b["o"], b["l"], ref, vref, eng.verify_workspace(n))
    torch.cuda.synchronize()
    att = eng.sign_attempts(ws_e, n)
    assert (att >= forced + 1).all()
    assert vref.cpu().numpy().tolist() == [0] * n
    sigs = torch.zeros_like(ref)
    ws = eng.sign_workspace(n)
    res = torch.full((n,), 7, dtype=torch.uint8, device=eng.device)
    print(f"[tuning] graphs MAYO_{v} n={n} forced={forced} shared={shared}", flush=True)
    gs = eng.sign_graph(b["pesk"], b["k"], b["m"], b["o"], b["l"], sigs, ws, stages=stages)
    gv = eng.verify_graph(b["pepk"], b["k"], b["m"], b["o"], b["l"], sigs, res, eng.verify_workspace(n))
    try:
        for rep in range(3):
            sigs.zero_()
            res.fill_(7)
            torch.cuda.synchronize()
            gs.launch()
            gv.launch()
            torch.cuda.synchronize()
            assert torch.equal(sigs, ref), f"sign replay {rep}"
            assert np.array_equal(eng.sign_attempts(ws, n), att), f"attempts replay {rep}"
            assert torch.equal(res, vref), f"verify replay {rep}"
            # eager work between replays on the same buffers
            eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], sigs, ws, stages=stages)
            torch.cuda.synchronize()
            assert torch.equal(sigs, ref)
    finally:
        gs.close()
        gv.close()
    esk = b["esk"].cpu().numpy()
    r = ref.cpu().numpy()
    for i in (0, n // 2, n - 1):
        m = b["flat"][b["offs"][i]:b["offs"][i] + b["lens"][i]].tobytes()
        osig, oatt = O.sign(v, esk[b["kidx"][i]].tobytes(), m, fail_first=forced)
        assert r[i].tobytes() == osig and int(att[i]) == oatt, i


@pytest.mark.parametrize("forced", (0, 1), ids=("plain", "forced-retries"))
def test_two_part_stage2_matches_one_part(engine_factory, restore_config, forced):
    """From 16,384 items on, sign_chunks = 2 runs the linear-algebra pass as
    two halves on two streams, each with its own failure list and retry
    pass: the signatures and attempt counts equal the one-part launch's, also
    when every item goes through the retry pass, and also replayed from a
    graph captured across both streams."""
    import torch
    v = 1
    eng = restore_config(engine_factory(v))
    z = eng.sizes
    n = 20000
    b = _batch(eng, n, 8, 90 + forced, maxlen=300)
    stages = 3 | (forced << 8)
    eng.set_launch_config(sign_chunks=1)
    ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=eng.device)
    ws1 = eng.sign_workspace(n)
    eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], ref, ws1, stages=stages)
    torch.cuda.synchronize()
    att = eng.sign_attempts(ws1, n)
    assert (att >= forced + 1).all()
    for order in (0, 1):
        eng.set_launch_config(sign_chunks=2, sign_order=order)
        print(f"[tuning] two-part stage 2 n={n} forced={forced} order={order}", flush=True)
        out = torch.full_like(ref, 0x5A)
        ws2 = eng.sign_workspace(n)
        eng.sign_prepared(b["pesk"], b["k"], b["m"], b["o"], b["l"], out, ws2, stages=stages)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), order
        assert np.array_equal(eng.sign_attempts(ws2, n), att), order
    eng.set_launch_config(sign_chunks=2, sign_order=0)
    gsig = torch.zeros_like(ref)
    gws = eng.sign_workspace(n)
    g = eng.sign_graph(b["pesk"], b["k"], b["m"], b["o"], b["l"], gsig, gws, stages=stages)
    try:
        for rep in range(2):
            gsig.zero_()
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert torch.equal(gsig, ref), f"graph replay {rep}"
            assert np.array_equal(eng.sign_attempts(gws, n), att)
    finally:
        g.close()
    esk = b["esk"].cpu().numpy()
    r = ref.cpu().numpy()
    for i in (0, n // 2 - 1, n // 2, n - 1):
        m = b["flat"][b["offs"][i]:b["offs"][i] + b["lens"][i]].tobytes()
        osig, oatt = O.sign(v, esk[b["kidx"][i]].tobytes(), m, fail_first=forced)
        assert r[i].tobytes() == osig and int(att[i]) == oatt, i
