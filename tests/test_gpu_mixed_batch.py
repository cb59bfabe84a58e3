"""GPU: one batch that mixes the four parameter sets (MixedBatch). Every
item's signature equals the one its set's single-set batch gives (and the
oracle's, sampled); verification of the mixed batch accepts them and
rejects mutated ones, in batch order."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_mixed_set_batch_equals_single_set_batches(engine_factory):
    import torch
    from cumayo_amd.mixed import MixedBatch
    sets = (1, 2, 3, 5)
    engines, prepared, tables = {}, {}, {}
    for v in sets:
        eng = engine_factory(v)
        esk, epk = eng.keypair_tables(4, rand_index=60 + v)
        engines[v], tables[v] = eng, esk
        prepared[v] = eng.prepare_tables(esk, epk)
    n = 700
    rng = np.random.default_rng(2024)
    set_of = rng.choice(sets, size=n, p=[0.4, 0.2, 0.2, 0.2]).astype(np.int32)
    lens = rng.integers(0, 900, n).astype(np.int32)
    lens[:3] = [0, 136, 137]
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = rng.integers(0, 4, n).astype(np.int32)
    dev = engines[1].device
    d_msgs = torch.from_numpy(flat).to(dev)
    mb = MixedBatch(engines, prepared)
    print(f"[mixed-batch] n={n} sets {np.bincount(set_of).tolist()}", flush=True)
    out = mb.sign(set_of, kidx, d_msgs, offs, lens)
    torch.cuda.synchronize()
    assert sorted(out) == list(sets)
    seen = np.zeros(n, bool)
    for v, (pos, sigs) in out.items():
        eng = engines[v]
        assert np.all(set_of[pos] == v) and np.all(np.diff(pos) > 0)
        seen[pos] = True
        # the same items as a single-set batch
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        ref = torch.zeros_like(sigs)
        eng.sign_prepared(prepared[v][0], t(kidx[pos]), d_msgs, t(offs[pos]), t(lens[pos]), ref,
                          eng.sign_workspace(pos.size))
        torch.cuda.synchronize()
        assert torch.equal(sigs, ref), f"MAYO_{v}"
        esk = tables[v].cpu().numpy()
        g = sigs.cpu().numpy()
        for k in (0, pos.size - 1):
            i = pos[k]
            m = flat[offs[i]:offs[i] + lens[i]].tobytes()
            assert g[k].tobytes() == O.sign(v, esk[kidx[i]].tobytes(), m)[0], (v, i)
    assert seen.all()
    # verification of the mixed batch, with every fifth signature mutated
    bad = {}
    expect = np.zeros(n, np.uint8)
    for v, (pos, sigs) in out.items():
        b = sigs.clone()
        b[::5, 1] ^= 0x40
        bad[v] = b
        expect[pos[::5]] = 1
    res = mb.verify(set_of, kidx, d_msgs, offs, lens, {v: s for v, (p, s) in out.items()})
    res_bad = mb.verify(set_of, kidx, d_msgs, offs, lens, bad)
    torch.cuda.synchronize()
    assert res.cpu().numpy().tolist() == [0] * n
    assert np.array_equal(res_bad.cpu().numpy(), expect)


def test_mixed_set_graph_replays_equal_eager(engine_factory):
    """One HIP graph over a four-set batch (MixedGraph: each set's pipeline
    captured by its library, composed as children of one graph): every
    replay, with the program's own eager work and synchronisations between
    replays, gives the eager MixedBatch bytes; the graph reads the messages
    at replay time; a verification graph accepts the signatures and rejects
    mutated ones."""
    import torch
    from cumayo_amd.mixed import MixedBatch, MixedGraph
    sets = (1, 2, 3, 5)
    engines, prepared = {}, {}
    for v in sets:
        eng = engine_factory(v)
        esk, epk = eng.keypair_tables(3, rand_index=80 + v)
        engines[v], prepared[v] = eng, eng.prepare_tables(esk, epk)
    n = 1200
    rng = np.random.default_rng(77)
    set_of = rng.choice(sets, size=n, p=[0.4, 0.2, 0.2, 0.2]).astype(np.int32)
    lens = rng.integers(0, 700, n).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = rng.integers(0, 3, n).astype(np.int32)
    dev = engines[1].device
    d_msgs = torch.from_numpy(flat).to(dev)
    mb = MixedBatch(engines, prepared)
    print(f"[mixed-graph] n={n} sets {np.bincount(set_of).tolist()}", flush=True)
    ref = mb.sign(set_of, kidx, d_msgs, offs, lens)
    torch.cuda.synchronize()
    g = MixedGraph(mb, "sign", set_of, kidx, d_msgs, offs, lens)
    for rep in range(3):
        for _, out in g.results().values():
            out.fill_(0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for v, (pos, out) in g.results().items():
            assert np.array_equal(pos, ref[v][0])
            assert torch.equal(out, ref[v][1]), (rep, v)
        # eager work between replays
        again = mb.sign(set_of, kidx, d_msgs, offs, lens)
        torch.cuda.synchronize()
        assert all(torch.equal(again[v][1], ref[v][1]) for v in sets)
    # messages rewritten in place: the next replay signs the new bytes
    flipped = flat.copy()
    flipped[::3] ^= 0x5A
    d_msgs.copy_(torch.from_numpy(flipped).to(dev))
    ref2 = mb.sign(set_of, kidx, d_msgs, offs, lens)
    torch.cuda.synchronize()
    g.launch()
    torch.cuda.synchronize()
    for v, (pos, out) in g.results().items():
        assert torch.equal(out, ref2[v][1]), v
        assert not torch.equal(out, ref[v][1]), v
    d_msgs.copy_(torch.from_numpy(flat).to(dev))
    torch.cuda.synchronize()
    # verification graphs: valid signatures, and every seventh one mutated
    sigs = {v: s for v, (p, s) in ref.items()}
    bad = {}
    expect_bad = {}
    for v, (pos, s) in ref.items():
        b = s.clone()
        b[::7, 2] ^= 0x21
        bad[v] = b
        e = np.zeros(pos.size, np.uint8)
        e[::7] = 1
        expect_bad[v] = e
    gv = MixedGraph(mb, "verify", set_of, kidx, d_msgs, offs, lens, sigs)
    gb = MixedGraph(mb, "verify", set_of, kidx, d_msgs, offs, lens, bad)
    for rep in range(3):
        for graph in (gv, gb):
            for _, out in graph.results().values():
                out.fill_(0xEE)
        torch.cuda.synchronize()
        gv.launch()
        gb.launch()
        torch.cuda.synchronize()
        for v, (pos, out) in gv.results().items():
            assert out.cpu().numpy().tolist() == [0] * pos.size, (rep, v)
        for v, (pos, out) in gb.results().items():
            assert np.array_equal(out.cpu().numpy(), expect_bad[v]), (rep, v)
        res = mb.verify(set_of, kidx, d_msgs, offs, lens, sigs)
        torch.cuda.synchronize()
        assert res.cpu().numpy().tolist() == [0] * n
    for graph in (g, gv, gb):
        graph.close()


def test_mixed_set_graph_with_two_part_launch(engine_factory):
    """The composed graph with MAYO_1's part above the two-part threshold
    (16,384 items) and the two-part launch in force for it: replays equal
    the eager bytes, and the verification graph accepts every signature."""
    import torch
    from cumayo_amd.mixed import MixedBatch, MixedGraph
    sets = (1, 2, 3, 5)
    engines, prepared = {}, {}
    for v in sets:
        eng = engine_factory(v)
        esk, epk = eng.keypair_tables(4, rand_index=90 + v)
        engines[v], prepared[v] = eng, eng.prepare_tables(esk, epk)
    saved = engines[1].launch_config()
    engines[1].set_launch_config(sign_chunks=2)
    try:
        n = 48000
        rng = np.random.default_rng(4242)
        set_of = rng.choice(sets, size=n, p=[0.4, 0.2, 0.2, 0.2]).astype(np.int32)
        assert (set_of == 1).sum() >= 16384
        lens = rng.integers(0, 300, n).astype(np.int32)
        offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
        flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
        kidx = rng.integers(0, 4, n).astype(np.int32)
        dev = engines[1].device
        d_msgs = torch.from_numpy(flat).to(dev)
        mb = MixedBatch(engines, prepared)
        print(f"[mixed-graph-2part] n={n} sets {np.bincount(set_of).tolist()}", flush=True)
        ref = mb.sign(set_of, kidx, d_msgs, offs, lens)
        torch.cuda.synchronize()
        print("[mixed-graph-2part] eager done, composing", flush=True)
        g = MixedGraph(mb, "sign", set_of, kidx, d_msgs, offs, lens)
        gv = MixedGraph(mb, "verify", set_of, kidx, d_msgs, offs, lens, {v: s for v, (p, s) in ref.items()})
        for rep in range(2):
            for _, out in g.results().values():
                out.fill_(0)
            torch.cuda.synchronize()
            g.launch()
            gv.launch()
            torch.cuda.synchronize()
            print(f"[mixed-graph-2part] replay {rep} done", flush=True)
            for v, (pos, out) in g.results().items():
                assert torch.equal(out, ref[v][1]), (rep, v)
            for v, (pos, out) in gv.results().items():
                assert out.cpu().numpy().tolist() == [0] * pos.size, (rep, v)
            again = mb.sign(set_of, kidx, d_msgs, offs, lens)
            torch.cuda.synchronize()
            assert all(torch.equal(again[v][1], ref[v][1]) for v in sets)
        g.close()
        gv.close()
    finally:
        engines[1].set_launch_config(**saved)
