"""Host-side logic that runs without a GPU: splitting a mixed-set batch by
parameter set, and the benchmark's seeded workload generators (lengths in
range, offsets packed, items reproducible from their index)."""
import hashlib

import numpy as np

import bench
from cumayo_amd.mixed import split_by_set


def test_split_by_set_partitions_the_batch_in_order():
    rng = np.random.default_rng(5)
    set_of = rng.choice([1, 2, 3, 5], size=1000)
    parts = split_by_set(set_of, (1, 2, 3, 5, 7))
    assert 7 not in parts
    allpos = np.sort(np.concatenate(list(parts.values())))
    assert np.array_equal(allpos, np.arange(1000))
    for v, pos in parts.items():
        assert np.all(set_of[pos] == v) and np.all(np.diff(pos) > 0)
    assert split_by_set(np.array([], dtype=np.int32), (1, 2)) == {}


def test_uniform_workload_is_packed_and_reproducible():
    flat, offs, lens, kidx = bench.make_workload(300, 7, 32, 1024, tag="unit")
    assert lens.min() >= 32 and lens.max() <= 1024
    assert offs[0] == 0 and np.array_equal(offs[1:], np.cumsum(lens[:-1]))
    assert flat.size == int(lens.sum())
    assert np.array_equal(kidx, np.arange(300) % 7)
    # item j is SHAKE256("mayo-bench/<tag>/msg" || LE64(j)), independent of the batch size
    j = 123
    m = hashlib.shake_256(b"mayo-bench/unit/msg" + j.to_bytes(8, "little")).digest(int(lens[j]))
    assert flat[offs[j]:offs[j] + lens[j]].tobytes() == m
    flat2, offs2, lens2, _ = bench.make_workload(150, 7, 32, 1024, tag="unit")
    assert np.array_equal(lens2, lens[:150])


def test_log_uniform_workload_bounds():
    flat, offs, lens, kidx = bench.log_uniform_workload(400, 4, 32, 16384, tag="unit")
    assert lens.min() >= 32 and lens.max() <= 16384
    assert flat.size == int(lens.sum())
    # log-uniform: about half the lengths below the geometric mean of the bounds
    below = float((lens < np.sqrt(32 * 16384)).mean())
    assert 0.35 < below < 0.65
