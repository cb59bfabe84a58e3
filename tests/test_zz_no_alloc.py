"""The no-allocation signing entry point (cumayo_sign_prepared_noalloc),
called through the C ABI: its signatures equal cumayo_sign_prepared's on
the same inputs, on two calls in a row (the second smaller, so the shared
device workspace holds the first call's contents when it starts), for every
parameter set; a batch above CUMAYO_NOALLOC_MAX_SIGN is refused with
hipErrorInvalidValue and leaves the outputs untouched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 2048  # CUMAYO_NOALLOC_MAX_SIGN
HIP_ERROR_INVALID_VALUE = 1


def _batch(eng, n, nkeys, seed, maxlen):
    import torch
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, maxlen, n).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    flat = rng.integers(0, 256, int(lens.sum()) + 1, dtype=np.uint8)
    kidx = rng.integers(0, nkeys, n).astype(np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(eng.device)
    return t(kidx), t(flat), t(offs), t(lens)


def _noalloc(eng, pesk, b, sigs, stream=0):
    k, m, o, l = b
    return eng.lib.cumayo_sign_prepared_noalloc(pesk.data_ptr(), k.data_ptr(), m.data_ptr(), o.data_ptr(),
                                                l.data_ptr(), sigs.data_ptr(), ctypes.c_size_t(k.numel()),
                                                ctypes.c_void_p(stream))


@pytest.mark.parametrize("v", (1, 2, 3, 5), ids=lambda v: f"MAYO_{v}")
def test_noalloc_sign_matches_workspace_entry_point(engine_factory, v):
    import torch
    eng = engine_factory(v)
    z = eng.sizes
    esk, _ = eng.keypair_tables(8, rand_index=3)
    pesk, _ = eng.prepare_tables(esk, None)
    for call, (n, seed, maxlen) in enumerate(((NOALLOC_MAX, 10 + v, 1200), (333, 20 + v, 300))):
        print(f"[noalloc] MAYO_{v} call {call} n={n}", flush=True)
        b = _batch(eng, n, 8, seed, maxlen)
        ref = torch.zeros((n, z.sig_bytes), dtype=torch.uint8, device=eng.device)
        eng.sign_prepared(pesk, *b, ref, eng.sign_workspace(n))
        out = torch.full_like(ref, 0x5A)
        torch.cuda.synchronize()
        assert _noalloc(eng, pesk, b, out) == 0
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"MAYO_{v} call {call}"
    # beyond the maximum: refused up front, outputs untouched
    n = NOALLOC_MAX + 1
    b = _batch(eng, n, 8, 99, 64)
    out = torch.full((n, z.sig_bytes), 0x5A, dtype=torch.uint8, device=eng.device)
    assert _noalloc(eng, pesk, b, out) == HIP_ERROR_INVALID_VALUE
    assert eng.lib.cumayo_last_error() == HIP_ERROR_INVALID_VALUE
    torch.cuda.synchronize()
    assert bool((out == 0x5A).all())
