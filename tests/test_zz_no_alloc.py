"""rl_trace_segments_noalloc (segment tracing through the library's static device scratch, no
allocation) against rl_trace_segments, both called through the C-ABI: two calls in a row per
mode, the second with fewer segments, and both bit-identical to the allocating entry point (and
so to the oracle, which rl_trace_segments matches bit for bit)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from relight_amd import _capi as A
from relight_amd import Relight, scenes

pytestmark = pytest.mark.gpu


def _segments(spec, n, seed):
    rng = np.random.default_rng(seed)
    allp = np.concatenate([m.positions for m in spec.meshes])
    lo, hi = allp.min(0), allp.max(0)
    return (np.ascontiguousarray(rng.uniform(lo, hi, (n, 3)), np.float32),
            np.ascontiguousarray(rng.uniform(lo, hi, (n, 3)), np.float32))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _trace(lib, fn, ctx, scene, a, b, mode):
    n = a.shape[0]
    if mode == A.RL_TRACE_ANY:
        flags = np.full(n, 7, np.uint8)
        st = getattr(lib, fn)(ctx, scene, _fp(a), _fp(b), n, mode, flags.ctypes.data_as(C.POINTER(C.c_uint8)), None)
        assert st == A.RL_SUCCESS, lib.rl_last_error()
        return flags
    hits = (A.rl_hit * n)()
    st = getattr(lib, fn)(ctx, scene, _fp(a), _fp(b), n, mode, None, hits)
    assert st == A.RL_SUCCESS, lib.rl_last_error()
    arr = np.ctypeslib.as_array(hits)
    return np.frombuffer(arr.tobytes(), np.uint8)  # every field, byte for byte


@pytest.mark.parametrize("mode", [A.RL_TRACE_ANY, A.RL_TRACE_CLOSEST])
def test_noalloc_trace_equals_allocating_trace(mode):
    print(f"[gpu] no-alloc trace mode={mode}", flush=True)
    r = Relight()
    lib = r._lib
    spec = scenes.city(num_boxes=600, tiles=(2, 2), side_range=(64, 128))
    scene, lms, ids = scenes.build_gpu(r, spec)
    ctx, sh = r._h, scene._h
    # First call at the maximum size, second smaller: stale contents of the shared scratch (the
    # first call's counters and records) must not leak into the second.
    for n, seed in ((A.RL_NOALLOC_MAX_SEGMENTS, 11), (20000, 12)):
        a, b = _segments(spec, n, seed)
        want = _trace(lib, "rl_trace_segments", ctx, sh, a, b, mode)
        want_rays = r.stats()["rays"]
        got = _trace(lib, "rl_trace_segments_noalloc", ctx, sh, a, b, mode)
        assert np.array_equal(got, want), f"n={n}: {np.count_nonzero(got != want)} bytes differ"
        assert r.stats()["rays"] == want_rays == n
    if mode == A.RL_TRACE_ANY:
        # Against the oracle as well (a subset keeps the CPU side short).
        a, b = _segments(spec, 4000, 13)
        got = _trace(lib, "rl_trace_segments_noalloc", ctx, sh, a, b, mode)
        ref = O.OracleScene.from_spec(spec).trace(a, b)
        assert np.array_equal(got, ref["flags"].astype(np.uint8))
    # Beyond the maximum: rejected with a message, before any device work.
    n = A.RL_NOALLOC_MAX_SEGMENTS + 1
    a = np.zeros((n, 3), np.float32)
    st = lib.rl_trace_segments_noalloc(ctx, sh, _fp(a), _fp(a), n, mode, None, None)
    assert st == A.RL_INVALID_ARGUMENT
    assert b"at most 65536" in lib.rl_last_error()
    del scene, lms
    r.close()
