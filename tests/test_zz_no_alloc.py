"""The allocation-free distance entry point against the allocating one.

spk_distance_noalloc keeps its buffers in the library's static device array;
two calls in a row (the second on a smaller grid, so stale contents of the
first would show) must equal spk_distance bit for bit, in every metric and
source mode, and a grid larger than the array must be rejected with
SPK_EUNSUPPORTED.  Both entry points are called through the C-ABI.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import spk_oracle as O

pytestmark = pytest.mark.gpu


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def test_distance_noalloc_matches_spk_distance():
    from psacking_amd import _lib
    lib = _lib.load()
    ctx = C.c_void_p()
    assert lib.spk_ctx_create(0, C.byref(ctx)) == 0, _lib.last_error()
    try:
        rng = np.random.default_rng(61)
        for shape in [(240, 123, 100), (37, 29, 41)]:
            grid = (rng.integers(1, 4, shape) * (rng.random(shape) < 0.02)).astype(np.int32)
            grid[0, 0, 0] = 1
            for metric, sources in [(1, 1), (0, 0), (1, 0), (0, 1)]:
                print("[case] distance noalloc", shape, metric, sources, flush=True)
                ref = np.empty(shape, np.int32)
                got = np.full(shape, -7, np.int32)
                assert lib.spk_distance(ctx, _ptr(grid), *shape, metric, sources, _ptr(ref)) == 0
                assert lib.spk_distance_noalloc(ctx, _ptr(grid), *shape, metric, sources,
                                                _ptr(got)) == 0, _lib.last_error()
                np.testing.assert_array_equal(got, ref)
            np.testing.assert_array_equal(
                got, O.distance_field(grid, "l1", "nonzero"))
        big = np.zeros((256, 256, 128), np.int32)
        out = np.empty_like(big)
        rc = lib.spk_distance_noalloc(ctx, _ptr(big), 256, 256, 128, 0, 0, _ptr(out))
        assert rc == 6 and "static scratch" in _lib.last_error()
    finally:
        lib.spk_ctx_destroy(ctx)
