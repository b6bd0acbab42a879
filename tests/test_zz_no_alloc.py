"""The allocation-free terrain entry point against the allocating one.

skt_terrain_generate_noalloc computes into the library's fixed device scratch
array instead of allocating; its outputs must equal skt_terrain_generate's
bit for bit, on two calls in a row (the second with fewer columns, so the
scratch still holds the first call's values beyond them).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    import skewtcloudsim_amd as skt
    L = skt.load_library()
    fn = L.skt_terrain_generate_noalloc
    i, f, p = ctypes.c_int32, ctypes.c_float, ctypes.c_void_p
    fn.restype = i
    fn.argtypes = [i, i, i, i, i, f, p, p]
    return L


def _gen(fn, nx, nz, seed, octaves, maxh):
    noise = np.full(nx * nz, np.nan, np.float32)
    heights = np.full(nx * nz, -7, np.int32)
    rc = fn(0, seed, nx, nz, octaves, maxh, noise.ctypes.data_as(ctypes.c_void_p),
            heights.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, rc
    return noise, heights


@pytest.mark.parametrize("calls", [((64, 48, 100, 12, 6.0), (33, 17, 7, 4, 3.0)),
                                   ((256, 256, 100, 12, 20.0), (8, 5, 3, 12, 0.2))])
def test_noalloc_terrain_matches_allocating_entry_point(calls):
    L = _lib()
    for nx, nz, seed, octaves, maxh in calls:
        print(f"noalloc terrain {nx}x{nz} seed {seed}", flush=True)
        n0, h0 = _gen(L.skt_terrain_generate, nx, nz, seed, octaves, maxh)
        n1, h1 = _gen(L.skt_terrain_generate_noalloc, nx, nz, seed, octaves, maxh)
        assert np.isfinite(n0).all()
        assert np.array_equal(n0.view(np.uint32), n1.view(np.uint32))
        assert np.array_equal(h0, h1)


def test_noalloc_terrain_rejects_oversize_and_bad_args():
    L = _lib()
    p = ctypes.c_void_p
    out = np.zeros(4, np.float32)
    fn = L.skt_terrain_generate_noalloc
    # (1 << 20) + 1 columns: rejected on the host before any device work
    assert fn(0, 1, (1 << 20) + 1, 1, 4, 1.0, out.ctypes.data_as(p), None) == 1
    assert fn(0, 1, 0, 4, 4, 1.0, out.ctypes.data_as(p), None) == 1
    assert fn(0, 1, 2, 2, 4, 1.0, None, None) == 1
    # the maximum itself is accepted
    n, h = _gen(fn, 1 << 10, 1 << 10, 5, 2, 1.0)
    assert np.isfinite(n).all()
