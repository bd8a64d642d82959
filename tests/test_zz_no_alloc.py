"""cfd_lbm_voxelize_surface_no_alloc (working memory in a static per-GPU device array, caller-owned
device input and output) against cfd_lbm_voxelize_surface (host input and output, library-allocated
working memory), through the C ABI, bit for bit, on two calls in a row, the second with a smaller
grid and surface, so that the second reuses what the first left in the array."""
import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu


def _no_alloc(hip, dev, tris, shape, lo, hi):
    t = dev.upload(np.ascontiguousarray(tris, np.float64).reshape(-1))
    nx, ny, nz = shape
    mask = dev.upload(np.full((nz, ny, nx), 0xA5, np.uint8))  # stale bytes must be overwritten
    hip.bindings.voxelize_surface_no_alloc(t, len(tris), shape, mask, lo, hi)
    return dev.download(mask, (nz, ny, nx), np.uint8)


def test_no_alloc_matches_allocating_entry_point_twice(hip, dev):
    from cfd_solvers_amd import geometry as G

    big = G.sphere_surface((.5, .5, .5), .37, 64, 128)             # 16128 triangles
    small = G.torus_surface((.45, .5, .52), .3, .12, 24, 12)       # smaller surface, ragged grid
    for tris, shape, lo, hi in ((big, (96, 80, 72), (0, 0, 0), (1, 1, 1)),
                                (small, (41, 35, 27), (-.05, 0, .02), (1.0, 1.1, .97))):
        report("voxelize_no_alloc", triangles=len(tris), shape=shape)
        want = hip.voxelize_surface(tris, shape, lo, hi)
        got = _no_alloc(hip, dev, tris, shape, lo, hi)
        assert want.any() and not want.all()
        assert np.array_equal(got, want), f"{int((got != want).sum())} cells differ"


def test_no_alloc_rejects_what_it_cannot_hold(hip, dev):
    from cfd_solvers_amd import geometry as G

    tris = G.sphere_surface((.5, .5, .5), .3, 8, 16)
    t = dev.upload(np.ascontiguousarray(tris, np.float64).reshape(-1))
    mask = dev.alloc(8)
    # (2056/8)^2 = 66049 tiles > 65536: rejected on the host before any launch.
    with pytest.raises(RuntimeError, match="at most 65536"):
        hip.bindings.voxelize_surface_no_alloc(t, len(tris), (1, 2056, 2056), mask)
    # A non-finite vertex is found on the device (the kernels read only the caller's buffer).
    bad = tris.copy()
    bad[3, 1, 2] = np.nan
    tb = dev.upload(np.ascontiguousarray(bad, np.float64).reshape(-1))
    m2 = dev.alloc(6 * 6 * 6)
    with pytest.raises(RuntimeError, match="finite"):
        hip.bindings.voxelize_surface_no_alloc(tb, len(bad), (6, 6, 6), m2)
    # The array is usable again after a rejected call.
    want = hip.voxelize_surface(tris, (6, 6, 6))
    m3 = dev.alloc(6 * 6 * 6)
    hip.bindings.voxelize_surface_no_alloc(t, len(tris), (6, 6, 6), m3)
    assert np.array_equal(dev.download(m3, (6, 6, 6), np.uint8), want)
