"""GPU ray-parity voxelisation (cfd_lbm_voxelize_surface / cfd_lbm_set_solid_surface) against the
reference's masks (tests/golden/voxel_masks.npz, from src/lbm/voxelize.cpp:5 compiled by
oracle/Makefile) and the oracle's restatement, bit for bit; at full size through the analytic
inside test.  The binning splits a tile's triangles into rounds of 64, so cases with long tile lists
and with more than 64 crossings per line are included."""
from pathlib import Path

import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["cube", "cube_moved", "ahmed", "sphere", "torus", "grazing", "prism"]


@pytest.fixture(scope="module")
def masks():
    return dict(np.load(GOLDEN / "voxel_masks.npz", allow_pickle=False))


@pytest.mark.parametrize("case", CASES)
def test_reference_masks(hip, masks, case):
    lo, hi = masks[f"{case}_box"]
    shape = tuple(int(v) for v in masks[f"{case}_grid"])
    report("voxelize_fixture", case=case, shape=shape)
    m = hip.voxelize_surface(masks[f"{case}_triangles"], shape, lo, hi)
    assert np.array_equal(m, masks[f"{case}_mask"])


def random_surface(rng, G, O, kind):
    axis, angle, shift = rng.normal(size=3), rng.uniform(0, 2 * np.pi), rng.uniform(.35, .65, size=3)
    if kind == "sphere":
        base = G.sphere_surface((0, 0, 0), rng.uniform(.2, .45), 14, 28)
    elif kind == "torus":
        base = G.torus_surface((0, 0, 0), .3, .12, 24, 12)
    else:  # a box that pokes out of the domain
        base = G.box_surface(-rng.uniform(.2, .7, 3), rng.uniform(.2, .7, 3))
    return O.transform_surface(base, shift, axis, angle)


@pytest.mark.parametrize("seed,kind,shape", [
    (0, "sphere", (40, 37, 29)), (1, "torus", (300, 24, 17)), (2, "box", (33, 65, 9)),
    (3, "sphere", (1, 50, 50)), (4, "torus", (64, 1, 64)), (5, "box", (17, 19, 1)),
])
def test_random_surfaces_match_oracle(hip, oracle_mod, seed, kind, shape):
    from cfd_solvers_amd import geometry as G

    rng = np.random.default_rng(seed)
    tris = random_surface(rng, G, oracle_mod, kind)
    lo, hi = (-.05, 0.0, .02), (1.0, 1.1, .97)
    report("voxelize_random", seed=seed, kind=kind, shape=shape, triangles=len(tris))
    m = hip.voxelize_surface(tris, shape, lo, hi)
    want = oracle_mod.voxelize(tris, shape, lo, hi)
    assert np.array_equal(m, want), f"{int((m != want).sum())} cells differ"


def test_dense_mesh_many_rounds(hip, oracle_mod):
    """9024 triangles on a coarse grid: each tile lists hundreds of triangles (many 64-triangle
    rounds), the parity of later rounds is XORed into the first."""
    from cfd_solvers_amd import geometry as G

    tris = oracle_mod.transform_surface(G.sphere_surface((0, 0, 0), .41, 48, 96), (.5, .5, .5), (1, 1, 0), .3)
    shape = (30, 26, 22)
    report("voxelize_dense", triangles=len(tris))
    m = hip.voxelize_surface(tris, shape)
    assert np.array_equal(m, oracle_mod.voxelize(tris, shape))
    _, _, pairs = hip.voxelize_timing()
    assert pairs > 64 * 9  # several rounds per tile


def test_many_crossings_per_line(hip, oracle_mod):
    """40 boxes in a row along x and 10 nested spheres: up to 80 crossings on one line."""
    from cfd_solvers_amd import geometry as G

    boxes = [G.box_surface((.02 + .024 * i, .3, .3), (.034 + .024 * i, .7, .7)) for i in range(40)]
    shells = [G.sphere_surface((.5, .5, .5), .05 + .04 * i, 8, 16) for i in range(10)]
    tris = np.concatenate(boxes + shells)
    shape = (200, 24, 24)
    report("voxelize_crossings", triangles=len(tris))
    m = hip.voxelize_surface(tris, shape)
    want = oracle_mod.voxelize(tris, shape)
    assert np.array_equal(m, want)
    assert want[12, 12].sum() > 40  # the middle line crosses every box and shell


def test_large_grid_analytic(hip):
    """512^3 cells, 16128-triangle sphere: every cell farther than 0.6 cells from the facetted sphere
    matches the analytic inside test; the run is deterministic (atomic binning order aside)."""
    from cfd_solvers_amd import geometry as G

    n, R = 512, .37
    tris = G.sphere_surface((.5, .5, .5), R, 64, 128)
    report("voxelize_large", n=n, triangles=len(tris))
    m = hip.voxelize_surface(tris, (n, n, n))
    b_ms, l_ms, pairs = hip.voxelize_timing()
    report("voxelize_large_timing", bin_ms=b_ms, lines_ms=l_ms, pairs=pairs)
    c = (np.arange(n) + .5) / n - .5
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    dist = np.sqrt(r2) - R
    chord = R * (1 - np.cos(np.pi / 64))  # facets lie within this of the sphere
    far = np.abs(dist) > chord + .6 / n
    assert np.array_equal(m[far], (dist[far] < 0).astype(np.uint8))
    assert np.array_equal(hip.voxelize_surface(tris, (n, n, n)), m)
    assert l_ms > 0 and b_ms > 0


def test_set_solid_surface_feeds_the_solver(hip, oracle_mod):
    """The device-resident set_solid_surface equals set_solid_mask(oracle mask): identical fields
    after 30 steps; keep_flags keeps a channel's inlet, outlet and walls."""
    from cfd_solvers_amd import geometry as G

    nx, ny, nz = 48, 26, 22
    tris = G.read_ascii_stl(GOLDEN / "ahmed_body.stl")
    mask = oracle_mod.voxelize(tris, (nx, ny, nz), (-.1, 0, 0), (1.1, 1, 1))
    report("set_solid_surface", shape=(nx, ny, nz))
    a = hip.EsotericPullHipSolver("d3q19", nx, ny, nz, 0.6)
    b = hip.EsotericPullHipSolver("d3q19", nx, ny, nz, 0.6)
    a.set_solid_surface(tris, (-.1, 0, 0), (1.1, 1, 1))
    b.set_solid_mask(mask)
    assert np.array_equal(a.flags(), np.where(mask != 0, hip.WALL, hip.FLUID).astype(np.uint8))
    for s in (a, b):
        s.initialize_taylor_green(0.02)
        s.step(30)
    ma, mb = a.download_macroscopic(), b.download_macroscopic()
    for f in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(getattr(ma, f), getattr(mb, f)), f
    channel = G.channel_flags(nx, ny, nz, inlet_outlet=True)
    a.set_flags(channel)
    a.set_solid_surface(tris, (-.1, 0, 0), (1.1, 1, 1), keep_flags=True)
    b.set_flags(np.where(mask != 0, hip.WALL, channel).astype(np.uint8))
    assert np.array_equal(a.flags(), b.flags())
    for s in (a, b):
        s.set_inlet_outlet(0.03, 1.0)
        s.initialize_uniform()
        s.step(30)
    ma, mb = a.download_macroscopic(), b.download_macroscopic()
    for f in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(getattr(ma, f), getattr(mb, f)), f
    assert np.abs(ma.ux).max() > 0
    a.close()
    b.close()
