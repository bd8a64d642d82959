"""Ray-parity voxelisation (src/lbm/voxelize.cpp:4-6) on the CPU: the oracle's restatement pinned
against the reference's own tests and data, the masks of the reference build (tests/golden/
voxel_masks.npz, written by make_voxel_masks.py from oracle/_ref), the live reference build where
it exists, and closed forms: grid-aligned boxes whose faces pass through cell centres (the ray
nudges decide), the half-space test of convex polyhedra, and the analytic torus.  Also the host
side of the HIP library that runs without a GPU: transform_surface and input validation.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from cfd_solvers_amd import geometry as G

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["cube", "cube_moved", "ahmed", "sphere", "torus", "grazing", "prism"]


@pytest.fixture(scope="module")
def masks():
    return dict(np.load(GOLDEN / "voxel_masks.npz", allow_pickle=False))


def centres(shape, lo=(0, 0, 0), hi=(1, 1, 1)):
    """Cell centres (z, y, x grids) in voxelize.cpp's arithmetic, without the ray nudges."""
    nx, ny, nz = shape
    d = [(hi[a] - lo[a]) / n for a, n in enumerate(shape)]
    x = lo[0] + (np.arange(nx) + .5) * d[0]
    y = lo[1] + (np.arange(ny) + .5) * d[1]
    z = lo[2] + (np.arange(nz) + .5) * d[2]
    return np.meshgrid(z, y, x, indexing="ij"), d


def test_reference_cube_tests():
    """tests/test_completion.cpp:885-901: the cube [.25,.75]^3 on 8^3; (4,4,4) inside, (0,0,0) not;
    moving it by 0.1 in x changes the mask.  Closed form: exactly the 4^3 cells i, j, k in 2..5."""
    cube = G.box_surface((.25, .25, .25), (.75, .75, .75))
    m = O.voxelize(cube, (8, 8, 8))
    assert m[4, 4, 4] == 1 and m[0, 0, 0] == 0
    want = np.zeros((8, 8, 8), np.uint8)
    want[2:6, 2:6, 2:6] = 1
    assert np.array_equal(m, want)
    moved = O.transform_surface(cube, (.1, 0, 0), (0, 0, 1), 0.0)
    m2 = O.voxelize(moved, (8, 8, 8))
    assert not np.array_equal(m2, m)
    want2 = np.zeros((8, 8, 8), np.uint8)
    want2[2:6, 2:6, 3:7] = 1  # x centres .4375 .. .8125 lie in (.35, .85)
    assert np.array_equal(m2, want2)


def test_ahmed_body_reference_test():
    """tests/test_real_3d_model.cpp:29-39: 18 triangles; a bounded solid on 32 x 20 x 20."""
    tris = G.read_ascii_stl(GOLDEN / "ahmed_body.stl")
    assert tris.shape == (18, 3, 3)
    m = O.voxelize(tris, (32, 20, 20))
    assert 0 < int(m.sum()) < m.size


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference_masks(masks, case):
    lo, hi = masks[f"{case}_box"]
    m = O.voxelize(masks[f"{case}_triangles"], tuple(masks[f"{case}_grid"]), lo, hi)
    assert np.array_equal(m, masks[f"{case}_mask"])


def test_transform_matches_reference(masks):
    a = masks["xform_args"]
    out = O.transform_surface(masks["xform_in"], a[0:3], a[3:6], a[6])
    assert np.array_equal(out, masks["xform_out"])


def test_library_transform_matches_reference(masks):
    """cfd_lbm_transform_surface is host arithmetic: it runs here, without a GPU."""
    import cfd_solvers_amd as hip

    a = masks["xform_args"]
    out = hip.transform_surface(masks["xform_in"], a[0:3], a[3:6], a[6])
    assert np.array_equal(out, masks["xform_out"])
    same = hip.transform_surface(masks["xform_in"], (1, 2, 3), (0, 0, 0), 0.4)  # zero axis: no rotation
    assert np.array_equal(same, masks["xform_in"] + np.array([1.0, 2.0, 3.0]))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_live_reference(seed):
    rng = np.random.default_rng(seed)
    axis, angle = rng.normal(size=3), rng.uniform(0, 2 * np.pi)
    shift = rng.uniform(.4, .6, size=3)
    surfaces = [G.sphere_surface((0, 0, 0), rng.uniform(.15, .35), 12, 24),
                G.torus_surface((0, 0, 0), .28, .1, 20, 10),
                G.box_surface(-rng.uniform(.1, .3, 3), rng.uniform(.1, .3, 3))]
    for s in surfaces:
        t = O.ref_transform_surface(s, shift, axis, angle)
        assert np.array_equal(O.transform_surface(s, shift, axis, angle), t)
        shape = tuple(int(v) for v in rng.integers(9, 30, size=3))
        assert np.array_equal(O.voxelize(t, shape), O.ref_voxelize(t, shape))


def test_grazing_box_closed_form(masks):
    """Faces through cell centres: a centre on a low x face is inside (the crossing at distance 0 is
    not counted, the far face is), one on a high x face outside; in y and z the nudges 1e-10 dx and
    2e-10 dx move centres on a face to the +y / +z side.  Without the nudges the rays graze the
    shared edges and the parity breaks."""
    lo, hi = (.3125, .3125, .1875), (.6875, .6875, .8125)
    m = O.voxelize(G.box_surface(lo, hi), (8, 8, 8))
    (z, y, x), _ = centres((8, 8, 8))
    want = (x >= lo[0]) & (x < hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) & (z < hi[2])
    assert np.array_equal(m, want.astype(np.uint8))
    assert int(m.sum()) == 3 * 3 * 5


def test_diagonal_prism_closed_form():
    """A prism whose face lies in the plane y = z, through the centres with j == k: the nudge
    (1e-10 dx in y, 2e-10 dx in z) puts them on the z > y side, outside this prism (z < y).  Swapped
    or equal nudges would put them inside or on the face."""
    tris = G.prism_surface(.3, .7, [(.2, .2), (.8, .8), (.8, .2)])
    m = O.voxelize(tris, (8, 8, 8))
    (z, y, x), _ = centres((8, 8, 8))
    want = (x > .3) & (x < .7) & (z > .2) & (y < .8) & (z < y)
    assert np.array_equal(m, want.astype(np.uint8))
    assert int(m.sum()) == 24


def convex_inside(tris, pts, band):
    """Inside test of a closed convex mesh with outward faces: every face plane has the point on its
    inner side.  Also returns the cells within `band` of some face plane (left out)."""
    a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1)[:, None]
    signed = pts @ n.T - np.einsum("fk,fk->f", a, n)[None]
    return np.all(signed < 0, axis=1), np.any(np.abs(signed) < band, axis=1)


@pytest.mark.parametrize("kind", ["box", "sphere"])
def test_convex_halfspace(kind):
    """Parity equals the half-space test of a convex polyhedron (a different algorithm) on every cell
    farther than 1e-9 from the surface's planes."""
    shape = (37, 29, 33)
    if kind == "box":
        base = G.box_surface((-.3, -.2, -.25), (.3, .2, .25))
    else:
        base = G.sphere_surface((0, 0, 0), .32, 10, 20)
    tris = O.transform_surface(base, (.51, .49, .5), (.3, -1, .6), 0.8)
    m = O.voxelize(tris, shape)
    (z, y, x), _ = centres(shape)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    inside, near = convex_inside(tris, pts, 1e-9)
    keep = ~near
    assert keep.sum() > 0.99 * keep.size
    assert np.array_equal(m.ravel()[keep], inside[keep].astype(np.uint8))
    assert 0 < m.sum() < m.size


def test_torus_analytic():
    """A torus: parity equals (rho - R)^2 + z^2 < r^2 away from the facetted surface."""
    R, r = .3, .12
    tris = G.torus_surface((.5, .5, .5), R, r, 48, 24)
    shape = (40, 40, 24)
    m = O.voxelize(tris, shape, (0, 0, .2), (1, 1, .8))
    (z, y, x), _ = centres(shape, (0, 0, .2), (1, 1, .8))
    d = np.sqrt((np.hypot(x - .5, y - .5) - R) ** 2 + (z - .5) ** 2) - r
    far = np.abs(d) > 0.01  # the facets deviate from the torus by < 3e-3
    assert np.array_equal(m[far], (d[far] < 0).astype(np.uint8))
    assert m.sum() > 0


def test_read_ascii_stl(tmp_path):
    good = tmp_path / "t.stl"
    good.write_text("solid t\nfacet normal 0 0 1\nouter loop\nvertex 0 0 0\nvertex 1 0 0\nvertex 0 1e-1 0\n"
                    "endloop\nendfacet\nendsolid t\n")
    t = G.read_ascii_stl(good)
    assert t.shape == (1, 3, 3) and t[0, 2, 1] == 0.1
    bad = tmp_path / "b.stl"
    bad.write_text("solid b\nvertex 0 0 0\nvertex 1 0 0\nendsolid\n")
    with pytest.raises(RuntimeError, match="malformed ASCII STL"):
        G.read_ascii_stl(bad)
    with pytest.raises(RuntimeError, match="cannot open STL"):
        G.read_ascii_stl(tmp_path / "missing.stl")


def test_voxelize_input_errors():
    """Validation runs before any device call (voxelize.cpp:5's message for the same inputs)."""
    import cfd_solvers_amd as hip

    cube = G.box_surface((.25, .25, .25), (.75, .75, .75))
    with pytest.raises(RuntimeError, match="invalid voxelization inputs"):
        hip.voxelize_surface(np.zeros((0, 3, 3)), (8, 8, 8))
    with pytest.raises(RuntimeError, match="invalid voxelization inputs"):
        hip.voxelize_surface(cube, (8, 0, 8))
    with pytest.raises(RuntimeError, match="invalid voxelization inputs"):
        hip.voxelize_surface(cube, (8, 8, 8), (0, 0, 0), (1, 0, 1))
    bad = cube.copy()
    bad[3, 1, 2] = np.nan
    with pytest.raises(RuntimeError, match="finite"):
        hip.voxelize_surface(bad, (8, 8, 8))
    with pytest.raises(RuntimeError, match="invalid voxelization inputs"):
        O.voxelize(np.zeros((0, 3, 3)), (8, 8, 8))
