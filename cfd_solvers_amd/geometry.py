"""Seeded synthetic geometries for the bounded lattice (uint8 flag fields, numpy, shape (nz, ny, nx)).

* ``channel_flags``: walls on y = 0, ny-1 and z = 0, nz-1, optionally a velocity inlet at x = 0 and a
  pressure outlet at x = nx-1 (the duct of the inlet/outlet cases), otherwise periodic in x.
* ``random_spheres``: random sequential addition of non-overlapping spheres with radii drawn
  uniformly from [r_min, r_max] until the fluid fraction of the channel interior reaches a target,
  or until a given count is placed (inside an optional centre box), with numpy's PCG64 seeded by
  ``seed``.  Spheres stay clear of the channel walls; in x they wrap periodically unless the
  channel has an inlet and outlet.
* ``rasterize``: cells whose centre lies within a sphere become walls.
* ``porosity_halo``: a partial-bounce-back blend that ramps from ``gamma_max`` at a sphere's surface
  to zero ``width`` cells out (the halo of a porous coating around the solid cluster).
* ``sphere_cluster_duct``: the flow-past-a-cluster case: an n^3 duct with inlet and outlet, 64
  spheres of radius U[10, 30] (seed 5) centred in the middle 40 % of each axis, and their halo.

Triangle surfaces (float64 arrays of shape (count, 3, 3): triangle, vertex a/b/c, x/y/z) feed the
GPU voxeliser (``voxelize_surface`` / ``EsotericPullHipSolver.set_solid_surface``, the reference's
src/lbm/voxelize.cpp):

* ``read_ascii_stl``: the reference's ASCII STL reader (src/io/mesh_io.cpp:18): every ``vertex``
  keyword followed by three numbers, grouped in threes.
* ``box_surface``, ``sphere_surface``, ``torus_surface``, ``prism_surface``: closed, consistently
  oriented meshes of a box, a UV sphere, a torus and a triangle extruded along x.
"""
from __future__ import annotations

import numpy as np

FLUID, INLET, OUTLET, WALL = 0x00, 0x01, 0x02, 0x80


def channel_flags(nx, ny, nz, inlet_outlet=False):
    f = np.zeros((nz, ny, nx), np.uint8)
    f[:, 0, :] = f[:, -1, :] = WALL
    f[0, :, :] = f[-1, :, :] = WALL
    if inlet_outlet:
        f[1:-1, 1:-1, 0] = INLET
        f[1:-1, 1:-1, -1] = OUTLET
    return f


def random_spheres(shape, r_min, r_max, seed, porosity=None, count=None, box=None, periodic_x=True,
                   max_tries=1_000_000):
    """(centres (k, 3) as x, y, z floats; radii (k,)).  Stops at fluid fraction <= porosity of the
    channel interior (cells 1..n-2 in y, z) or at `count` spheres.  `box` = ((x0, x1), (y0, y1),
    (z0, z1)) bounds the centres."""
    nz, ny, nx = shape
    if (porosity is None) == (count is None):
        raise ValueError("give exactly one of porosity, count")
    rng = np.random.default_rng(seed)
    interior = nx * (ny - 2) * (nz - 2)
    solid_target = None if porosity is None else (1.0 - porosity) * interior
    cap = count if count is not None else int(interior / (4.0 / 3.0 * np.pi * r_min ** 3)) + 1
    C = np.empty((cap, 3))
    R = np.empty(cap)
    k = 0
    solid = 0.0
    for _ in range(max_tries):
        if count is not None and k >= count:
            break
        if solid_target is not None and solid >= solid_target:
            break
        r = rng.uniform(r_min, r_max)
        lo = np.array([0.0 if periodic_x else r + 1.0, r + 1.0, r + 1.0])
        hi = np.array([nx if periodic_x else nx - 1.0 - r, ny - 1.0 - r, nz - 1.0 - r])
        if box is not None:
            lo = np.maximum(lo, [b[0] for b in box])
            hi = np.minimum(hi, [b[1] for b in box])
        if np.any(hi <= lo):
            continue
        c = rng.uniform(lo, hi)
        if k:
            d = C[:k] - c
            if periodic_x:
                d[:, 0] -= nx * np.round(d[:, 0] / nx)
            if np.any(np.einsum("ij,ij->i", d, d) < (R[:k] + r) ** 2):
                continue
        C[k], R[k] = c, r
        k += 1
        solid += 4.0 / 3.0 * np.pi * r ** 3
    else:
        raise RuntimeError("random sphere packing did not reach its target")
    return C[:k].copy(), R[:k].copy()


def rasterize(shape, centres, radii, periodic_x=True):
    """Boolean mask of the cells whose centre (x, y, z integer positions) lies inside a sphere."""
    nz, ny, nx = shape
    pad = int(np.ceil(max(radii))) + 1 if (periodic_x and len(radii)) else 0
    ext = np.zeros((nz, ny, nx + 2 * pad), bool)
    for (cx, cy, cz), r in zip(centres, radii):
        x0, x1 = int(np.floor(cx - r)), int(np.ceil(cx + r))
        if not periodic_x:
            x0, x1 = max(x0, 0), min(x1, nx - 1)
        y0, y1 = max(int(np.floor(cy - r)), 0), min(int(np.ceil(cy + r)), ny - 1)
        z0, z1 = max(int(np.floor(cz - r)), 0), min(int(np.ceil(cz + r)), nz - 1)
        zs, ys, xs = np.arange(z0, z1 + 1), np.arange(y0, y1 + 1), np.arange(x0, x1 + 1)
        inside = ((zs - cz)[:, None, None] ** 2 + (ys - cy)[None, :, None] ** 2
                  + (xs - cx)[None, None, :] ** 2) <= r * r
        ext[z0:z1 + 1, y0:y1 + 1, x0 + pad:x1 + pad + 1] |= inside
    mask = ext[:, :, pad:pad + nx].copy()
    if pad:
        mask[:, :, nx - pad:] |= ext[:, :, :pad]  # x < 0 wraps to the end
        mask[:, :, :pad] |= ext[:, :, pad + nx:]  # x >= nx wraps to the start
    return mask


def porosity_halo(shape, centres, radii, width, gamma_max):
    """Blend field: gamma_max * (1 - s / width) at distance s in (0, width] outside the nearest sphere
    surface, 0 elsewhere (and inside the spheres, which are walls)."""
    nz, ny, nx = shape
    g = np.zeros(shape, np.float32)
    for (cx, cy, cz), r in zip(centres, radii):
        R = r + width
        x0, x1 = max(int(np.floor(cx - R)), 0), min(int(np.ceil(cx + R)), nx - 1)
        y0, y1 = max(int(np.floor(cy - R)), 0), min(int(np.ceil(cy + R)), ny - 1)
        z0, z1 = max(int(np.floor(cz - R)), 0), min(int(np.ceil(cz + R)), nz - 1)
        zs, ys, xs = np.arange(z0, z1 + 1), np.arange(y0, y1 + 1), np.arange(x0, x1 + 1)
        dist = np.sqrt((zs - cz)[:, None, None] ** 2 + (ys - cy)[None, :, None] ** 2 + (xs - cx)[None, None, :] ** 2)
        s = dist - r
        val = np.where((s > 0) & (s <= width), gamma_max * (1.0 - s / width), 0.0).astype(np.float32)
        sub = g[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1]
        np.maximum(sub, val, out=sub)
    return g


def sphere_cluster_duct(n, seed=5, count=64, r_min=10, r_max=30, halo_width=4.0, halo_max=0.5):
    """Flat (cells,) arrays of the sphere-cluster duct: uint8 flags (walls, inlet x = 0, outlet
    x = n-1, the spheres as walls), the bool body of the spheres (the drag body) and the float32
    partial-bounce-back halo (zero outside fluid cells)."""
    shape = (n, n, n)
    lo, hi = 0.3 * n, 0.7 * n
    centres, radii = random_spheres(shape, r_min, r_max, seed=seed, count=count, box=((lo, hi),) * 3,
                                    periodic_x=False)
    flags = channel_flags(n, n, n, inlet_outlet=True)
    body = rasterize(shape, centres, radii, periodic_x=False)
    flags[body] = WALL
    halo = porosity_halo(shape, centres, radii, halo_width, halo_max)
    halo[flags != 0] = 0.0
    return flags.reshape(-1), body.reshape(-1), halo.reshape(-1)


def read_ascii_stl(path):
    """Triangles of an ASCII STL file as a (count, 3, 3) float64 array (src/io/mesh_io.cpp:18):
    the three numbers after each ``vertex`` keyword, grouped in threes; a trailing incomplete
    triangle is an error."""
    try:
        text = open(path, "r", encoding="utf-8", errors="replace").read()
    except OSError as e:
        raise RuntimeError("cannot open STL") from e
    words = text.split()
    verts = []
    i = 0
    while i < len(words):
        if words[i] == "vertex":
            if i + 3 >= len(words):
                raise RuntimeError("malformed ASCII STL")
            verts.append([float(w) for w in words[i + 1:i + 4]])
            i += 4
        else:
            i += 1
    if len(verts) % 3:
        raise RuntimeError("malformed ASCII STL")
    return np.asarray(verts, np.float64).reshape(-1, 3, 3)


def box_surface(lo, hi):
    """12 outward-oriented triangles of the axis-aligned box [lo, hi]."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    p = np.array([[x0, y0, z0], [x1, y0, z0], [x0, y1, z0], [x1, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x0, y1, z1], [x1, y1, z1]], np.float64)
    faces = [(0, 3, 1), (0, 2, 3), (4, 5, 7), (4, 7, 6), (0, 1, 5), (0, 5, 4),
             (2, 6, 7), (2, 7, 3), (0, 4, 6), (0, 6, 2), (1, 3, 7), (1, 7, 5)]
    return p[np.array(faces)]


def prism_surface(x0, x1, triangle_yz):
    """8 triangles of the prism x0 <= x <= x1 over the (y, z) triangle ((ay, az), (by, bz), (cy, cz))."""
    (ay, az), (by, bz), (cy, cz) = triangle_yz
    a0, b0, c0 = [x0, ay, az], [x0, by, bz], [x0, cy, cz]
    a1, b1, c1 = [x1, ay, az], [x1, by, bz], [x1, cy, cz]
    return np.array([(a0, c0, b0), (a1, b1, c1), (a0, b0, b1), (a0, b1, a1), (b0, c0, c1), (b0, c1, b1),
                     (c0, a0, a1), (c0, a1, c1)], np.float64)


def sphere_surface(centre, radius, n_theta=16, n_phi=32):
    """Closed UV sphere: n_phi segments around z, n_theta bands from pole to pole."""
    c = np.asarray(centre, np.float64)
    th = np.linspace(0.0, np.pi, n_theta + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, n_phi + 1)[:-1]
    ring = lambda t: np.stack([np.sin(t) * np.cos(ph), np.sin(t) * np.sin(ph), np.full_like(ph, np.cos(t))], -1)
    tris = []
    top, bottom = c + radius * np.array([0.0, 0.0, 1.0]), c + radius * np.array([0.0, 0.0, -1.0])
    for b in range(n_theta):
        r0 = c + radius * ring(th[b])
        r1 = c + radius * ring(th[b + 1])
        for k in range(n_phi):
            k1 = (k + 1) % n_phi
            if b == 0:
                tris.append([top, r1[k], r1[k1]])
            elif b == n_theta - 1:
                tris.append([r0[k], bottom, r0[k1]])
            else:
                tris.append([r0[k], r1[k], r1[k1]])
                tris.append([r0[k], r1[k1], r0[k1]])
    return np.asarray(tris, np.float64)


def torus_surface(centre, major, minor, n_major=32, n_minor=16):
    """Closed torus around the z axis through `centre`."""
    c = np.asarray(centre, np.float64)
    u = np.linspace(0.0, 2.0 * np.pi, n_major + 1)[:-1]
    v = np.linspace(0.0, 2.0 * np.pi, n_minor + 1)[:-1]
    U, V = np.meshgrid(u, v, indexing="ij")
    pts = np.stack([(major + minor * np.cos(V)) * np.cos(U), (major + minor * np.cos(V)) * np.sin(U),
                    minor * np.sin(V)], -1) + c
    tris = []
    for a in range(n_major):
        a1 = (a + 1) % n_major
        for b in range(n_minor):
            b1 = (b + 1) % n_minor
            tris.append([pts[a, b], pts[a1, b], pts[a1, b1]])
            tris.append([pts[a, b], pts[a1, b1], pts[a, b1]])
    return np.asarray(tris, np.float64)
