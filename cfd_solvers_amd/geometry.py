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

The reference builds wall masks from STL voxelisation (src/lbm/voxelize.cpp); these generators
stand in for it with reproducible random packings.
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
