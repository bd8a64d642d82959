"""Writes tests/golden/*.npz from the reference's own LBM code (oracle/_ref, built by oracle/Makefile
from the reference sources).  Run in a container that has the reference tree:

    make -C oracle all && python tests/golden/make_golden.py

Each fixture holds inputs (the configuration, as scalars) and the reference's outputs.  The
configurations are the ones the reference's own regression tests use (tests/test_main.cpp:312-499)
plus a few larger ones for the parity tests.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def dims(q, nx, ny, nz):
    return (nx, ny, 1 if q == 9 else nz)


def index_maps():
    """Load slots of EsotericPullSolver::load_cell at both parities (esoteric_pull.hpp:349-365).

    Slot identifiers d*N + n are planted as float values (exact below 2^24) in the raw storage;
    logical_populations() then returns, for every (d, n), the identifier of the slot it read."""
    out = {}
    for q in (9, 19, 27):
        nx, ny, nz = dims(q, 5, 4, 3)
        r = O.RefEsoteric(q, nx, ny, nz, 0.7)
        ids = np.arange(q * r.cells, dtype=np.float32)
        r.set_raw(ids)
        out[f"q{q}_parity0"] = r.logical().astype(np.int64)
        r.init_uniform()
        r.step(1)  # odd parity
        r.set_raw(ids)
        out[f"q{q}_parity1"] = r.logical().astype(np.int64)
        out[f"q{q}_shape"] = np.array([nx, ny, nz])
    np.savez_compressed(OUT / "index_maps.npz", **out)


def parity_runs():
    """tests/test_main.cpp:341-372 (TG 0.017, tau 0.73, 7 and 8 steps) and :404-434 (Guo forcing,
    60 steps) on 18x14x10 (nz = 1 for D2Q9): the reference's esoteric and two-grid outputs."""
    out = {}
    for q in (9, 19, 27):
        nx, ny, nz = dims(q, 18, 14, 10)
        out[f"q{q}_shape"] = np.array([nx, ny, nz])
        r = O.RefEsoteric(q, nx, ny, nz, 0.73)
        r.init_taylor_green(0.017)
        out[f"q{q}_init_logical"] = r.logical()
        for steps in (7, 8):
            r.init_taylor_green(0.017)
            r.step(steps)
            out[f"q{q}_tg{steps}"] = np.stack(r.macroscopic())
            out[f"q{q}_tg{steps}_scalars"] = np.array([r.mass(), r.kinetic_energy(), r.max_speed()])
            p = O.RefPull(q, nx, ny, nz, 0.73)
            p.init_taylor_green(0.017)
            p.step(steps)
            out[f"q{q}_tg{steps}_twogrid"] = np.stack(p.macroscopic())
        accel = (2.0e-6, -1.0e-6, 0.5e-6 if q != 9 else 0.0)
        r = O.RefEsoteric(q, nx, ny, nz, 0.73, accel)
        r.init_uniform()
        r.step(60)
        out[f"q{q}_forced60"] = np.stack(r.macroscopic())
        out[f"q{q}_forced60_scalars"] = np.array([r.mass(), r.kinetic_energy(), r.max_speed()])
        out[f"q{q}_forced_accel"] = np.array(accel)
    np.savez_compressed(OUT / "parity_runs.npz", **out)


def taylor_green_32():
    """D3Q19 32^3, tau 0.6, Taylor-Green amplitude 0.05: rho, u after 110 steps (10 + 100)."""
    r = O.RefEsoteric(19, 32, 32, 32, 0.6)
    r.init_taylor_green(0.05)
    m0 = r.mass()
    r.step(110)
    np.savez_compressed(OUT / "tg_d3q19_32.npz", fields=np.stack(r.macroscopic()),
                        scalars=np.array([m0, r.mass(), r.kinetic_energy(), r.max_speed()]))


def bounded_defect():
    """The reference's bounded path with an all-fluid mask vs its periodic path after one step
    (D3Q19 16x12x10, TG 0.02, tau 0.7).  The odd-parity repack at esoteric_pull.hpp:483 reads the
    wrong cell, so the two differ; the restatement fixes that line and makes them identical."""
    a = O.RefEsoteric(19, 16, 12, 10, 0.7)
    b = O.RefEsoteric(19, 16, 12, 10, 0.7)
    for s in (a, b):
        s.init_taylor_green(0.02)
    b.set_mask(np.zeros(a.cells, dtype=np.uint8))
    a.step(1)
    b.step(1)
    pa, pb = a.macroscopic(), b.macroscopic()
    du = max(float(np.abs(x - y).max()) for x, y in zip(pa[1:], pb[1:]))
    umax = float(np.max(np.abs(np.stack(pa[1:]))))
    np.savez_compressed(OUT / "bounded_defect.npz", periodic=np.stack(pa), masked_zero=np.stack(pb),
                        max_du=np.array(du), max_u=np.array(umax))


def d2q9_walls():
    """The reference's two-grid D2Q9Solver on its own wall tests: Poiseuille channel
    (tests/test_main.cpp:449-477) and lid-driven cavity (:479-499)."""
    out = {}
    s = O.RefD2Q9(24, 14, 0.8, 1.0e-5, 0.0)
    for x in range(24):
        s.set_solid(x, 0)
        s.set_solid(x, 13)
    s.init_uniform()
    s.step(5000)
    out["poiseuille"] = np.stack(s.fields())
    c = O.RefD2Q9(48, 48, 0.8)
    for x in range(48):
        c.set_solid(x, 0)
        c.set_wall_velocity(x, 47, 0.05, 0.0)
    for y in range(48):
        c.set_solid(0, y)
        c.set_solid(47, y)
    c.init_uniform()
    c.step(6000)
    out["cavity"] = np.stack(c.fields())
    np.savez_compressed(OUT / "d2q9_walls.npz", **out)


def precision_fp64():
    """PrecisionPullSolver<D, double> (two-grid FP64): D2Q9 Taylor-Green A 0.001, tau 0.8, 20 steps
    on 16^2, 32^2, 64^2 (tests/test_main.cpp:526-565) and D3Q19 16^3 A 0.02, tau 0.7, 20 steps."""
    out = {}
    for n in (16, 32, 64):
        p = O.RefPrecision(9, n, n, 1, 0.8)
        p.init_taylor_green(0.001)
        p.step(20)
        out[f"d2q9_{n}"] = np.stack(p.macroscopic())
    p = O.RefPrecision(19, 16, 16, 16, 0.7)
    p.init_taylor_green(0.02)
    p.step(20)
    out["d3q19_16"] = np.stack(p.macroscopic())
    np.savez_compressed(OUT / "precision_fp64.npz", **out)


def local_acceleration():
    """set_local_body_acceleration (esoteric_pull.hpp:144-151; applied in the collision at :414-420
    and in compute_macroscopic at :180-209) with a seeded random field on top of a uniform
    acceleration, Taylor-Green 0.03, tau 0.7, 9 steps (both parities): the inputs, the fields at
    step 0 and step 9 and the raw storage at step 9."""
    out = {}
    for q, n in ((9, (16, 12, 1)), (19, (10, 8, 6)), (27, (8, 6, 5))):
        accel = (1.0e-6, -0.5e-6, 2.0e-6 if q != 9 else 0.0)
        rng = np.random.default_rng(14 + q)
        cells = n[0] * n[1] * n[2]
        local = rng.uniform(-4.0e-6, 4.0e-6, (3, cells)).astype(np.float32)
        if q == 9:
            local[2] = 0.0
        r = O.RefEsoteric(q, *n, 0.7, accel)
        r.init_taylor_green(0.03)
        r.set_local_acceleration(*local)
        out[f"q{q}_shape"] = np.array(n)
        out[f"q{q}_accel"] = np.array(accel)
        out[f"q{q}_local"] = local
        out[f"q{q}_fields0"] = np.stack(r.macroscopic())
        r.step(9)
        out[f"q{q}_fields9"] = np.stack(r.macroscopic())
        out[f"q{q}_raw9"] = r.raw()
    np.savez_compressed(OUT / "local_acceleration.npz", **out)


FIXTURES = {"index_maps": index_maps, "parity_runs": parity_runs, "taylor_green_32": taylor_green_32,
            "bounded_defect": bounded_defect, "d2q9_walls": d2q9_walls, "precision_fp64": precision_fp64,
            "local_acceleration": local_acceleration}

if __name__ == "__main__":
    # python tests/golden/make_golden.py [fixture ...]   (default: all of them)
    if not O.ref_available():
        sys.exit("oracle/_ref/libref_lbm.so is not built (make -C oracle all, with the reference tree present)")
    for name in sys.argv[1:] or list(FIXTURES):
        FIXTURES[name]()
    for f in sorted(OUT.glob("*.npz")):
        print(f.name, f.stat().st_size)
