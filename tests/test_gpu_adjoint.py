"""The HIP discrete adjoint (in-place, slot space, two-level checkpointing) against
  * the oracle's textbook adjoint in logical space (reverse streaming + transposed collision),
  * itself under different checkpoint intervals (bitwise), and
  * central finite differences of the FP64 forward at the full config-4 size.

Config 4 (D3Q19 128^3 duct, velocity inlet peak 0.05, pressure outlet, walls in y/z, per-cell
partial bounce-back design field gamma ~ U[0.3, 0.7] i.i.d. (seed 4), T = 2000, J = mean pressure
drop over the last 200 states, snapshots every 100 steps).  The design field enters the collision
as the blend fraction n_s = N_MAX * gamma: the raw fraction gamma in every cell makes the duct
impermeable (the inlet density grows without bound), see DESIGN.md.
"""
import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu

N_MAX = 0.01


def duct(nx, ny, nz):
    f = np.zeros((nz, ny, nx), np.uint8)
    f[:, 0, :] = f[:, -1, :] = 0x80
    f[0, :, :] = f[-1, :, :] = 0x80
    f[1:-1, 1:-1, 0] = 0x01
    f[1:-1, 1:-1, -1] = 0x02
    return f.reshape(-1)


def design(flags, seed=4):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.3, 0.7, flags.size)
    return np.where(flags == 0, g, 0.0)


def gpu_duct(hip, n, blend, u_peak=0.05, storage="fp64"):
    nx, ny, nz = n
    s = hip.EsotericPullHipSolver(19, nx, ny, nz, 0.6, storage=storage)
    flags = duct(nx, ny, nz)
    s.set_flags(flags)
    s.set_inlet_outlet(u_peak, 1.0)
    s.set_porosity(blend.astype(np.float32))
    s.initialize_uniform()
    return s


def oracle_duct(oracle_mod, n, blend, u_peak=0.05, storage="fp64"):
    nx, ny, nz = n
    o = oracle_mod.Oracle(19, nx, ny, nz, float(np.float32(0.6)), storage=storage)
    o.set_flags(duct(nx, ny, nz))
    o.set_inlet_outlet(float(np.float32(u_peak)), 1.0)
    o.set_gamma(blend.astype(np.float32).astype(np.float64))
    o.init_uniform()
    return o


@pytest.mark.parametrize("storage,tol", [("fp64", 1e-9), ("fp32", 2e-4)])
def test_adjoint_matches_logical_space_oracle(hip, oracle_mod, storage, tol):
    report("adjoint_vs_oracle", storage=storage)
    n = (16, 8, 8)
    flags = duct(*n)
    blend = 0.05 * design(flags)  # a stronger medium than config 4: larger, clearer gradients
    T, W, every = 40, 8, 7
    g = gpu_duct(hip, n, blend, storage=storage)
    o = oracle_duct(oracle_mod, n, blend, storage="fp64")
    J, grad, du = g.adjoint_run(T, W, every, 1, n[0] - 2)
    Jo, grado, duo = o.adjoint_run(T, W, 1, n[0] - 2)
    assert abs(J - Jo) <= tol * abs(Jo)
    scale = np.abs(grado).max()
    assert np.abs(grad - grado).max() <= tol * scale, np.abs(grad - grado).max() / scale
    assert abs(du - duo) <= tol * abs(duo)
    assert g.time_step == T
    assert np.all(grad.reshape(-1)[flags != 0] == 0.0)


def test_adjoint_periodic_porous_matches_oracle(hip, oracle_mod):
    """No flags: periodic lattice, uniform-random PBB, Taylor-Green start, objective planes x=2, 9."""
    report("adjoint_periodic")
    n = (12, 10, 8)
    rng = np.random.default_rng(8)
    blend = rng.uniform(0.0, 0.05, int(np.prod(n)))
    g = hip.EsotericPullHipSolver(19, *n, 0.7, storage="fp64")
    g.set_porosity(blend.astype(np.float32))
    g.initialize_taylor_green(0.03, 1e-3, 2)
    o = oracle_mod.Oracle(19, *n, float(np.float32(0.7)), storage="fp64")
    o.set_gamma(blend.astype(np.float32).astype(np.float64))
    o.init_taylor_green(float(np.float32(0.03)), float(np.float32(1e-3)), 2)
    J, grad, _ = g.adjoint_run(25, 5, 4, 2, 9)
    Jo, grado, _ = o.adjoint_run(25, 5, 2, 9)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert np.abs(grad - grado).max() <= 1e-9 * np.abs(grado).max()


def test_adjoint_checkpoint_interval_is_exact(hip):
    """Recomputed states are bit-identical to the original ones, so the gradient does not depend on
    the snapshot interval (including no recomputation at all)."""
    report("adjoint_intervals")
    n = (16, 8, 8)
    blend = 0.05 * design(duct(*n))
    results = []
    for every in (1, 6, 40):
        g = gpu_duct(hip, n, blend, storage="fp32")
        results.append(g.adjoint_run(40, 8, every, 1, n[0] - 2))
    for J, grad, du in results[1:]:
        assert J == results[0][0] and du == results[0][2]
        assert np.array_equal(grad, results[0][1])


def forward_objective(hip, n, blend, u_peak, T, W, storage="fp64"):
    s = gpu_duct(hip, n, blend, u_peak, storage)
    s.step(T - W)
    J = 0.0
    for _ in range(W):
        s.step(1)
        J += s.objective_pressure_drop(1, n[0] - 2)
    s.close()
    return J / W


def test_config4_gradient_vs_finite_differences(hip):
    """128^3, T = 2000, window 200, snapshots every 100: the FP64 adjoint gradient matches central
    finite differences (design step h = 1e-3) on 16 seeded cells and u_peak to 1e-4; the FP32
    adjoint matches the FP64 one."""
    report("config4_fd", n=128)
    n = (128, 128, 128)
    T, W, every = 2000, 200, 100
    flags = duct(*n)
    gamma = design(flags)
    blend = (N_MAX * gamma).astype(np.float32)
    s64 = gpu_duct(hip, n, blend)
    J64, g64, du64 = s64.adjoint_run(T, W, every, 1, n[0] - 2)
    s64.close()
    s32 = gpu_duct(hip, n, blend, storage="fp32")
    J32, g32, du32 = s32.adjoint_run(T, W, every, 1, n[0] - 2)
    fwd_ms, rev_ms = s32.adjoint_timing()
    s32.close()
    report("config4_adjoint", J64=J64, J32=J32, du64=du64, du32=du32, fwd_ms=fwd_ms, rev_ms=rev_ms)
    assert abs(J32 - J64) <= 1e-4 * abs(J64)
    fluid = np.flatnonzero(flags == 0)
    cells = np.random.default_rng(44).choice(fluid, 16, replace=False)
    h = 1e-3
    worst = 0.0
    for c in cells:
        plus, minus = blend.copy(), blend.copy()
        plus[c] = np.float32(N_MAX * (gamma[c] + h))
        minus[c] = np.float32(N_MAX * (gamma[c] - h))
        dj = forward_objective(hip, n, plus, 0.05, T, W) - forward_objective(hip, n, minus, 0.05, T, W)
        fd_blend = dj / (float(plus[c]) - float(minus[c]))  # dJ/dn_s with the float-rounded step
        adj = g64.reshape(-1)[c]
        rel = abs(adj - fd_blend) / abs(fd_blend)
        worst = max(worst, rel)
        report("config4_fd_cell", cell=int(c), adjoint=adj, fd=fd_blend, rel=rel,
               fp32=g32.reshape(-1)[c])
        assert rel <= 1e-4
        assert abs(g32.reshape(-1)[c] - adj) <= 1e-3 * abs(adj)
    up, um = np.float32(0.05 + h), np.float32(0.05 - h)
    fd_u = (forward_objective(hip, n, blend, float(up), T, W)
            - forward_objective(hip, n, blend, float(um), T, W)) / (float(up) - float(um))
    report("config4_fd_upeak", adjoint=du64, fd=fd_u, fp32=du32, worst_cell_rel=worst)
    assert abs(du64 - fd_u) <= 1e-4 * abs(fd_u)
    assert abs(du32 - du64) <= 1e-3 * abs(du64)
