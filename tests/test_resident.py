"""The resident coupling modules (include/cfd/solvers/lbm/resident_multiphysics_sycl.hpp:54-408) as
restated in oracle/, pinned on the CPU.  The reference modules are SYCL-only (not buildable here);
they are pinned by what their mathematics fixes:

* D3Q7 thermal step: one Fourier mode evolves by the 7x7 transfer matrix
  M(k) = ((1 - omega) I + omega E(u)) diag(exp(-i k.c)), E_dj = w_d (1 + 4 c_d.u), evaluated in
  complex128 -- a different algorithm from the real-space streaming;
* Boussinesq: hot fluid accelerates against gravity, linearly in T - T_ref;
* volume of fluid: zero flow leaves the fill unchanged, uniform flow is the first-order upwind
  update and conserves volume, Courant 1 is a shift by one cell; a sphere's interface normals are
  unit vectors pointing inward and its curvature is 2/R; surface tension points inward;
* particles: Stokes relaxation v_n = U (1 - (1 - dt/tau)^n) (+ gravity), periodic wrap, the
  two-way reaction sums to -drag with its centroid at the particle, trilinear interpolation is
  exact for linear fields;
* Q-criterion of a linear field A x equals (|Omega|^2 - |S|^2)/2 of A.
"""
import numpy as np
import pytest

import oracle as O

C7 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
W7 = np.array([0.25] + [0.125] * 6)


def coords(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x, y, z


@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.05, -0.03, 0.02)])
def test_thermal_fourier_mode(u):
    nx, ny, nz, steps, tau = 16, 8, 4, 25, 0.58
    omega = np.float32(1.0) / np.float32(tau)
    r = O.Resident(nx, ny, nz)
    x, y, z = coords(nx, ny, nz)
    k = 2 * np.pi * np.array([2 / nx, 1 / ny, 1 / nz])
    phase = k[0] * x + k[1] * y + k[2] * z
    T0, A = 1.0, 0.1
    g = r.thermal_equilibrium(T0 + A * np.cos(phase))
    vel = np.stack([np.full((nz, ny, nx), c) for c in u])
    T = r.thermal_step(g, vel, omega, steps)
    T = r.thermal_scalar(T)
    uu = np.array(u, float)
    E = W7[:, None] * (1 + 4 * C7 @ uu)[:, None] * np.ones((1, 7))
    M = ((1 - float(omega)) * np.eye(7) + float(omega) * E) @ np.diag(np.exp(-1j * (C7 @ k)))
    lam = np.ones(7) @ np.linalg.matrix_power(M, steps) @ W7
    want = T0 + A * np.real(np.exp(1j * phase) * lam)
    assert np.abs(T - want).max() < 2e-6
    assert abs(lam) < 0.9  # the mode has decayed measurably: the check is not trivial
    if any(uu):
        assert abs(np.angle(lam)) > 1e-3  # and moved with the flow


def test_thermal_conserves_uniform_scalar():
    """The reference's chain test (tests/test_v0156_gpu_lbm_multiphysics.cpp:32-41): 300 K, 3 steps."""
    r = O.Resident(10, 8, 6)
    g = r.thermal_equilibrium(np.full(r.shape, 300.0))
    rng = np.random.default_rng(0)
    vel = rng.uniform(-0.05, 0.05, (3,) + r.shape)
    T = r.thermal_scalar(r.thermal_step(g, vel, 1 / (0.5 + 4 * 0.02), 3))
    assert abs(T.astype(np.float64).sum() - 300.0 * r.cells) / (300.0 * r.cells) < 2e-5


def test_boussinesq_against_gravity():
    r = O.Resident(6, 5, 4)
    rng = np.random.default_rng(1)
    T = 300.0 + rng.uniform(-5, 5, r.shape)
    a0 = rng.normal(size=(3,) + r.shape)
    a = r.boussinesq(T, a0, 300.0, 3e-3, (0.0, -1.0, 0.0))
    d = a - a0.astype(np.float32)
    assert np.allclose(d[0], 0) and np.allclose(d[2], 0)
    hot = T > 300.0
    assert np.all(d[1][hot] > 0) and np.all(d[1][~hot] <= 0)  # buoyancy opposes gravity for hot fluid
    assert np.allclose(d[1], 3e-3 * (T - 300.0), rtol=1e-4, atol=1e-7)
    same = r.boussinesq(np.full(r.shape, 300.0), a0, 300.0, 3e-3, (0.0, -1.0, 0.0))
    assert np.array_equal(same, a0.astype(np.float32))


def test_vof_half_space_and_zero_flow():
    r = O.Resident(10, 8, 6)
    phi = r.half_space(3.5)
    want = np.where(np.arange(10) < 3, 1.0, np.where(np.arange(10) == 3, 0.5, 0.0))
    assert np.array_equal(phi, np.broadcast_to(want, r.shape).astype(np.float32))
    out = r.advect(phi, np.zeros((3,) + r.shape), 0.2)
    assert np.array_equal(out, phi)
    assert abs(out.sum() - phi.sum()) < 2e-4 * r.cells  # the reference's chain-test bound


def test_vof_uniform_flow_is_upwind_and_conservative():
    r = O.Resident(32, 4, 4)
    x = np.arange(32)
    phi = np.broadcast_to(0.5 + 0.3 * np.sin(2 * np.pi * x / 32), r.shape).astype(np.float32)
    U, dt = 0.3, 0.7
    vel = np.stack([np.full(r.shape, U), np.zeros(r.shape), np.zeros(r.shape)])
    out = r.advect(phi, vel, dt)
    p = phi.astype(np.float64)
    want = p - dt * U * (p - np.roll(p, 1, axis=2))
    assert np.abs(out - want).max() < 1e-6
    assert abs(out.astype(np.float64).sum() - p.sum()) < 1e-5 * p.sum()
    shifted = r.advect(phi, np.stack([np.ones(r.shape), np.zeros(r.shape), np.zeros(r.shape)]), 1.0)
    assert np.abs(shifted - np.roll(phi, 1, axis=2)).max() < 1e-6  # Courant 1: one cell per step


def sphere_fill(n, R, width=1.5):
    x, y, z = coords(n, n, n)
    c = (n - 1) / 2
    d = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    return (0.5 * (1 - np.tanh((d - R) / width))).astype(np.float32), d


def test_vof_sphere_normals_curvature_and_tension():
    n, R = 40, 10.0
    r = O.Resident(n, n, n)
    phi, d = sphere_fill(n, R)
    geo = r.geometry(phi)
    band = np.abs(d - R) < 1.0
    norm = np.sqrt(geo[0] ** 2 + geo[1] ** 2 + geo[2] ** 2)
    assert np.abs(norm[band] - 1).max() < 1e-6
    x, y, z = coords(n, n, n)
    c = (n - 1) / 2
    radial = np.stack([x - c, y - c, z - c]) / np.maximum(d, 1e-9)
    inward = -(geo[:3] * radial).sum(axis=0)
    assert inward[band].min() > 0.95
    kappa = geo[3][band]
    assert abs(np.median(kappa) - 2 / R) < 0.1 * (2 / R)
    a = r.surface_tension(phi, np.zeros((3,) + r.shape), 0.072, 1.0)
    along = (a * radial).sum(axis=0)
    assert along[band].max() <= 1e-9 and along[band].min() < 0  # pulls inward (Laplace pressure)
    assert np.abs(a.astype(np.float64).sum(axis=(1, 2, 3))).max() < 1e-3 * np.abs(a).sum()  # no net force


def test_particle_stokes_relaxation_and_wrap():
    nx, ny, nz = 12, 10, 8
    r = O.Resident(nx, ny, nz)
    U, dt, tau, steps = 0.04, 0.05, 0.5, 40
    vel = np.stack([np.full(r.shape, U), np.zeros(r.shape), np.zeros(r.shape)])
    p = np.array([[11.99, 3.5, 2.75, 0.0, 0.0, 0.0, 1.0, tau]])
    xs = [p[0, 0]]
    vx = 0.0
    for i in range(steps):
        p, _ = r.particles_advance(p, vel, np.zeros((3,) + r.shape), dt, False)
        vx = vx + dt * (U - vx) / tau
        xs.append(xs[-1] + dt * vx)
    assert abs(p[0, 3] - U * (1 - (1 - dt / tau) ** steps)) < 1e-6
    assert abs(p[0, 0] - (xs[-1] % nx)) < 1e-4 and p[0, 0] < 1.0  # wrapped through x = nx
    assert p[0, 1] == np.float32(3.5) and p[0, 2] == np.float32(2.75)
    rest, _ = r.particles_advance(np.array([[2.25, 3.5, 2.75, 0, 0, 0, 1.0, 0.5]]), np.zeros((3,) + r.shape),
                                  np.zeros((3,) + r.shape), dt, True)
    assert np.array_equal(rest, np.array([[2.25, 3.5, 2.75, 0, 0, 0, 1.0, 0.5]]))
    g = (0.0, -1.0, 0.0)
    q = np.array([[5.0, 5.0, 4.0, 0, 0, 0, 2.0, tau]])
    for i in range(steps):
        q, _ = r.particles_advance(q, np.zeros((3,) + r.shape), np.zeros((3,) + r.shape), dt, False, g)
    assert abs(q[0, 4] - (-tau * (1 - (1 - dt / tau) ** steps))) < 1e-6  # terminal velocity g tau


def test_particle_two_way_reaction_and_linear_interpolation():
    nx, ny, nz = 12, 10, 8
    r = O.Resident(nx, ny, nz)
    x, y, z = coords(nx, ny, nz)
    vel = np.stack([0.01 + 0.002 * x, -0.003 * y, 0.001 * z + 0.004 * x]).astype(np.float32)
    px, py, pz, m, tau = 4.3, 5.6, 3.2, 2.0, 0.7
    p = np.array([[px, py, pz, 0.01, 0.0, -0.02, m, tau]])
    out, a = r.particles_advance(p, vel, np.zeros((3,) + r.shape), 0.05, True)
    u_at = np.array([0.01 + 0.002 * px, -0.003 * py, 0.001 * pz + 0.004 * px])
    drag = m * (u_at - p[0, 3:6]) / tau
    total = a.astype(np.float64).sum(axis=(1, 2, 3))
    assert np.abs(total + drag).max() < 1e-6
    for comp in range(3):
        if abs(drag[comp]) > 1e-6:
            centre = np.array([(a[comp] * c).sum() / a[comp].sum() for c in (x, y, z)])
            assert np.abs(centre - [px, py, pz]).max() < 1e-4
    assert np.abs(out[0, 3:6] - (p[0, 3:6] + 0.05 * drag / m)).max() < 1e-6


def test_q_criterion_of_linear_fields():
    nx, ny, nz = 12, 11, 10
    r = O.Resident(nx, ny, nz)
    x, y, z = coords(nx, ny, nz)
    rng = np.random.default_rng(3)
    for A in (np.array([[0, -0.02, 0], [0.02, 0, 0], [0, 0, 0]]), rng.normal(scale=0.01, size=(3, 3))):
        vel = np.einsum("ij,jzyx->izyx", A, np.stack([x, y, z]).astype(float)).astype(np.float32)
        q = r.q_criterion(vel)
        S, Om = (A + A.T) / 2, (A - A.T) / 2
        want = 0.5 * ((Om ** 2).sum() - (S ** 2).sum())
        inner = q[1:-1, 1:-1, 1:-1]
        assert np.abs(inner - want).max() < 1e-5 * max(abs(want), 1e-4)
    assert want != 0
