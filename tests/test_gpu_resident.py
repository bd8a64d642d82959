"""Device-resident coupling modules on the GPU (resident_multiphysics_sycl.hpp:54-408): the reference's
chain test (tests/test_v0156_gpu_lbm_multiphysics.cpp:25-71) and parity with the oracle's
restatement on the flow's own device views -- bit for bit for the thermal lattice, the fill
advection, interface geometry, surface tension, one-way particles and the Q-criterion; to float
rounding for the atomically summed two-way particle force."""
import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu


def flow_with_motion(hip, nx=20, ny=16, nz=12, steps=5):
    f = hip.EsotericPullHipSolver("d3q19", nx, ny, nz, 0.72)
    f.initialize_taylor_green(0.03, 0.01, 7)
    f.step(steps)
    return f


def velocity(flow, macro):
    return np.stack([flow.device_download(p).reshape(flow.shape) for p in macro[1:]])


def test_reference_chain(hip):
    nx, ny, nz = 10, 8, 6
    report("resident_chain", shape=(nx, ny, nz))
    flow = hip.EsotericPullHipSolver("d3q19", nx, ny, nz, 0.72)
    flow.initialize_uniform(1.0, 0.0, 0.0, 0.0)
    force = flow.device_local_acceleration(True)
    macro = flow.refresh_device_macroscopic()
    thermal = hip.ResidentThermalD3Q7(flow, 0.02)
    thermal.initialize_uniform(300.0)
    heat0 = thermal.total_scalar()
    thermal.transfer_stats(reset=True)
    thermal.step(macro, 3)
    thermal.add_boussinesq_acceleration(force, 300.0, 3e-3, (0.0, -1.0, 0.0))
    heat1 = thermal.total_scalar()
    assert abs(heat1 - heat0) / heat0 < 2e-5
    assert thermal.transfer_stats()["host_transfer_bytes"] == 0
    surface = hip.ResidentFreeSurface(flow)
    surface.initialize_half_space_x(3.5)
    v0 = surface.volume()
    surface.advect(macro, 0.2)
    assert abs(surface.volume() - v0) < 2e-4 * nx * ny * nz
    surface.add_surface_tension_acceleration(force, 0.072, 1.0)
    particles = hip.ResidentParticles(flow, 1)
    particles.upload([[2.25, 3.5, 2.75, 0, 0, 0, 1.0, 0.5]])
    particles.transfer_stats(reset=True)
    particles.advance(macro, force, 0.05, True)
    assert particles.transfer_stats()["host_transfer_bytes"] == 0
    diagnostics = hip.ResidentFlowDiagnostics(flow)
    assert diagnostics.q_criterion(macro)
    flow.transfer_stats(reset=True)
    flow.step(2)
    flow.wait()
    assert flow.transfer_stats()["host_to_device_bytes"] + flow.transfer_stats()["device_to_host_bytes"] == 0
    for m in (diagnostics, particles, surface, thermal):
        m.close()
    flow.close()


def test_thermal_matches_oracle(hip, oracle_mod):
    flow = flow_with_motion(hip)
    macro = flow.refresh_device_macroscopic()
    u = velocity(flow, macro)
    r = oracle_mod.Resident(flow.nx, flow.ny, flow.nz)
    rng = np.random.default_rng(4)
    T = (300.0 + rng.uniform(-2, 2, flow.shape)).astype(np.float32)
    D = 0.015
    thermal = hip.ResidentThermalD3Q7(flow, D)
    thermal.initialize_fields(T)
    thermal.step(macro, 20)
    got = thermal.scalar()
    tau = np.float32(0.5) + np.float32(4.0) * np.float32(D)
    want = r.thermal_scalar(r.thermal_step(r.thermal_equilibrium(T), u, np.float32(1.0) / tau, 20))
    report("thermal_parity", max_abs=float(np.abs(got - want).max()))
    assert np.array_equal(got, want)
    force = flow.device_local_acceleration(True)
    thermal.add_boussinesq_acceleration(force, 300.0, 2e-3, (0.1, -1.0, 0.3))
    a = np.stack([flow.device_download(p).reshape(flow.shape) for p in force])
    assert np.array_equal(a, r.boussinesq(want, np.zeros((3,) + flow.shape), 300.0, 2e-3, (0.1, -1.0, 0.3)))
    assert thermal.resident_bytes() == 15 * flow.cells * 4
    thermal.close()
    flow.close()


def test_surface_matches_oracle(hip, oracle_mod):
    flow = flow_with_motion(hip, 24, 20, 18)
    macro = flow.refresh_device_macroscopic()
    u = velocity(flow, macro)
    r = oracle_mod.Resident(flow.nx, flow.ny, flow.nz)
    z, y, x = np.meshgrid(*(np.arange(n) for n in flow.shape), indexing="ij")
    d = np.sqrt((x - 11.3) ** 2 + (y - 9.6) ** 2 + (z - 8.2) ** 2)
    phi = (0.5 * (1 - np.tanh((d - 6.0) / 1.5))).astype(np.float32)
    s = hip.ResidentFreeSurface(flow)
    s.upload_fill(phi)
    want = phi
    for _ in range(6):
        s.advect(macro, 0.5)
        want = r.advect(want, u, 0.5)
    assert np.array_equal(s.fill(), want)
    geo = s.refresh_geometry()
    got_geo = np.stack([flow.device_download(p).reshape(flow.shape) for p in geo])
    assert np.array_equal(got_geo, r.geometry(want))
    force = flow.device_local_acceleration(True)
    s.add_surface_tension_acceleration(force, 0.072, 1.3)
    a = np.stack([flow.device_download(p).reshape(flow.shape) for p in force])
    assert np.array_equal(a, r.surface_tension(want, np.zeros((3,) + flow.shape), 0.072, 1.3))
    assert s.volume() == pytest.approx(float(want.astype(np.float64).sum()), rel=1e-12)
    s.close()
    flow.close()


def test_particles_match_oracle(hip, oracle_mod):
    flow = flow_with_motion(hip)
    macro = flow.refresh_device_macroscopic()
    u = velocity(flow, macro)
    r = oracle_mod.Resident(flow.nx, flow.ny, flow.nz)
    rng = np.random.default_rng(5)
    n = 300
    p = np.column_stack([rng.uniform(-3, flow.nx + 3, n), rng.uniform(0, flow.ny, n), rng.uniform(0, flow.nz, n),
                         rng.normal(scale=0.01, size=(n, 3)), rng.uniform(0.5, 2, n), rng.uniform(0.2, 1, n)])
    parts = hip.ResidentParticles(flow, n)
    parts.upload(p)
    force = flow.device_local_acceleration(True)
    want = np.column_stack([p[:, :6].astype(np.float32), p[:, 6:].astype(np.float32)]).astype(np.float64)
    zero = np.zeros((3,) + flow.shape)
    for _ in range(10):
        parts.advance(macro, force, 0.05, False, (0.0, -0.01, 0.0))
        want, _ = r.particles_advance(want, u, zero, 0.05, False, (0.0, -0.01, 0.0))
    assert np.array_equal(parts.download(), want)
    assert not np.any(np.stack([flow.device_download(q) for q in force]))  # one-way: no reaction
    parts.advance(macro, force, 0.05, True)
    got_a = np.stack([flow.device_download(q).reshape(flow.shape) for q in force])
    want_p, want_a = r.particles_advance(want, u, zero, 0.05, True)
    assert np.array_equal(parts.download(), want_p)
    scale = np.abs(want_a).max()
    report("particles_two_way", max_rel=float(np.abs(got_a - want_a).max() / scale))
    assert np.abs(got_a - want_a).max() <= 1e-5 * scale  # atomics sum in another order
    parts.close()
    one = hip.ResidentParticles(flow, 1)
    one.upload([[5.3, 7.1, 4.4, 0.0, 0.01, 0.0, 1.0, 0.4]])
    flow.device_local_acceleration(True)
    one.advance(macro, force, 0.05, True)
    got_a = np.stack([flow.device_download(q).reshape(flow.shape) for q in force])
    _, want_a = r.particles_advance(np.array([[5.3, 7.1, 4.4, 0.0, 0.01, 0.0, 1.0, 0.4]]), u, zero, 0.05, True)
    assert np.array_equal(got_a, want_a)
    one.close()
    flow.close()


def test_q_criterion_matches_oracle(hip, oracle_mod):
    flow = flow_with_motion(hip)
    macro = flow.refresh_device_macroscopic()
    u = velocity(flow, macro)
    d = hip.ResidentFlowDiagnostics(flow)
    q = flow.device_download(d.q_criterion(macro)).reshape(flow.shape)
    want = oracle_mod.Resident(flow.nx, flow.ny, flow.nz).q_criterion(u)
    assert np.array_equal(q, want)
    assert np.abs(want).max() > 0
    d.close()
    flow.close()


@pytest.mark.parametrize("shape", [(300, 37, 11), (65, 5, 3), (3, 129, 4), (64, 128, 6), (256, 128, 71),
                                   (200, 100, 130)])
def test_stencils_on_ragged_shapes(hip, oracle_mod, shape):
    """Every stencil kernel bit for bit against the oracle on extents that leave partial blocks
    and tiles in x and y.  The 3-D launch (normals, curvature, tension, thermal step) takes 64-256
    wide blocks of several rows per thread; 64 x 128 x 6 holds a multiple of 8 of them per plane
    (the per-XCD remap).  The z-marching tiles (fill advection, Q-criterion; 64 x 4 cells) march
    several planes per block in the last two shapes: 256 x 128 x 71 in 15 chunks of 5 planes (the
    last of 1), 1920 blocks, a multiple of 8 (the per-XCD block placement); 200 x 100 x 130 in 19
    chunks of 7 planes (the last of 4), 1900 blocks."""
    nx, ny, nz = shape
    report("resident_ragged", shape=shape)
    flow = hip.EsotericPullHipSolver("d3q19", nx, ny, nz, 0.7)
    flow.initialize_taylor_green(0.04, 0.01, 3)
    flow.step(3)
    macro = flow.refresh_device_macroscopic()
    u = velocity(flow, macro)
    r = oracle_mod.Resident(nx, ny, nz)
    rng = np.random.default_rng(nx)
    T = (300.0 + rng.uniform(-2, 2, flow.shape)).astype(np.float32)
    thermal = hip.ResidentThermalD3Q7(flow, 0.02)
    thermal.initialize_fields(T)
    thermal.step(macro, 3)
    tau = np.float32(0.5) + np.float32(4.0) * np.float32(0.02)
    want_t = r.thermal_scalar(r.thermal_step(r.thermal_equilibrium(T), u, np.float32(1.0) / tau, 3))
    assert np.array_equal(thermal.scalar(), want_t)
    phi = rng.uniform(0, 1, flow.shape).astype(np.float32)
    s = hip.ResidentFreeSurface(flow)
    s.upload_fill(phi)
    s.advect(macro, 0.5)
    want = r.advect(phi, u, 0.5)
    assert np.array_equal(s.fill(), want)
    geo = s.refresh_geometry()
    assert np.array_equal(np.stack([flow.device_download(p).reshape(flow.shape) for p in geo]), r.geometry(want))
    force = flow.device_local_acceleration(True)
    s.add_surface_tension_acceleration(force, 0.05, 1.1)
    a = np.stack([flow.device_download(p).reshape(flow.shape) for p in force])
    assert np.array_equal(a, r.surface_tension(want, np.zeros((3,) + flow.shape), 0.05, 1.1))
    d = hip.ResidentFlowDiagnostics(flow)
    assert np.array_equal(flow.device_download(d.q_criterion(macro)).reshape(flow.shape), r.q_criterion(u))
    for m in (d, s, thermal):
        m.close()
    flow.close()


def test_device_acceleration_drives_the_flow(hip):
    """Buoyancy written into the device acceleration planes steers the next steps exactly as the same
    field set from the host does."""
    a = flow_with_motion(hip)
    b = flow_with_motion(hip)
    macro = a.refresh_device_macroscopic()
    thermal = hip.ResidentThermalD3Q7(a, 0.02)
    z, y, x = np.meshgrid(*(np.arange(n) for n in a.shape), indexing="ij")
    thermal.initialize_fields((300.0 + 3.0 * np.sin(2 * np.pi * x / a.nx)).astype(np.float32))
    thermal.step(macro, 4)
    force = a.device_local_acceleration(True)
    thermal.add_boussinesq_acceleration(force, 300.0, 1e-4, (0.0, -1.0, 0.0))
    host_a = [a.device_download(p).reshape(a.shape) for p in force]
    assert np.abs(host_a[1]).max() > 0
    b.set_local_body_acceleration(*host_a)
    a.step(10)
    b.step(10)
    ma, mb = a.download_macroscopic(), b.download_macroscopic()
    for f in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(getattr(ma, f), getattr(mb, f)), f
    thermal.close()
    a.close()
    b.close()


def test_flow_close_closes_its_modules(hip):
    """The modules run on the flow solver's stream: closing the flow first closes them, and a closed
    module reports a null handle instead of touching the destroyed stream."""
    flow = hip.EsotericPullHipSolver("d3q19", 8, 8, 8, 0.7)
    thermal = hip.ResidentThermalD3Q7(flow, 0.01)
    surface = hip.ResidentFreeSurface(flow)
    flow.close()
    assert thermal._h is None and surface._h is None
    with pytest.raises(RuntimeError, match="null module handle"):
        thermal.total_scalar()
    surface.close()  # closing again is harmless
    with pytest.raises(RuntimeError, match="null solver handle"):
        hip.ResidentParticles(flow, 1)


def test_module_errors(hip):
    f64 = hip.EsotericPullHipSolver("d3q19", 8, 8, 8, 0.7, storage="fp64")
    with pytest.raises(RuntimeError, match="FP32 device views"):
        hip.ResidentThermalD3Q7(f64, 0.01)
    f64.close()
    flow = hip.EsotericPullHipSolver("d3q19", 8, 8, 8, 0.7)
    with pytest.raises(RuntimeError, match="invalid resident thermal lattice"):
        hip.ResidentThermalD3Q7(flow, 0.0)
    s = hip.ResidentFreeSurface(flow)
    with pytest.raises(RuntimeError, match=r"\[0,1\]"):
        s.initialize_uniform(1.5)
    p = hip.ResidentParticles(flow, 2)
    with pytest.raises(RuntimeError, match="count mismatch"):
        p.upload([[1, 1, 1, 0, 0, 0, 1, 1]])
    with pytest.raises(RuntimeError, match="positive"):
        p.upload([[1, 1, 1, 0, 0, 0, 1, 1], [1, 1, 1, 0, 0, 0, 0, 1]])
    s.close()
    p.close()
    flow.close()
