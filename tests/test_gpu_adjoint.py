"""The HIP discrete adjoint (in-place, slot space, two-level checkpointing) against
  * the oracle's textbook adjoint in logical space (reverse streaming + transposed collision),
  * itself under different checkpoint intervals (bitwise), and
  * central finite differences of the FP64 forward at the full config-4 size.

Config 4 (D3Q19 128^3 duct, velocity inlet peak 0.05, pressure outlet, walls in y/z, per-cell
partial bounce-back design field gamma ~ U[0.3, 0.7] i.i.d. (seed 4), T = 2000, J = mean pressure
drop over the last 200 states, snapshots every 100 steps).  The design field enters the collision
as the blend fraction n_s = N_MAX * gamma with N_MAX = 0.01: the raw fraction gamma in every cell
makes the duct impermeable (at 48^3 the inlet density grows past 6000 within 1000 steps).
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


def test_fp64x_adjoint_matches_exact_weight_oracle(hip, oracle_mod):
    """FP64X (the lattice weights' exact values) against the oracle's FP64 textbook adjoint with the
    same exact weights, to 1e-9; the reference-weight FP64 gradient differs from it by more."""
    report("adjoint_vs_oracle", storage="fp64x")
    n = (16, 8, 8)
    flags = duct(*n)
    blend = 0.05 * design(flags)
    T, W, every = 40, 8, 7
    g = gpu_duct(hip, n, blend, storage="fp64x")
    J, grad, du = g.adjoint_run(T, W, every, 1, n[0] - 2)
    with oracle_mod.exact_fp64_weights():
        o = oracle_duct(oracle_mod, n, blend, storage="fp64")
        Jo, grado, duo = o.adjoint_run(T, W, 1, n[0] - 2)
    assert abs(J - Jo) <= 1e-9 * abs(Jo)
    scale = np.abs(grado).max()
    assert np.abs(grad - grado).max() <= 1e-9 * scale
    assert abs(du - duo) <= 1e-9 * abs(duo)
    g64 = gpu_duct(hip, n, blend, storage="fp64")
    J64, _, _ = g64.adjoint_run(T, W, every, 1, n[0] - 2)
    report("adjoint_fp64x", J_rel_to_fp64=abs(J64 - J) / abs(J))
    assert J64 != J


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


@pytest.mark.parametrize("n,storage,tol", [((67, 6, 5), "fp64", 1e-9), ((130, 5, 4), "fp32", 2e-4),
                                           ((3 + 64, 9, 3), "fp32", 2e-4)])
def test_adjoint_on_ragged_ducts(hip, oracle_mod, n, storage, tol):
    """The adjoint kernels on extents that leave partial tiles and 64-cell segments (one fluid row
    in z for the last case), against the FP64 logical-space oracle."""
    report("adjoint_ragged", shape=n, storage=storage)
    flags = duct(*n)
    blend = 0.05 * design(flags, seed=n[0])
    T, W, every = 21, 6, 5
    g = gpu_duct(hip, n, blend, storage=storage)
    o = oracle_duct(oracle_mod, n, blend, storage="fp64")
    J, grad, du = g.adjoint_run(T, W, every, 1, n[0] - 2)
    Jo, grado, duo = o.adjoint_run(T, W, 1, n[0] - 2)
    assert abs(J - Jo) <= tol * abs(Jo)
    assert np.abs(grad - grado).max() <= tol * np.abs(grado).max()
    assert abs(du - duo) <= tol * abs(duo)
    assert np.abs(grado).max() > 0 and np.all(grad.reshape(-1)[flags != 0] == 0.0)


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


def test_progress_lines_leave_the_gradient_unchanged(hip, capfd, monkeypatch):
    """CFD_LBM_PROGRESS_STEPS=k makes an adjoint run print a line on stderr after every k forward
    sweep steps and every k transposed steps, each after the stream has finished them.  The run's
    objective and gradients are the same bits as without the lines."""
    report("adjoint_progress")
    n = (16, 8, 8)
    blend = 0.05 * design(duct(*n))
    monkeypatch.delenv("CFD_LBM_PROGRESS_STEPS", raising=False)
    quiet = gpu_duct(hip, n, blend, storage="fp32").adjoint_run_binomial(12, 4, 3)
    capfd.readouterr()
    monkeypatch.setenv("CFD_LBM_PROGRESS_STEPS", "5")
    loud = gpu_duct(hip, n, blend, storage="fp32").adjoint_run_binomial(12, 4, 3)
    err = capfd.readouterr().err
    monkeypatch.delenv("CFD_LBM_PROGRESS_STEPS")
    lines = [l for l in err.splitlines() if l.startswith("[adjoint]")]
    assert [l.split(":")[0] for l in lines] == ["[adjoint] forward sweep"] * 2 + ["[adjoint] reverse"] * 2
    assert lines[0].startswith("[adjoint] forward sweep: 5 of 12 steps")
    assert lines[-1].startswith("[adjoint] reverse: 10 of 12 steps")
    assert quiet[0] == loud[0] and quiet[2] == loud[2] and quiet[3] == loud[3]
    assert np.array_equal(quiet[1], loud[1])


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


# ---------------------------------------------------------------- binomial checkpointing, drag, FP16
def cube_duct(n=(16, 10, 10)):
    nx, ny, nz = n
    f = duct(nx, ny, nz).reshape(nz, ny, nx)
    body = np.zeros_like(f)
    body[4:6, 4:6, 6:8] = 1
    f[body != 0] = 0x80
    return f.reshape(-1), body.reshape(-1)


def gpu_cube(hip, blend, storage, n=(16, 10, 10), u_peak=0.05):
    flags, body = cube_duct(n)
    s = hip.EsotericPullHipSolver(19, *n, 0.6, storage=storage)
    s.set_flags(flags)
    s.set_drag_body(body, 0)
    s.set_inlet_outlet(u_peak, 1.0)
    s.set_porosity(blend.astype(np.float32))
    s.initialize_uniform()
    return s


def oracle_cube(oracle_mod, blend, storage, n=(16, 10, 10), u_peak=0.05):
    flags, body = cube_duct(n)
    o = oracle_mod.Oracle(19, *n, float(np.float32(0.6)), storage=storage)
    o.set_flags(flags)
    o.set_drag_body(body, 0)
    o.set_inlet_outlet(float(np.float32(u_peak)), 1.0)
    o.set_gamma(blend.astype(np.float32).astype(np.float64))
    o.init_uniform()
    return o


def test_binomial_matches_two_level_bitwise(hip):
    """Binomial plans with 1, 2, 3 and 'as many as fit' buffers run the same transposed steps on
    bit-identical recomputed states: J and both gradients equal the two-level result bit for bit, and
    the forward-step count is the plan's optimum."""
    report("binomial_vs_two_level")
    n = (16, 8, 8)
    blend = 0.05 * design(duct(*n))
    T, W = 40, 8
    ref = gpu_duct(hip, n, blend, storage="fp32").adjoint_run(T, W, 1, 1, n[0] - 2)
    for snaps in (1, 2, 3, 0):
        g = gpu_duct(hip, n, blend, storage="fp32")
        J, grad, du, fw = g.adjoint_run_binomial(T, W, snaps, "pressure_drop", 1, n[0] - 2)
        assert J == ref[0] and du == ref[2]
        assert np.array_equal(grad, ref[1])
        expected = hip.checkpoint_plan(T, snaps if snaps else T - 1)[0]
        assert fw == expected, (snaps, fw, expected)
        assert g.time_step == T


@pytest.mark.parametrize("storage", ["fp32", "fp16s"])
def test_objective_run_equals_adjoint_objective(hip, storage):
    """cfd_lbm_objective_run (steps + objective, no adjoint) returns the adjoint runs' J bit for bit,
    for the pressure drop and the drag, and leaves the same state."""
    report("objective_run", storage=storage)
    n = (16, 8, 8)
    blend = 0.05 * design(duct(*n))
    T, W = 40, 8
    a = gpu_duct(hip, n, blend, storage=storage)
    J, _, _, _ = a.adjoint_run_binomial(T, W, 3, "pressure_drop", 1, n[0] - 2, want_gamma=False)
    b = gpu_duct(hip, n, blend, storage=storage)
    assert b.objective_run(T, W, "pressure_drop", 1, n[0] - 2) == J
    assert b.time_step == a.time_step == T
    ma, mb = a.download_macroscopic(), b.download_macroscopic()
    assert all(np.array_equal(getattr(ma, k), getattr(mb, k)) for k in ("rho", "ux", "uy", "uz"))
    flags, body = cube_duct()
    cb = 0.05 * design(flags)
    c = gpu_cube(hip, cb, storage)
    Jd, _, _, _ = c.adjoint_run_binomial(30, 6, 2, "drag", want_gamma=False)
    d = gpu_cube(hip, cb, storage)
    assert d.objective_run(30, 6, "drag") == Jd
    with pytest.raises(RuntimeError, match="window"):
        d.objective_run(5, 6, "drag")


def test_drag_objective_matches_oracle(hip, oracle_mod):
    report("drag_objective")
    flags, _ = cube_duct()
    blend = 0.05 * design(flags)
    g = gpu_cube(hip, blend, "fp32")
    o = oracle_cube(oracle_mod, blend, "fp32")
    for steps in (0, 1, 7):
        g.step(steps)
        o.step(steps)
        jg, jo = g.objective_drag(), o.objective(kind="drag")
        assert jg == pytest.approx(jo, rel=1e-12, abs=1e-15)
    assert g.objective_drag() > 0.0


@pytest.mark.parametrize("storage,tol", [("fp64", 1e-9), ("fp32", 2e-4), ("fp16s", 2e-4)])
def test_drag_adjoint_binomial_matches_oracle(hip, oracle_mod, storage, tol):
    """The drag objective's adjoint under binomial checkpointing against the oracle's logical-space
    adjoint on the same (bit-identical) forward states; FP16 storage keeps lambda in FP32."""
    report("drag_adjoint", storage=storage)
    flags, _ = cube_duct()
    blend = 0.05 * design(flags)
    T, W = 30, 6
    g = gpu_cube(hip, blend, storage)
    o = oracle_cube(oracle_mod, blend, storage)
    J, grad, du, fw = g.adjoint_run_binomial(T, W, 3, "drag")
    Jo, grado, duo = o.adjoint_run(T, W, kind="drag")
    assert J == pytest.approx(Jo, rel=tol)
    scale = np.abs(grado).max()
    assert np.abs(grad - grado).max() <= tol * scale
    assert du == pytest.approx(duo, rel=tol)
    assert fw == hip.checkpoint_plan(T, 3)[0]


def test_fp16_pressure_adjoint_matches_oracle(hip, oracle_mod):
    report("adjoint_fp16")
    n = (16, 8, 8)
    blend = 0.05 * design(duct(*n))
    T, W = 40, 8
    g = gpu_duct(hip, n, blend, storage="fp16s")
    o = oracle_duct(oracle_mod, n, blend, storage="fp16s")
    J, grad, du = g.adjoint_run(T, W, 7, 1, n[0] - 2)
    Jo, grado, duo = o.adjoint_run(T, W, 1, n[0] - 2)
    assert J == pytest.approx(Jo, rel=1e-9)
    assert np.abs(grad - grado).max() <= 2e-4 * np.abs(grado).max()
    assert du == pytest.approx(duo, rel=2e-4)


TORCH_CHILD = r"""
import json, sys
import numpy as np
import torch  # first: the library then shares torch's HIP runtime

torch.ones(1, device="cuda:0").sum().item()
sys.path.insert(0, sys.argv[1])
import cfd_solvers_amd as hip
from cfd_solvers_amd.autograd import LbmAdjointProblem, lbm_objective

n = (16, 8, 8)
f = np.zeros((8, 8, 16), np.uint8)
f[:, 0, :] = f[:, -1, :] = 0x80
f[0] = f[-1] = 0x80
f[1:-1, 1:-1, 0] = 0x01
f[1:-1, 1:-1, -1] = 0x02
flags = f.reshape(-1)
rng = np.random.default_rng(4)
blend = np.where(flags == 0, 0.05 * rng.uniform(0.3, 0.7, flags.size), 0.0).astype(np.float32)
prob = LbmAdjointProblem(flags=flags, shape=n, storage="fp64", steps=40, window=8, checkpoint_every=0,
                         snapshots=4, x_in=1, x_out=14)
dev = torch.device("cuda:0")
gamma = torch.tensor(blend, device=dev, requires_grad=True)
u_peak = torch.tensor(0.05, dtype=torch.float32, device=dev, requires_grad=True)
J = lbm_objective(gamma, u_peak, prob)
J.backward()
s = hip.EsotericPullHipSolver(19, *n, 0.6, storage="fp64")
s.set_flags(flags)
s.set_inlet_outlet(0.05, 1.0)
s.set_porosity(blend)
s.initialize_uniform()
ref_J, ref_grad, ref_du = s.adjoint_run(40, 8, 1, 1, 14)
v = np.where(flags == 0, np.random.default_rng(12).standard_normal(flags.size), 0.0)
gp = (blend + 1e-3 * 0.05 * v).astype(np.float32)
gm = (blend - 1e-3 * 0.05 * v).astype(np.float32)
with torch.no_grad():
    Jp = lbm_objective(torch.tensor(gp, device=dev), u_peak.detach(), prob).item()
    Jm = lbm_objective(torch.tensor(gm, device=dev), u_peak.detach(), prob).item()
    J_nograd = lbm_objective(gamma, u_peak, prob).item()
J_detached = lbm_objective(gamma.detach(), u_peak.detach(), prob)
# A gamma outside [0, 1] is refused on the device and the field in use stays as it was.
bad = torch.tensor(blend, device=dev)
bad[5] = 1.5
torch.cuda.synchronize()
solver = prob.solver()
try:
    solver.set_porosity_device(bad.data_ptr())
    refused = False
except RuntimeError as e:
    refused = "[0, 1]" in str(e)
solver.initialize_uniform()
J_after = solver.objective_run(prob.steps, prob.window, "pressure_drop", prob.x_in, 14)
# The host path (a CPU gamma) gives the same J and gradient.
gamma_cpu = torch.tensor(blend, requires_grad=True)
J_cpu = lbm_objective(gamma_cpu, u_peak.detach().cpu(), prob)
J_cpu.backward()
print(json.dumps({
    "J": J.item(), "ref_J": ref_J, "grad_device": str(gamma.grad.device),
    "grad_err": float((gamma.grad.double().cpu().numpy() - ref_grad.reshape(-1).astype(np.float32)).__abs__().max()),
    "grad_scale": float(np.abs(ref_grad).max()), "du": u_peak.grad.item(), "ref_du": ref_du,
    "dJ": Jp - Jm, "directional": float(np.dot(ref_grad.reshape(-1), gp.astype(np.float64) - gm.astype(np.float64))),
    "J_nograd": J_nograd, "J_detached": J_detached.item(), "detached_has_graph": J_detached.requires_grad,
    "refused": refused, "J_after": J_after, "J_cpu": J_cpu.item(),
    "cpu_grad_equal": bool(torch.equal(gamma_cpu.grad, gamma.grad.cpu())),
}))
"""


def test_torch_autograd_function():
    """LbmObjective on CUDA tensors: J.backward() returns the adjoint gradients (computed into a device
    tensor), a directional central difference of J along a random design perturbation matches
    <dJ/dgamma, v> (FP64 storage), evaluations without a gradient run the objective alone and return
    the same J, a gamma outside [0, 1] uploaded from device memory is refused, and CPU tensors give
    the same J and gradient through the host path.  Runs in a child process that imports torch first, as a torch user would: the
    library then shares torch's HIP runtime (this process has already started the library's own)."""
    import json
    import subprocess
    import sys
    from pathlib import Path

    report("torch_autograd")
    root = Path(__file__).resolve().parents[1]
    out = subprocess.run([sys.executable, "-c", TORCH_CHILD, str(root)], capture_output=True, text=True,
                         timeout=240, cwd=root)
    assert out.returncode == 0, out.stderr[-3000:]
    r = json.loads(out.stdout.strip().splitlines()[-1])
    report("torch_autograd_result", **r)
    assert r["J"] == r["ref_J"]
    assert r["grad_device"] == "cuda:0"
    assert r["grad_err"] <= 1e-6 * r["grad_scale"]
    assert r["du"] == pytest.approx(r["ref_du"], rel=1e-6)
    assert r["dJ"] == pytest.approx(r["directional"], rel=1e-5)
    # Objective-only evaluations (no grad mode, or no input requiring a gradient) return J bit for bit.
    assert r["J_nograd"] == r["J"] and r["J_detached"] == r["J"] and not r["detached_has_graph"]
    assert r["refused"] and r["J_after"] == r["J"]
    assert r["J_cpu"] == r["J"] and r["cpu_grad_equal"]
