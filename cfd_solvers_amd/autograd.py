"""``torch.autograd.Function`` over the gfx950 discrete adjoint.

``lbm_objective(gamma, u_peak, problem)`` runs ``problem.steps`` lattice-Boltzmann steps on the GPU
from a uniform rest state with the partial-bounce-back field ``gamma`` (one value per cell, in
[0, 1]) and the inlet peak velocity ``u_peak``, and returns the objective J (FP64, averaged over
the last ``problem.window`` states).  When a gradient is wanted (grad mode on and an input that
requires it), the same C-ABI call also runs the hand-written adjoint, so ``J.backward()`` hands back
dJ/dgamma and dJ/du_peak without further GPU work; otherwise (``torch.no_grad()``, line searches,
finite differences) only the steps and the objective run (``cfd_lbm_objective_run``), the same J
bit for bit at the cost of the forward steps alone.  A ``gamma`` on the GPU stays there: the field
is read from device memory and, with binomial checkpointing, dJ/dgamma is written into a device
tensor; no host round trip.  PyTorch is only the caller's interface here: the forward steps, the
transposed steps and the checkpointing all run in ``libcfd_lbm_hip.so``.

The reference has no differentiable solver; the gradients are checked against central finite
differences and the FP64 build of the same adjoint.

One HIP runtime per process: torch's wheel bundles its own ``libamdhip64.so.7``.  Import torch
before the library is first used (importing this module does so); the library then binds to
torch's runtime by soname and both share the device.  A process that has already run the library
on the GPU under another runtime cannot start torch's afterwards ("No HIP GPUs are available").
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from .bindings import EsotericPullHipSolver, LATTICES


@dataclass
class LbmAdjointProblem:
    """One differentiable flow problem: geometry, physics, horizon, objective and checkpointing.

    ``checkpoint_every > 0`` selects two-level checkpointing (a snapshot every that many steps);
    ``checkpoint_every == 0`` selects binomial checkpointing with ``snapshots`` state buffers
    (0: as many as fit in device memory).  ``objective`` is ``"pressure_drop"`` (planes ``x_in``,
    ``x_out``) or ``"drag"`` (``body``: nonzero on the wall cells of the body, along ``drag_axis``).
    """

    flags: np.ndarray
    shape: tuple  # (nx, ny, nz)
    lattice: str = "d3q19"
    tau: float = 0.6
    storage: str = "fp32"
    rho_out: float = 1.0
    steps: int = 100
    window: int = 10
    checkpoint_every: int = 0
    snapshots: int = 0
    objective: str = "pressure_drop"
    x_in: int = 1
    x_out: Optional[int] = None
    body: Optional[np.ndarray] = None
    drag_axis: int = 0
    _solver: Optional[EsotericPullHipSolver] = field(default=None, repr=False)

    def solver(self) -> EsotericPullHipSolver:
        if self._solver is None:
            nx, ny, nz = self.shape
            s = EsotericPullHipSolver(LATTICES[self.lattice], nx, ny, nz, self.tau, storage=self.storage)
            s.set_flags(np.ascontiguousarray(self.flags, dtype=np.uint8).reshape(-1))
            if self.objective == "drag":
                if self.body is None:
                    raise ValueError("the drag objective needs a body mask")
                s.set_drag_body(self.body, self.drag_axis)
            self._solver = s
        return self._solver

    def _prepare(self, gamma, u_peak: float) -> EsotericPullHipSolver:
        s = self.solver()
        s.set_inlet_outlet(float(u_peak), self.rho_out)
        if isinstance(gamma, torch.Tensor) and gamma.is_cuda:
            g = gamma.detach().to(torch.float32).contiguous().reshape(-1)
            if g.numel() != s.cells:
                raise ValueError("gamma must have one value per cell")
            torch.cuda.current_stream(g.device).synchronize()  # the library reads it on its own stream
            s.set_porosity_device(g.data_ptr())
        else:
            if isinstance(gamma, torch.Tensor):
                gamma = gamma.detach().cpu().numpy()
            s.set_porosity(np.ascontiguousarray(gamma, dtype=np.float32).reshape(-1))
        s.initialize_uniform()
        return s

    def _x_out(self) -> int:
        return self.shape[0] - 2 if self.x_out is None else self.x_out

    def evaluate(self, gamma, u_peak: float) -> float:
        """J alone: the forward steps and the objective, no adjoint (the J of ``run`` bit for bit)."""
        s = self._prepare(gamma, u_peak)
        return s.objective_run(self.steps, self.window, self.objective, self.x_in, self._x_out())

    def run(self, gamma, u_peak: float):
        """J, dJ/dgamma (float64, shape of gamma; a device tensor when gamma is one and the plan is
        binomial, else a numpy array) and dJ/du_peak from one forward + adjoint run."""
        s = self._prepare(gamma, u_peak)
        x_out = self._x_out()
        if self.checkpoint_every > 0:
            if self.objective != "pressure_drop":
                raise ValueError("two-level checkpointing runs the pressure-drop objective; use checkpoint_every=0")
            J, grad, du = s.adjoint_run(self.steps, self.window, self.checkpoint_every, self.x_in, x_out)
        elif isinstance(gamma, torch.Tensor) and gamma.is_cuda:
            grad = torch.empty(s.cells, dtype=torch.float64, device=gamma.device)
            torch.cuda.current_stream(grad.device).synchronize()
            J, du, _ = s.adjoint_run_binomial_device(self.steps, self.window, grad.data_ptr(), self.snapshots,
                                                     self.objective, self.x_in, x_out)
            return J, grad.reshape(gamma.shape), du
        else:
            J, grad, du, _ = s.adjoint_run_binomial(self.steps, self.window, self.snapshots, self.objective,
                                                    self.x_in, x_out)
        return J, grad.reshape(np.shape(gamma)), du

    def close(self):
        if self._solver is not None:
            self._solver.close()
            self._solver = None


class LbmObjective(torch.autograd.Function):
    """J(gamma, u_peak) for an ``LbmAdjointProblem``; differentiable in both inputs."""

    @staticmethod
    def forward(ctx, gamma: torch.Tensor, u_peak: torch.Tensor, problem: LbmAdjointProblem) -> torch.Tensor:
        u = float(u_peak.detach().cpu())
        if not any(ctx.needs_input_grad[:2]):
            J = problem.evaluate(gamma, u)
            ctx.save_for_backward(torch.zeros_like(gamma), torch.zeros_like(u_peak))
            return torch.tensor(J, dtype=torch.float64, device=gamma.device)
        J, grad, du = problem.run(gamma, u)
        if not isinstance(grad, torch.Tensor):
            grad = torch.from_numpy(grad)
        ctx.save_for_backward(grad.to(gamma.device, gamma.dtype),
                              torch.tensor(du, dtype=u_peak.dtype, device=u_peak.device))
        return torch.tensor(J, dtype=torch.float64, device=gamma.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        dgamma, dpeak = ctx.saved_tensors
        g = grad_out.to(dgamma.dtype)
        return g * dgamma, grad_out.to(dpeak.dtype) * dpeak, None


def lbm_objective(gamma: torch.Tensor, u_peak: torch.Tensor, problem: LbmAdjointProblem) -> torch.Tensor:
    """J(gamma, u_peak).  Without grad mode, or when neither input requires a gradient, only the
    forward steps run and the result carries no graph."""
    if not torch.is_grad_enabled() or not (gamma.requires_grad or u_peak.requires_grad):
        J = problem.evaluate(gamma, float(u_peak.detach().cpu()))
        return torch.tensor(J, dtype=torch.float64, device=gamma.device)
    return LbmObjective.apply(gamma, u_peak, problem)
