"""``torch.autograd.Function`` over the gfx950 discrete adjoint.

``LbmObjective.apply(gamma, u_peak, problem)`` runs ``problem.steps`` lattice-Boltzmann steps on the
GPU from a uniform rest state with the partial-bounce-back field ``gamma`` (one value per cell, in
[0, 1]) and the inlet peak velocity ``u_peak``, and returns the objective J (FP64, averaged over
the last ``problem.window`` states).  The same C-ABI call also runs the hand-written adjoint, so
``J.backward()`` hands back dJ/dgamma and dJ/du_peak without further GPU work.  PyTorch is only the
caller's interface here: the forward steps, the transposed steps and the checkpointing all run in
``libcfd_lbm_hip.so``.

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

    def run(self, gamma: np.ndarray, u_peak: float):
        """J, dJ/dgamma (float64, shape of gamma) and dJ/du_peak from one forward + adjoint run."""
        s = self.solver()
        s.set_inlet_outlet(float(u_peak), self.rho_out)
        s.set_porosity(np.ascontiguousarray(gamma, dtype=np.float32).reshape(-1))
        s.initialize_uniform()
        x_out = self.shape[0] - 2 if self.x_out is None else self.x_out
        if self.checkpoint_every > 0:
            if self.objective != "pressure_drop":
                raise ValueError("two-level checkpointing runs the pressure-drop objective; use checkpoint_every=0")
            J, grad, du = s.adjoint_run(self.steps, self.window, self.checkpoint_every, self.x_in, x_out)
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
        g = gamma.detach().to("cpu", torch.float32).numpy()
        J, grad, du = problem.run(g, float(u_peak.detach().cpu()))
        ctx.save_for_backward(torch.from_numpy(grad).to(gamma.device, gamma.dtype),
                              torch.tensor(du, dtype=u_peak.dtype, device=u_peak.device))
        return torch.tensor(J, dtype=torch.float64, device=gamma.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        dgamma, dpeak = ctx.saved_tensors
        g = grad_out.to(dgamma.dtype)
        return g * dgamma, grad_out.to(dpeak.dtype) * dpeak, None


def lbm_objective(gamma: torch.Tensor, u_peak: torch.Tensor, problem: LbmAdjointProblem) -> torch.Tensor:
    return LbmObjective.apply(gamma, u_peak, problem)
