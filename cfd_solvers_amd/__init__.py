"""MI355X (gfx950) esoteric-pull lattice-Boltzmann backend for cfd_solvers.

The compute path is ``libcfd_lbm_hip.so`` (hand-written HIP kernels behind the C ABI declared
in ``include/cfd_lbm/c_api.h``); this package is the ctypes host layer over it, mirroring the
reference's ``python/cfd_solvers`` bindings.  Build the library with
``python -m cfd_solvers_amd.build``.
"""
from .bindings import (EsotericPullHipSolver, MacroscopicFields, SolverLibrary, load, checkpoint_plan, FLUID,
                       VELOCITY_INLET, PRESSURE_OUTLET, WALL, LATTICES, STORAGE, OBJECTIVES, voxelize_surface,
                       voxelize_timing, transform_surface, ResidentThermalD3Q7, ResidentFreeSurface,
                       ResidentParticles, ResidentFlowDiagnostics)

__all__ = ["EsotericPullHipSolver", "MacroscopicFields", "SolverLibrary", "load", "checkpoint_plan", "FLUID",
           "VELOCITY_INLET", "PRESSURE_OUTLET", "WALL", "LATTICES", "STORAGE", "OBJECTIVES", "voxelize_surface",
           "voxelize_timing", "transform_surface", "ResidentThermalD3Q7", "ResidentFreeSurface",
           "ResidentParticles", "ResidentFlowDiagnostics"]

# torch.autograd.Function over the adjoint: `from cfd_solvers_amd.autograd import LbmObjective`
# (imported on demand so that the ctypes layer does not need torch).
