"""ctypes bindings for the gfx950 LBM C ABI (include/cfd_lbm/c_api.h).

Mirrors the reference's dependency-free ctypes layer (python/cfd_solvers/bindings.py:5-30):
``load(path=None)`` honours ``$CFD_SOLVERS_LIBRARY`` and raises ``RuntimeError`` with the C
ABI's error text whenever an entry point returns non-zero.  When no path is given the in-tree
``libcfd_lbm_hip.so`` is loaded; if it has not been built, loading fails loudly -- there is no
CPU fallback.

``EsotericPullHipSolver`` follows the method names of the reference's solver classes
(include/cfd/solvers/lbm/esoteric_pull.hpp:104-278, esoteric_pull_sycl.hpp:18-475), with numpy
arrays of shape (nz, ny, nx) for fields and (q, cells) for populations.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
DEFAULT_LIBRARY = PKG / "libcfd_lbm_hip.so"

LATTICES = {"d2q9": 9, "d3q19": 19, "d3q27": 27}
STORAGE = {"fp32": 0, "fp16s": 1, "fp64": 2, "fp32s": 3, "fp64x": 4}
FLUID, VELOCITY_INLET, PRESSURE_OUTLET, WALL = 0x00, 0x01, 0x02, 0x80

_ERR = 1024
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

# Every exported symbol and its argument types (the error buffer pair is appended).
_SIGNATURES = {
    "cfd_lbm_create": [C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_float, C.c_float,
                       C.c_float, C.c_int, C.POINTER(C.c_void_p)],
    "cfd_lbm_destroy": [C.c_void_p],
    "cfd_lbm_initialize_uniform": [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float],
    "cfd_lbm_initialize_taylor_green": [C.c_void_p, C.c_float, C.c_float, C.c_uint64],
    "cfd_lbm_initialize_fields": [C.c_void_p, _f32p, _f32p, _f32p, _f32p],
    "cfd_lbm_upload_populations": [C.c_void_p, _f32p],
    "cfd_lbm_download_populations": [C.c_void_p, _f32p],
    "cfd_lbm_download_raw_slots": [C.c_void_p, _f32p],
    "cfd_lbm_set_flags": [C.c_void_p, C.c_void_p],
    "cfd_lbm_set_wall_velocity": [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float],
    "cfd_lbm_set_porosity": [C.c_void_p, C.c_void_p],
    "cfd_lbm_set_inlet_outlet": [C.c_void_p, C.c_float, C.c_float],
    "cfd_lbm_set_local_acceleration": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "cfd_lbm_step": [C.c_void_p, C.c_uint64],
    "cfd_lbm_step_timed": [C.c_void_p, C.c_uint64, C.POINTER(C.c_float)],
    "cfd_lbm_synchronize": [C.c_void_p],
    "cfd_lbm_download_macroscopic": [C.c_void_p, _f32p, _f32p, _f32p, _f32p],
    "cfd_lbm_download_macroscopic_f64": [C.c_void_p, _f64p, _f64p, _f64p, _f64p],
    "cfd_lbm_refresh_device_macroscopic": [C.c_void_p] + [C.POINTER(C.c_void_p)] * 4,
    "cfd_lbm_total_mass": [C.c_void_p, C.POINTER(C.c_double)],
    "cfd_lbm_diagnostics": [C.c_void_p] + [C.POINTER(C.c_double)] * 3,
    "cfd_lbm_time_step": [C.c_void_p, C.POINTER(C.c_uint64)],
    "cfd_lbm_cells": [C.c_void_p, C.POINTER(C.c_size_t)],
    "cfd_lbm_population_bytes": [C.c_void_p, C.POINTER(C.c_size_t)],
    "cfd_lbm_resident_bytes": [C.c_void_p, C.POINTER(C.c_size_t)],
    "cfd_lbm_device_name": [C.c_void_p, C.c_char_p, C.c_size_t],
    "cfd_lbm_stream": [C.c_void_p, C.POINTER(C.c_void_p)],
    "cfd_lbm_transfer_stats": [C.c_void_p, C.POINTER(C.c_uint64), C.c_int],
    "cfd_lbm_index_map": [C.c_void_p, C.c_int, _i64p, _i64p],
    "cfd_lbm_objective_pressure_drop": [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)],
    "cfd_lbm_set_porosity_device": [C.c_void_p, C.c_void_p],
    "cfd_lbm_adjoint_run_binomial_device": [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                            C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_double),
                                            C.POINTER(C.c_uint64)],
    "cfd_lbm_halo_message_bytes": [C.c_void_p, C.POINTER(C.c_size_t)],
    "cfd_lbm_halo_pack": [C.c_void_p, C.c_void_p, C.c_void_p],
    "cfd_lbm_halo_unpack": [C.c_void_p, C.c_void_p, C.c_void_p],
    "cfd_lbm_objective_run": [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_double)],
    "cfd_lbm_adjoint_run": [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                            C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_double)],
    "cfd_lbm_adjoint_timing": [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "cfd_lbm_set_drag_body": [C.c_void_p, C.c_void_p, C.c_int],
    "cfd_lbm_set_smagorinsky": [C.c_void_p, C.c_int],
    "cfd_lbm_objective_drag": [C.c_void_p, C.POINTER(C.c_double)],
    "cfd_lbm_adjoint_run_binomial": [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                     C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_double),
                                     C.POINTER(C.c_uint64)],
    "cfd_lbm_checkpoint_plan": [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint64)],
    "cfd_lbm_voxelize_surface": [_f64p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _f64p, _f64p, C.c_void_p],
    "cfd_lbm_voxelize_surface_no_alloc": [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _f64p,
                                          _f64p, C.c_void_p, C.c_void_p],
    "cfd_lbm_voxelize_timing": [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
    "cfd_lbm_transform_surface": [_f64p, C.c_size_t, _f64p, _f64p, C.c_double, _f64p],
    "cfd_lbm_set_solid_surface": [C.c_void_p, _f64p, C.c_size_t, _f64p, _f64p, C.c_int],
    "cfd_lbm_device_local_acceleration": [C.c_void_p, C.c_int] + [C.POINTER(C.c_void_p)] * 3,
    "cfd_lbm_device_download": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t],
    "cfd_lbm_device_upload": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t],
    "cfd_lbm_thermal_create": [C.c_void_p, C.c_float, C.POINTER(C.c_void_p)],
    "cfd_lbm_thermal_destroy": [C.c_void_p],
    "cfd_lbm_thermal_initialize_uniform": [C.c_void_p, C.c_float],
    "cfd_lbm_thermal_initialize_fields": [C.c_void_p, _f32p],
    "cfd_lbm_thermal_step": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64],
    "cfd_lbm_thermal_refresh_device_scalar": [C.c_void_p, C.POINTER(C.c_void_p)],
    "cfd_lbm_thermal_download_scalar": [C.c_void_p, _f32p],
    "cfd_lbm_thermal_add_boussinesq": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                                       C.c_double, C.c_double, C.c_double],
    "cfd_lbm_thermal_total_scalar": [C.c_void_p, C.POINTER(C.c_double)],
    "cfd_lbm_thermal_resident_bytes": [C.c_void_p, C.POINTER(C.c_size_t)],
    "cfd_lbm_thermal_transfer_stats": [C.c_void_p, C.POINTER(C.c_uint64), C.c_int],
    "cfd_lbm_surface_create": [C.c_void_p, C.POINTER(C.c_void_p)],
    "cfd_lbm_surface_destroy": [C.c_void_p],
    "cfd_lbm_surface_initialize_uniform": [C.c_void_p, C.c_float],
    "cfd_lbm_surface_initialize_half_space_x": [C.c_void_p, C.c_float],
    "cfd_lbm_surface_upload_fill": [C.c_void_p, _f32p],
    "cfd_lbm_surface_download_fill": [C.c_void_p, _f32p],
    "cfd_lbm_surface_advect": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float],
    "cfd_lbm_surface_fill_device": [C.c_void_p, C.POINTER(C.c_void_p)],
    "cfd_lbm_surface_refresh_geometry": [C.c_void_p] + [C.POINTER(C.c_void_p)] * 4,
    "cfd_lbm_surface_add_surface_tension": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float],
    "cfd_lbm_surface_volume": [C.c_void_p, C.POINTER(C.c_double)],
    "cfd_lbm_surface_transfer_stats": [C.c_void_p, C.POINTER(C.c_uint64), C.c_int],
    "cfd_lbm_particles_create": [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)],
    "cfd_lbm_particles_destroy": [C.c_void_p],
    "cfd_lbm_particles_upload": [C.c_void_p, _f64p, C.c_size_t],
    "cfd_lbm_particles_download": [C.c_void_p, _f64p, C.c_size_t],
    "cfd_lbm_particles_advance": [C.c_void_p] + [C.c_void_p] * 6 + [C.c_float, C.c_int, C.c_double, C.c_double,
                                                                    C.c_double],
    "cfd_lbm_particles_transfer_stats": [C.c_void_p, C.POINTER(C.c_uint64), C.c_int],
    "cfd_lbm_flow_diagnostics_create": [C.c_void_p, C.POINTER(C.c_void_p)],
    "cfd_lbm_flow_diagnostics_destroy": [C.c_void_p],
    "cfd_lbm_q_criterion": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)],
    "cfd_lbm_get_flags": [C.c_void_p, C.c_void_p],
    "cfd_lbm_save_checkpoint": [C.c_void_p, C.c_char_p, C.c_int],
    "cfd_lbm_load_checkpoint": [C.c_void_p, C.c_char_p],
}

OBJECTIVES = {"pressure_drop": 0, "drag": 1}


class SolverLibrary:
    """Loaded C ABI; each call raises RuntimeError on a non-zero return code."""

    def __init__(self, path):
        self.path = Path(path)
        self._lib = C.CDLL(os.fspath(self.path))
        self._lib.cfd_lbm_version.restype = C.c_char_p
        for name, args in _SIGNATURES.items():
            fn = getattr(self._lib, name)
            fn.argtypes = list(args) + [C.c_char_p, C.c_size_t]
            fn.restype = C.c_int

    @property
    def version(self) -> str:
        return self._lib.cfd_lbm_version().decode()

    def call(self, name, *args):
        err = C.create_string_buffer(_ERR)
        rc = getattr(self._lib, name)(*args, err, _ERR)
        if rc:
            raise RuntimeError(f"{name}: {err.value.decode('utf-8', errors='replace')}")


_loaded: dict[str, SolverLibrary] = {}


def load(path=None) -> SolverLibrary:
    path = path or os.environ.get("CFD_SOLVERS_LIBRARY") or DEFAULT_LIBRARY
    path = Path(path)
    if not path.exists():
        raise RuntimeError(f"HIP LBM library not found at {path}; build it with `python -m cfd_solvers_amd.build`")
    key = os.fspath(path.resolve())
    if key not in _loaded:
        _loaded[key] = SolverLibrary(path)
    return _loaded[key]


@dataclass
class MacroscopicFields:
    rho: np.ndarray
    ux: np.ndarray
    uy: np.ndarray
    uz: np.ndarray


class EsotericPullHipSolver:
    """Device-resident esoteric-pull solver (EsotericPullSolver / EsotericPullSyclSolver API)."""

    def __init__(self, lattice="d3q19", nx=64, ny=64, nz=64, tau=0.6, acceleration=(0.0, 0.0, 0.0),
                 storage="fp32", library=None):
        self.lib = load(library)
        self.q = LATTICES[lattice] if isinstance(lattice, str) else int(lattice)
        self.storage = storage
        nz = 1 if self.q == 9 else nz
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.shape = (nz, ny, nx)
        h = C.c_void_p()
        self.lib.call("cfd_lbm_create", self.q, nx, ny, nz, tau, *map(float, acceleration),
                      STORAGE[storage], C.byref(h))
        self._h = h
        self.cells = nx * ny * nz
        self._modules = weakref.WeakSet()  # resident modules on this solver's stream

    def close(self):
        if getattr(self, "_h", None):
            for m in list(getattr(self, "_modules", ())):
                m.close()  # they run on this solver's stream: destroy them first
            self.lib.call("cfd_lbm_destroy", self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # initial conditions -------------------------------------------------------------
    def initialize_uniform(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0):
        self.lib.call("cfd_lbm_initialize_uniform", self._h, rho, ux, uy, uz)

    def initialize_taylor_green(self, amplitude=0.03, noise=0.0, seed=0):
        self.lib.call("cfd_lbm_initialize_taylor_green", self._h, amplitude, noise, seed)

    def initialize_fields(self, rho, ux, uy, uz):
        """initialize(initializer) (esoteric_pull.hpp:312-333) with host fields of `cells` values each."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (rho, ux, uy, uz)]
        if any(a.size != self.cells for a in arrs):
            raise ValueError("initial LBM field size mismatch: rho, ux, uy, uz need cells values each")
        self.lib.call("cfd_lbm_initialize_fields", self._h, *arrs)

    def upload_populations(self, logical):
        a = np.ascontiguousarray(logical, dtype=np.float32).reshape(-1)
        if a.size != self.q * self.cells:
            raise ValueError("populations must have q*cells values")
        self.lib.call("cfd_lbm_upload_populations", self._h, a)

    def logical_populations(self):
        out = np.empty(self.q * self.cells, dtype=np.float32)
        self.lib.call("cfd_lbm_download_populations", self._h, out)
        return out.reshape(self.q, self.cells)

    def raw_storage(self):
        """The in-place slot array, shape (q, cells), in the reference's order (raw_storage,
        esoteric_pull.hpp:276-278), each slot as a full float population."""
        out = np.empty(self.q * self.cells, dtype=np.float32)
        self.lib.call("cfd_lbm_download_raw_slots", self._h, out)
        return out.reshape(self.q, self.cells)

    # boundaries and forcing ---------------------------------------------------------
    def set_flags(self, flags):
        if flags is None:
            self.lib.call("cfd_lbm_set_flags", self._h, None)
            return
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        if f.size != self.cells:
            raise ValueError("LBM flag field size mismatch")
        self.lib.call("cfd_lbm_set_flags", self._h, f.ctypes.data)

    def flags(self):
        """The solver's flag field (uint8, shape (nz, ny, nx)); all FLUID when none are set."""
        out = np.empty(self.cells, np.uint8)
        self.lib.call("cfd_lbm_get_flags", self._h, out.ctypes.data)
        return out.reshape(self.shape)

    def set_solid_mask(self, mask):
        m = np.asarray(mask).reshape(-1)
        self.set_flags(np.where(m != 0, WALL, FLUID).astype(np.uint8))

    def clear_solid_mask(self):
        self.set_flags(None)

    def set_solid_surface(self, triangles, minimum=(0.0, 0.0, 0.0), maximum=(1.0, 1.0, 1.0), keep_flags=False):
        """set_solid_mask(voxelize_surface(triangles, grid)) on the device (examples/real_lbm3d_model.cpp:
        146-151): the grid is this lattice's nx, ny, nz over [minimum, maximum]; cells inside the
        closed surface become walls at rest.  keep_flags keeps the other cells' current flags."""
        t = _triangles(triangles)
        self.lib.call("cfd_lbm_set_solid_surface", self._h, t, len(t) // 9, _vec3(minimum), _vec3(maximum),
                      1 if keep_flags else 0)

    def set_wall_velocity(self, index, u):
        self.lib.call("cfd_lbm_set_wall_velocity", self._h, index, *map(float, u))

    def set_porosity(self, gamma):
        if gamma is None:
            self.lib.call("cfd_lbm_set_porosity", self._h, None)
            return
        g = np.ascontiguousarray(gamma, dtype=np.float32).reshape(-1)
        if g.size != self.cells:
            raise ValueError("porosity field size mismatch")
        self.lib.call("cfd_lbm_set_porosity", self._h, g.ctypes.data)

    def set_inlet_outlet(self, u_peak, rho_out=1.0):
        self.lib.call("cfd_lbm_set_inlet_outlet", self._h, u_peak, rho_out)

    def set_local_body_acceleration(self, ax, ay, az):
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (ax, ay, az)]
        if any(a.size != self.cells for a in arrs):
            raise ValueError("local LBM acceleration field size mismatch")
        self._lacc = arrs
        self.lib.call("cfd_lbm_set_local_acceleration", self._h, *(a.ctypes.data for a in arrs))

    def clear_local_body_acceleration(self):
        self.lib.call("cfd_lbm_set_local_acceleration", self._h, None, None, None)

    # stepping -----------------------------------------------------------------------
    def step(self, count=1):
        self.lib.call("cfd_lbm_step", self._h, count)

    def step_timed(self, count=1) -> float:
        ms = C.c_float()
        self.lib.call("cfd_lbm_step_timed", self._h, count, C.byref(ms))
        return ms.value

    def wait(self):
        self.lib.call("cfd_lbm_synchronize", self._h)

    # results ------------------------------------------------------------------------
    def download_macroscopic(self, f64=False) -> MacroscopicFields:
        dt = np.float64 if f64 else np.float32
        out = [np.empty(self.cells, dtype=dt) for _ in range(4)]
        self.lib.call("cfd_lbm_download_macroscopic_f64" if f64 else "cfd_lbm_download_macroscopic",
                      self._h, *out)
        return MacroscopicFields(*(o.reshape(self.shape) for o in out))

    compute_macroscopic = download_macroscopic

    def refresh_device_macroscopic(self):
        ptrs = [C.c_void_p() for _ in range(4)]
        self.lib.call("cfd_lbm_refresh_device_macroscopic", self._h, *(C.byref(p) for p in ptrs))
        return tuple(p.value for p in ptrs)

    def save_checkpoint(self, path, raw=True):
        """CFDCHK3 checkpoint (distributed_pull.hpp:176-203) plus a trailer with the parity and,
        with raw, the exact device state."""
        self.lib.call("cfd_lbm_save_checkpoint", self._h, os.fsencode(path), int(bool(raw)))

    def load_checkpoint(self, path):
        self.lib.call("cfd_lbm_load_checkpoint", self._h, os.fsencode(path))

    def device_local_acceleration(self, clear=False):
        """(ax, ay, az) device pointers of the per-cell acceleration planes
        (esoteric_pull_sycl.hpp:349-358); allocated zeroed on first use."""
        ptrs = [C.c_void_p() for _ in range(3)]
        self.lib.call("cfd_lbm_device_local_acceleration", self._h, int(clear), *(C.byref(p) for p in ptrs))
        return tuple(p.value for p in ptrs)

    def device_download(self, pointer, dtype=np.float32, count=None):
        """Copy of `count` values (default: one per cell) at a device pointer of this solver."""
        count = self.cells if count is None else int(count)
        out = np.empty(count, dtype)
        self.lib.call("cfd_lbm_device_download", self._h, pointer, out.ctypes.data, out.nbytes)
        return out

    def device_upload(self, pointer, values):
        v = np.ascontiguousarray(values)
        self.lib.call("cfd_lbm_device_upload", self._h, pointer, v.ctypes.data, v.nbytes)

    def total_mass(self) -> float:
        m = C.c_double()
        self.lib.call("cfd_lbm_total_mass", self._h, C.byref(m))
        return m.value

    mass = total_mass

    def diagnostics(self):
        m, ke, vmax = C.c_double(), C.c_double(), C.c_double()
        self.lib.call("cfd_lbm_diagnostics", self._h, C.byref(m), C.byref(ke), C.byref(vmax))
        return m.value, ke.value, vmax.value

    def kinetic_energy(self) -> float:
        return self.diagnostics()[1]

    def max_speed(self) -> float:
        return self.diagnostics()[2]

    @property
    def time_step(self) -> int:
        v = C.c_uint64()
        self.lib.call("cfd_lbm_time_step", self._h, C.byref(v))
        return v.value

    def population_bytes(self) -> int:
        v = C.c_size_t()
        self.lib.call("cfd_lbm_population_bytes", self._h, C.byref(v))
        return v.value

    def two_lattice_population_bytes(self) -> int:
        return 2 * self.population_bytes()

    def resident_bytes(self) -> int:
        v = C.c_size_t()
        self.lib.call("cfd_lbm_resident_bytes", self._h, C.byref(v))
        return v.value

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        self.lib.call("cfd_lbm_device_name", self._h, buf, 256)
        return buf.value.decode()

    def stream(self) -> int:
        p = C.c_void_p()
        self.lib.call("cfd_lbm_stream", self._h, C.byref(p))
        return p.value or 0

    def transfer_stats(self, reset=False):
        v = (C.c_uint64 * 4)()
        self.lib.call("cfd_lbm_transfer_stats", self._h, v, int(reset))
        return {"host_to_device_bytes": v[0], "device_to_host_bytes": v[1],
                "device_to_device_bytes": v[2], "synchronization_points": v[3]}

    def objective_pressure_drop(self, x_in, x_out):
        v = C.c_double()
        self.lib.call("cfd_lbm_objective_pressure_drop", self._h, x_in, x_out, C.byref(v))
        return v.value

    def objective_run(self, steps, window, objective="pressure_drop", x_in=1, x_out=None):
        """J of adjoint_run / adjoint_run_binomial from the current state, without the adjoint."""
        J = C.c_double()
        x_out = self.shape[-1] - 2 if x_out is None else x_out
        self.lib.call("cfd_lbm_objective_run", self._h, OBJECTIVES[objective], steps, window, x_in, x_out, C.byref(J))
        return J.value

    def adjoint_run(self, steps, window, checkpoint_every, x_in, x_out, want_gamma=True):
        """J, dJ/dgamma (shape (nz, ny, nx), FP64) and dJ/du_peak; the state advances by `steps`."""
        J, du = C.c_double(), C.c_double()
        grad = np.zeros(self.cells) if want_gamma else None
        self.lib.call("cfd_lbm_adjoint_run", self._h, steps, window, checkpoint_every, x_in, x_out, C.byref(J),
                      grad.ctypes.data if grad is not None else None, C.byref(du))
        return J.value, (grad.reshape(self.shape) if grad is not None else None), du.value

    def set_smagorinsky(self, enabled=True):
        """The reference's smagorinsky_les switch (esoteric_pull.hpp:72-97, 422-425)."""
        self.lib.call("cfd_lbm_set_smagorinsky", self._h, int(enabled))

    def set_drag_body(self, body, axis=0):
        """Registers the wall cells of the drag body (nonzero entries of `body`, shape cells); None clears."""
        if body is None:
            self.lib.call("cfd_lbm_set_drag_body", self._h, None, axis)
            return
        b = np.ascontiguousarray(np.asarray(body).reshape(-1) != 0, dtype=np.uint8)
        if b.size != self.cells:
            raise ValueError("drag body must have one entry per cell")
        self.lib.call("cfd_lbm_set_drag_body", self._h, b.ctypes.data, axis)

    def objective_drag(self):
        v = C.c_double()
        self.lib.call("cfd_lbm_objective_drag", self._h, C.byref(v))
        return v.value

    def adjoint_run_binomial(self, steps, window, snapshots=0, objective="pressure_drop", x_in=1, x_out=None,
                             want_gamma=True):
        """Binomial checkpointing: J, dJ/dgamma (shape (nz, ny, nx)), dJ/du_peak, forward steps executed."""
        J, du, fw = C.c_double(), C.c_double(), C.c_uint64()
        grad = np.zeros(self.cells) if want_gamma else None
        x_out = self.shape[-1] - 2 if x_out is None else x_out
        self.lib.call("cfd_lbm_adjoint_run_binomial", self._h, OBJECTIVES[objective], steps, window, snapshots, x_in,
                      x_out, C.byref(J), grad.ctypes.data if grad is not None else None, C.byref(du), C.byref(fw))
        return J.value, (grad.reshape(self.shape) if grad is not None else None), du.value, fw.value

    def halo_message_bytes(self):
        """Bytes of one z-slab halo message (crossing pairs x nx x ny populations)."""
        b = C.c_size_t()
        self.lib.call("cfd_lbm_halo_message_bytes", self._h, C.byref(b))
        return b.value

    def halo_pack(self, up_device, down_device):
        """After a step: this slab's up and down messages into two device buffers (addresses)."""
        self.lib.call("cfd_lbm_halo_pack", self._h, C.c_void_p(int(up_device)), C.c_void_p(int(down_device)))

    def halo_unpack(self, from_below_device, from_above_device):
        """The up message of the slab below and the down message of the slab above, into this slab."""
        self.lib.call("cfd_lbm_halo_unpack", self._h, C.c_void_p(int(from_below_device)),
                      C.c_void_p(int(from_above_device)))

    def set_porosity_device(self, gamma_device):
        """Porosity from device memory: gamma_device is the device address (int) of cells floats."""
        self.lib.call("cfd_lbm_set_porosity_device", self._h, C.c_void_p(int(gamma_device)))

    def adjoint_run_binomial_device(self, steps, window, grad_device, snapshots=0, objective="pressure_drop", x_in=1,
                                    x_out=None):
        """adjoint_run_binomial with dJ/dgamma written to the device address grad_device (cells
        doubles): J, dJ/du_peak, forward steps executed."""
        J, du, fw = C.c_double(), C.c_double(), C.c_uint64()
        x_out = self.shape[-1] - 2 if x_out is None else x_out
        self.lib.call("cfd_lbm_adjoint_run_binomial_device", self._h, OBJECTIVES[objective], steps, window, snapshots,
                      x_in, x_out, C.byref(J), C.c_void_p(int(grad_device)), C.byref(du), C.byref(fw))
        return J.value, du.value, fw.value

    def adjoint_timing(self):
        fw, rv = C.c_float(), C.c_float()
        self.lib.call("cfd_lbm_adjoint_timing", self._h, C.byref(fw), C.byref(rv))
        return fw.value, rv.value

    def index_map(self, parity):
        load = np.empty(self.q * self.cells, dtype=np.int64)
        store = np.empty(self.q * self.cells, dtype=np.int64)
        self.lib.call("cfd_lbm_index_map", self._h, parity, load, store)
        return load.reshape(self.q, self.cells), store.reshape(self.q, self.cells)


def checkpoint_plan(steps, snapshots, library=None):
    """Binomial plan for `steps` steps with `snapshots` free state buffers (host only):
    (forward steps by the closed form, repetition number, forward steps by walking the plan)."""
    lib = load(library)
    fw, r, walked = C.c_uint64(), C.c_uint64(), C.c_uint64()
    lib.call("cfd_lbm_checkpoint_plan", steps, snapshots, C.byref(fw), C.byref(r), C.byref(walked))
    return fw.value, r.value, walked.value


def _triangles(triangles):
    t = np.ascontiguousarray(triangles, dtype=np.float64)
    if t.ndim == 0 or t.size % 9:
        raise ValueError("triangles must hold 9 coordinates each (shape (count, 3, 3) or (count, 9))")
    return t.reshape(-1)


def _vec3(v):
    a = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    if a.size != 3:
        raise ValueError("expected three coordinates")
    return a


def voxelize_surface(triangles, shape, minimum=(0.0, 0.0, 0.0), maximum=(1.0, 1.0, 1.0), library=None):
    """voxelize_surface (src/lbm/voxelize.cpp:5) on the GPU: the ray-parity inside mask of a closed
    triangle surface (triangles: (count, 3, 3) vertex coordinates) on the grid of shape (nx, ny, nz)
    spanning [minimum, maximum]; uint8 array of shape (nz, ny, nx), equal to the reference's CPU
    mask bit for bit."""
    nx, ny, nz = (int(n) for n in shape)
    t = _triangles(triangles)
    mask = np.empty((nz, ny, nx), np.uint8)
    load(library).call("cfd_lbm_voxelize_surface", t, len(t) // 9, nx, ny, nz, _vec3(minimum), _vec3(maximum),
                       mask.ctypes.data)
    return mask


def voxelize_surface_no_alloc(triangles_device, count, shape, mask_device, minimum=(0.0, 0.0, 0.0),
                              maximum=(1.0, 1.0, 1.0), stream=None, library=None):
    """cfd_lbm_voxelize_surface_no_alloc: the mask of voxelize_surface written into caller-owned
    device memory without device allocation.  triangles_device, mask_device: device addresses (ints)
    of count x 9 float64 vertex coordinates and nx*ny*nz bytes; stream: a hipStream_t as an int
    (None: the null stream).  Returns nothing; the mask is written when the call returns."""
    nx, ny, nz = (int(n) for n in shape)
    load(library).call("cfd_lbm_voxelize_surface_no_alloc", C.c_void_p(int(triangles_device)), int(count), nx, ny,
                       nz, _vec3(minimum), _vec3(maximum), C.c_void_p(int(mask_device)),
                       C.c_void_p(int(stream) if stream else None))


def voxelize_timing(library=None):
    """(binning ms, line kernel ms, triangle-tile pairs) of this thread's last voxelisation."""
    b, l, n = C.c_double(), C.c_double(), C.c_uint64()
    load(library).call("cfd_lbm_voxelize_timing", C.byref(b), C.byref(l), C.byref(n))
    return b.value, l.value, n.value


def transform_surface(triangles, translation, axis, angle, library=None):
    """transform_surface (src/lbm/voxelize.cpp:6): rotate about `axis` by `angle` radians, then
    translate; returns a new (count, 3, 3) float64 array."""
    t = _triangles(triangles)
    out = np.empty_like(t)
    load(library).call("cfd_lbm_transform_surface", t, len(t) // 9, _vec3(translation), _vec3(axis), float(angle), out)
    return out.reshape(-1, 3, 3)


def _stats(lib, name, handle, reset):
    v = (C.c_uint64 * 4)()
    lib.call(name, handle, v, int(reset))
    return {"host_to_device_bytes": v[0], "device_to_host_bytes": v[1], "device_to_device_bytes": v[2],
            "synchronization_points": v[3], "host_transfer_bytes": v[0] + v[1]}


class _Module:
    """A device-resident module on a flow solver's device and stream
    (include/cfd/solvers/lbm/resident_multiphysics_sycl.hpp).  Field arguments are device pointers
    (ints), e.g. from ``flow.refresh_device_macroscopic()`` and ``flow.device_local_acceleration()``."""

    _prefix = ""

    def __init__(self, flow, *args):
        self.flow = flow
        self.lib = flow.lib
        h = C.c_void_p()
        self.lib.call(f"cfd_lbm_{self._prefix}_create", flow._h, *args, C.byref(h))
        self._h = h
        flow._modules.add(self)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.call(f"cfd_lbm_{self._prefix}_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class ResidentThermalD3Q7(_Module):
    """ResidentThermalD3Q7Sycl (resident_multiphysics_sycl.hpp:54-198)."""

    _prefix = "thermal"

    def __init__(self, flow, diffusivity=0.01):
        super().__init__(flow, float(diffusivity))

    def initialize_uniform(self, temperature):
        self.lib.call("cfd_lbm_thermal_initialize_uniform", self._h, float(temperature))

    def initialize_fields(self, temperature):
        t = np.ascontiguousarray(temperature, np.float32).reshape(-1)
        if t.size != self.flow.cells:
            raise ValueError("thermal field size mismatch")
        self.lib.call("cfd_lbm_thermal_initialize_fields", self._h, t)

    def step(self, macroscopic, count=1):
        """macroscopic: (rho, ux, uy, uz) device pointers from flow.refresh_device_macroscopic()."""
        _, ux, uy, uz = macroscopic
        self.lib.call("cfd_lbm_thermal_step", self._h, ux, uy, uz, int(count))

    def refresh_device_scalar(self):
        p = C.c_void_p()
        self.lib.call("cfd_lbm_thermal_refresh_device_scalar", self._h, C.byref(p))
        return p.value

    def scalar(self):
        out = np.empty(self.flow.cells, np.float32)
        self.lib.call("cfd_lbm_thermal_download_scalar", self._h, out)
        return out.reshape(self.flow.shape)

    def add_boussinesq_acceleration(self, force, reference_temperature, beta, gravity):
        ax, ay, az = force
        self.lib.call("cfd_lbm_thermal_add_boussinesq", self._h, ax, ay, az, float(reference_temperature), float(beta),
                      *map(float, gravity))

    def total_scalar(self):
        v = C.c_double()
        self.lib.call("cfd_lbm_thermal_total_scalar", self._h, C.byref(v))
        return v.value

    def resident_bytes(self):
        v = C.c_size_t()
        self.lib.call("cfd_lbm_thermal_resident_bytes", self._h, C.byref(v))
        return v.value

    def transfer_stats(self, reset=False):
        return _stats(self.lib, "cfd_lbm_thermal_transfer_stats", self._h, reset)


class ResidentFreeSurface(_Module):
    """ResidentFreeSurfaceSycl3D (resident_multiphysics_sycl.hpp:200-324)."""

    _prefix = "surface"

    def initialize_uniform(self, fill):
        self.lib.call("cfd_lbm_surface_initialize_uniform", self._h, float(fill))

    def initialize_half_space_x(self, interface_x):
        self.lib.call("cfd_lbm_surface_initialize_half_space_x", self._h, float(interface_x))

    def upload_fill(self, fill):
        f = np.ascontiguousarray(fill, np.float32).reshape(-1)
        if f.size != self.flow.cells:
            raise ValueError("fill field size mismatch")
        self.lib.call("cfd_lbm_surface_upload_fill", self._h, f)

    def fill(self):
        out = np.empty(self.flow.cells, np.float32)
        self.lib.call("cfd_lbm_surface_download_fill", self._h, out)
        return out.reshape(self.flow.shape)

    def advect(self, macroscopic, dt):
        _, ux, uy, uz = macroscopic
        self.lib.call("cfd_lbm_surface_advect", self._h, ux, uy, uz, float(dt))

    def fill_device(self):
        p = C.c_void_p()
        self.lib.call("cfd_lbm_surface_fill_device", self._h, C.byref(p))
        return p.value

    def refresh_geometry(self):
        """(normal_x, normal_y, normal_z, curvature) device pointers."""
        ptrs = [C.c_void_p() for _ in range(4)]
        self.lib.call("cfd_lbm_surface_refresh_geometry", self._h, *(C.byref(p) for p in ptrs))
        return tuple(p.value for p in ptrs)

    def add_surface_tension_acceleration(self, force, surface_tension, density=1.0):
        ax, ay, az = force
        self.lib.call("cfd_lbm_surface_add_surface_tension", self._h, ax, ay, az, float(surface_tension),
                      float(density))

    def volume(self):
        v = C.c_double()
        self.lib.call("cfd_lbm_surface_volume", self._h, C.byref(v))
        return v.value

    def transfer_stats(self, reset=False):
        return _stats(self.lib, "cfd_lbm_surface_transfer_stats", self._h, reset)


class ResidentParticles(_Module):
    """ResidentParticlesSycl3D (resident_multiphysics_sycl.hpp:326-380); particles as (count, 8)
    float64 rows: position xyz, velocity xyz, mass, response time (ImmersedBoundaryParticle)."""

    _prefix = "particles"

    def __init__(self, flow, count):
        self.count = int(count)
        super().__init__(flow, self.count)

    def upload(self, particles):
        p = np.ascontiguousarray(particles, np.float64).reshape(-1, 8)
        self.lib.call("cfd_lbm_particles_upload", self._h, p.reshape(-1), len(p))

    def download(self):
        out = np.empty((self.count, 8), np.float64)
        self.lib.call("cfd_lbm_particles_download", self._h, out.reshape(-1), self.count)
        return out

    def advance(self, macroscopic, force, dt, two_way, gravity=(0.0, 0.0, 0.0)):
        _, ux, uy, uz = macroscopic
        ax, ay, az = force
        self.lib.call("cfd_lbm_particles_advance", self._h, ux, uy, uz, ax, ay, az, float(dt), int(bool(two_way)),
                      *map(float, gravity))

    def transfer_stats(self, reset=False):
        return _stats(self.lib, "cfd_lbm_particles_transfer_stats", self._h, reset)


class ResidentFlowDiagnostics(_Module):
    """ResidentFlowDiagnosticsSycl3D (resident_multiphysics_sycl.hpp:382-408)."""

    _prefix = "flow_diagnostics"

    def q_criterion(self, macroscopic):
        """Device pointer to the Q-criterion field (cells floats)."""
        _, ux, uy, uz = macroscopic
        p = C.c_void_p()
        self.lib.call("cfd_lbm_q_criterion", self._h, ux, uy, uz, C.byref(p))
        return p.value
