"""ctypes access to the CPU checker libraries.  TEST INFRASTRUCTURE ONLY.

``Oracle`` drives ``oracle/liblbm_oracle.so`` (the CPU restatement of the reference's
esoteric-pull path, ``lbm_oracle.cpp``); ``Ref*`` classes drive ``oracle/_ref/libref_lbm*.so``,
built from the reference's own sources by ``oracle/Makefile``.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this package.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liblbm_oracle.so"
REF_SO = HERE / "_ref" / "libref_lbm.so"
REF_NATIVE_SO = HERE / "_ref" / "libref_lbm_native.so"

STORAGE = {"fp32": 0, "fp16s": 1, "fp64": 2, "fp32s": 3}
FLUID, INLET, OUTLET, WALL = 0x00, 0x01, 0x02, 0x80

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


REF_CFD_BENCH = HERE / "_ref" / "cfd-bench-hip"


def build(reference: str | None = None, quiet: bool = True) -> None:
    """Run oracle/Makefile (the restatement always; _ref only where the reference tree exists).
    With the reference tree and the HIP library present, also builds _ref/cfd-bench-hip: the
    reference's apps/cfd-bench.cpp with INTEGRATION.md's --backend hip insertions, linked to the
    library (oracle/integration.py)."""
    cmd = ["make", "-C", str(HERE), "all"]
    if reference:
        cmd.append(f"REF={reference}")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL if quiet else None)
    if reference and (Path(reference) / "apps" / "cfd-bench.cpp").exists():
        from . import integration

        if (integration.HIP_LIB_DIR / "libcfd_lbm_hip.so").exists():
            (HERE / "_ref").mkdir(exist_ok=True)
            r = integration.compile_cfd_bench(reference, REF_CFD_BENCH)
            if r.returncode != 0:
                raise RuntimeError("building the reference's cfd-bench with --backend hip failed:\n" + r.stderr[-2000:])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            build()
        L = C.CDLL(str(ORACLE_SO))
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.POINTER(C.c_void_p)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_flags.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_wall_velocity.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_set_gamma.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_local_acceleration.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_inlet_outlet.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_init_fields.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_init_uniform.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_init_taylor_green.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_uint64]
        L.orc_step.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_macroscopic.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        for name in ("orc_mass", "orc_kinetic_energy", "orc_max_speed"):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = C.c_double
        L.orc_logical.argtypes = [C.c_void_p, _dp]
        L.orc_set_logical.argtypes = [C.c_void_p, _dp]
        L.orc_raw.argtypes = [C.c_void_p, _dp]
        for name in ("orc_load_map", "orc_store_map", "orc_bb_load_map"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, _i64p]
        L.orc_time_step.argtypes = [C.c_void_p]
        L.orc_time_step.restype = C.c_uint64
        L.orc_set_time_step.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_cells.argtypes = [C.c_void_p]
        L.orc_cells.restype = C.c_size_t
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_uniform01.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_uniform01.restype = C.c_double
        L.orc_splitmix64.argtypes = [_u64p, _u64p, C.c_size_t]
        L.orc_uniform01_cells.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, _dp]
        L.orc_voxelize.argtypes = [_dp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _dp, _dp, _u8p]
        S3 = [C.c_size_t] * 3
        L.orc_thermal_fill.argtypes = [_fp, C.c_void_p, C.c_float, C.c_size_t]
        L.orc_thermal_step.argtypes = [_fp, _fp, _fp, _fp, _fp] + S3 + [C.c_float]
        L.orc_thermal_scalar.argtypes = [_fp, _fp, C.c_size_t]
        L.orc_boussinesq.argtypes = [_fp, _fp, _fp, _fp, C.c_size_t] + [C.c_float] * 5
        L.orc_vof_half_space.argtypes = [_fp] + S3 + [C.c_float]
        L.orc_vof_advect.argtypes = [_fp, _fp, _fp, _fp, _fp] + S3 + [C.c_float]
        L.orc_vof_geometry.argtypes = [_fp, _fp] + S3
        L.orc_vof_tension.argtypes = [_fp, _fp, _fp, _fp, _fp] + S3 + [C.c_float, C.c_float]
        L.orc_particles_advance.argtypes = [_fp, C.c_size_t, _fp, _fp, _fp, _fp, _fp, _fp] + S3 + [
            C.c_float, C.c_int, C.c_float, C.c_float, C.c_float]
        L.orc_q_criterion.argtypes = [_fp, _fp, _fp, _fp] + S3
        L.orc_transform_surface.argtypes = [_dp, C.c_size_t, _dp, _dp, C.c_double, _dp]
        _lib = L
    return _lib


def _check(rc: int, L=None, prefix="orc") -> None:
    if rc != 0:
        L = L or lib()
        raise RuntimeError(getattr(L, f"{prefix}_last_error")().decode())


class Oracle:
    """CPU restatement of EsotericPullSolver<D> (fp32 / fp16-shifted / fp64 storage)."""

    def __init__(self, q, nx, ny, nz, tau, accel=(0.0, 0.0, 0.0), storage="fp32"):
        self.L = lib()
        self.q = q
        self.shape = (nz if q != 9 else 1, ny, nx)
        h = C.c_void_p()
        _check(self.L.orc_create(q, STORAGE[storage], nx, ny, nz, tau, *accel, C.byref(h)))
        self.h = h
        self.cells = int(self.L.orc_cells(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_flags(self, flags):
        if flags is None:
            _check(self.L.orc_set_flags(self.h, None))
            return
        self._flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        assert self._flags.size == self.cells
        _check(self.L.orc_set_flags(self.h, self._flags.ctypes.data))

    def set_wall_velocity(self, index, u):
        _check(self.L.orc_set_wall_velocity(self.h, index, *map(float, u)))

    def set_smagorinsky(self, enabled=True):
        self.L.orc_set_smagorinsky.argtypes = [C.c_void_p, C.c_int]
        _check(self.L.orc_set_smagorinsky(self.h, int(enabled)))

    def set_gamma(self, gamma):
        if gamma is None:
            _check(self.L.orc_set_gamma(self.h, None))
            return
        g = np.ascontiguousarray(gamma, dtype=np.float64).reshape(-1)
        _check(self.L.orc_set_gamma(self.h, g.ctypes.data))

    def set_local_acceleration(self, ax, ay=None, az=None):
        if ax is None:
            _check(self.L.orc_set_local_acceleration(self.h, None, None, None))
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (ax, ay, az)]
        _check(self.L.orc_set_local_acceleration(self.h, *(a.ctypes.data for a in arrs)))

    def set_inlet_outlet(self, u_peak, rho_out=1.0):
        _check(self.L.orc_set_inlet_outlet(self.h, u_peak, rho_out))

    def init_fields(self, rho, ux, uy, uz):
        a = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in (rho, ux, uy, uz)]
        _check(self.L.orc_init_fields(self.h, *a))

    def init_uniform(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0):
        _check(self.L.orc_init_uniform(self.h, rho, ux, uy, uz))

    def init_taylor_green(self, amplitude=0.03, noise=0.0, seed=0):
        _check(self.L.orc_init_taylor_green(self.h, amplitude, noise, seed))

    def step(self, count=1):
        _check(self.L.orc_step(self.h, count))

    def macroscopic(self):
        out = [np.empty(self.cells) for _ in range(4)]
        _check(self.L.orc_macroscopic(self.h, *out))
        return [o.reshape(self.shape) for o in out]

    def mass(self):
        return self.L.orc_mass(self.h)

    def kinetic_energy(self):
        return self.L.orc_kinetic_energy(self.h)

    def max_speed(self):
        return self.L.orc_max_speed(self.h)

    def logical(self):
        out = np.empty(self.q * self.cells)
        _check(self.L.orc_logical(self.h, out))
        return out.reshape(self.q, self.cells)

    def set_logical(self, pops):
        a = np.ascontiguousarray(pops, dtype=np.float64).reshape(-1)
        _check(self.L.orc_set_logical(self.h, a))

    def raw(self):
        out = np.empty(self.q * self.cells)
        _check(self.L.orc_raw(self.h, out))
        return out.reshape(self.q, self.cells)

    def _map(self, fn, parity):
        out = np.empty(self.q * self.cells, dtype=np.int64)
        _check(fn(self.h, parity, out))
        return out.reshape(self.q, self.cells)

    def load_map(self, parity):
        return self._map(self.L.orc_load_map, parity)

    def store_map(self, parity):
        return self._map(self.L.orc_store_map, parity)

    def bb_load_map(self, parity):
        return self._map(self.L.orc_bb_load_map, parity)

    @property
    def time_step(self):
        return int(self.L.orc_time_step(self.h))

    @time_step.setter
    def time_step(self, t):
        self.L.orc_set_time_step(self.h, t)


@contextlib.contextmanager
def exact_fp64_weights():
    """FP64 oracles created and stepped inside use the lattice weights' exact values instead of the
    reference's FP32 constants (a precision-study switch with no reference counterpart)."""
    L = lib()
    L.orc_set_exact_fp64_weights.argtypes = [C.c_int]
    L.orc_set_exact_fp64_weights(1)
    try:
        yield
    finally:
        L.orc_set_exact_fp64_weights(0)


def f32_to_f16_bits(values):
    L = lib()
    return np.array([L.orc_f32_to_f16(float(v)) for v in np.asarray(values, dtype=np.float32)], dtype=np.uint16)


def f16_bits_to_f32(bits):
    L = lib()
    return np.array([L.orc_f16_to_f32(int(b)) for b in np.asarray(bits, dtype=np.uint16)], dtype=np.float32)


def splitmix64(values):
    """The SplitMix64 output function of the oracle, elementwise over uint64 values."""
    a = np.ascontiguousarray(values, dtype=np.uint64).reshape(-1)
    out = np.empty_like(a)
    lib().orc_splitmix64(a, out, a.size)
    return out


def uniform01_cells(seed, comp, n):
    """The oracle's noise variates uniform01(seed, cell, comp) for cells 0..n-1."""
    out = np.empty(n)
    lib().orc_uniform01_cells(seed, comp, n, out)
    return out


def uniform01(seed, cell, comp):
    return lib().orc_uniform01(seed, cell, comp)


def _tri_args(triangles):
    t = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1)
    assert t.size % 9 == 0
    return t, t.size // 9


def _vox(fn, L, prefix, triangles, shape, lo, hi):
    nx, ny, nz = shape
    t, n = _tri_args(triangles)
    mask = np.zeros(nx * ny * nz, np.uint8)
    _check(fn(t, n, nx, ny, nz, np.asarray(lo, np.float64), np.asarray(hi, np.float64), mask), L, prefix)
    return mask.reshape(nz, ny, nx)


def _xform(fn, L, prefix, triangles, translation, axis, angle):
    t, n = _tri_args(triangles)
    out = np.empty_like(t)
    _check(fn(t, n, np.asarray(translation, np.float64), np.asarray(axis, np.float64), float(angle), out), L, prefix)
    return out.reshape(-1, 3, 3)


def voxelize(triangles, shape, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Ray-parity mask (nz, ny, nx) of voxelize.cpp:5, restated (every cell against every triangle)."""
    L = lib()
    return _vox(L.orc_voxelize, L, "orc", triangles, shape, lo, hi)


def transform_surface(triangles, translation, axis, angle):
    L = lib()
    return _xform(L.orc_transform_surface, L, "orc", triangles, translation, axis, angle)


def ref_voxelize(triangles, shape, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), native=False):
    """The reference's own cfd::lbm::voxelize_surface, compiled from src/lbm/voxelize.cpp."""
    L = ref_lib(native)
    return _vox(L.ref_voxelize, L, "ref", triangles, shape, lo, hi)


def ref_checkpoint_tg(shape, tau, amplitude, steps, path):
    """The reference's DistributedPullBlock<D3Q19> (one brick): Taylor-Green start, `steps` virtual
    steps, CFDCHK3 file at `path` (distributed_pull.hpp:176-203)."""
    L = ref_lib()
    _check(L.ref_dist_tg_checkpoint(*shape, tau, amplitude, steps, os.fsencode(path)), L, "ref")


def ref_checkpoint_restart(shape, tau, path_in, steps, path_out):
    """The reference block restored from `path_in` (:205-245), `steps` steps, saved to `path_out`."""
    L = ref_lib()
    _check(L.ref_dist_restart(*shape, tau, os.fsencode(path_in), steps, os.fsencode(path_out)), L, "ref")


def ref_transform_surface(triangles, translation, axis, angle):
    L = ref_lib()
    return _xform(L.ref_transform_surface, L, "ref", triangles, translation, axis, angle)


class Resident:
    """The resident coupling modules (resident_multiphysics_sycl.hpp:54-408) restated on float numpy
    arrays of shape (nz, ny, nx), periodic.  Velocity and acceleration fields are (3, nz, ny, nx)."""

    def __init__(self, nx, ny, nz):
        self.L = lib()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (nz, ny, nx)
        self.cells = nx * ny * nz

    def _f(self, a, n=1):
        a = np.ascontiguousarray(a, np.float32).reshape(-1)
        assert a.size == n * self.cells
        return a

    # D3Q7 thermal
    def thermal_equilibrium(self, temperature):
        g = np.empty(7 * self.cells, np.float32)
        t = self._f(temperature)
        self.L.orc_thermal_fill(g, t.ctypes.data, 0.0, self.cells)
        return g

    def thermal_step(self, g, u, omega, steps=1):
        ux, uy, uz = (self._f(c) for c in u)
        a, b = np.ascontiguousarray(g, np.float32).copy(), np.empty(7 * self.cells, np.float32)
        for _ in range(steps):
            self.L.orc_thermal_step(a, b, ux, uy, uz, self.nx, self.ny, self.nz, float(omega))
            a, b = b, a
        return a

    def thermal_scalar(self, g):
        t = np.empty(self.cells, np.float32)
        self.L.orc_thermal_scalar(np.ascontiguousarray(g, np.float32), t, self.cells)
        return t.reshape(self.shape)

    def boussinesq(self, temperature, accel, t_ref, beta, gravity):
        a = [self._f(c).copy() for c in accel]
        self.L.orc_boussinesq(self._f(temperature), *a, self.cells, float(t_ref), float(beta), *map(float, gravity))
        return np.stack([c.reshape(self.shape) for c in a])

    # volume of fluid
    def half_space(self, x):
        phi = np.empty(self.cells, np.float32)
        self.L.orc_vof_half_space(phi, self.nx, self.ny, self.nz, float(x))
        return phi.reshape(self.shape)

    def advect(self, phi, u, dt):
        out = np.empty(self.cells, np.float32)
        ux, uy, uz = (self._f(c) for c in u)
        self.L.orc_vof_advect(self._f(phi), out, ux, uy, uz, self.nx, self.ny, self.nz, float(dt))
        return out.reshape(self.shape)

    def geometry(self, phi):
        geo = np.empty(4 * self.cells, np.float32)
        self.L.orc_vof_geometry(self._f(phi), geo, self.nx, self.ny, self.nz)
        return geo.reshape((4,) + self.shape)

    def surface_tension(self, phi, accel, sigma, density=1.0):
        curvature = self.geometry(phi)[3]
        a = [self._f(c).copy() for c in accel]
        self.L.orc_vof_tension(self._f(phi), self._f(curvature), *a, self.nx, self.ny, self.nz, float(sigma),
                               float(density))
        return np.stack([c.reshape(self.shape) for c in a])

    # particles: rows (x, y, z, vx, vy, vz, mass, response time)
    def particles_advance(self, particles, u, accel, dt, two_way, gravity=(0.0, 0.0, 0.0)):
        p = np.ascontiguousarray(particles, np.float64).reshape(-1, 8)
        s = np.ascontiguousarray(p.T.astype(np.float32)).reshape(-1)
        ux, uy, uz = (self._f(c) for c in u)
        a = [self._f(c).copy() for c in accel]
        self.L.orc_particles_advance(s, len(p), ux, uy, uz, *a, self.nx, self.ny, self.nz, float(dt), int(two_way),
                                     *map(float, gravity))
        return s.reshape(8, -1).T.astype(np.float64), np.stack([c.reshape(self.shape) for c in a])

    def q_criterion(self, u):
        q = np.empty(self.cells, np.float32)
        ux, uy, uz = (self._f(c) for c in u)
        self.L.orc_q_criterion(ux, uy, uz, q, self.nx, self.ny, self.nz)
        return q.reshape(self.shape)


# ---------------------------------------------------------------- compiled reference (oracle/_ref)
_ref = {}


def ref_available(native=False) -> bool:
    return (REF_NATIVE_SO if native else REF_SO).exists()


def ref_lib(native=False):
    key = "native" if native else "exact"
    if key not in _ref:
        path = REF_NATIVE_SO if native else REF_SO
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (needs the reference tree; see oracle/Makefile)")
        L = C.CDLL(str(path))
        L.ref_last_error.restype = C.c_char_p
        L.ref_eso_create.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.POINTER(C.c_void_p)]
        L.ref_eso_create_les.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_int, C.POINTER(C.c_void_p)]
        L.ref_eso_destroy.argtypes = [C.c_void_p]
        L.ref_eso_init_tg.argtypes = [C.c_void_p, C.c_float]
        L.ref_eso_init_uniform.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float]
        L.ref_eso_step.argtypes = [C.c_void_p, C.c_size_t]
        L.ref_eso_macroscopic.argtypes = [C.c_void_p, _fp, _fp, _fp, _fp]
        L.ref_eso_mass.argtypes = [C.c_void_p]
        L.ref_eso_mass.restype = C.c_double
        L.ref_eso_kinetic_energy.argtypes = [C.c_void_p]
        L.ref_eso_kinetic_energy.restype = C.c_double
        L.ref_eso_max_speed.argtypes = [C.c_void_p]
        L.ref_eso_max_speed.restype = C.c_float
        L.ref_eso_logical.argtypes = [C.c_void_p, _fp]
        L.ref_eso_raw.argtypes = [C.c_void_p, _fp]
        L.ref_eso_set_raw.argtypes = [C.c_void_p, _fp]
        L.ref_eso_set_mask.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_eso_set_local_acceleration.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_eso_time_step.argtypes = [C.c_void_p]
        L.ref_eso_time_step.restype = C.c_uint64
        L.ref_pull_create.argtypes = L.ref_eso_create.argtypes
        L.ref_pull_destroy.argtypes = [C.c_void_p]
        L.ref_pull_init_tg.argtypes = [C.c_void_p, C.c_float]
        L.ref_pull_init_uniform.argtypes = [C.c_void_p]
        L.ref_pull_step.argtypes = [C.c_void_p, C.c_size_t]
        L.ref_pull_macroscopic.argtypes = [C.c_void_p, _fp, _fp, _fp, _fp]
        L.ref_prec_create.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                      C.POINTER(C.c_void_p)]
        L.ref_prec_destroy.argtypes = [C.c_void_p]
        L.ref_prec_init_tg.argtypes = [C.c_void_p, C.c_double]
        L.ref_prec_step.argtypes = [C.c_void_p, C.c_size_t]
        L.ref_prec_macroscopic.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.ref_d2q9_create.argtypes = [C.c_size_t, C.c_size_t, C.c_float, C.c_float, C.c_float,
                                      C.POINTER(C.c_void_p)]
        L.ref_d2q9_destroy.argtypes = [C.c_void_p]
        L.ref_d2q9_set_solid.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t]
        L.ref_d2q9_set_wall_velocity.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_float, C.c_float]
        L.ref_d2q9_init_uniform.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float]
        L.ref_d2q9_init_tg.argtypes = [C.c_void_p, C.c_float]
        L.ref_d2q9_step.argtypes = [C.c_void_p, C.c_size_t]
        L.ref_d2q9_fields.argtypes = [C.c_void_p, _fp, _fp, _fp]
        L.ref_d2q9_mass.argtypes = [C.c_void_p]
        L.ref_d2q9_mass.restype = C.c_double
        L.ref_voxelize.argtypes = [_dp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _dp, _dp, _u8p]
        L.ref_dist_tg_checkpoint.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_float, C.c_size_t,
                                             C.c_char_p]
        L.ref_dist_restart.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_char_p, C.c_size_t,
                                       C.c_char_p]
        L.ref_transform_surface.argtypes = [_dp, C.c_size_t, _dp, _dp, C.c_double, _dp]
        _ref[key] = L
    return _ref[key]


class RefEsoteric:
    """The reference's EsotericPullSolver<D>, compiled from its sources."""

    def __init__(self, q, nx, ny, nz, tau, accel=(0.0, 0.0, 0.0), native=False, smagorinsky=False):
        self.L = ref_lib(native)
        self.q = q
        h = C.c_void_p()
        _check(self.L.ref_eso_create_les(q, nx, ny, nz, tau, *accel, int(smagorinsky), C.byref(h)), self.L, "ref")
        self.h = h
        self.cells = nx * ny * (nz if q != 9 else 1)
        self.shape = (nz if q != 9 else 1, ny, nx)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_eso_destroy(self.h)
            self.h = None

    def init_taylor_green(self, a):
        _check(self.L.ref_eso_init_tg(self.h, a), self.L, "ref")

    def init_uniform(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0):
        _check(self.L.ref_eso_init_uniform(self.h, rho, ux, uy, uz), self.L, "ref")

    def step(self, n=1):
        _check(self.L.ref_eso_step(self.h, n), self.L, "ref")

    def macroscopic(self):
        out = [np.empty(self.cells, dtype=np.float32) for _ in range(4)]
        _check(self.L.ref_eso_macroscopic(self.h, *out), self.L, "ref")
        return [o.reshape(self.shape) for o in out]

    def mass(self):
        return self.L.ref_eso_mass(self.h)

    def kinetic_energy(self):
        return self.L.ref_eso_kinetic_energy(self.h)

    def max_speed(self):
        return self.L.ref_eso_max_speed(self.h)

    def logical(self):
        out = np.empty(self.q * self.cells, dtype=np.float32)
        _check(self.L.ref_eso_logical(self.h, out), self.L, "ref")
        return out.reshape(self.q, self.cells)

    def raw(self):
        out = np.empty(self.q * self.cells, dtype=np.float32)
        _check(self.L.ref_eso_raw(self.h, out), self.L, "ref")
        return out.reshape(self.q, self.cells)

    def set_raw(self, values):
        a = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        assert a.size == self.q * self.cells
        _check(self.L.ref_eso_set_raw(self.h, a), self.L, "ref")

    def set_local_acceleration(self, ax, ay=None, az=None):
        """set_local_body_acceleration (esoteric_pull.hpp:144-151); ax None clears it (:153-157)."""
        if ax is None:
            self._lacc = None
            _check(self.L.ref_eso_set_local_acceleration(self.h, None, None, None), self.L, "ref")
            return
        self._lacc = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (ax, ay, az)]
        assert all(a.size == self.cells for a in self._lacc)
        _check(self.L.ref_eso_set_local_acceleration(self.h, *(a.ctypes.data for a in self._lacc)), self.L, "ref")

    def set_mask(self, mask):
        if mask is None:
            _check(self.L.ref_eso_set_mask(self.h, None), self.L, "ref")
            return
        self._mask = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        _check(self.L.ref_eso_set_mask(self.h, self._mask.ctypes.data), self.L, "ref")

    @property
    def time_step(self):
        return int(self.L.ref_eso_time_step(self.h))


class RefPull:
    """The reference's two-grid OneStepPullSolver<D>."""

    def __init__(self, q, nx, ny, nz, tau, accel=(0.0, 0.0, 0.0)):
        self.L = ref_lib()
        h = C.c_void_p()
        _check(self.L.ref_pull_create(q, nx, ny, nz, tau, *accel, C.byref(h)), self.L, "ref")
        self.h = h
        self.cells = nx * ny * (nz if q != 9 else 1)
        self.shape = (nz if q != 9 else 1, ny, nx)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_pull_destroy(self.h)
            self.h = None

    def init_taylor_green(self, a):
        _check(self.L.ref_pull_init_tg(self.h, a), self.L, "ref")

    def init_uniform(self):
        _check(self.L.ref_pull_init_uniform(self.h), self.L, "ref")

    def step(self, n=1):
        _check(self.L.ref_pull_step(self.h, n), self.L, "ref")

    def macroscopic(self):
        out = [np.empty(self.cells, dtype=np.float32) for _ in range(4)]
        _check(self.L.ref_pull_macroscopic(self.h, *out), self.L, "ref")
        return [o.reshape(self.shape) for o in out]


class RefPrecision:
    """The reference's PrecisionPullSolver<D, double> (two-grid, periodic, unforced FP64)."""

    def __init__(self, q, nx, ny, nz, tau):
        self.L = ref_lib()
        h = C.c_void_p()
        _check(self.L.ref_prec_create(q, nx, ny, nz, tau, C.byref(h)), self.L, "ref")
        self.h = h
        self.cells = nx * ny * (nz if q != 9 else 1)
        self.shape = (nz if q != 9 else 1, ny, nx)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_prec_destroy(self.h)
            self.h = None

    def init_taylor_green(self, a):
        _check(self.L.ref_prec_init_tg(self.h, a), self.L, "ref")

    def step(self, n=1):
        _check(self.L.ref_prec_step(self.h, n), self.L, "ref")

    def macroscopic(self):
        out = [np.empty(self.cells) for _ in range(4)]
        _check(self.L.ref_prec_macroscopic(self.h, *out), self.L, "ref")
        return [o.reshape(self.shape) for o in out]


class RefD2Q9:
    """The reference's two-grid D2Q9Solver (halfway bounce-back, moving walls)."""

    def __init__(self, nx, ny, tau, ax=0.0, ay=0.0):
        self.L = ref_lib()
        h = C.c_void_p()
        _check(self.L.ref_d2q9_create(nx, ny, tau, ax, ay, C.byref(h)), self.L, "ref")
        self.h = h
        self.nx, self.ny = nx, ny

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_d2q9_destroy(self.h)
            self.h = None

    def set_solid(self, x, y):
        _check(self.L.ref_d2q9_set_solid(self.h, x, y), self.L, "ref")

    def set_wall_velocity(self, x, y, ux, uy):
        _check(self.L.ref_d2q9_set_wall_velocity(self.h, x, y, ux, uy), self.L, "ref")

    def init_uniform(self, rho=1.0, ux=0.0, uy=0.0):
        _check(self.L.ref_d2q9_init_uniform(self.h, rho, ux, uy), self.L, "ref")

    def init_taylor_green(self, a):
        _check(self.L.ref_d2q9_init_tg(self.h, a), self.L, "ref")

    def step(self, n=1):
        _check(self.L.ref_d2q9_step(self.h, n), self.L, "ref")

    def fields(self):
        n = self.nx * self.ny
        out = [np.empty(n, dtype=np.float32) for _ in range(3)]
        _check(self.L.ref_d2q9_fields(self.h, *out), self.L, "ref")
        return [o.reshape(self.ny, self.nx) for o in out]

    def mass(self):
        return self.L.ref_d2q9_mass(self.h)


def _adjoint_bindings(L):
    L.orc_objective.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.orc_set_drag_body.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.orc_adjoint_run.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                  C.POINTER(C.c_double), _dp, C.POINTER(C.c_double)]


_KINDS = {"pressure_drop": 0, "drag": 1}


def _objective(self, x_in=0, x_out=0, kind="pressure_drop"):
    _adjoint_bindings(self.L)
    out = C.c_double()
    _check(self.L.orc_objective(self.h, _KINDS[kind], x_in, x_out, C.byref(out)))
    return out.value


def _set_drag_body(self, body, axis=0):
    _adjoint_bindings(self.L)
    b = np.ascontiguousarray(np.asarray(body).reshape(-1) != 0, dtype=np.uint8)
    _check(self.L.orc_set_drag_body(self.h, b.ctypes.data, axis))


def _adjoint_run(self, steps, window, x_in=0, x_out=0, kind="pressure_drop"):
    """Forward `steps` steps, J over the last `window` states, and dJ/dgamma (per cell), dJ/du_peak."""
    _adjoint_bindings(self.L)
    J, du = C.c_double(), C.c_double()
    grad = np.zeros(self.cells)
    _check(self.L.orc_adjoint_run(self.h, _KINDS[kind], steps, window, x_in, x_out, C.byref(J), grad, C.byref(du)))
    return J.value, grad.reshape(self.shape), du.value


Oracle.objective = _objective
Oracle.set_drag_body = _set_drag_body
Oracle.adjoint_run = _adjoint_run
