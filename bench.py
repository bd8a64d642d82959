#!/usr/bin/env python3
"""Benchmark of the gfx950 esoteric-pull stream-collide (one GPU).

    python bench.py [--steps K] [--warmup W]
                    [--legs fp32,fp16,fp32s,fp64,d3q27,d2q9,adjoint,voxelize,resident,shard,config1,config5]
                    [--config5-steps T] [--config5-storages fp16s,fp32] [--no-cpu]

A step is one fused in-place stream-collide launch over the whole lattice, inputs resident in
HBM.  The headline workload is a D3Q19 256^3 lid-driven cavity (walls on five faces, lid
u = (0.1, 0, 0), tau 0.55, uint8 flags) with FP32 population storage; further legs run the same
geometry with FP16 and FP32 shifted-population storage, FP64 (the reference for the velocity
error of the others) and D3Q27, the d2q9 leg a 4096^2 D2Q9 cavity (FP32 and FP16), and the adjoint leg runs the
checkpointed discrete adjoint of a 128^3 duct (pressure-drop objective, 2000 steps), and the
voxelize leg the GPU STL voxelisation of a 16128-triangle sphere on 512^3 cells, and the resident
leg the device-resident coupling modules (thermal, fill, particles, Q-criterion) at 256^3, and the shard
leg the kernel and its halo pack/unpack on one GPU's share of a 768^3 lattice split in eight
z-slabs, and config1 the
cache-resident 64^3 Taylor-Green correctness case, and the config5
leg forward + adjoint of a 768^3 duct past a sphere cluster (drag objective, binomial
checkpointing) with FP16 and FP32 storage, on a horizon truncated to 200 steps by default, with the
plan's cost at the specified 10^4 steps projected from the measured step times
(`--config5-steps 10000` runs it: ~10 min with FP16 storage).  Prints progress on stderr and ONE JSON
line on stdout.  MLUPS counts every lattice cell per step, as the reference's cfd-bench does
(apps/cfd-bench.cpp:111-117), and `fluid_mlups` beside it counts only the cells a step updates
(wall cells return at once); the roofline uses algorithmic HBM bytes per launch (each fluid
population read and written once; the bounded kernel's metadata -- one 4-byte span per 64-cell row
segment and a 4-byte link descriptor per cell inside a span -- is overhead, not algorithmic, and
shows in the measured `traffic`) over the kernel time measured with HIP events on the solver's
stream.  Two probes run in the same process: a plain 2 GiB copy
and the in-place esoteric-pull access pattern alone (the bandwidth ceiling of an in-place
stream-collide), with the headline kernel's fraction of the latter.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "D3Q19 MLUPS fwd FP32/FP16-store + adjoint, % of 8 TB/s HBM roofline"
PEAK_GBS = 8000.0
BYTES_PER_POP = {"fp32": 4, "fp16s": 2, "fp64": 8, "fp32s": 4}
ADJ_BYTES = 3 * 19 * 4 + 1 + 4 + 16
TRAFFIC_FILE = ROOT / "profiles" / "traffic_stream_collide_d3q19_256_cavity.json"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cavity_flags(nx, ny, nz):
    """Lid y = ny-1 moving (wall-velocity slot 1); walls x = 0, nx-1, y = 0, z = 0, nz-1 at rest.
    Side walls are written after the lid, as apps/cfd-solve.cpp:258-265 does in 2-D."""
    f = np.zeros((nz, ny, nx), dtype=np.uint8)
    f[:, -1, :] = 0x81
    f[:, :, 0] = f[:, :, -1] = 0x80
    f[:, 0, :] = 0x80
    f[0, :, :] = f[-1, :, :] = 0x80
    return f.reshape(-1)


def run_leg(hip, name, lattice, storage, n, tau, steps, warmup):
    log(f"leg {name}: {lattice} {n}^3 {storage} cavity, warmup {warmup}, steps {steps}")
    q = {"d3q19": 19, "d3q27": 27}[lattice]
    s = hip.EsotericPullHipSolver(lattice, n, n, n, tau, storage=storage)
    flags = cavity_flags(n, n, n)
    s.set_flags(flags)
    s.set_wall_velocity(1, (0.1, 0.0, 0.0))
    s.initialize_uniform()
    mass0 = s.total_mass()
    s.step(warmup)
    s.wait()
    t0 = time.perf_counter()
    gpu_ms = s.step_timed(steps)  # events on the solver's stream; synchronises at the end
    s.wait()
    wall = time.perf_counter() - t0
    # After the timed region: the same kernel over a longer run.  Launch times fall over the first
    # tens of launches of a fresh lattice, so a short timed region reads a few % slow; this figure
    # is reported beside the timed one, which remains the one the roofline uses.
    steady_steps = max(200, steps)
    steady_ms = s.step_timed(steady_steps) / steady_steps
    mass1, ke, vmax = s.diagnostics()
    m = s.download_macroscopic()
    fluid_mask = flags == 0
    velocity = np.stack([m.ux.reshape(-1)[fluid_mask], m.uy.reshape(-1)[fluid_mask],
                         m.uz.reshape(-1)[fluid_mask]]).astype(np.float64)
    cells = n ** 3
    fluid = int(np.count_nonzero(flags == 0))
    bytes_launch = fluid * 2 * q * BYTES_PER_POP[storage]
    launch_s = gpu_ms * 1e-3 / steps
    gbs = bytes_launch / launch_s / 1e9
    res = {
        "lattice": lattice.upper(), "storage": storage, "n": n, "cells": cells, "fluid_cells": fluid,
        "steps": steps, "warmup": warmup, "seconds": wall, "ms_per_step": wall * 1e3 / steps,
        "mlups": cells * steps / wall / 1e6, "fluid_mlups": fluid * steps / wall / 1e6, "kernel_ms": gpu_ms / steps,
        "kernel_mlups": cells / launch_s / 1e6, "algorithmic_bytes_per_launch": bytes_launch,
        "bytes_per_fluid_cell_update": 2 * q * BYTES_PER_POP[storage],
        "achieved_gbs": gbs, "roofline_frac": gbs / PEAK_GBS,
        "steady_state": {"steps": steady_steps, "kernel_ms": steady_ms,
                         "gbs": bytes_launch / (steady_ms * 1e-3) / 1e9,
                         "roofline_frac": bytes_launch / (steady_ms * 1e-3) / 1e9 / PEAK_GBS},
        "mass_rel_drift": abs(mass1 - mass0) / mass0, "max_speed": vmax,
        "finite": bool(np.isfinite(mass1) and np.isfinite(vmax)),
    }
    log(f"leg {name}: {res['mlups']:.0f} MLUPS, kernel {res['kernel_ms']:.3f} ms, {gbs:.0f} GB/s "
        f"({100 * gbs / PEAK_GBS:.1f}% of 8 TB/s), mass drift {res['mass_rel_drift']:.2e}; "
        f"over {steady_steps} more steps {steady_ms:.3f} ms ({100 * res['steady_state']['roofline_frac']:.1f}%)")
    s.close()
    return res, velocity


def d2q9_leg(hip, n=4096, steps=20, warmup=5, storages=("fp32", "fp16s")):
    """D2Q9 lid-driven cavity on n x n cells (16.8 M at 4096^2, as many as the 256^3 legs): lid
    y = n-1 moving at (0.1, 0), walls x = 0, n-1 and y = 0 at rest, tau 0.55, FP32 and FP16S
    storage.  Algorithmic bytes: each fluid cell's 9 populations read and written once."""
    out = {"workload": f"D2Q9 {n}^2 lid-driven cavity, lid u=(0.1,0), tau 0.55"}
    f = np.zeros((n, n), dtype=np.uint8)
    f[-1, :] = 0x81
    f[:, 0] = f[:, -1] = 0x80
    f[0, :] = 0x80
    flags = f.reshape(-1)
    fluid = int(np.count_nonzero(flags == 0))
    for storage in storages:
        s = hip.EsotericPullHipSolver("d2q9", n, n, 1, 0.55, storage=storage)
        s.set_flags(flags)
        s.set_wall_velocity(1, (0.1, 0.0, 0.0))
        s.initialize_uniform()
        s.step(warmup)
        s.wait()
        gpu_ms = s.step_timed(steps)
        steady_ms = s.step_timed(200) / 200
        mass, _, vmax = s.diagnostics()
        s.close()
        bytes_launch = fluid * 2 * 9 * BYTES_PER_POP[storage]
        kernel_ms = gpu_ms / steps
        res = {"kernel_ms": kernel_ms, "kernel_mlups": n * n / (kernel_ms * 1e3),
               "algorithmic_bytes_per_launch": bytes_launch, "gbs": bytes_launch / (kernel_ms * 1e-3) / 1e9,
               "roofline_frac": bytes_launch / (kernel_ms * 1e-3) / 1e9 / PEAK_GBS,
               "steady_state": {"steps": 200, "kernel_ms": steady_ms,
                                "roofline_frac": bytes_launch / (steady_ms * 1e-3) / 1e9 / PEAK_GBS},
               "finite": bool(np.isfinite(mass) and np.isfinite(vmax))}
        out[storage] = res
        log(f"d2q9 {storage}: {res['kernel_mlups']:.0f} MLUPS, kernel {kernel_ms:.3f} ms, {res['gbs']:.0f} GB/s "
            f"({100 * res['roofline_frac']:.1f}% of 8 TB/s); over 200 more steps {steady_ms:.3f} ms "
            f"({100 * res['steady_state']['roofline_frac']:.1f}%)")
    return out


def config1_leg(hip, n=64, steps=100, warmup=10):
    """Config 1: D3Q19 64^3 fully periodic, FP32, tau 0.6, Taylor-Green A 0.05 plus U(-1e-4, 1e-4)
    noise on u (seed 1), 10 warmup + 100 timed steps.  The 19.9 MB state stays in the caches, so
    its MLUPS is not a roofline figure (flagged cache_resident)."""
    s = hip.EsotericPullHipSolver("d3q19", n, n, n, 0.6)
    s.initialize_taylor_green(0.05, 1e-4, 1)
    s.step(warmup)
    s.wait()
    gpu_ms = s.step_timed(steps)
    out = {"workload": f"D3Q19 {n}^3 periodic Taylor-Green A 0.05 + noise 1e-4 (seed 1), FP32, tau 0.6",
           "kernel_ms": gpu_ms / steps, "mlups": n ** 3 * steps / (gpu_ms * 1e-3) / 1e6, "cache_resident": True,
           "mass": s.total_mass()}
    log(f"config1: {out['mlups']:.0f} MLUPS (cache-resident, not a roofline figure), kernel {gpu_ms / steps * 1e3:.1f} us")
    s.close()
    return out


class _DeviceBuffers:
    """hipMalloc / hipFree through the HIP runtime the solver library loaded (by soname)."""

    def __init__(self):
        import ctypes as C

        self.C = C
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        if self.rt.hipMalloc(self.C.byref(p), int(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.live.append(p.value)
        return p.value

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live.clear()


def shard_leg(hip, shape=(768, 768, 96), steps=100, warmup=10, storages=("fp16s", "fp32"), exchanges=50):
    """The kernel at one GPU's share of a z-slab decomposition: 768 x 768 x 96 cells is one of eight
    slabs of a 768^3 lattice, periodic Taylor-Green start (A 0.02, tau 0.6).  The slab runs the
    periodic kernel; after each step a distributed run packs the crossing populations of its two
    faces into an up and a down message (5 planes of 768 x 768 each), exchanges them with the
    neighbouring slabs and unpacks what it receives (cfd_lbm_halo_pack / cfd_lbm_halo_unpack).  On
    one GPU the slab is its own neighbour; the pack + unpack pair is timed here (wall time of
    `exchanges` pairs between two synchronisations, launch gaps included, so an upper bound), the
    transfer between GPUs is not."""
    nx, ny, nz = shape
    cells = nx * ny * nz
    out = {"workload": f"D3Q19 {nx}x{ny}x{nz} periodic Taylor-Green slab (one GPU's share of 768^3 on 8)"}
    bufs = _DeviceBuffers()
    for storage in storages:
        s = hip.EsotericPullHipSolver("d3q19", nx, ny, nz, 0.6, storage=storage)
        s.initialize_taylor_green(0.02)
        s.step(warmup)
        s.wait()
        gpu_ms = s.step_timed(steps)
        nbytes = cells * 2 * 19 * BYTES_PER_POP[storage]
        launch_s = gpu_ms * 1e-3 / steps
        msg = s.halo_message_bytes()
        up, down = bufs.alloc(msg), bufs.alloc(msg)
        s.halo_pack(up, down)
        s.halo_unpack(up, down)
        s.wait()
        t0 = time.perf_counter()
        for _ in range(exchanges):
            s.halo_pack(up, down)
            s.halo_unpack(up, down)
        s.wait()
        pair_ms = (time.perf_counter() - t0) * 1e3 / exchanges
        out[storage] = {"kernel_ms": gpu_ms / steps, "mlups": cells / launch_s / 1e6,
                        "algorithmic_bytes_per_launch": nbytes, "gbs": nbytes / launch_s / 1e9,
                        "roofline_frac": nbytes / launch_s / 1e9 / PEAK_GBS, "finite": bool(np.isfinite(s.total_mass())),
                        "halo": {"message_bytes": msg, "messages_per_step": 2,
                                 "pack_unpack_ms_upper_bound": pair_ms,
                                 "fraction_of_step": pair_ms / (gpu_ms / steps)}}
        log(f"shard {storage}: {out[storage]['mlups']:.0f} MLUPS, kernel {gpu_ms / steps:.3f} ms, "
            f"{out[storage]['gbs']:.0f} GB/s ({100 * out[storage]['roofline_frac']:.1f}% of 8 TB/s); "
            f"halo pack + unpack of 2 x {msg / 1e6:.1f} MB <= {pair_ms * 1e3:.0f} us per step")
        s.close()
        bufs.free()
    return out


def duct_flags(n):
    f = np.zeros((n, n, n), dtype=np.uint8)
    f[:, 0, :] = f[:, -1, :] = 0x80
    f[0, :, :] = f[-1, :, :] = 0x80
    f[1:-1, 1:-1, 0] = 0x01
    f[1:-1, 1:-1, -1] = 0x02
    return f.reshape(-1)


def adjoint_leg(hip, n=128, steps=2000, window=200, every=100, kernel_steps=100):
    """Config-4 workload: D3Q19 n^3 duct, velocity inlet (peak 0.05), pressure outlet, walls in y/z,
    partial bounce-back blend 0.01 * U[0.3, 0.7] (seed 4); J = mean pressure drop over the last
    `window` of `steps` states; snapshots every `every` steps.  A second, short run with a snapshot
    every step (no recomputation) times the transposed-step kernel alone."""
    log(f"leg adjoint: D3Q19 {n}^3 duct, T {steps}, window {window}, snapshots every {every}")
    flags = duct_flags(n)
    rng = np.random.default_rng(4)
    blend = np.where(flags == 0, 0.01 * rng.uniform(0.3, 0.7, flags.size), 0.0).astype(np.float32)

    def solver():
        s = hip.EsotericPullHipSolver("d3q19", n, n, n, 0.6, storage="fp32")
        s.set_flags(flags)
        s.set_inlet_outlet(0.05, 1.0)
        s.set_porosity(blend)
        s.initialize_uniform()
        return s

    s = solver()
    t0 = time.perf_counter()
    J, grad, du = s.adjoint_run(steps, window, every, 1, n - 2)
    wall = time.perf_counter() - t0
    fwd_ms, rev_ms = s.adjoint_timing()
    s.close()
    # The same problem with binomial checkpointing in as many state buffers as fit in HBM.
    s = solver()
    Jb, gradb, dub, fw_b = s.adjoint_run_binomial(steps, window, 0, "pressure_drop", 1, n - 2)
    fwd_b, rev_b = s.adjoint_timing()
    s.close()
    same = bool(Jb == J and dub == du and np.array_equal(gradb, grad))
    del grad, gradb
    s = solver()
    s.adjoint_run(kernel_steps, 1, 1, 1, n - 2, want_gamma=False)
    _, kern_ms = s.adjoint_timing()
    s.close()
    cells = n ** 3
    fluid = int(np.count_nonzero(flags == 0)) + int(np.count_nonzero((flags == 1) | (flags == 2)))
    adj_step_ms = kern_ms / kernel_steps
    adj_bytes = fluid * ADJ_BYTES  # lambda read + write, f read, flag, gamma, FP64 gradient read + write
    res = {
        "workload": f"D3Q19 {n}^3 duct, inlet peak 0.05, outlet rho 1, PBB blend 0.01*U[0.3,0.7], FP32",
        "steps": steps, "window": window, "objective": J, "du_peak": du,
        "two_level": {"snapshot_every": every, "wall_seconds": wall, "forward_ms": fwd_ms, "reverse_ms": rev_ms,
                      "reverse_over_forward": rev_ms / fwd_ms, "forward_steps": 2 * steps - (steps + every - 1) // every},
        "binomial": {"forward_ms": fwd_b, "reverse_ms": rev_b, "reverse_over_forward": rev_b / fwd_b,
                     "forward_steps": fw_b, "identical_to_two_level": same},
        "forward_mlups": cells * steps / (fwd_ms * 1e3), "reverse_mlups": cells * steps / (rev_ms * 1e3),
        "adjoint_kernel_ms": adj_step_ms, "adjoint_kernel_mlups": cells / (adj_step_ms * 1e3),
        "adjoint_bytes_per_cell_update": ADJ_BYTES, "adjoint_kernel_gbs": adj_bytes / (adj_step_ms * 1e-3) / 1e9,
    }
    res["adjoint_kernel_roofline_frac"] = res["adjoint_kernel_gbs"] / PEAK_GBS
    log(f"leg adjoint: two-level forward {fwd_ms:.1f} ms, reverse {rev_ms:.1f} ms ({rev_ms / fwd_ms:.2f}x); "
        f"binomial reverse {rev_b:.1f} ms ({rev_b / fwd_b:.2f}x, {fw_b} forward steps, identical {same}); "
        f"transposed step {adj_step_ms * 1e3:.1f} us = {res['adjoint_kernel_mlups']:.0f} MLUPS, "
        f"{res['adjoint_kernel_gbs']:.0f} GB/s")
    return res


def config5_leg(hip, n=768, steps=200, storages=("fp16s", "fp32"), kernel_steps=4, horizon=10000):
    """Flow past a cluster: D3Q19 n^3 duct (inlet peak 0.04, outlet rho 1), 64 non-overlapping spheres
    of radius U[10, 30] (seed 5) with centres in the middle 40 % of each axis, a partial-bounce-back
    halo ramping from 0.5 at the sphere surfaces to 0 four cells out, tau 0.52, impulsive start
    (u = (0.04, 0, 0) everywhere, so that the cluster carries drag from the first step); J = drag
    along x averaged over all `steps` states; binomial checkpointing with as many state buffers as fit.
    The horizon is truncated to `steps` (the specified T is `horizon` = 10^4 steps, which takes
    minutes per storage); the cost of the plan at T = `horizon` with the same buffers is reported
    from the closed form and the measured per-step times.  The transposed step's time is the long
    run's reverse phase less its recomputed steps at the forward sweep's rate; a second, short run
    per storage with a buffer for every state (`kernel_steps` steps, no recomputation) times it
    alone as a check (its first launches run slow).  The FP16-storage run is compared with the
    FP32-storage run of the same problem."""
    from cfd_solvers_amd import geometry as G

    t0 = time.perf_counter()
    flags, body, halo = G.sphere_cluster_duct(n, seed=5)
    halo_cells = np.flatnonzero(halo > 0)
    fluid = int(np.count_nonzero(flags == 0)) + int(np.count_nonzero((flags == 1) | (flags == 2)))
    log(f"leg config5: geometry {time.perf_counter() - t0:.1f} s, {int(body.sum())} body cells, "
        f"{halo_cells.size} halo cells")
    cells = n ** 3
    out = {"workload": f"D3Q19 {n}^3 duct, 64-sphere cluster + PBB halo, tau 0.52, inlet 0.04, drag",
           "cells": cells, "fluid_cells": fluid, "steps": steps, "horizon_specified": horizon,
           "truncated": steps < horizon, "spheres": 64, "body_cells": int(body.sum()),
           "halo_cells": int(halo_cells.size)}
    grads = {}

    def solver(storage):
        s = hip.EsotericPullHipSolver("d3q19", n, n, n, 0.52, storage=storage)
        s.set_flags(flags)
        s.set_drag_body(body, 0)
        s.set_inlet_outlet(0.04, 1.0)
        s.set_porosity(halo)
        s.initialize_uniform(1.0, 0.04, 0.0, 0.0)  # impulsive start: the inflow everywhere
        return s

    if steps >= 2000:  # minutes per run: the library reports its progress every 1000 steps
        os.environ.setdefault("CFD_LBM_PROGRESS_STEPS", "1000")
    for storage in storages:
        elem = BYTES_PER_POP[storage]
        log(f"leg config5: {storage} {n}^3, T {steps}")
        s = solver(storage)
        t0 = time.perf_counter()
        J, grad, du, fw = s.adjoint_run_binomial(steps, steps, 0, "drag")
        wall = time.perf_counter() - t0
        fwd_ms, rev_ms = s.adjoint_timing()
        s.close()
        grads[storage] = grad.reshape(-1)[halo_cells]
        del grad
        # The same problem with a buffer for every state: the reverse phase is transposed steps only.
        s = solver(storage)
        s.adjoint_run_binomial(kernel_steps, kernel_steps, kernel_steps - 1, "drag", want_gamma=False)
        kfwd_ms, krev_ms = s.adjoint_timing()
        s.close()
        step_ms = fwd_ms / steps
        # The transposed step from the long run: its reverse phase less the recomputed steps at the
        # forward sweep's rate, per step (the short run alone carries the first launches' ramp).
        adj_ms = (rev_ms - (fw - steps) * step_ms) / steps
        alone_ms = krev_ms / kernel_steps
        # Buffers of the truncated run: solve the plan's closed form backwards for the snapshot
        # count that gives the measured forward steps, then cost the specified horizon with it.
        snaps = next((k for k in range(1, 4096) if hip.checkpoint_plan(steps, k)[0] == fw), None)
        adj_bytes = fluid * (2 * 19 * 4 + 19 * elem + 4 + 16)  # lambda r/w (FP32), f read, gamma, FP64 grad r/w
        res = {"objective": J, "du_peak": du, "forward_steps": fw, "wall_seconds": wall,
               "forward_ms": fwd_ms, "reverse_ms": rev_ms, "forward_step_ms": step_ms,
               "forward_mlups": cells * steps / (fwd_ms * 1e3), "adjoint_mlups": cells * steps / (rev_ms * 1e3),
               "reverse_over_forward": rev_ms / fwd_ms, "total_over_forward": (fwd_ms + rev_ms) / fwd_ms,
               "snapshots": snaps,
               "transposed_step": {"ms": adj_ms, "over_forward_step": adj_ms / step_ms,
                                   "ms_alone_short_run": alone_ms, "short_run_steps": kernel_steps,
                                   "algorithmic_bytes": adj_bytes,
                                   "gbs": adj_bytes / (adj_ms * 1e-3) / 1e9,
                                   "roofline_frac": adj_bytes / (adj_ms * 1e-3) / 1e9 / PEAK_GBS}}
        if snaps is not None:
            hf, hr, _ = hip.checkpoint_plan(horizon, snaps)
            proj_rev = (hf - horizon) * step_ms + horizon * adj_ms
            res["plan_at_horizon"] = {"steps": horizon, "snapshots": snaps, "forward_steps": hf, "repetitions": hr,
                                      "projected_reverse_over_forward": proj_rev / (horizon * step_ms),
                                      "projected_seconds": (hf * step_ms + horizon * adj_ms) * 1e-3}
        out[storage] = res
        log(f"leg config5 {storage}: J {J:.6e}, du {du:.6e}, forward {fwd_ms:.0f} ms, reverse {rev_ms:.0f} ms "
            f"({rev_ms / fwd_ms:.2f}x, {fw} forward steps); transposed step {adj_ms:.2f} ms = "
            f"{adj_ms / step_ms:.2f}x a forward step, {res['transposed_step']['gbs']:.0f} GB/s; "
            f"T={horizon}: {res.get('plan_at_horizon')}")
    if "fp16s" in grads and "fp32" in grads:
        a, b = grads["fp16s"], grads["fp32"]
        out["fp16_vs_fp32"] = {
            "objective_rel": abs(out["fp16s"]["objective"] - out["fp32"]["objective"]) / abs(out["fp32"]["objective"]),
            "du_peak_rel": (abs(out["fp16s"]["du_peak"] - out["fp32"]["du_peak"]) / abs(out["fp32"]["du_peak"])
                            if out["fp32"]["du_peak"] != 0 else None),
            "grad_rel_l2_halo": float(np.linalg.norm(a - b) / np.linalg.norm(b))}
        log(f"leg config5: FP16 vs FP32 {out['fp16_vs_fp32']}")
    return out


def voxelize_leg(hip, n=512, reps=5):
    """GPU STL voxelisation (cfd_lbm_voxelize_surface, src/lbm/voxelize.cpp:5): a closed
    16128-triangle sphere (radius 0.37) on an n^3 grid over [0, 1]^3.  Kernel times are HIP events
    on the call's stream (binning: count + scan + fill; lines: the per-line parity kernel); the
    wall time includes the triangle upload and the mask download over PCIe.  The reference's CPU
    figure for the same mesh is taken in cpu_baseline()."""
    from cfd_solvers_amd import geometry as G

    tris = G.sphere_surface((.5, .5, .5), .37, 64, 128)
    hip.voxelize_surface(tris, (n, n, n))  # warmup
    kernel, lines, bins, wall = [], [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        m = hip.voxelize_surface(tris, (n, n, n))
        wall.append((time.perf_counter() - t0) * 1e3)
        b, l, pairs = hip.voxelize_timing()
        bins.append(b)
        lines.append(l)
        kernel.append(b + l)
    k_ms, l_ms, b_ms = (float(np.median(v)) for v in (kernel, lines, bins))
    cells = n ** 3
    res = {"workload": f"{len(tris)}-triangle sphere r=0.37 on {n}^3 cells in [0,1]^3", "cells": cells,
           "triangles": len(tris), "solid_cells": int(m.sum()), "tile_pairs": int(pairs),
           "bin_ms": b_ms, "lines_ms": l_ms, "kernel_ms": k_ms, "wall_ms": float(np.median(wall)),
           "gcells_per_s": cells / (k_ms * 1e-3) / 1e9,
           "mask_write_gbs": cells / (l_ms * 1e-3) / 1e9}
    log(f"voxelize {n}^3, {len(tris)} triangles: bin {b_ms:.3f} ms + lines {l_ms:.3f} ms "
        f"({res['gcells_per_s']:.1f} Gcells/s), wall {res['wall_ms']:.1f} ms incl. PCIe")
    return res


def resident_leg(hip, n=256, reps=50, particles=1 << 20):
    """Device-resident coupling modules on an n^3 D3Q19 Taylor-Green flow (FP32 views): D3Q7 thermal
    step, Boussinesq coupling, fill advection, Q-criterion and a one-way particle push.  Each figure
    is the wall time of `reps` back-to-back launches on the flow's stream between two
    synchronisations (launch gaps included), so it bounds the kernel time from above.  Algorithmic
    bytes per cell: thermal 7 reads + 7 writes + 3 velocity reads = 68 B; Boussinesq 56 B (one pass:
    7 population reads, the temperature write, acceleration read/write); advection fill read + write + 3 velocities = 20 B;
    interface geometry (normals: fill read + 3 writes; curvature: 3 reads + 1 write) 32 B; surface
    tension (geometry, then fill and curvature reads + acceleration read/write) 64 B;
    Q-criterion 3 velocities + 1 write = 16 B; per particle 8 state reads + 6 writes = 56 B plus
    the 8-cell velocity gather."""
    flow = hip.EsotericPullHipSolver("d3q19", n, n, n, 0.6)
    flow.initialize_taylor_green(0.02)
    flow.step(2)
    macro = flow.refresh_device_macroscopic()
    force = flow.device_local_acceleration(True)
    thermal = hip.ResidentThermalD3Q7(flow, 0.02)
    thermal.initialize_uniform(300.0)
    surface = hip.ResidentFreeSurface(flow)
    surface.initialize_half_space_x(n / 2 + 0.5)
    diag = hip.ResidentFlowDiagnostics(flow)
    rng = np.random.default_rng(0)
    pts = np.column_stack([rng.uniform(0, n, (particles, 3)), np.zeros((particles, 3)),
                           np.ones(particles), np.full(particles, 0.5)])
    parts = hip.ResidentParticles(flow, particles)
    parts.upload(pts)
    cells = n ** 3

    def timed(fn, count):
        fn()
        flow.wait()
        t0 = time.perf_counter()
        for _ in range(count):
            fn()
        flow.wait()
        return (time.perf_counter() - t0) * 1e3 / count

    out = {"workload": f"D3Q19 {n}^3 Taylor-Green flow, FP32 views; {particles} particles", "cells": cells,
           "particles": particles}
    for name, fn, nbytes in (
            ("thermal_step", lambda: thermal.step(macro, 1), 68 * cells),
            ("boussinesq", lambda: thermal.add_boussinesq_acceleration(force, 300.0, 1e-4, (0, -1, 0)), 56 * cells),
            ("fill_advect", lambda: surface.advect(macro, 0.1), 20 * cells),
            ("interface_geometry", lambda: surface.refresh_geometry(), 32 * cells),
            ("surface_tension", lambda: surface.add_surface_tension_acceleration(force, 0.072, 1.0), 64 * cells),
            ("q_criterion", lambda: diag.q_criterion(macro), 16 * cells),
            ("particles_one_way", lambda: parts.advance(macro, force, 0.05, False), 56 * particles)):
        ms = timed(fn, reps)
        out[name] = {"ms": ms, "algorithmic_bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
                     "frac_of_peak": nbytes / (ms * 1e-3) / 1e9 / PEAK_GBS}
        log(f"resident {name}: {ms * 1e3:.1f} us per launch, {out[name]['gbs']:.0f} GB/s algorithmic")
    for m in (parts, diag, surface, thermal):
        m.close()
    flow.close()
    return out


def copy_probe():
    import ctypes as C

    lib = C.CDLL(str(ROOT / "cfd_solvers_amd" / "libhbm_probe.so"))
    lib.hbm_probe_copy.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_float)]
    ms = C.c_float()
    nbytes = 1 << 31
    if lib.hbm_probe_copy(nbytes, 20, C.byref(ms)) != 0:
        return None
    gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
    log(f"copy probe: 2 GiB copy {ms.value:.3f} ms -> {gbs:.0f} GB/s (read + write)")
    return {"bytes_per_copy": 2 * nbytes, "ms": ms.value, "gbs": gbs, "frac_of_peak": gbs / PEAK_GBS}


def inplace_probe(n=256, repeats=20):
    """The in-place esoteric-pull access pattern alone (hbm_probe_inplace: D3Q19, periodic, the
    row-blocked layout of the solver's defaults -- 8 rows per block for FP32, 32 for FP16 --
    streaming stores, trivial arithmetic): the bandwidth ceiling of an in-place stream-collide in
    that layout.  19 reads + 19 writes per cell."""
    import ctypes as C

    lib = C.CDLL(str(ROOT / "cfd_solvers_amd" / "libhbm_probe.so"))
    lib.hbm_probe_inplace.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_float)]
    out = {}
    for name, elem, rs in (("fp32", 4, 3), ("fp16", 2, 5)):
        ms = C.c_float()
        if lib.hbm_probe_inplace(elem, n, n, rs, repeats, C.byref(ms)) != 0:
            return None
        nbytes = n ** 3 * 19 * 2 * elem
        gbs = nbytes / (ms.value * 1e-3) / 1e9
        out[name] = {"ms": ms.value, "bytes": nbytes, "gbs": gbs, "frac_of_peak": gbs / PEAK_GBS}
        log(f"in-place pattern probe {name}: {ms.value:.3f} ms -> {gbs:.0f} GB/s ({100 * gbs / PEAK_GBS:.1f}% of 8 TB/s)")
    return out


def host_cores():
    """The host cores this process may use: its CPU affinity set, capped by a cgroup CPU quota
    where one is set (cpu.max "quota period"), with what was read."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else max(1, min(affinity, int(quota)))
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), "")
    except OSError:
        pass
    return {"usable": usable, "affinity": affinity, "cpu_count": os.cpu_count(), "cgroup_quota_cpus": quota,
            "model": model, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(target_seconds=15.0, voxelize_mesh=None, resident=False, n=256):
    """The reference's EsotericPullSolver<D3Q19> (oracle/_ref, built from the reference sources
    like its Release build, run on this host) on every host core the process may use (one OpenMP
    thread per core, OMP_PROC_BIND=close, OMP_PLACES=cores): 256^3 periodic Taylor-Green -- the
    reference's cfd-bench workload and protocol (apps/cfd-bench.cpp:174-191), not the GPU
    headline's cavity -- on a bounded number of steps.  Falls back to the CPU restatement (kind
    "port") where the reference build is absent."""
    cores = host_cores()
    threads = cores["usable"]
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    import oracle

    # libgomp reads OMP_NUM_THREADS once, when it is first loaded; set the calling thread's team
    # size directly as well, in the libgomp instance the checker libraries link.
    try:
        import ctypes as C

        gomp = C.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(threads)
        threads = int(gomp.omp_get_max_threads())
    except OSError:
        pass
    if oracle.ref_available(native=True):
        kind = "reference"
        s = oracle.RefEsoteric(19, n, n, n, 0.6, native=True)
        s.init_taylor_green(0.02)
    else:
        kind = "port"
        s = oracle.Oracle(19, n, n, n, 0.6)
        s.init_taylor_green(0.02)
    s.mass()
    t0 = time.perf_counter()
    s.step(1)  # warmup, also sizes the sample
    one = time.perf_counter() - t0
    steps = int(max(1, min(50, target_seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    s.step(steps)
    dt = time.perf_counter() - t0
    s.mass()
    mlups = n ** 3 * steps / dt / 1e6
    log(f"cpu baseline ({kind}): {mlups:.1f} MLUPS on {threads} threads, {steps} steps in {dt:.1f} s")
    out = {"value": mlups, "unit": "MLUPS", "cores": threads, "kind": kind,
           "sample": f"D3Q19 {n}^3 periodic Taylor-Green (A 0.02, tau 0.6), 1 warmup + {steps} timed steps, "
                     f"FP32, {threads} OpenMP threads on {cores['model']}",
           "host": cores}
    if voxelize_mesh is not None and oracle.ref_available(native=True):
        # The reference's voxelize_surface (src/lbm/voxelize.cpp:5: every cell against every
        # triangle, one thread) on a small grid with the voxelize leg's mesh.
        vn = 40
        t0 = time.perf_counter()
        oracle.ref_voxelize(voxelize_mesh, (vn, vn, vn), native=True)
        vdt = time.perf_counter() - t0
        out["voxelize"] = {"value": vn ** 3 / vdt, "unit": "cells/s", "cores": 1, "kind": "reference",
                           "sample": f"{vn}^3 cells in [0,1]^3, {len(voxelize_mesh)}-triangle sphere, {vdt:.1f} s"}
        log(f"cpu baseline voxelize (reference): {vn ** 3 / vdt:.0f} cells/s on 1 thread ({vdt:.1f} s)")
    if resident:
        out["resident"] = resident_cpu_sample(oracle)
    return out


def resident_cpu_sample(oracle, n=160, particles=1 << 18, repeats=3):
    """The resident modules' CPU restatement (oracle/, one thread; the reference's modules are
    SYCL-only) on an n^3 periodic lattice with the resident leg's kind of fields: the fastest of
    `repeats` calls of each module, inputs converted beforehand.  Rates in cells/s (particles/s for
    the push)."""
    rng = np.random.default_rng(0)
    r = oracle.Resident(n, n, n)
    shape = (n, n, n)
    u = [np.ascontiguousarray(rng.uniform(-0.05, 0.05, shape), np.float32).reshape(-1) for _ in range(3)]
    g = r.thermal_equilibrium(np.full(shape, 300.0, np.float32))
    phi = np.ascontiguousarray(r.half_space(n / 2 + 0.5)).reshape(-1)
    zero = [np.zeros(n ** 3, np.float32) for _ in range(3)]
    pts = np.column_stack([rng.uniform(0, n, (particles, 3)), np.zeros((particles, 3)), np.ones(particles),
                           np.full(particles, 0.5)])
    cases = (("thermal_step", lambda: r.thermal_step(g, u, 1.0 / 0.58, 1), n ** 3),
             ("fill_advect", lambda: r.advect(phi, u, 0.1), n ** 3),
             ("interface_geometry", lambda: r.geometry(phi), n ** 3),
             ("surface_tension", lambda: r.surface_tension(phi, zero, 0.072, 1.0), n ** 3),
             ("q_criterion", lambda: r.q_criterion(u), n ** 3),
             ("particles_one_way", lambda: r.particles_advance(pts, u, zero, 0.05, False), particles))
    out = {"kind": "port", "cores": 1,
           "sample": f"{n}^3 periodic lattice, {particles} particles, fastest of {repeats} calls per module, one thread"}
    for name, fn, units in cases:
        dt = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            fn()
            dt = min(dt, time.perf_counter() - t0)
        out[name] = {"value": units / dt, "unit": "particles/s" if name.startswith("particles") else "cells/s"}
        log(f"cpu baseline resident {name} (port): {units / dt:.3g} {out[name]['unit']} ({dt:.2f} s)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--legs", default="fp32,fp16,fp32s,fp64,d3q27,d2q9,adjoint,voxelize,resident,shard,config1,config5")
    ap.add_argument("--config5-steps", type=int, default=200)
    ap.add_argument("--config5-storages", default="fp16s,fp32")
    ap.add_argument("--config5-n", type=int, default=768)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the z-slab halo "
              "messages exist (cfd_lbm_halo_pack / cfd_lbm_halo_unpack) but their transport between GPUs "
              "is not implemented.", file=sys.stderr)
        sys.exit(2)

    import cfd_solvers_amd as hip

    # The bandwidth probes run first: besides their own figures, they bring the GPU to its working
    # clocks before the legs' timed steps.
    probe = None if args.no_probe else copy_probe()
    inplace = None if args.no_probe else inplace_probe()
    legs = {}
    wanted = [l.strip() for l in args.legs.split(",") if l.strip()]
    table = {"fp32": ("d3q19", "fp32"), "fp16": ("d3q19", "fp16s"), "d3q27": ("d3q27", "fp32"),
             "fp64": ("d3q19", "fp64"), "fp32s": ("d3q19", "fp32s")}
    velocities = {}
    for name in wanted:
        if name == "adjoint":
            legs[name] = adjoint_leg(hip)
            continue
        if name == "config5":
            legs[name] = config5_leg(hip, n=args.config5_n, steps=args.config5_steps,
                                     storages=tuple(args.config5_storages.split(",")))
            continue
        if name == "voxelize":
            legs[name] = voxelize_leg(hip)
            continue
        if name == "resident":
            legs[name] = resident_leg(hip)
            continue
        if name == "shard":
            legs[name] = shard_leg(hip)
            continue
        if name == "config1":
            legs[name] = config1_leg(hip)
            continue
        if name == "d2q9":
            legs[name] = d2q9_leg(hip, steps=args.steps, warmup=args.warmup)
            continue
        lattice, storage = table[name]
        legs[name], vel = run_leg(hip, name, lattice, storage, args.n, 0.55, args.steps, args.warmup)
        if lattice == "d3q19":
            velocities[name] = vel
    # Relative L2 error of the velocity field against the FP64 build after the same steps.
    if "fp64" in velocities:
        ref = velocities["fp64"]
        for name, vel in velocities.items():
            if name != "fp64":
                err = float(np.sqrt(np.sum((vel - ref) ** 2) / np.sum(ref ** 2)))
                legs[name]["velocity_rel_l2_vs_fp64"] = err
                log(f"leg {name}: velocity relative L2 error vs FP64 {err:.2e}")
    velocities.clear()
    head = legs.get("fp32") or next((v for v in legs.values() if "achieved_gbs" in v), None)
    if head is None:  # opt-in runs of the adjoint / config5 legs only
        print(json.dumps({"metric": METRIC, "value": None, "unit": "MLUPS", "n_gpus": 1, "legs": legs}), flush=True)
        return

    traffic = None
    if TRAFFIC_FILE.exists() and args.n == 256:
        try:
            traffic = json.loads(TRAFFIC_FILE.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    result = {
        "metric": METRIC,
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"D3Q19 {args.n}^3 lid-driven cavity, walls on 5 faces, lid u=(0.1,0,0), "
                               f"tau 0.55, FP32 populations, uint8 flags, esoteric-pull in place",
                   "lattice": "D3Q19", "nx": args.n, "ny": args.n, "nz": args.n, "storage": "fp32",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": head["achieved_gbs"] / PEAK_GBS, "traffic": traffic,
                     "kernel": "stream_collide<19, float, bounded>",
                     "bytes_per_launch": head["algorithmic_bytes_per_launch"]},
        "legs": legs,
        "copy_probe": probe,
        "inplace_probe": inplace,
    }
    if inplace:
        result["roofline"]["frac_of_inplace_probe"] = head["achieved_gbs"] / inplace["fp32"]["gbs"]
        if "fp16" in legs:
            legs["fp16"]["frac_of_inplace_probe"] = legs["fp16"]["achieved_gbs"] / inplace["fp16"]["gbs"]
    if "fp16" in legs and "fp32" in legs:
        result["fp16_over_fp32"] = legs["fp16"]["mlups"] / legs["fp32"]["mlups"]
    if not args.no_cpu:
        mesh = None
        if "voxelize" in legs:
            from cfd_solvers_amd import geometry as G

            mesh = G.sphere_surface((.5, .5, .5), .37, 64, 128)
        result["cpu_baseline"] = cpu_baseline(voxelize_mesh=mesh, resident="resident" in legs)
        vox = result["cpu_baseline"].pop("voxelize", None)
        if vox is not None:
            legs["voxelize"]["cpu_reference"] = vox
            legs["voxelize"]["gpu_over_cpu_reference"] = legs["voxelize"]["gcells_per_s"] * 1e9 / vox["value"]
        res = result["cpu_baseline"].pop("resident", None)
        if res is not None:
            legs["resident"]["cpu_port"] = res
            cells = legs["resident"]["cells"]
            for name, v in res.items():
                if isinstance(v, dict) and name in legs["resident"]:
                    units = legs["resident"]["particles"] if name.startswith("particles") else cells
                    gpu_rate = units / (legs["resident"][name]["ms"] * 1e-3)
                    legs["resident"][name]["gpu_over_cpu_port"] = gpu_rate / v["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
