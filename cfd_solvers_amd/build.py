"""Builds libcfd_lbm_hip.so (HIP kernels + C ABI for gfx950) and the C++ apps, in tree.

    python -m cfd_solvers_amd.build [--verbose]

Every HIP source is compiled with ``--offload-arch=gfx950 -ffp-contract=off`` (the collision
arithmetic then matches the reference's FP32 arithmetic bit for bit) and with packed-FP32
instructions disabled.  The apps (cfd-bench, cfd-solve, cfd-distributed) are plain C++ linked to the library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcfd_lbm_hip.so"
PROBE = PKG / "libhbm_probe.so"
APPS = {"cfd-bench": PKG / "apps" / "cfd_bench.cpp", "cfd-solve": PKG / "apps" / "cfd_solve.cpp",
        "cfd-distributed": PKG / "apps" / "cfd_distributed.cpp"}

HIP_SOURCES = ["lbm_kernels.hip", "lbm_adjoint.hip", "voxelize.hip", "multiphysics.hip", "lbm_capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(map(str, cmd)), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd[:3]))} ... ({r.returncode})")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = [BUILD / (Path(s).stem + ".o") for s in HIP_SOURCES]

    def compile_one(src_obj):
        src, obj = src_obj
        _run([HIPCC, *HIP_FLAGS, "-c", CSRC / src, "-o", obj], verbose)

    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as pool:
        list(pool.map(compile_one, zip(HIP_SOURCES, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    probe_tmp = PROBE.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", CSRC / "hbm_probe.hip", "-o", probe_tmp], verbose)
    os.replace(probe_tmp, PROBE)
    for name, src in APPS.items():
        if src.exists():
            _run(["g++", "-O2", "-std=c++20", "-I", ROOT / "include", src, "-o", PKG / name,
                  f"-L{PKG}", "-lcfd_lbm_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="--verbose" in sys.argv)
    print(LIB)
