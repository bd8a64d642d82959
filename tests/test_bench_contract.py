"""bench.py's command-line contract without a GPU: one GPU only (--gpus N > 1 exits with a message,
before any device is touched), the CPU baseline runs on every host core the process may use and
says so, and the resident modules' CPU restatement sample reports a rate per module in the units
the resident leg uses."""
import os
import subprocess
import sys
from pathlib import Path

import oracle

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "4"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 4" in r.stderr
    assert r.stdout == ""


def test_resident_cpu_sample_units():
    import bench

    out = bench.resident_cpu_sample(oracle, n=24, particles=512, repeats=1)
    assert out["kind"] == "port" and out["cores"] == 1
    for name in ("thermal_step", "fill_advect", "interface_geometry", "surface_tension", "q_criterion"):
        assert out[name]["unit"] == "cells/s" and out[name]["value"] > 0
    assert out["particles_one_way"]["unit"] == "particles/s" and out["particles_one_way"]["value"] > 0


def test_cpu_baseline_uses_every_usable_core():
    import bench

    cores = bench.host_cores()
    assert cores["affinity"] == len(os.sched_getaffinity(0))
    if cores["cgroup_quota_cpus"] is None:
        assert cores["usable"] == cores["affinity"]
    else:
        assert cores["usable"] == max(1, min(cores["affinity"], int(cores["cgroup_quota_cpus"])))
    out = bench.cpu_baseline(target_seconds=0.05, n=16)
    assert out["cores"] == cores["usable"] and out["value"] > 0 and out["unit"] == "MLUPS"
    assert out["kind"] in ("reference", "port") and out["host"]["usable"] == cores["usable"]
    assert f"{cores['usable']} OpenMP threads" in out["sample"]
