"""The drop-in command-line surface on the GPU: cfd-bench with the reference's flags and output
schema (apps/cfd-bench.cpp:47-191) and --backend hip, the -hip cases of cfd-solve
(apps/cfd-solve.cpp:59-73), and the package's smoke entry point."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

from conftest import report

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
BENCH = ROOT / "cfd_solvers_amd" / "cfd-bench"
SOLVE = ROOT / "cfd_solvers_amd" / "cfd-solve"


def run(args, timeout=120):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout, cwd=ROOT)


@pytest.mark.parametrize("lattice,storage", [("d3q19", "fp32"), ("d3q19", "fp16s"), ("d3q27", "fp32"),
                                             ("d2q9", "fp32")])
def test_cfd_bench_json(lattice, storage):
    report("cfd_bench", lattice=lattice, storage=storage)
    nz = "1" if lattice == "d2q9" else "64"
    out = run([BENCH, "--lattice", lattice, "--backend", "hip", "--nx", 64, "--ny", 64, "--nz", nz,
               "--warmup", 2, "--steps", 10, "--storage", storage, "--json"])
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    report("cfd_bench_result", mlups=r["mlups"], mass_error_rel=r["mass_error_rel"])
    assert r["schema"] == "cfd_solvers.benchmark.v1"
    assert r["backend"] == "hip" and r["streaming"] == "in-place" and r["lattice"] == lattice
    assert r["mlups"] > 0 and r["steps"] == 10
    assert r["mass_error_rel"] < 1e-5


def test_cfd_bench_csv_and_rejections():
    out = run([BENCH, "--lattice", "d3q19", "--backend", "hip", "--nx", 32, "--ny", 32, "--nz", 32, "--warmup", 1,
               "--steps", 3, "--csv"])
    assert out.returncode == 0, out.stderr
    header, row = out.stdout.strip().splitlines()[-2:]
    assert header.startswith("lattice,backend,streaming,device,nx,ny,nz,q,steps,seconds,mlups")
    assert row.startswith("d3q19,hip,in-place,")
    assert run([BENCH, "--backend", "cpu"]).returncode == 2  # the CPU path stays the reference's
    assert run([BENCH, "--backend", "hip", "--json", "--csv"]).returncode != 0  # mutually exclusive
    assert run([BENCH, "--backend", "hip", "--tau", "0.5"]).returncode != 0  # tau must exceed 0.5


@pytest.mark.parametrize("case", ["lbm-d3q19-hip", "lbm-d3q27-hip", "lbm-d2q9-inplace-hip", "lbm-poiseuille-hip",
                                  "lbm-cavity-hip", "lbm-channel-io-hip"])
def test_cfd_solve_cases(case):
    report("cfd_solve", case=case)
    out = run([SOLVE, case], timeout=300)
    assert out.returncode == 0, out.stderr
    line = out.stdout.strip().splitlines()[-1]
    fields = dict(kv.split("=", 1) for kv in line.replace('"', "").split() if "=" in kv)
    report("cfd_solve_result", case=case, line=line[:160])
    for key, value in fields.items():
        try:
            v = float(value)
        except ValueError:
            continue
        assert v == v and abs(v) != float("inf"), f"{key}={value}"
    if case == "lbm-poiseuille-hip":
        assert float(fields["relative_L2"]) < 0.02
    if case in ("lbm-d3q19-hip", "lbm-d3q27-hip", "lbm-d2q9-inplace-hip"):
        assert float(fields["mass_error_rel"]) < 1e-5 and float(fields["MLUPS"]) > 0


def test_smoke_entry_point():
    report("smoke")
    out = subprocess.run([sys.executable, "-c", "import __graft_entry__ as g; g.smoke(); print('smoke ok')"],
                         capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "smoke ok" in out.stdout
