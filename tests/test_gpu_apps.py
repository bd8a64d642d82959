"""The drop-in command-line surface on the GPU: cfd-bench with the reference's flags and output
schema (apps/cfd-bench.cpp:47-191) and --backend hip, the -hip cases of cfd-solve
(apps/cfd-solve.cpp:59-73), cfd-distributed --backend hip on one rank with the reference's
checkpoint/restart files (apps/cfd-distributed.cpp:142-174), and the package's smoke entry point."""
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
DISTRIBUTED = ROOT / "cfd_solvers_amd" / "cfd-distributed"
GOLDEN = ROOT / "tests" / "golden"
CHK_BYTES = 116 + 4 * 19 * 12 * 10 * 8  # CFDCHK3 header + body of the 12x10x8 D3Q19 fixtures


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


REF_BENCH = ROOT / "oracle" / "_ref" / "cfd-bench-hip"


@pytest.mark.skipif(not REF_BENCH.exists(), reason="the reference's patched cfd-bench was not built (no reference tree)")
@pytest.mark.parametrize("lattice", ["d3q19", "d3q27", "d2q9"])
def test_reference_cfd_bench_drives_the_hip_backend(lattice):
    """The reference's own apps/cfd-bench.cpp with INTEGRATION.md's three insertions (built by
    oracle.build into oracle/_ref/) runs --backend hip on this GPU: the reference's output schema,
    backend "hip", and the same mass error, bit for bit, as this repository's cfd-bench on the same
    run (both step the FP32 lattice through the same library from Taylor-Green 0.02)."""
    report("reference_cfd_bench", lattice=lattice)
    nz = 1 if lattice == "d2q9" else 32
    args = ["--lattice", lattice, "--backend", "hip", "--nx", 32, "--ny", 32, "--nz", nz, "--warmup", 2,
            "--steps", 10, "--json"]
    ref = run([REF_BENCH, *args])
    assert ref.returncode == 0, ref.stderr
    r = json.loads(ref.stdout.strip().splitlines()[-1])
    ours = run([BENCH, *args])
    assert ours.returncode == 0, ours.stderr
    o = json.loads(ours.stdout.strip().splitlines()[-1])
    report("reference_cfd_bench_result", device=r["device"], mass_error_rel=r["mass_error_rel"])
    assert r["schema"] == "cfd_solvers.benchmark.v1" and r["backend"] == "hip" and r["streaming"] == "in-place"
    assert r["q"] == {"d3q19": 19, "d3q27": 27, "d2q9": 9}[lattice] and r["steps"] == 10
    assert "gfx950" in r["device"]
    assert r["mass_error_abs"] == o["mass_error_abs"] and r["mass_error_rel"] < 1e-5


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
        # The reference's gate is 1 % (tests/test_main.cpp:449-477, D2Q9); D3Q19 halfway
        # bounce-back at tau 0.8 and 30 fluid rows measures ~1.1e-3.
        assert float(fields["relative_L2"]) < 2e-3
    if case in ("lbm-d3q19-hip", "lbm-d3q27-hip", "lbm-d2q9-inplace-hip"):
        assert float(fields["mass_error_rel"]) < 1e-5 and float(fields["MLUPS"]) > 0


def test_cfd_distributed_restarts_from_the_reference_files(tmp_path):
    """The reference's DistributedPullBlock wrote CFDCHK3 files at steps 5 and 8 of its restart test
    (12x10x8, tau 0.71, Taylor-Green 0.02; tests/test_main.cpp:225-266).  Restarting the one-rank HIP
    run from the step-5 file and stepping 3 times writes the step-8 file byte for byte (header and
    body; a trailer with the parity follows)."""
    report("cfd_distributed_restart")
    (tmp_path / "ref.rank0.bin").write_bytes((GOLDEN / "cfdchk3_d3q19_12x10x8_step5.bin").read_bytes())
    out = run([DISTRIBUTED, "--backend", "hip", "--nx", 12, "--ny", 10, "--nz", 8, "--tau", "0.71", "--steps", 3,
               "--restart", tmp_path / "ref", "--checkpoint", tmp_path / "out"])
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    fields = dict(kv.split("=", 1) for kv in lines[-2].split() if "=" in kv)
    assert fields["backend"] == "hip-esoteric" and fields["ranks"] == "1" and fields["process_grid"] == "1x1x1"
    assert fields["global"] == "12x10x8" and fields["steps"] == "3" and fields["halo_bytes_per_step"] == "0"
    assert lines[-1].startswith("rank=0 local_rank=0/1 ") and lines[-1].endswith("brick=12x10x8 begin=0,0,0")
    saved = (tmp_path / "out.rank0.bin").read_bytes()
    want = (GOLDEN / "cfdchk3_d3q19_12x10x8_step8.bin").read_bytes()
    assert len(want) == CHK_BYTES and saved[:CHK_BYTES] == want
    assert abs(float(fields["mass"]) - 960.0) < 1e-3  # rho = 1 on 960 cells, conserved


def test_cfd_distributed_fresh_run_csv_and_rejections(tmp_path):
    out = run([DISTRIBUTED, "--backend", "hip", "--lattice", "d3q27", "--nx", 16, "--ny", 12, "--nz", 10,
               "--steps", 4, "--csv", "--checkpoint", tmp_path / "c"])
    assert out.returncode == 0, out.stderr
    row = out.stdout.strip().splitlines()[-1].split(",")
    assert len(row) == 16 and row[:10] == ["hip-esoteric", "d3q27", "1", "1", "1", "1", "16", "12", "10", "4"]
    assert float(row[10]) > 0 and float(row[11]) > 0 and abs(float(row[15]) - 1920.0) < 1e-3
    chk = (tmp_path / "c.rank0.bin").read_bytes()
    assert chk[:8] == b"CFDCHK3\0" and len(chk) > 116 + 4 * 27 * 1920
    again = run([DISTRIBUTED, "--backend", "hip", "--lattice", "d3q27", "--nx", 16, "--ny", 12, "--nz", 10,
                 "--steps", 1, "--csv", "--restart", tmp_path / "c"])
    assert again.returncode == 0, again.stderr
    assert run([DISTRIBUTED, "--backend", "cpu"]).returncode == 2  # the MPI CPU path stays the reference's
    assert run([DISTRIBUTED, "--backend", "hip", "--lattice", "d2q9"]).returncode == 1
    bad = run([DISTRIBUTED, "--backend", "hip", "--nx", 12, "--ny", 10, "--nz", 9, "--tau", "0.71", "--steps", 1,
               "--restart", GOLDEN / "cfdchk3_d3q19_12x10x8_step5"])
    assert bad.returncode == 1 and "cfd-distributed:" in bad.stderr  # no such rank file


def test_smoke_entry_point():
    report("smoke")
    out = subprocess.run([sys.executable, "-c", "import __graft_entry__ as g; g.smoke(); print('smoke ok')"],
                         capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "smoke ok" in out.stdout
