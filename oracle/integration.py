"""The reference's cfd-bench with INTEGRATION.md's `--backend hip` insertions.  TEST INFRASTRUCTURE ONLY.

``patched_cfd_bench(reference)`` returns the text of the reference's ``apps/cfd-bench.cpp`` with the
three C++ blocks that INTEGRATION.md marks ``<!-- patch cfd-bench include|runner|dispatch -->``
inserted where that document says: the header after the optional SYCL include (apps/cfd-bench.cpp:3-6),
``run_hip`` after ``run_sycl`` (:215-233), and the dispatch line before ``dispatch_backend``'s final
throw (:236-261).  ``build_cfd_bench_hip`` compiles it with ``-DCFD_HAS_HIP_LBM`` and links it to
``libcfd_lbm_hip.so``: the reference's own bench driving the gfx950 backend through the C++ front end.
The patched source is written to a temporary folder only; the binary goes where the caller says
(``oracle/_ref/`` for the GPU test).  Nothing here runs where the reference tree is absent.
"""
from __future__ import annotations

import re
import subprocess
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
INTEGRATION_MD = ROOT / "INTEGRATION.md"
HIP_LIB_DIR = ROOT / "cfd_solvers_amd"


def integration_blocks(text: str | None = None) -> dict[str, str]:
    """The C++ blocks of INTEGRATION.md that follow a ``<!-- patch cfd-bench NAME -->`` marker."""
    text = INTEGRATION_MD.read_text() if text is None else text
    blocks = {}
    for m in re.finditer(r"<!-- patch cfd-bench (\w+) -->\s*```cpp\n(.*?)```", text, re.S):
        blocks[m.group(1)] = m.group(2)
    missing = {"include", "runner", "dispatch"} - blocks.keys()
    if missing:
        raise ValueError(f"INTEGRATION.md lacks the cfd-bench patch blocks {sorted(missing)}")
    return blocks


def _insert_after(src: str, anchor: str, block: str) -> str:
    i = src.index(anchor) + len(anchor)
    return src[:i] + "\n" + block + src[i:]


def _insert_before(src: str, anchor: str, block: str) -> str:
    i = src.index(anchor)
    return src[:i] + block + "\n" + src[i:]


def patched_cfd_bench(reference: str | Path) -> str:
    src = (Path(reference) / "apps" / "cfd-bench.cpp").read_text()
    b = integration_blocks()
    # The SYCL include block: `#if defined(CFD_HAS_SYCL)` ... `#endif`, first occurrence.
    m = re.search(r"#if defined\(CFD_HAS_SYCL\)\n#include [^\n]*\n#endif\n", src)
    if m is None:
        raise ValueError("cfd-bench.cpp: optional SYCL include not found")
    src = src[:m.end()] + b["include"] + src[m.end():]
    # run_hip right after run_sycl: the `#endif` that closes the SYCL runner precedes dispatch_backend.
    anchor = re.search(r"\n#endif\n\s*\ntemplate<class Descriptor>\nint dispatch_backend", src)
    if anchor is None:
        raise ValueError("cfd-bench.cpp: run_sycl / dispatch_backend not found")
    cut = anchor.start() + len("\n#endif\n")
    src = src[:cut] + b["runner"] + src[cut:]
    final_throw = '    throw std::invalid_argument("backend must be cpu or sycl");'
    if src.count(final_throw) != 1:
        raise ValueError("cfd-bench.cpp: dispatch_backend's final throw not found")
    return _insert_before(src, final_throw, b["dispatch"].rstrip("\n"))


def compile_cfd_bench(reference: str | Path, out: str | Path, hip: bool = True) -> subprocess.CompletedProcess:
    """g++ the patched bench (hip=True: with -DCFD_HAS_HIP_LBM, linked to libcfd_lbm_hip.so) into `out`."""
    reference = Path(reference)
    with tempfile.TemporaryDirectory() as tmp:
        src = Path(tmp) / "cfd-bench.cpp"
        src.write_text(patched_cfd_bench(reference))
        cmd = ["g++", "-std=c++20", "-O2", "-fopenmp", "-DCFD_HAS_OPENMP", f"-I{reference / 'include'}",
               f"-I{ROOT / 'include'}", str(src), str(reference / "src" / "lbm" / "esoteric_pull.cpp"),
               "-o", str(out)]
        if hip:
            cmd[1:1] = ["-DCFD_HAS_HIP_LBM"]
            cmd += [f"-L{HIP_LIB_DIR}", "-lcfd_lbm_hip", f"-Wl,-rpath,{HIP_LIB_DIR}",
                    "-Wl,-rpath,$ORIGIN/../../cfd_solvers_amd"]
        return subprocess.run(cmd, capture_output=True, text=True)
