"""The C-ABI library loads and exports every entry point include/cfd_lbm/c_api.h declares.

Runs without a GPU: no compute call is made.  On a machine without a HIP device, creating a
solver fails cleanly with return code 1 and a message (the reference's error convention,
src/python/c_api.cpp:8-15) instead of crashing.
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "cfd_lbm" / "c_api.h"


def declared():
    text = HEADER.read_text()
    return re.findall(r"CFD_LBM_API\s+[\w\s\*]+?\b(cfd_lbm_\w+)\s*\(", text)


@pytest.fixture(scope="module")
def library():
    from cfd_solvers_amd import bindings

    if not bindings.DEFAULT_LIBRARY.exists():
        from cfd_solvers_amd import build

        build.build()
    return bindings.DEFAULT_LIBRARY


def test_header_declares_entry_points():
    names = declared()
    assert len(names) == len(set(names)) >= 25
    for required in ("cfd_lbm_create", "cfd_lbm_step", "cfd_lbm_set_flags", "cfd_lbm_download_macroscopic",
                     "cfd_lbm_total_mass", "cfd_lbm_index_map"):
        assert required in names


def test_every_declared_symbol_is_exported(library):
    lib = C.CDLL(str(library))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", str(library)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (cfd_lbm_\w+)", out))
    assert exported == set(declared()), f"exported but undeclared: {exported - set(declared())}"


def test_bindings_cover_the_header(library):
    from cfd_solvers_amd import bindings

    lib = bindings.load(library)
    assert set(bindings._SIGNATURES) | {"cfd_lbm_version"} == set(declared())
    assert "gfx950" in lib.version


def test_create_reports_errors_through_the_buffer(library):
    from cfd_solvers_amd import bindings

    lib = bindings.load(library)
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = lib._lib.cfd_lbm_create(19, 8, 8, 8, C.c_float(0.4), 0.0, 0.0, 0.0, 0, C.byref(h), err, 256)
    assert rc == 1 and b"tau" in err.value
    rc = lib._lib.cfd_lbm_create(19, 8, 8, 8, C.c_float(0.6), 0.0, 0.0, 0.0, 7, C.byref(h), err, 256)
    assert rc == 1 and b"storage" in err.value
    rc = lib._lib.cfd_lbm_create(19, 8, 8, 8, C.c_float(0.6), 0.0, 0.0, 0.0, 0, C.byref(h), err, 256)
    if rc == 0:  # a GPU is present
        assert lib._lib.cfd_lbm_destroy(h, err, 256) == 0
    else:
        assert rc == 1 and len(err.value) > 0
    # Truncation keeps the buffer NUL-terminated.
    small = C.create_string_buffer(4)
    rc = lib._lib.cfd_lbm_create(19, 8, 8, 8, C.c_float(0.4), 0.0, 0.0, 0.0, 0, C.byref(h), small, 4)
    assert rc == 1 and len(small.value) == 3


def test_product_path_has_no_cpu_fallback():
    """Nothing in the package imports or loads the oracle."""
    for src in (ROOT / "cfd_solvers_amd").rglob("*"):
        if src.suffix in {".py", ".hip", ".hpp", ".cpp", ".h"}:
            text = src.read_text()
            assert "oracle" not in text.replace("oracle/", ""), f"{src} mentions the oracle"
