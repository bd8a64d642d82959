import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))
        return cache[name]

    return get


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    if not oracle.ORACLE_SO.exists():
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail loudly (no fallback) if it is missing."""
    import cfd_solvers_amd

    lib = cfd_solvers_amd.load()
    return cfd_solvers_amd


def report(name, **kv):
    """One flushed line per GPU case so a box that dies mid-run still names the case."""
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kv.items()), flush=True)


class DeviceMemory:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself loaded (resolved by
    soname once the library is loaded), so that GPU tests need no second runtime (torch brings its
    own).  Frees everything it allocated on free()."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(int(nbytes), 1)) == 0
        self.live.append(p.value)
        return p.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live.clear()


@pytest.fixture
def dev(hip):
    hip.load()
    d = DeviceMemory()
    yield d
    d.free()
