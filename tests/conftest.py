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
