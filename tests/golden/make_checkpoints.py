"""Writes CFDCHK3 files of the reference's DistributedPullBlock<D3Q19> (one brick covering the
lattice, include/cfd/solvers/lbm/distributed_pull.hpp:176-203, compiled by oracle/Makefile) for the
configuration of its restart test (tests/test_main.cpp:225-266): 12 x 10 x 8 cells, tau 0.71,
Taylor-Green amplitude 0.02.

  cfdchk3_d3q19_12x10x8_step5.bin   after 5 steps
  cfdchk3_d3q19_12x10x8_step8.bin   after 8 steps (equal to restoring step 5 and running 3 more,
                                     which the script checks with the reference's own loader)

Run where the reference tree is present:

    python tests/golden/make_checkpoints.py
"""
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402

SHAPE, TAU, AMP = (12, 10, 8), 0.71, 0.02


def main():
    here = Path(__file__).parent
    a = here / "cfdchk3_d3q19_12x10x8_step5.bin"
    b = here / "cfdchk3_d3q19_12x10x8_step8.bin"
    O.ref_checkpoint_tg(SHAPE, TAU, AMP, 5, a)
    O.ref_checkpoint_tg(SHAPE, TAU, AMP, 8, b)
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "restart.bin"
        O.ref_checkpoint_restart(SHAPE, TAU, a, 3, c)
        assert c.read_bytes() == b.read_bytes(), "the reference's restart is not bitwise"
    print("written", a.stat().st_size, b.stat().st_size, flush=True)


if __name__ == "__main__":
    main()
