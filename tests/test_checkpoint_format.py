"""The CFDCHK3 checkpoint files of the reference's DistributedPullBlock (tests/golden/
cfdchk3_*.bin, written by make_checkpoints.py from include/cfd/solvers/lbm/distributed_pull.hpp:
176-203): the layout the HIP solver reads and writes, and -- where the reference build exists --
the reference's own bitwise restart (tests/test_main.cpp:225-266)."""
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
A = GOLDEN / "cfdchk3_d3q19_12x10x8_step5.bin"
B = GOLDEN / "cfdchk3_d3q19_12x10x8_step8.bin"


def parse(path):
    raw = path.read_bytes()
    assert raw[:8] == b"CFDCHK3\0"
    version, endian, q = struct.unpack_from("<3I", raw, 8)
    ext = struct.unpack_from("<9Q", raw, 20)
    step, = struct.unpack_from("<Q", raw, 92)
    tau, ax, ay, az = struct.unpack_from("<4f", raw, 100)
    n = ext[6] * ext[7] * ext[8]
    body = np.frombuffer(raw, np.float32, q * n, 116).reshape(q, ext[8], ext[7], ext[6])
    return dict(version=version, endian=endian, q=q, global_=ext[0:3], begin=ext[3:6], extent=ext[6:9], step=step,
                tau=tau, accel=(ax, ay, az), body=body, tail=raw[116 + 4 * q * n:])


@pytest.mark.parametrize("path,step", [(A, 5), (B, 8)])
def test_reference_file_layout(path, step):
    h = parse(path)
    assert (h["version"], h["endian"], h["q"]) == (1, 0x01020304, 19)
    assert h["global_"] == (12, 10, 8) and h["begin"] == (0, 0, 0) and h["extent"] == (12, 10, 8)
    assert h["step"] == step and h["tau"] == np.float32(0.71) and h["accel"] == (0.0, 0.0, 0.0)
    assert h["tail"] == b""
    mass = h["body"].astype(np.float64).sum()
    assert abs(mass - 12 * 10 * 8) < 1e-3  # density ~1 everywhere, conserved


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_reference_restart_is_bitwise(tmp_path):
    out = tmp_path / "r.bin"
    O.ref_checkpoint_restart((12, 10, 8), 0.71, A, 3, out)
    assert out.read_bytes() == B.read_bytes()
