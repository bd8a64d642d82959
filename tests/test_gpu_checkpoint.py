"""Checkpoint / restart of the HIP solver (cfd_lbm_save_checkpoint / cfd_lbm_load_checkpoint):
the reference's CFDCHK3 files (distributed_pull.hpp:176-245) restore into the esoteric solver and
continue bit for bit -- after 3 more steps the saved file equals the reference's own 8-step file
byte for byte -- and the solver's own checkpoints restart bitwise for every storage on a bounded
lattice with a moving lid (tests/test_main.cpp:225-266 is the reference's pattern)."""
import struct
from pathlib import Path

import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
A = GOLDEN / "cfdchk3_d3q19_12x10x8_step5.bin"
B = GOLDEN / "cfdchk3_d3q19_12x10x8_step8.bin"
BODY = 116 + 4 * 19 * 12 * 10 * 8


def test_reference_file_restores_and_continues(hip, tmp_path):
    report("checkpoint_reference")
    s = hip.EsotericPullHipSolver("d3q19", 12, 10, 8, float(np.float32(0.71)))
    s.load_checkpoint(A)
    assert s.time_step == 5
    s.step(3)
    out = tmp_path / "ours.bin"
    s.save_checkpoint(out, raw=False)
    data = out.read_bytes()
    assert data[:BODY] == B.read_bytes()
    assert data[BODY:BODY + 8] == b"CFDESO1\0"
    parity, convention, storage, raw = struct.unpack_from("<4I", data, BODY + 8)
    assert (parity, convention, storage, raw) == (0, 0, 0, 0)
    s.close()


def cavity(hip, storage):
    n = (20, 18, 16)
    f = np.zeros(n[::-1], np.uint8)
    f[:, -1, :] = 0x81
    f[:, 0, :] = 0x80
    f[0, :, :] = f[-1, :, :] = 0x80
    f[:, :, 0] = f[:, :, -1] = 0x80
    s = hip.EsotericPullHipSolver("d3q19", *n, 0.62, storage=storage)
    s.set_flags(f)
    s.set_wall_velocity(1, (0.08, 0.0, 0.01))
    return s, f


@pytest.mark.parametrize("storage", ["fp32", "fp16s", "fp32s", "fp64"])
def test_restart_is_bitwise(hip, tmp_path, storage):
    report("checkpoint_restart", storage=storage)
    a, flags = cavity(hip, storage)
    a.initialize_uniform()
    a.step(7)
    path = tmp_path / f"{storage}.bin"
    a.save_checkpoint(path)
    a.step(5)
    b, _ = cavity(hip, storage)
    b.load_checkpoint(path)
    assert b.time_step == 7
    b.step(5)
    ma, mb = a.download_macroscopic(), b.download_macroscopic()
    for f in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(getattr(ma, f), getattr(mb, f)), f
    assert np.array_equal(a.logical_populations(), b.logical_populations())
    if storage == "fp32":  # the float body alone restores the fluid bit for bit
        c, _ = cavity(hip, storage)
        a2, _ = cavity(hip, storage)
        a2.initialize_uniform()
        a2.step(7)
        a2.save_checkpoint(tmp_path / "body.bin", raw=False)
        c.load_checkpoint(tmp_path / "body.bin")
        a2.step(4)
        c.step(4)
        fluid = flags.reshape(-1) == 0
        m2, mc = a2.download_macroscopic(), c.download_macroscopic()
        for f in ("rho", "ux", "uy", "uz"):
            assert np.array_equal(getattr(m2, f).reshape(-1)[fluid], getattr(mc, f).reshape(-1)[fluid]), f
        for s in (a2, c):
            s.close()
    a.close()
    b.close()


def test_checkpoint_errors(hip, tmp_path):
    s = hip.EsotericPullHipSolver("d3q19", 12, 10, 8, 0.6)
    with pytest.raises(RuntimeError, match="configuration mismatch"):
        s.load_checkpoint(A)  # tau 0.71 in the file
    t = hip.EsotericPullHipSolver("d3q19", 12, 10, 9, float(np.float32(0.71)))
    with pytest.raises(RuntimeError, match="decomposition"):
        t.load_checkpoint(A)
    u, _ = cavity(hip, "fp32")
    walled = tmp_path / "walled.bin"
    u.save_checkpoint(walled)  # raw device state with the wall densities
    u.close()
    plain = hip.EsotericPullHipSolver("d3q19", 20, 18, 16, 0.62)
    with pytest.raises(RuntimeError, match="wall state does not match"):
        plain.load_checkpoint(walled)
    plain.close()
    w = hip.EsotericPullHipSolver("d3q19", 12, 10, 8, float(np.float32(0.71)))
    w.set_flags(np.where(np.arange(960).reshape(8, 10, 12) % 97 == 0, 0x80, 0).astype(np.uint8))
    with pytest.raises(RuntimeError, match="periodic lattice only"):
        w.load_checkpoint(A)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTACHK!" + A.read_bytes()[8:])
    with pytest.raises(RuntimeError, match="magic"):
        s.load_checkpoint(bad)
    with pytest.raises(RuntimeError, match="cannot open"):
        s.load_checkpoint(tmp_path / "missing.bin")
    for x in (s, t, w):
        x.close()
