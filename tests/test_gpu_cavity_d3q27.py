"""Config 2 at full size: D3Q27 256^3 lid-driven cavity (walls on five faces, lid y = ny-1 moving
with u = (0.1, 0, 0), tau 0.55, rho 1, u 0 at start, FP32, uint8 flags), 5000 steps.

At step 500 the HIP path is checked against the CPU restatement's run (tests/golden/
cavity_d3q27_256_step500.npz, written by make_cavity_d3q27.py): the SHA-256 of rho, ux, uy, uz over
all fluid cells and the centreline profiles ux(y), uy(x) must agree bit for bit.  The run then
continues to step 5000, where the centreline profiles must show the primary vortex (return flow
under the lid).  Mass drifts slowly (1.7e-4 of the total after 5000 steps): the moving-wall
correction -6 w rho (c . u_w) of a population bounced off the lid has no mirror term where its
partner direction hits a resting side wall at the lid's edges, so the bound is 1e-3.
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "cavity_d3q27_256_step500.npz"


def cavity_flags(n):
    f = np.zeros((n, n, n), np.uint8)
    f[:, -1, :] = 0x81
    f[:, :, 0] = f[:, :, -1] = 0x80
    f[:, 0, :] = 0x80
    f[0, :, :] = f[-1, :, :] = 0x80
    return f


def test_config2_d3q27_cavity_256(hip):
    fx = np.load(FIXTURE)
    n, mid = 256, 128
    flags = cavity_flags(n)
    fluid = flags == 0
    s = hip.EsotericPullHipSolver(27, n, n, n, 0.55, storage="fp32")
    s.set_flags(flags.reshape(-1))
    s.set_wall_velocity(1, (0.1, 0.0, 0.0))
    s.initialize_uniform()
    mass0 = s.total_mass()
    s.step(int(fx["steps"]))
    m = s.download_macroscopic()
    h = hashlib.sha256()
    for a in (m.rho, m.ux, m.uy, m.uz):
        h.update(np.ascontiguousarray(a.reshape(n, n, n)[fluid], dtype=np.float32).tobytes())
    report("config2_step500", digest=h.hexdigest()[:16], fixture=str(fx["digest"])[:16])
    assert np.array_equal(m.ux.reshape(n, n, n)[mid, :, mid], fx["ux_y"])
    assert np.array_equal(m.uy.reshape(n, n, n)[mid, mid, :], fx["uy_x"])
    assert h.hexdigest() == str(fx["digest"])
    assert s.total_mass() == pytest.approx(float(fx["mass"]), rel=1e-12)
    s.step(5000 - int(fx["steps"]))
    m = s.download_macroscopic()
    ux_y = m.ux.reshape(n, n, n)[mid, :, mid].astype(np.float64)
    uy_x = m.uy.reshape(n, n, n)[mid, mid, :].astype(np.float64)
    mass1 = s.total_mass()
    report("config2_step5000", ux_min=ux_y.min(), ux_min_y=int(ux_y.argmin()), ux_below_lid=ux_y[-2],
           uy_min=uy_x.min(), uy_max=uy_x.max(), mass_drift=abs(mass1 - mass0) / mass0)
    assert np.all(np.isfinite(ux_y)) and np.all(np.isfinite(uy_x))
    assert ux_y[-2] > 0.05          # dragged by the lid
    assert ux_y[1:n - 2].min() < -1e-3  # return flow under the lid layer (the vortex is ~100 cells deep)
    assert uy_x.max() > 0 > uy_x.min()  # up on one side, down on the other
    assert abs(mass1 - mass0) / mass0 < 1e-3
    s.close()
