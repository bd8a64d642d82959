"""Writes tests/golden/cavity_d3q27_256_step500.npz: the D3Q27 256^3 lid-driven cavity (walls on
five faces, lid y = ny-1 moving with u = (0.1, 0, 0), tau 0.55, rho 1 and u 0 at start, FP32
populations) after 500 steps of the CPU restatement in oracle/ (corrected three-pass bounded
step, moving-wall correction of src/lbm/d2q9.cpp:217-223).

The full fields are 268 MB, so the fixture keeps what pins them: the SHA-256 of the float32 bytes
of rho, ux, uy, uz over the fluid cells (in cell order), the centreline profiles ux(y) at
x = z = 128 and uy(x) at y = z = 128, and the FP64 mass.  The GPU test recomputes the same digest
from the HIP path.  Runs about 15 minutes on 8 cores:

    python tests/golden/make_cavity_d3q27.py
"""
import hashlib
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402

N, STEPS, TAU, LID = 256, 500, 0.55, 0.1


def cavity_flags(n):
    f = np.zeros((n, n, n), np.uint8)
    f[:, -1, :] = 0x81
    f[:, :, 0] = f[:, :, -1] = 0x80
    f[:, 0, :] = 0x80
    f[0, :, :] = f[-1, :, :] = 0x80
    return f


def digest(fields, fluid):
    h = hashlib.sha256()
    for a in fields:
        h.update(np.ascontiguousarray(a[fluid], dtype=np.float32).tobytes())
    return h.hexdigest()


def main():
    flags = cavity_flags(N)
    fluid = flags == 0
    o = O.Oracle(27, N, N, N, float(np.float32(TAU)), storage="fp32")
    o.set_flags(flags.reshape(-1))
    o.set_wall_velocity(1, (float(np.float32(LID)), 0.0, 0.0))
    o.init_uniform()
    t0 = time.time()
    for k in range(0, STEPS, 50):
        o.step(50)
        print(f"step {k + 50}: {time.time() - t0:.0f} s", flush=True)
    rho, ux, uy, uz = o.macroscopic()
    mid = N // 2
    np.savez_compressed(
        Path(__file__).with_name("cavity_d3q27_256_step500.npz"),
        digest=np.array(digest((rho, ux, uy, uz), fluid)),
        ux_y=ux[mid, :, mid].astype(np.float32), uy_x=uy[mid, mid, :].astype(np.float32),
        mass=np.array(o.mass()), steps=np.array(STEPS))
    print("written", flush=True)


if __name__ == "__main__":
    main()
