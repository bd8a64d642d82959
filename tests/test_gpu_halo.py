"""z-slab decomposition on one GPU (cfd_lbm_halo_pack / cfd_lbm_halo_unpack): a periodic lattice cut
into z-slabs, one solver per slab, exchanging the crossing populations after every step (the
selective halo of include/cfd/solvers/lbm/distributed_pull.hpp:326-395 and
include/cfd/distributed/selective_halo.hpp:15-21, here for the in-place esoteric slots), holds bit
for bit the state of the whole lattice: logical populations and macroscopic fields after an even and
an odd number of steps, for D3Q19 and D3Q27, FP32, FP16S and FP64, forced, on ragged extents.  The
messages travel between the slabs' own device buffers, as an RCCL send/receive between GPUs would
carry them; on one GPU one slab that is its own neighbour receives its own messages unchanged."""
import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu

ACC = (2e-6, -1e-6, 5e-7)


def whole_and_slabs(hip, lattice, storage, shape, cuts):
    nx, ny, nz = shape
    g = hip.EsotericPullHipSolver(lattice, nx, ny, nz, 0.7, ACC, storage=storage)
    g.initialize_taylor_green(0.03, 1e-3, 7)
    pops = g.logical_populations().reshape(-1, nz, ny, nx)
    g.upload_populations(pops)  # the whole lattice and the slabs start from the same stored values
    slabs, z0 = [], 0
    for z1 in list(cuts) + [nz]:
        s = hip.EsotericPullHipSolver(lattice, nx, ny, z1 - z0, 0.7, ACC, storage=storage)
        s.upload_populations(np.ascontiguousarray(pops[:, z0:z1]))
        slabs.append(s)
        z0 = z1
    return g, slabs


def run_slabs(slabs, dev, steps):
    nbytes = slabs[0].halo_message_bytes()
    up = [dev.alloc(nbytes) for _ in slabs]
    down = [dev.alloc(nbytes) for _ in slabs]
    k = len(slabs)
    for _ in range(steps):
        for s in slabs:
            s.step(1)
        for i, s in enumerate(slabs):
            s.halo_pack(up[i], down[i])
        for s in slabs:
            s.wait()  # every message is packed before any slot is overwritten
        for i, s in enumerate(slabs):
            s.halo_unpack(up[(i - 1) % k], down[(i + 1) % k])  # periodic in z: the ring of slabs
        for s in slabs:
            s.wait()


@pytest.mark.parametrize("lattice,storage,shape,cuts", [
    ("d3q19", "fp32", (20, 12, 14), (5,)),
    ("d3q19", "fp16s", (20, 12, 14), (9,)),
    ("d3q19", "fp64", (17, 9, 10), (4,)),
    ("d3q27", "fp32", (20, 12, 14), (6,)),
    ("d3q19", "fp32", (33, 10, 14), (2, 7)),      # three slabs, one of the minimum depth 2
    ("d3q27", "fp16s", (16, 8, 12), (3, 6, 9)),   # four slabs
])
@pytest.mark.parametrize("steps", [7, 8])
def test_slabs_equal_the_whole_lattice(hip, dev, lattice, storage, shape, cuts, steps):
    report("halo_slabs", lattice=lattice, storage=storage, shape=shape, cuts=cuts, steps=steps)
    g, slabs = whole_and_slabs(hip, lattice, storage, shape, cuts)
    q = g.q
    per = {"d3q19": 5, "d3q27": 9}[lattice]
    width = {"fp32": 4, "fp16s": 2, "fp64": 8}[storage]
    assert slabs[0].halo_message_bytes() == per * shape[0] * shape[1] * width
    g.step(steps)
    run_slabs(slabs, dev, steps)
    want = g.logical_populations().reshape(q, shape[2], shape[1], shape[0])
    got = np.concatenate([s.logical_populations().reshape(q, -1, shape[1], shape[0]) for s in slabs], axis=1)
    assert np.array_equal(got, want), f"{int((got != want).sum())} populations differ"
    mw = g.download_macroscopic()
    for name in ("rho", "ux", "uy", "uz"):
        part = np.concatenate([getattr(s.download_macroscopic(), name) for s in slabs], axis=0)
        assert np.array_equal(part, getattr(mw, name)), name
    for s in slabs + [g]:
        s.close()


def test_one_slab_is_its_own_neighbour(hip, dev):
    report("halo_self")
    a, _ = whole_and_slabs(hip, "d3q19", "fp32", (20, 12, 10), ())
    b, _ = whole_and_slabs(hip, "d3q19", "fp32", (20, 12, 10), ())
    a.step(5)
    run_slabs([b], dev, 5)
    assert np.array_equal(a.logical_populations(), b.logical_populations())


def test_halo_rejections(hip, dev):
    report("halo_rejections")
    buf = dev.alloc(1 << 16)
    s = hip.EsotericPullHipSolver("d3q19", 8, 8, 8, 0.6)
    s.initialize_uniform()
    with pytest.raises(RuntimeError, match="no step"):
        s.halo_pack(buf, buf)
    flags = np.zeros(512, np.uint8)
    flags[:64] = 0x80
    s.set_flags(flags)
    s.step(1)
    with pytest.raises(RuntimeError, match="without flags"):
        s.halo_pack(buf, buf)
    d2 = hip.EsotericPullHipSolver("d2q9", 8, 8, 1, 0.6)
    with pytest.raises(RuntimeError, match="3-D"):
        d2.halo_message_bytes()
