"""The drop-in boundary compiled against the reference's own code (CPU; skipped without the
reference tree).

1. The C++ front end (include/cfd_lbm/esoteric_pull_hip.hpp) with the reference's own types:
   InPlaceLbmConfig (include/cfd/solvers/lbm/esoteric_pull.hpp:20-29, float tau), D3Q19Descriptor /
   D3Q27Descriptor (descriptors.hpp:59-146), D2Q9InPlaceDescriptor, cfd::io::TriangleSurface
   (include/cfd/io/mesh_io.hpp:9-10) and VoxelGrid3D (include/cfd/solvers/lbm/voxelize.hpp:8).
   Every member is instantiated for the three lattices; the program is compiled with -Wall -Wextra
   -Werror and linked to libcfd_lbm_hip.so.
2. The reference's apps/cfd-bench.cpp with INTEGRATION.md's three insertions (the blocks are read
   from INTEGRATION.md itself): compiled with -DCFD_HAS_HIP_LBM, linked to libcfd_lbm_hip.so, its
   CPU backend still runs and prints the reference's output, and `--backend hip` reaches the HIP
   library (without a GPU it fails with the runtime's message and exit code 1, not a crash).
"""
import json
import shutil
import subprocess
from pathlib import Path

import pytest

from oracle import integration

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
LIB = ROOT / "cfd_solvers_amd" / "libcfd_lbm_hip.so"

pytestmark = [
    pytest.mark.skipif(not (REF / "apps" / "cfd-bench.cpp").exists(), reason="reference tree absent"),
    pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available"),
    pytest.mark.skipif(not LIB.exists(), reason="libcfd_lbm_hip.so not built"),
]

FRONT_END = r"""
#include "cfd/io/mesh_io.hpp"
#include "cfd/solvers/lbm/descriptors.hpp"
#include "cfd/solvers/lbm/esoteric_pull.hpp"
#include "cfd/solvers/lbm/voxelize.hpp"
#include "cfd_lbm/esoteric_pull_hip.hpp"

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

static_assert(std::is_same_v<decltype(cfd::lbm::InPlaceLbmConfig{}.tau), float>);

template <class D> double exercise() {
    using cfd::lbm::EsotericPullHipSolver;
    using cfd::lbm::HipPopulationStorage;
    const cfd::lbm::InPlaceLbmConfig config{8, 8, D::dimensions == 2 ? 1U : 8U, 0.6F, 1e-6F, 0.0F, 0.0F, false};
    EsotericPullHipSolver<D> s(config, HipPopulationStorage::fp16_shifted);
    EsotericPullHipSolver<D> plain(config);
    const std::vector<float> cells(s.cells(), 0.0F);
    const std::vector<std::uint8_t> mask(s.cells(), 0);
    cfd::io::TriangleSurface surface;
    surface.triangles.push_back({{0, 0, 0}, {1, 0, 0}, {0, 1, 0}});
    s.initialize_uniform();
    s.initialize_taylor_green(0.02F, 0.0F, 1);
    s.initialize_fields(cells, cells, cells, cells);
    s.set_solid_mask(mask);
    s.set_flags(mask);
    s.clear_solid_mask();
    s.set_wall_velocity(1, 0.1F, 0.0F, 0.0F);
    s.set_porosity(cells);
    s.clear_porosity();
    s.set_inlet_outlet(0.05F);
    s.set_local_body_acceleration(cells, cells, cells);
    s.clear_local_body_acceleration();
    s.set_solid_surface(surface, cfd::fem::Point3{0, 0, 0}, cfd::fem::Point3{1, 1, 1});
    s.set_solid_surface(surface, cfd::fem::Point3{0, 0, 0}, cfd::fem::Point3{1, 1, 1}, true);
    s.step(2);
    s.wait();
    s.save_checkpoint("state.bin");
    s.load_checkpoint("state.bin");
    const auto m = s.compute_macroscopic();
    const auto d = s.download_macroscopic();
    const auto v = s.refresh_device_macroscopic();
    const auto pops = s.logical_populations();
    s.upload_populations(pops);
    const auto raw = s.raw_storage();
    const auto acc = s.device_local_acceleration(true);
    const auto stats = s.transfer_stats();
    s.reset_transfer_stats();
    (void)raw; (void)acc; (void)v; (void)stats; (void)d; (void)s.queue();
    return s.mass() + s.total_mass() + s.kinetic_energy() + s.max_speed() + static_cast<double>(m.rho.size()) +
           static_cast<double>(s.time_step() + s.population_bytes() + s.two_lattice_population_bytes() +
                               s.resident_bytes() + s.device_name().size() + plain.cells()) +
           (s.has_local_body_acceleration() ? 1.0 : 0.0) + static_cast<double>(s.config().nx);
}

int main(int argc, char**) {
    if (argc < 2) return 0;  // linked, not run: the solver needs a GPU
    cfd::io::TriangleSurface surface;
    const cfd::lbm::VoxelGrid3D grid{4, 4, 4, {0, 0, 0}, {1, 1, 1}};
    const auto mask = cfd::lbm::voxelize_surface_hip(surface, grid);
    return static_cast<int>(exercise<cfd::lbm::D3Q19Descriptor>() + exercise<cfd::lbm::D3Q27Descriptor>() +
                            exercise<cfd::lbm::D2Q9InPlaceDescriptor>()) +
           static_cast<int>(mask.size());
}
"""


def test_front_end_with_reference_types(tmp_path):
    src = tmp_path / "front_end.cpp"
    src.write_text(FRONT_END)
    exe = tmp_path / "front_end"
    r = subprocess.run(["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", f"-I{REF / 'include'}",
                        f"-I{ROOT / 'include'}", str(src), "-o", str(exe), f"-L{LIB.parent}", "-lcfd_lbm_hip",
                        f"-Wl,-rpath,{LIB.parent}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    # The program loads the HIP library and exits before touching a device.
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, run.stderr


def test_integration_blocks_are_complete():
    blocks = integration.integration_blocks()
    assert '#include "cfd_lbm/esoteric_pull_hip.hpp"' in blocks["include"]
    assert "int run_hip(const Options& options)" in blocks["runner"]
    assert 'options.backend == "hip"' in blocks["dispatch"]
    patched = integration.patched_cfd_bench(REF)
    # The header sits before the anonymous namespace, the runner inside it, after run_sycl.
    assert patched.index("esoteric_pull_hip.hpp") < patched.index("namespace {")
    assert patched.index("int run_sycl(") < patched.index("int run_hip(") < patched.index("int dispatch_backend(")


def test_patched_reference_cfd_bench(tmp_path):
    exe = tmp_path / "cfd-bench"
    r = integration.compile_cfd_bench(REF, exe)
    assert r.returncode == 0, r.stderr[-4000:]
    common = ["--lattice", "d3q19", "--nx", "8", "--ny", "8", "--nz", "8", "--warmup", "1", "--steps", "2", "--json"]
    cpu = subprocess.run([str(exe), *common, "--threads", "1"], capture_output=True, text=True, timeout=120)
    assert cpu.returncode == 0, cpu.stderr
    row = json.loads(cpu.stdout)
    assert row["schema"] == "cfd_solvers.benchmark.v1" and row["backend"] == "cpu"
    assert row["streaming"] == "in-place" and row["q"] == 19
    hip = subprocess.run([str(exe), *common, "--backend", "hip"], capture_output=True, text=True, timeout=120)
    if hip.returncode == 0:  # a GPU is visible
        row = json.loads(hip.stdout)
        assert row["backend"] == "hip" and row["mass_error_rel"] < 1e-6
    else:
        assert hip.returncode == 1, (hip.returncode, hip.stderr)
        assert "hip" in hip.stderr.lower(), hip.stderr
    # The unpatched bench still rejects the backend the patch adds.
    plain = tmp_path / "cfd-bench-plain"
    r = integration.compile_cfd_bench(REF, plain, hip=False)
    assert r.returncode == 0, r.stderr[-4000:]
    rej = subprocess.run([str(plain), *common, "--backend", "hip"], capture_output=True, text=True, timeout=60)
    assert rej.returncode != 0 and "backend must be cpu or sycl" in rej.stderr
