"""The C++ front end (include/cfd_lbm/esoteric_pull_hip.hpp, the reference's EsotericPullSolver<D>
surface of include/cfd/solvers/lbm/esoteric_pull.hpp:104-278 plus the device extras of
esoteric_pull_sycl.hpp) is a class template: a member is only checked when it is instantiated.
This instantiates every public member, the surface voxelisation and the checkpoint calls with
stand-in types of the reference's shapes (descriptor, config, TriangleSurface, VoxelGrid3D) and
compiles the result with g++ -fsyntax-only.  No GPU, no link."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

SOURCE = r"""
#include "cfd_lbm/esoteric_pull_hip.hpp"
#include <vector>

struct Point { double x, y, z; };
struct Triangle { Point a, b, c; };
struct Surface { std::vector<Triangle> triangles; };                     // cfd::io::TriangleSurface
struct Grid { std::size_t nx, ny, nz; Point minimum, maximum; };          // VoxelGrid3D
struct D3Q19 { static constexpr int q = 19; static constexpr int dimensions = 3; };
struct D2Q9 { static constexpr int q = 9; static constexpr int dimensions = 2; };
struct Config {                                                           // InPlaceLbmConfig
    std::size_t nx = 8, ny = 8, nz = 8;
    double tau = 0.6, acceleration_x = 0, acceleration_y = 0, acceleration_z = 0;
    bool smagorinsky_les = false;
};

template <class D> double exercise() {
    using cfd::lbm::EsotericPullHipSolver;
    using cfd::lbm::HipPopulationStorage;
    EsotericPullHipSolver<D> s(Config{}, HipPopulationStorage::fp16_shifted);
    const std::vector<float> cells(s.cells(), 0.0F);
    const std::vector<std::uint8_t> mask(s.cells(), 0);
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
    s.set_solid_surface(Surface{}, Point{0, 0, 0}, Point{1, 1, 1});
    s.set_solid_surface(Surface{}, Point{0, 0, 0}, Point{1, 1, 1}, true);
    s.step(2);
    s.wait();
    s.save_checkpoint("state.bin");
    s.save_checkpoint("state.bin", false);
    s.load_checkpoint("state.bin");
    const auto m = s.compute_macroscopic();
    const auto d = s.download_macroscopic();
    const auto v = s.refresh_device_macroscopic();
    const auto pops = s.logical_populations();
    s.upload_populations(pops);
    const auto raw = s.raw_storage();
    const auto acc = s.device_local_acceleration(true);
    (void)s.device_local_acceleration();
    (void)raw; (void)acc; (void)s.queue();
    const auto stats = s.transfer_stats();
    s.reset_transfer_stats();
    (void)v; (void)stats; (void)d;
    return s.mass() + s.total_mass() + s.kinetic_energy() + s.max_speed() + static_cast<double>(m.rho.size()) +
           static_cast<double>(s.time_step() + s.population_bytes() + s.two_lattice_population_bytes() +
                               s.resident_bytes() + s.device_name().size()) +
           (s.has_local_body_acceleration() ? 1.0 : 0.0) + (s.stream() ? 1.0 : 0.0) +
           static_cast<double>(s.config().nx) + static_cast<double>(static_cast<int>(s.storage()));
}

int main() {
    const auto mask = cfd::lbm::voxelize_surface_hip(Surface{}, Grid{4, 4, 4, {0, 0, 0}, {1, 1, 1}});
    return static_cast<int>(exercise<D3Q19>() + exercise<D2Q9>()) + static_cast<int>(mask.size());
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_every_member_compiles(tmp_path):
    src = tmp_path / "front_end.cpp"
    src.write_text(SOURCE)
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I",
                        str(ROOT / "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
