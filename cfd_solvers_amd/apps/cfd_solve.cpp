// cfd-solve, LBM cases of the HIP backend.  Case names and sizes follow the reference's
// dispatcher (apps/cfd-solve.cpp:59-73, 966-989): the `-sycl` in-place cases become `-hip`, and the
// 2-D wall cases (lbm-poiseuille, lbm-cavity, lbm-channel-io at :225-300) get 3-D HIP
// counterparts on the esoteric-pull kernels.  Output is one key=value line per case, as the
// reference prints (:200-220).
#include "cfd_lbm/esoteric_pull_hip.hpp"

#include <chrono>
#include <cmath>
#include <iostream>
#include <string_view>
#include <vector>

namespace {

struct D2Q9 {
    static constexpr int q = 9;
    static constexpr int dimensions = 2;
};
struct D3Q19 {
    static constexpr int q = 19;
    static constexpr int dimensions = 3;
};
struct D3Q27 {
    static constexpr int q = 27;
    static constexpr int dimensions = 3;
};

void print_list() {
    std::cout << "Available solvers (HIP backend):\n"
              << "  lbm-d2q9-inplace-hip   single-grid periodic BGK D2Q9 (gfx950)\n"
              << "  lbm-d3q19-hip          single-grid periodic BGK D3Q19 (gfx950)\n"
              << "  lbm-d3q27-hip          single-grid periodic BGK D3Q27 (gfx950)\n"
              << "  lbm-poiseuille-hip     force-driven D3Q19 channel, halfway bounce-back walls\n"
              << "  lbm-cavity-hip         D3Q27 moving-lid cavity (5 walls + lid)\n"
              << "  lbm-channel-io-hip     D3Q19 duct with velocity inlet and pressure outlet\n";
}

template <class D> int run_inplace(cfd::lbm::HipLbmConfig config, std::size_t steps) {
    cfd::lbm::EsotericPullHipSolver<D> solver(config);
    solver.initialize_taylor_green(0.02F);
    const double mass0 = solver.total_mass();
    const auto t0 = std::chrono::steady_clock::now();
    solver.step(steps);
    solver.wait();
    const auto t1 = std::chrono::steady_clock::now();
    const double seconds = std::chrono::duration<double>(t1 - t0).count();
    const double mass1 = solver.total_mass();
    const double mlups = static_cast<double>(solver.cells()) * static_cast<double>(steps) / (seconds * 1.0e6);
    std::cout << "backend=hip streaming=in-place device=\"" << solver.device_name() << "\""
              << " cells=" << solver.cells()
              << " population_MiB=" << static_cast<double>(solver.population_bytes()) / (1024.0 * 1024.0)
              << " mass_error_abs=" << std::abs(mass1 - mass0)
              << " mass_error_rel=" << std::abs(mass1 - mass0) / std::abs(mass0)
              << " kinetic_energy=" << solver.kinetic_energy() << " max_speed=" << solver.max_speed()
              << " MLUPS=" << mlups << '\n';
    return 0;
}

std::size_t idx(std::size_t x, std::size_t y, std::size_t z, std::size_t nx, std::size_t ny) {
    return (z * ny + y) * nx + x;
}

int run_poiseuille() {
    const std::size_t nx = 16, ny = 32, nz = 8;
    const float tau = 0.8F, a = 5.0e-6F;
    cfd::lbm::EsotericPullHipSolver<D3Q19> solver(cfd::lbm::HipLbmConfig{nx, ny, nz, tau, a});
    std::vector<std::uint8_t> flags(nx * ny * nz, CFD_LBM_FLUID);
    for (std::size_t z = 0; z < nz; ++z)
        for (std::size_t x = 0; x < nx; ++x) flags[idx(x, 0, z, nx, ny)] = flags[idx(x, ny - 1, z, nx, ny)] = CFD_LBM_WALL;
    solver.set_flags(flags);
    solver.initialize_uniform();
    solver.step(12000);
    const auto m = solver.compute_macroscopic();
    const double nu = (static_cast<double>(tau) - 0.5) / 3.0, height = static_cast<double>(ny - 2);
    double e2 = 0.0, x2 = 0.0;
    for (std::size_t y = 1; y + 1 < ny; ++y) {
        double ux = 0.0;
        for (std::size_t z = 0; z < nz; ++z)
            for (std::size_t x = 0; x < nx; ++x) ux += m.ux[idx(x, y, z, nx, ny)];
        ux /= static_cast<double>(nx * nz);
        const double yw = static_cast<double>(y) - 0.5;
        const double exact = static_cast<double>(a) * yw * (height - yw) / (2.0 * nu);
        e2 += (ux - exact) * (ux - exact);
        x2 += exact * exact;
    }
    std::cout << "case=poiseuille-hip relative_L2=" << std::sqrt(e2 / x2) << " max_speed=" << solver.max_speed()
              << '\n';
    return 0;
}

int run_cavity() {
    const std::size_t n = 64;
    cfd::lbm::EsotericPullHipSolver<D3Q27> solver(cfd::lbm::HipLbmConfig{n, n, n, 0.55F});
    std::vector<std::uint8_t> flags(n * n * n, CFD_LBM_FLUID);
    for (std::size_t z = 0; z < n; ++z)
        for (std::size_t y = 0; y < n; ++y)
            for (std::size_t x = 0; x < n; ++x) {
                auto& f = flags[idx(x, y, z, n, n)];
                if (y == n - 1) f = CFD_LBM_WALL | 1;
                if (x == 0 || x == n - 1 || y == 0 || z == 0 || z == n - 1) f = CFD_LBM_WALL;
            }
    solver.set_flags(flags);
    solver.set_wall_velocity(1, 0.1F, 0.0F, 0.0F);
    solver.initialize_uniform();
    solver.step(2000);
    const auto m = solver.compute_macroscopic();
    const std::size_t c = idx(n / 2, n / 2, n / 2, n, n);
    std::cout << "case=lid-driven-cavity-hip center_ux=" << m.ux[c] << " center_uy=" << m.uy[c]
              << " near_lid_ux=" << m.ux[idx(n / 2, n - 2, n / 2, n, n)] << " max_speed=" << solver.max_speed()
              << '\n';
    return 0;
}

int run_channel_io() {
    const std::size_t nx = 96, ny = 24, nz = 24;
    cfd::lbm::EsotericPullHipSolver<D3Q19> solver(cfd::lbm::HipLbmConfig{nx, ny, nz, 0.8F});
    std::vector<std::uint8_t> flags(nx * ny * nz, CFD_LBM_FLUID);
    for (std::size_t z = 0; z < nz; ++z)
        for (std::size_t y = 0; y < ny; ++y)
            for (std::size_t x = 0; x < nx; ++x) {
                auto& f = flags[idx(x, y, z, nx, ny)];
                if (y == 0 || y == ny - 1 || z == 0 || z == nz - 1) f = CFD_LBM_WALL;
                else if (x == 0) f = CFD_LBM_VELOCITY_INLET;
                else if (x == nx - 1) f = CFD_LBM_PRESSURE_OUTLET;
            }
    solver.set_flags(flags);
    solver.set_inlet_outlet(0.04F, 1.0F);
    solver.initialize_uniform();
    solver.step(6000);
    const auto m = solver.compute_macroscopic();
    auto flux = [&](std::size_t x) {
        double s = 0.0;
        for (std::size_t z = 1; z + 1 < nz; ++z)
            for (std::size_t y = 1; y + 1 < ny; ++y) s += m.rho[idx(x, y, z, nx, ny)] * m.ux[idx(x, y, z, nx, ny)];
        return s;
    };
    std::cout << "case=channel-io-hip inlet_flux=" << flux(0) << " mid_flux=" << flux(nx / 2)
              << " outlet_flux=" << flux(nx - 1) << " inlet_rho_centre=" << m.rho[idx(0, ny / 2, nz / 2, nx, ny)]
              << '\n';
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc < 2 || std::string_view(argv[1]) == "--list") {
            print_list();
            return 0;
        }
        const std::string_view s = argv[1];
        if (s == "lbm-d2q9-inplace-hip") return run_inplace<D2Q9>({256, 256, 1, 0.60F}, 200);
        if (s == "lbm-d3q19-hip") return run_inplace<D3Q19>({64, 64, 64, 0.60F}, 100);
        if (s == "lbm-d3q27-hip") return run_inplace<D3Q27>({56, 56, 56, 0.60F}, 100);
        if (s == "lbm-poiseuille-hip") return run_poiseuille();
        if (s == "lbm-cavity-hip") return run_cavity();
        if (s == "lbm-channel-io-hip") return run_channel_io();
        std::cerr << "unknown solver: " << s << " (the reference build provides the CPU/SYCL cases)\n";
        return 2;
    } catch (const std::exception& e) {
        std::cerr << "cfd-solve: " << e.what() << '\n';
        return 1;
    }
}
