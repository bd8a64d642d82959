// cfd-bench with the HIP backend: the reference's LBM benchmark CLI (apps/cfd-bench.cpp:47-85)
// with its flags, protocol (:174-191, :215-233) and output schema (:101-171), where
// `--backend hip` runs the gfx950 esoteric-pull kernels through the C ABI.  The CPU and SYCL
// backends belong to the reference build; asking this binary for them exits with code 2, as the
// reference does for an SYCL backend it was not compiled with.
//
// Added flags: --storage fp32|fp16s|fp64|fp32s|fp64x (population storage of the HIP backend).
#include "cfd_lbm/esoteric_pull_hip.hpp"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <string_view>

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

struct Options {
    std::string lattice{"d3q19"};
    std::string backend{"cpu"};
    std::string streaming{"in-place"};
    std::string storage{"fp32"};
    std::size_t nx{64}, ny{64}, nz{64};
    std::size_t warmup{10}, steps{100};
    unsigned threads{0};
    float tau{0.6F};
    bool csv{false}, json{false};
};

std::size_t parse_size(const char* text, std::string_view option) {
    const auto value = std::strtoull(text, nullptr, 10);
    if (value == 0ULL) throw std::invalid_argument(std::string(option) + " must be > 0");
    return static_cast<std::size_t>(value);
}

float parse_tau(const char* text, std::string_view option) {
    const float value = std::strtof(text, nullptr);
    if (!(value > 0.5F)) throw std::invalid_argument(std::string(option) + " must be > 0.5");
    return value;
}

Options parse_options(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string_view arg = argv[i];
        const auto value = [&](std::string_view name) -> const char* {
            if (i + 1 >= argc) throw std::invalid_argument(std::string(name) + " requires a value");
            return argv[++i];
        };
        if (arg == "--lattice") o.lattice = value(arg);
        else if (arg == "--backend") o.backend = value(arg);
        else if (arg == "--streaming") o.streaming = value(arg);
        else if (arg == "--storage") o.storage = value(arg);
        else if (arg == "--nx") o.nx = parse_size(value(arg), arg);
        else if (arg == "--ny") o.ny = parse_size(value(arg), arg);
        else if (arg == "--nz") o.nz = parse_size(value(arg), arg);
        else if (arg == "--warmup") o.warmup = parse_size(value(arg), arg);
        else if (arg == "--steps") o.steps = parse_size(value(arg), arg);
        else if (arg == "--threads") o.threads = static_cast<unsigned>(parse_size(value(arg), arg));
        else if (arg == "--tau") o.tau = parse_tau(value(arg), arg);
        else if (arg == "--csv") o.csv = true;
        else if (arg == "--json") o.json = true;
        else if (arg == "--help" || arg == "-h") {
            std::cout << "Usage: cfd-bench [options]\n"
                      << "  --lattice d2q9|d3q19|d3q27\n"
                      << "  --backend cpu|sycl|hip   (this build runs hip)\n"
                      << "  --streaming in-place|pull|both   (pull/both are CPU-only)\n"
                      << "  --storage fp32|fp16s|fp64|fp32s|fp64x  (hip population storage)\n"
                      << "  --nx N --ny N --nz N\n"
                      << "  --warmup N --steps N --tau T\n"
                      << "  --threads N      OpenMP thread count (CPU)\n"
                      << "  --csv            CSV output\n"
                      << "  --json           newline-delimited JSON output\n";
            std::exit(0);
        } else {
            throw std::invalid_argument("unknown option: " + std::string(arg));
        }
    }
    if (o.csv && o.json) throw std::invalid_argument("--csv and --json are mutually exclusive");
    return o;
}

std::string json_escape(std::string_view text) {
    std::string out;
    for (char c : text) {
        if (c == '\\' || c == '"') {
            out.push_back('\\');
            out.push_back(c);
        } else if (c == '\n') out += "\\n";
        else out.push_back(c);
    }
    return out;
}

cfd::lbm::HipPopulationStorage storage_of(const std::string& s) {
    if (s == "fp32") return cfd::lbm::HipPopulationStorage::fp32;
    if (s == "fp16s") return cfd::lbm::HipPopulationStorage::fp16_shifted;
    if (s == "fp64") return cfd::lbm::HipPopulationStorage::fp64;
    if (s == "fp32s") return cfd::lbm::HipPopulationStorage::fp32_shifted;
    if (s == "fp64x") return cfd::lbm::HipPopulationStorage::fp64_exact_weights;
    throw std::invalid_argument("storage must be fp32, fp16s, fp64, fp32s or fp64x");
}

std::size_t bytes_per_population(const std::string& s) {
    return s == "fp16s" ? 2U : (s == "fp64" || s == "fp64x" ? 8U : 4U);
}

template <class D> int run_hip(const Options& o) {
    if (o.streaming != "in-place") throw std::invalid_argument("HIP benchmark supports only --streaming in-place");
    cfd::lbm::HipLbmConfig config{o.nx, o.ny, o.nz, o.tau};
    cfd::lbm::EsotericPullHipSolver<D> solver(config, storage_of(o.storage));
    solver.initialize_taylor_green(0.02F);
    const double mass0 = solver.total_mass();
    solver.step(o.warmup);
    solver.wait();
    const auto begin = std::chrono::steady_clock::now();
    solver.step(o.steps);
    solver.wait();
    const auto end = std::chrono::steady_clock::now();
    const double seconds = std::chrono::duration<double>(end - begin).count();
    const double mass1 = solver.total_mass();

    const std::size_t cells = solver.cells();
    const double mlups = static_cast<double>(cells) * static_cast<double>(o.steps) / seconds / 1.0e6;
    // Each update reads and writes every population once (apps/cfd-bench.cpp:114-117), at the
    // storage width actually used.
    const double bytes_per_update = 2.0 * D::q * static_cast<double>(bytes_per_population(o.storage));
    const double gbps = mlups * bytes_per_update / 1000.0;
    const double allocated_mib = static_cast<double>(solver.population_bytes()) / (1024.0 * 1024.0);
    const double single_mib = allocated_mib;
    const double mass_error = std::abs(mass1 - mass0);
    const double mass_rel = mass0 != 0.0 ? mass_error / std::abs(mass0) : 0.0;
    const std::string device = solver.device_name();
    const std::size_t nz = D::dimensions == 2 ? 1 : o.nz;

    if (o.csv) {
        std::cout << "lattice,backend,streaming,device,nx,ny,nz,q,steps,seconds,mlups,estimated_ddf_gbps,"
                     "allocated_ddf_mib,single_grid_ddf_mib,mass_error_abs,mass_error_rel,storage\n";
        std::cout << o.lattice << ",hip,in-place,\"" << device << "\"," << o.nx << ',' << o.ny << ',' << nz << ','
                  << D::q << ',' << o.steps << ',' << std::setprecision(9) << seconds << ',' << mlups << ','
                  << gbps << ',' << allocated_mib << ',' << single_mib << ',' << mass_error << ',' << mass_rel
                  << ',' << o.storage << '\n';
        return 0;
    }
    if (o.json) {
        std::cout << std::setprecision(12) << "{\"schema\":\"cfd_solvers.benchmark.v1\","
                  << "\"lattice\":\"" << json_escape(o.lattice) << "\",\"backend\":\"hip\","
                  << "\"streaming\":\"in-place\",\"device\":\"" << json_escape(device) << "\","
                  << "\"nx\":" << o.nx << ",\"ny\":" << o.ny << ",\"nz\":" << nz << ",\"q\":" << D::q
                  << ",\"steps\":" << o.steps << ",\"seconds\":" << seconds << ",\"mlups\":" << mlups
                  << ",\"estimated_ddf_gbps\":" << gbps << ",\"allocated_ddf_mib\":" << allocated_mib
                  << ",\"single_grid_ddf_mib\":" << single_mib << ",\"mass_error_abs\":" << mass_error
                  << ",\"mass_error_rel\":" << mass_rel << ",\"storage\":\"" << o.storage
                  << "\",\"roofline_fraction\":" << gbps / 8000.0 << "}\n";
        return 0;
    }
    std::cout << "lattice=" << o.lattice << " backend=hip streaming=in-place device=\"" << device << "\"\n"
              << "grid=" << o.nx << 'x' << o.ny << 'x' << nz << " q=" << D::q << " steps=" << o.steps
              << " storage=" << o.storage << '\n'
              << std::fixed << std::setprecision(3) << "time_s=" << seconds << " MLUPS=" << mlups
              << " estimated_DDF_GB/s=" << gbps << " roofline_%=" << 100.0 * gbps / 8000.0 << '\n'
              << "allocated_DDF_MiB=" << allocated_mib << " single_grid_DDF_MiB=" << single_mib << '\n'
              << std::scientific << "mass_error_abs=" << mass_error << " mass_error_rel=" << mass_rel << '\n';
    return 0;
}

template <class D> int dispatch(const Options& o) {
    if (o.backend == "hip") return run_hip<D>(o);
    if (o.backend == "cpu" || o.backend == "sycl") {
        std::cerr << o.backend << " backend is provided by the reference build; this binary runs --backend hip\n";
        return 2;
    }
    throw std::invalid_argument("backend must be cpu, sycl, or hip");
}

}  // namespace

int main(int argc, char** argv) {
    try {
        Options o = parse_options(argc, argv);
        if (o.lattice == "d2q9") {
            o.nz = 1;
            return dispatch<D2Q9>(o);
        }
        if (o.lattice == "d3q19") return dispatch<D3Q19>(o);
        if (o.lattice == "d3q27") return dispatch<D3Q27>(o);
        throw std::invalid_argument("lattice must be d2q9, d3q19, or d3q27");
    } catch (const std::exception& error) {
        std::cerr << "cfd-bench: " << error.what() << '\n';
        return 1;
    }
}
