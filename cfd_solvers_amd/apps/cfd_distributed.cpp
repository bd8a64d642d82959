// cfd-distributed with the HIP backend on one GPU: the reference's distributed LBM CLI
// (apps/cfd-distributed.cpp:48-86) with its flags, its run protocol (run_cpu / run_sycl, :142-233:
// restart from PREFIX.rank<r>.bin or Taylor-Green 0.02, timed steps, checkpoint to
// PREFIX.rank<r>.bin) and its report lines (:94-139, :262-270), where `--backend hip` runs the
// periodic lattice through the gfx950 esoteric-pull kernels as rank 0 of a one-rank job.  The
// checkpoint files are the reference's CFDCHK3 layout (distributed_pull.hpp:176-245): the
// reference's DistributedPullBlock restarts from them and this binary restarts from its files.
// There are no halo messages on one rank; the z-slab exchange a multi-GPU run would add is not
// part of this build.  The CPU and SYCL backends belong to the reference's MPI build: asking this
// binary for them exits with code 2.
#include "cfd_lbm/esoteric_pull_hip.hpp"

#include <unistd.h>

#include <chrono>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <string_view>

namespace {

struct D3Q19 {
    static constexpr int q = 19;
    static constexpr int dimensions = 3;
};
struct D3Q27 {
    static constexpr int q = 27;
    static constexpr int dimensions = 3;
};

struct Options {
    std::size_t nx{128}, ny{128}, nz{128};
    std::size_t steps{100};
    float tau{0.7F};
    std::string lattice{"d3q19"};
    std::string backend{"cpu"};
    bool gpu_aware_mpi{false};
    bool csv{false};
    std::string checkpoint_prefix;
    std::string restart_prefix;
};

std::size_t parse_size(std::string_view value, std::string_view name) {
    const unsigned long long parsed = std::stoull(std::string(value));
    if (parsed == 0) throw std::invalid_argument(std::string(name) + " must be positive");
    return static_cast<std::size_t>(parsed);
}

Options parse_options(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string_view arg = argv[i];
        auto value = [&](std::string_view name) -> std::string_view {
            if (i + 1 >= argc) throw std::invalid_argument(std::string(name) + " requires a value");
            return argv[++i];
        };
        if (arg == "--nx") o.nx = parse_size(value(arg), arg);
        else if (arg == "--ny") o.ny = parse_size(value(arg), arg);
        else if (arg == "--nz") o.nz = parse_size(value(arg), arg);
        else if (arg == "--steps") o.steps = parse_size(value(arg), arg);
        else if (arg == "--tau") o.tau = std::stof(std::string(value(arg)));
        else if (arg == "--lattice") o.lattice = std::string(value(arg));
        else if (arg == "--backend") o.backend = std::string(value(arg));
        else if (arg == "--gpu-aware-mpi") o.gpu_aware_mpi = true;
        else if (arg == "--csv") o.csv = true;
        else if (arg == "--checkpoint") o.checkpoint_prefix = std::string(value(arg));
        else if (arg == "--restart") o.restart_prefix = std::string(value(arg));
        else if (arg == "--help") {
            std::cout << "Usage: cfd-distributed [--nx N --ny N --nz N --steps N --tau T] "
                         "[--lattice d3q19|d3q27] [--backend cpu|sycl|hip] [--gpu-aware-mpi] [--csv] "
                         "[--checkpoint PREFIX] [--restart PREFIX]\n\n"
                         "The hip backend runs the periodic lattice on one GPU as rank 0 of a one-rank job "
                         "(no halo exchange); --gpu-aware-mpi has nothing to stage there.\n";
            std::exit(0);
        } else {
            throw std::invalid_argument("unknown option: " + std::string(arg));
        }
    }
    if (o.lattice != "d3q19" && o.lattice != "d3q27") throw std::invalid_argument("--lattice must be d3q19 or d3q27");
    if (o.backend != "cpu" && o.backend != "sycl" && o.backend != "hip")
        throw std::invalid_argument("--backend must be cpu, sycl or hip");
    return o;
}

std::string rank_path(const std::string& prefix, int rank) { return prefix + ".rank" + std::to_string(rank) + ".bin"; }

struct Config {
    std::size_t nx, ny, nz;
    double tau;
    double acceleration_x = 0.0, acceleration_y = 0.0, acceleration_z = 0.0;
    bool smagorinsky_les = false;
};

template <class Descriptor> int run_hip(const Options& o, std::string_view lattice) {
    cfd::lbm::EsotericPullHipSolver<Descriptor> solver(Config{o.nx, o.ny, o.nz, static_cast<double>(o.tau)});
    if (!o.restart_prefix.empty()) solver.load_checkpoint(rank_path(o.restart_prefix, 0));
    else solver.initialize_taylor_green(0.02F);
    solver.wait();
    const auto start = std::chrono::steady_clock::now();
    solver.step(o.steps);
    solver.wait();
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    if (!o.checkpoint_prefix.empty()) solver.save_checkpoint(rank_path(o.checkpoint_prefix, 0));

    const double mass = solver.total_mass();
    const double mlups = static_cast<double>(solver.cells()) * static_cast<double>(o.steps) /
                         (seconds > 1.0e-30 ? seconds : 1.0e-30) / 1.0e6;
    const char* backend = "hip-esoteric";
    if (o.csv) {
        std::cout << backend << ',' << lattice << ",1,1,1,1," << o.nx << ',' << o.ny << ',' << o.nz << ',' << o.steps
                  << ',' << std::fixed << std::setprecision(6) << seconds << ',' << std::setprecision(3) << mlups
                  << ",0,0," << std::setprecision(6) << 0.0 << ',' << std::setprecision(9) << mass << '\n';
    } else {
        std::cout << "rank=0 device=\"" << solver.device_name() << "\"\n";
        std::cout << "backend=" << backend << " lattice=" << lattice << " ranks=1 process_grid=1x1x1"
                  << " global=" << o.nx << 'x' << o.ny << 'x' << o.nz << " steps=" << o.steps
                  << " seconds=" << std::fixed << std::setprecision(6) << seconds << " mlups=" << std::setprecision(3)
                  << mlups << " halo_messages_per_step=0 halo_bytes_per_step=0 halo_gib=" << std::setprecision(6)
                  << 0.0 << " mass=" << std::setprecision(9) << mass << '\n';
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Options o = parse_options(argc, argv);
        if (o.backend != "hip") {
            std::cerr << "cfd-distributed: the " << o.backend
                      << " backend belongs to the reference's MPI build; this binary runs --backend hip\n";
            return 2;
        }
        const int rc = o.lattice == "d3q19" ? run_hip<D3Q19>(o, "d3q19") : run_hip<D3Q27>(o, "d3q27");
        if (!o.csv) {
            char host[256] = {};
            if (gethostname(host, sizeof(host) - 1) != 0) host[0] = '\0';
            std::cout << "rank=0 local_rank=0/1 node=" << host << " brick=" << o.nx << 'x' << o.ny << 'x' << o.nz
                      << " begin=0,0,0\n";
        }
        return rc;
    } catch (const std::exception& e) {
        std::cerr << "cfd-distributed: " << e.what() << '\n';
        return 1;
    }
}
