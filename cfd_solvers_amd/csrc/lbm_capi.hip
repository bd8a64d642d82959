// lbm_capi.hip -- device-resident solver object behind the C ABI of include/cfd_lbm/c_api.h.
//
// Mirrors EsotericPullSyclSolver<D> (include/cfd/solvers/lbm/esoteric_pull_sycl.hpp:18-475):
// the solver owns one padded q-plane population array in HBM, an in-order HIP stream, lazily
// allocated device scratch for macroscopic fields and reductions, and a host/device transfer
// ledger (include/cfd/core/device_residency.hpp:11-42).  Error convention of
// src/python/c_api.cpp:8-15: 0 ok, 1 std::exception, 2 unknown, message in a caller buffer.
#include "../../include/cfd_lbm/c_api.h"
#include "checkpoint_schedule.hpp"
#include "lbm_adjoint.hpp"
#include "lbm_kernels.hpp"
#include "solver_view.hpp"
#include "voxelize.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace {

using cfd_hip::Geometry;
using cfd_hip::KernelOps;
using cfd_hip::LaunchState;
using cfd_hip::Physics;

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct TransferStats {
    uint64_t h2d = 0, d2h = 0, d2d = 0, syncs = 0;
};

// A timing event, destroyed on every way out of the scope that made it.
struct TimingEvent {
    hipEvent_t e = nullptr;
    TimingEvent() { hip_check(hipEventCreate(&e), "hipEventCreate"); }
    ~TimingEvent() { (void)hipEventDestroy(e); }
    TimingEvent(const TimingEvent&) = delete;
    TimingEvent& operator=(const TimingEvent&) = delete;
};

template <class T> struct DeviceBuffer {
    T* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { reset(); }
    void allocate(size_t n_bytes) {
        reset();
        void* p = nullptr;
        hip_check(hipMalloc(&p, n_bytes), "hipMalloc");
        ptr = static_cast<T*>(p);
        bytes = n_bytes;
    }
    void reset() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    explicit operator bool() const { return ptr != nullptr; }
};

}  // namespace

namespace {
// Zero fill by a kernel rather than hipMemsetAsync, so that callers may capture it in a graph.
__global__ void __launch_bounds__(256) zero_words_kernel(uint32_t* p, uint64_t words) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < words) p[i] = 0U;
}
}  // namespace

namespace {
// Porosity from device memory: gamma (cells floats) into the solver's plane of `real_bytes`-wide
// reals; `bad` counts the values outside [0, 1] (NaN included).
__global__ void __launch_bounds__(256) porosity_from_device(const float* __restrict__ src, void* __restrict__ dst,
                                                            uint32_t real_bytes, uint64_t cells, uint32_t* bad) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const float g = src[i];
    if (!(g >= 0.0F && g <= 1.0F)) atomicAdd(bad, 1U);
    if (real_bytes == 4) static_cast<float*>(dst)[i] = g;
    else static_cast<double*>(dst)[i] = static_cast<double>(g);
}
}  // namespace

struct cfd_lbm_solver {
    int q = 19, storage = 0, device = 0;
    uint64_t pop_elems = 0;  // elements of the population array (layout of Geometry)
    size_t nx = 0, ny = 0, nz = 0, cells = 0;
    float tau = 0.6F, ax = 0, ay = 0, az = 0;
    size_t elem = 4, real = 4;
    Geometry geo{};
    hipStream_t stream = nullptr;
    const KernelOps* ops = nullptr;
    DeviceBuffer<unsigned char> f, flags, links, link_vel, segments, solid_rho, gamma, lacc, macro, partial, reduce_out;
    float wall_u[8][3] = {};
    float u_peak = 0.0F, rho_out = 1.0F;
    bool les = false;  // Smagorinsky LES (esoteric_pull.hpp:72-97, InPlaceLbmConfig::smagorinsky_les)
    uint64_t t = 0;
    uint32_t step_lds = 0, step_lds_periodic = 0, adjoint_lds = 0;
    std::vector<uint8_t> host_flags;  // copy of the flags (empty: periodic)
    float last_forward_ms = 0.0F, last_reverse_ms = 0.0F;
    uint64_t last_forward_steps = 0;
    mutable TransferStats stats;
    // Host copy of the Physics<Real> passed by value to every launch.
    alignas(16) unsigned char physics[sizeof(Physics<double>)] = {};

    cfd_lbm_solver(int lattice, size_t nx_, size_t ny_, size_t nz_, float tau_, float ax_, float ay_, float az_,
                   int storage_) {
        if (lattice != 9 && lattice != 19 && lattice != 27)
            throw std::invalid_argument("lattice must be D2Q9 (9), D3Q19 (19) or D3Q27 (27)");
        if (storage_ < 0 || storage_ > 4)
            throw std::invalid_argument("storage must be FP32 (0), FP16S (1), FP64 (2), FP32S (3) or FP64X (4)");
        if (nx_ < 2 || ny_ < 2) throw std::invalid_argument("LBM grid requires nx, ny >= 2");
        if (lattice == 9) nz_ = 1;
        else if (nz_ < 2) throw std::invalid_argument("3-D LBM grid requires nz >= 2");
        if (!(tau_ > 0.5F)) throw std::invalid_argument("LBM tau must be > 0.5 for positive viscosity");
        q = lattice;
        storage = storage_;
        nx = nx_;
        ny = ny_;
        nz = nz_;
        cells = nx * ny * nz;
        if (cells >= (size_t{1} << 32) || nz > 65535U * 256U)
            throw std::invalid_argument("lattice exceeds 2^32 cells (32-bit cell indices)");
        tau = tau_;
        ax = ax_;
        ay = ay_;
        az = az_;
        elem = storage == 1 ? 2 : (storage == 2 || storage == 4 ? 8 : 4);
        real = storage == 2 || storage == 4 ? 8 : 4;

        // Launch geometry: 256-thread tiles, bx a power of two in [64, 256] covering x, so that a
        // wavefront's 64 lanes lie in one row segment (the bounded step reads the segment's span
        // once per wavefront).
        uint32_t bx = 64;
        while (bx < nx && bx < 256) bx <<= 1;
        if (const char* v = std::getenv("CFD_LBM_BX")) {  // tuning: tile width 64, 128 or 256
            const long w = std::atol(v);
            if (w == 64 || w == 128 || w == 256) bx = static_cast<uint32_t>(w);
        }
        geo.nx = static_cast<uint32_t>(nx);
        geo.ny = static_cast<uint32_t>(ny);
        geo.nz = static_cast<uint32_t>(nz);
        geo.bx = bx;
        geo.by = 256U / bx;
        geo.tiles_x = static_cast<uint32_t>((nx + bx - 1) / bx);
        geo.tiles_y = static_cast<uint32_t>((ny + geo.by - 1) / geo.by);
        const uint64_t tiles = static_cast<uint64_t>(geo.tiles_x) * geo.tiles_y * nz;
        if (tiles >= (uint64_t{1} << 31)) throw std::invalid_argument("lattice too large for one launch");
        geo.tiles = static_cast<uint32_t>(tiles);
        geo.xcd_remap = 0U;
        if (const char* v = std::getenv("CFD_LBM_XCD_REMAP"))
            geo.xcd_remap = (std::atoi(v) != 0 && geo.tiles % 8U == 0U && geo.tiles >= 64U) ? 1U : 0U;
        // Population layout (Geometry in lbm_device.hpp).  Kernels address a population as a plane
        // base plus a 32-bit byte offset, reaching one plane further for bounce-back; the adjoint
        // keeps lambda in the same layout with `real`-sized elements.  Blocks of 2^k x-rows holding
        // up to block_cells cells each (measured +3.5 % on the in-place pattern at 256^3 FP32) when the whole
        // array fits those offsets, plain planes otherwise; CFD_LBM_LAYOUT=planes forces plain
        // planes, CFD_LBM_BLOCK_CELLS sets the block size.
        const size_t widest = std::max(elem, real);
        const uint64_t rows = static_cast<uint64_t>(ny) * nz;
        // Measured at 256^3 (same-box sweeps of 1024-16384): 2048 cells per block is fastest for FP32,
        // FP32S, FP64 and D3Q27 (1.4-2.1 % over 4096), 8192 for FP16S (0.8 %).
        uint64_t block_cells = storage == 1 ? 8192 : 2048;
        // Rows wider than that still get a block each rather than plain planes (4096^2 D2Q9 cavity,
        // FP32: 63.6 % of 8 TB/s in plain planes, 70.2 % in one-row blocks).
        block_cells = std::max<uint64_t>(block_cells, nx);
        if (const char* v = std::getenv("CFD_LBM_BLOCK_CELLS")) block_cells = std::strtoull(v, nullptr, 10);
        bool blocked = block_cells >= nx;
        if (const char* v = std::getenv("CFD_LBM_LAYOUT")) blocked = blocked && std::string(v) != "planes";
        uint32_t rs = 0;
        while ((static_cast<uint64_t>(nx) << (rs + 1)) <= block_cells && (uint64_t{1} << (rs + 1)) <= rows) ++rs;
        const uint64_t block = (uint64_t{1} << rs) * nx;
        const uint64_t blocks = (rows + (uint64_t{1} << rs) - 1) >> rs;
        blocked = blocked && (blocks * q + 1) * block * widest < (uint64_t{1} << 32);
        if (blocked) {
            geo.row_shift = rs;
            geo.stride = block;
            geo.block_elems = static_cast<uint32_t>(q * block);
            pop_elems = blocks * q * block;
        } else {
            // Plain planes start on 256-element boundaries and are staggered by 4352 elements so that
            // they never sit a power of two apart (measured +2 % on 256^3); CFD_LBM_PLANE_PAD overrides.
            uint64_t pad = 4352;
            if (const char* v = std::getenv("CFD_LBM_PLANE_PAD")) pad = std::strtoull(v, nullptr, 10);
            geo.row_shift = 31;
            geo.block_elems = 0;
            geo.stride = (static_cast<uint64_t>(cells) + pad + 255U) / 256U * 256U;
            pop_elems = static_cast<uint64_t>(q) * geo.stride;
            if (2U * geo.stride * elem > (uint64_t{1} << 32))
                throw std::invalid_argument("lattice too large: two population planes must fit in 4 GiB");
        }
        // Resident wavefronts per SIMD of the step kernel, bounded and periodic, capped by reserving
        // LDS per block (160 KiB per CU; a 256-thread block puts one wave on each of the 4 SIMDs).
        // Measured on MI355X:
        //  * row blocks: FP32 D3Q19 5 (256^3: 4 % over 8), FP32 D3Q27 4, FP16 uncapped;
        //  * plain planes, bounded: FP32 3 and FP16 5 (768^3 cavity: 12.4 ms against 14.1 at 4,
        //    FP16 6.66 against 7.24 uncapped; 768^3 sphere cluster: 12.6 against 13.4, FP16 6.78
        //    against 7.24);
        //  * plain planes, periodic: FP32 D3Q19 4 (512^3: 3.85 ms against 4.01 at 3), FP32 D3Q27 3
        //    (5.15 against 5.77 at 4), FP16 5 (2.06 against 2.10 uncapped).
        // CFD_LBM_WAVES (3..8) overrides both.
        const bool fp32_arith = storage == 0 || storage == 3;
        long waves_bounded = 8, waves_periodic = 8;
        if (blocked) {
            if (fp32_arith) waves_bounded = waves_periodic = lattice == 27 ? 4 : 5;
        } else if (fp32_arith) {
            waves_bounded = 3;
            waves_periodic = lattice == 27 ? 3 : 4;
        } else if (storage == 1) {
            waves_bounded = waves_periodic = 5;
        }
        if (const char* v = std::getenv("CFD_LBM_WAVES")) waves_bounded = waves_periodic = std::atol(v);
        auto lds_cap = [](long w) { return w >= 3 && w < 8 ? static_cast<uint32_t>(163840 / w - 1024) : 0U; };
        step_lds = lds_cap(waves_bounded);
        step_lds_periodic = lds_cap(waves_periodic);
        // The transposed step (adjoint) runs 6 (FP16) or 5 (FP32) waves per SIMD by its registers.
        // On HBM-bound lattices fewer measured faster: FP16 at 5 (768^3 sphere cluster: 20.6 ->
        // 19.6 ms; on 128^3, which the caches partly hold, slower: 76 -> 79 us), FP32 in plain planes
        // at 3 (768^3 sphere cluster, reverse phase of 200 steps: 14.00 / 13.25 s against 14.07 /
        // 13.45 s in two runs).  CFD_LBM_ADJOINT_WAVES (3..8) overrides.
        long adjoint_waves = 8;
        if (storage == 1 && cells >= (size_t{1} << 24)) adjoint_waves = 5;
        else if (fp32_arith && !blocked) adjoint_waves = 3;
        if (const char* v = std::getenv("CFD_LBM_ADJOINT_WAVES")) adjoint_waves = std::atol(v);
        adjoint_lds = lds_cap(adjoint_waves);

        ops = cfd_hip::kernel_ops(q, storage);
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (count < 1) throw std::runtime_error("no HIP device available");
        hip_check(hipGetDevice(&device), "hipGetDevice");
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        try {
            f.allocate(pop_elems * elem);
            initialize_uniform(1.0F, 0.0F, 0.0F, 0.0F);
        } catch (...) {
            (void)hipStreamDestroy(stream);
            throw;
        }
    }

    ~cfd_lbm_solver() {
        (void)hipSetDevice(device);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    void use_device() const { hip_check(hipSetDevice(device), "hipSetDevice"); }
    bool odd() const { return (t & 1U) != 0U; }
    // 32-bit byte offsets cover an array of this layout with elements of `bytes` (plus one plane).
    bool offsets_fit(size_t bytes) const {
        const uint64_t span = geo.row_shift == 31 ? 2U * geo.stride : pop_elems + geo.stride;
        return span * bytes <= (uint64_t{1} << 32);
    }
    bool bounded() const { return static_cast<bool>(flags); }
    uint32_t seg_x() const { return static_cast<uint32_t>((nx + 63) / 64); }
    size_t vel_words() const { return static_cast<size_t>((q - 1 + 9) / 10); }

    template <class Real> void build_physics() {
        Physics<Real> p{};
        p.omega = Real(1) / static_cast<Real>(tau);
        p.tau = static_cast<Real>(tau);
        p.les = les ? 1U : 0U;
        p.ax = static_cast<Real>(ax);
        p.ay = static_cast<Real>(ay);
        p.az = static_cast<Real>(az);
        p.u_peak = static_cast<Real>(u_peak);
        p.rho_out = static_cast<Real>(rho_out);
        for (int k = 0; k < 8; ++k)
            for (int c = 0; c < 3; ++c) p.wall_u[k][c] = static_cast<Real>(wall_u[k][c]);
        p.flags = flags.ptr;
        p.gamma = reinterpret_cast<const Real*>(gamma.ptr);
        p.lax = reinterpret_cast<const Real*>(lacc.ptr);
        p.solid_rho = reinterpret_cast<const Real*>(solid_rho.ptr);
        p.links = reinterpret_cast<const uint32_t*>(links.ptr);
        p.link_vel = reinterpret_cast<const uint32_t*>(link_vel.ptr);
        p.segments = reinterpret_cast<const uint32_t*>(segments.ptr);
        p.seg_x = seg_x();
        p.cells = static_cast<uint32_t>(cells);
        std::memcpy(physics, &p, sizeof(p));
    }

    LaunchState state() {
        if (real == 8) build_physics<double>();
        else build_physics<float>();
        return LaunchState{f.ptr, geo, physics, stream, static_cast<uint32_t>(cells), step_lds, step_lds_periodic,
                           adjoint_lds};
    }

    void sync() const {
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        ++stats.syncs;
    }

    void h2d(void* dst, const void* src, size_t bytes) {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
        stats.h2d += bytes;
    }
    void d2h(void* dst, const void* src, size_t bytes) {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
        stats.d2h += bytes;
    }

    // Host float array -> device Real array (converted on the host; float -> double is exact).
    void upload_real(DeviceBuffer<unsigned char>& dst, const float* const* parts, size_t nparts) {
        const size_t n = cells * nparts;
        if (!dst || dst.bytes != n * real) dst.allocate(n * real);
        if (real == 4) {
            for (size_t k = 0; k < nparts; ++k) h2d(dst.ptr + k * cells * 4, parts[k], cells * 4);
            sync();
        } else {
            std::vector<double> tmp(n);
            for (size_t k = 0; k < nparts; ++k)
                for (size_t i = 0; i < cells; ++i) tmp[k * cells + i] = static_cast<double>(parts[k][i]);
            h2d(dst.ptr, tmp.data(), n * 8);
            sync();
        }
    }

    // After the plain layout has been (re)written: freeze wall densities, prepare bounce-back slots.
    void prepare_walls() {
        if (!bounded()) return;
        const LaunchState st = state();
        ops->freeze(st, odd(), flags.ptr, nullptr, solid_rho.ptr);
        ops->bounce_back(st, odd(), true, flags.ptr);
    }

    void run_init(int mode, double rho, double ux, double uy, double uz, double amplitude, double noise,
                  uint64_t seed, const void* tables, const void* fields) {
        t = 0;
        unsigned char args[256];
        if (cfd_hip::init_args_size(storage) > sizeof(args)) throw std::logic_error("init args buffer too small");
        cfd_hip::fill_init_args(storage, args, mode, rho, ux, uy, uz, amplitude, noise, seed, tables, fields);
        const LaunchState st = state();
        ops->initialize(st, args, flags.ptr);
        prepare_walls();
    }

    void initialize_uniform(float rho, float ux, float uy, float uz) {
        if (!(rho > 0.0F) || !std::isfinite(rho) || !std::isfinite(ux) || !std::isfinite(uy) || !std::isfinite(uz))
            throw std::invalid_argument("LBM uniform state must be finite with positive density");
        use_device();
        run_init(0, rho, ux, uy, uz, 0.0, 0.0, 0, nullptr, nullptr);
    }

    template <class Real> std::vector<Real> taylor_green_tables() const {
        // esoteric_pull.hpp:121-137: the trigonometric factors of x, y and z, evaluated on the
        // host with the same libm calls as the reference; the device multiplies them.
        const Real two_pi = Real(2) * static_cast<Real>(3.14159265358979323846);
        std::vector<Real> tab(2 * nx + 2 * ny + nz);
        for (size_t x = 0; x < nx; ++x) {
            const Real xf = (static_cast<Real>(x) + Real(0.5)) / static_cast<Real>(nx);
            tab[x] = std::sin(two_pi * xf);
            tab[nx + x] = std::cos(two_pi * xf);
        }
        for (size_t y = 0; y < ny; ++y) {
            const Real yf = (static_cast<Real>(y) + Real(0.5)) / static_cast<Real>(ny);
            tab[2 * nx + y] = std::sin(two_pi * yf);
            tab[2 * nx + ny + y] = std::cos(two_pi * yf);
        }
        for (size_t z = 0; z < nz; ++z) {
            if (q == 9) {
                tab[2 * nx + 2 * ny + z] = Real(1);
            } else {
                const Real zf = (static_cast<Real>(z) + Real(0.5)) / static_cast<Real>(nz);
                tab[2 * nx + 2 * ny + z] = std::cos(two_pi * zf);
            }
        }
        return tab;
    }

    void initialize_taylor_green(float amplitude, float noise, uint64_t seed) {
        if (!std::isfinite(amplitude) || !std::isfinite(noise))
            throw std::invalid_argument("Taylor-Green amplitude and noise must be finite");
        use_device();
        DeviceBuffer<unsigned char> tables;
        if (real == 4) {
            const auto tab = taylor_green_tables<float>();
            tables.allocate(tab.size() * 4);
            h2d(tables.ptr, tab.data(), tab.size() * 4);
        } else {
            const auto tab = taylor_green_tables<double>();
            tables.allocate(tab.size() * 8);
            h2d(tables.ptr, tab.data(), tab.size() * 8);
        }
        run_init(1, 1.0, 0.0, 0.0, 0.0, amplitude, noise, seed, tables.ptr, nullptr);
        sync();
    }

    void initialize_fields(const float* rho, const float* ux, const float* uy, const float* uz) {
        if (!rho || !ux || !uy || !uz) throw std::invalid_argument("initialize_fields requires rho, ux, uy, uz");
        use_device();
        DeviceBuffer<unsigned char> fields;
        const float* parts[4] = {rho, ux, uy, uz};
        upload_real(fields, parts, 4);
        run_init(2, 0, 0, 0, 0, 0, 0, 0, nullptr, fields.ptr);
        sync();
    }

    void upload_populations(const float* logical) {
        if (!logical) throw std::invalid_argument("upload_populations requires q*cells values");
        use_device();
        DeviceBuffer<unsigned char> buf;
        std::vector<const float*> parts(q);
        for (int d = 0; d < q; ++d) parts[d] = logical + static_cast<size_t>(d) * cells;
        upload_real(buf, parts.data(), q);
        const LaunchState st = state();
        ops->upload(st, odd(), buf.ptr);
        prepare_walls();
        sync();
    }

    void download_populations(float* logical) {
        if (!logical) throw std::invalid_argument("download_populations requires q*cells values");
        use_device();
        DeviceBuffer<unsigned char> buf;
        buf.allocate(static_cast<size_t>(q) * cells * real);
        const LaunchState st = state();
        ops->download(st, odd(), bounded(), buf.ptr);
        const size_t n = static_cast<size_t>(q) * cells;
        if (real == 4) {
            d2h(logical, buf.ptr, n * 4);
            sync();
        } else {
            std::vector<double> tmp(n);
            d2h(tmp.data(), buf.ptr, n * 8);
            sync();
            for (size_t i = 0; i < n; ++i) logical[i] = static_cast<float>(tmp[i]);
        }
    }

    // The in-place slot array as full populations in the order d * cells + n (raw_storage).
    void download_raw_slots(float* out) {
        if (!out) throw std::invalid_argument("download_raw_slots requires q*cells values");
        use_device();
        DeviceBuffer<unsigned char> buf;
        buf.allocate(static_cast<size_t>(q) * cells * real);
        ops->slots(state(), buf.ptr);
        const size_t n = static_cast<size_t>(q) * cells;
        if (real == 4) {
            d2h(out, buf.ptr, n * 4);
            sync();
        } else {
            std::vector<double> tmp(n);
            d2h(tmp.data(), buf.ptr, n * 8);
            sync();
            for (size_t i = 0; i < n; ++i) out[i] = static_cast<float>(tmp[i]);
        }
    }

    void set_flags(const uint8_t* host_flags) {
        use_device();
        DeviceBuffer<unsigned char> fresh;
        if (host_flags) {
            for (size_t i = 0; i < cells; ++i) {
                const uint8_t v = host_flags[i];
                if (v != 0x00 && v != 0x01 && v != 0x02 && (v & 0x80U) == 0U)
                    throw std::invalid_argument("cell flags must be 0 (fluid), 1 (inlet), 2 (outlet) or 0x80|k (wall)");
                if ((v & 0x80U) != 0U && (v & 0x78U) != 0U)
                    throw std::invalid_argument("wall flags use bits 0-2 for the wall-velocity index only");
            }
            fresh.allocate(cells);
            h2d(fresh.ptr, host_flags, cells);
        }
        if (host_flags) host_flags_assign(host_flags);
        else this->host_flags.clear();
        install_flags(fresh);
    }

    // Makes the device flags in `fresh` (empty: periodic, no flags) the solver's flags: walls that
    // stop being walls are thawed, new walls frozen, links and the segment spans rebuilt.  The host
    // copy (host_flags) must already hold the same values.
    void install_flags(DeviceBuffer<unsigned char>& fresh) {
        const bool have = static_cast<bool>(fresh);
        clear_drag_body();  // the body's fluid neighbours depend on the flags
        LaunchState st = state();
        if (bounded()) {
            ops->bounce_back(st, odd(), false, flags.ptr);  // back to the plain layout
            ops->thaw(st, odd(), flags.ptr, fresh.ptr, solid_rho.ptr);
        }
        if (!have) {
            sync();
            flags.reset();
            links.reset();
            link_vel.reset();
            segments.reset();
            solid_rho.reset();
            return;
        }
        if (!segments) segments.allocate(static_cast<size_t>(seg_x()) * ny * nz * sizeof(uint32_t));
        ops->segments(st, fresh.ptr, seg_x(), reinterpret_cast<uint32_t*>(segments.ptr));
        if (gamma) ops->porous_segments(st, gamma.ptr, seg_x(), reinterpret_cast<uint32_t*>(segments.ptr));
        if (!links) links.allocate(cells * sizeof(uint32_t));
        if (!link_vel) link_vel.allocate(cells * vel_words() * sizeof(uint32_t));
        ops->links(st, fresh.ptr, reinterpret_cast<uint32_t*>(links.ptr), reinterpret_cast<uint32_t*>(link_vel.ptr));
        if (!solid_rho) solid_rho.allocate(cells * real);
        ops->freeze(st, odd(), fresh.ptr, flags.ptr, solid_rho.ptr);
        ops->bounce_back(st, odd(), true, fresh.ptr);
        sync();
        std::swap(flags.ptr, fresh.ptr);
        std::swap(flags.bytes, fresh.bytes);
    }

    void host_flags_assign(const uint8_t* v) { this->host_flags.assign(v, v + cells); }

    // ------------------------------------------------------------------ checkpoint / restart
    // File: the CFDCHK3 layout of DistributedPullBlock::save_checkpoint
    // (include/cfd/solvers/lbm/distributed_pull.hpp:176-203) for one brick covering the lattice,
    // then a trailer the reference's loader does not read:
    //   "CFDESO1\0", u32 parity (time_step & 1), u32 body convention (0: the two-grid block's
    //   outgoing populations, 1: this solver's logical incoming populations), u32 storage,
    //   u32 raw (1: the device state follows), and with raw: u32 row_shift, u32 block_elems,
    //   u64 stride, u64 population bytes, the population array, u64 wall-density bytes, those bytes.
    // Periodic lattices write the outgoing convention, f_out(d, x) = f_in(d, x + c_d), so that
    // the reference's block restores them; lattices with flags write the logical populations.
    template <class V> static void put(std::ofstream& o, const V& v) { o.write(reinterpret_cast<const char*>(&v), sizeof(V)); }
    template <class V> static V get(std::ifstream& in) {
        V v{};
        in.read(reinterpret_cast<char*>(&v), sizeof(V));
        return v;
    }

    // f_out(d, x) = f_in(d, x + c_d) (dir = +1) or its inverse (dir = -1), periodic.
    template <int Q> void shift_streams(const float* in, float* out, int dir) const {
        using L = cfd_hip::Lattice<Q>;
        for (int d = 0; d < Q; ++d)
            for (size_t z = 0; z < nz; ++z)
                for (size_t y = 0; y < ny; ++y)
                    for (size_t x = 0; x < nx; ++x) {
                        const auto wrap = [](size_t v, int delta, size_t n) {
                            const long long s = static_cast<long long>(v) + delta + static_cast<long long>(n);
                            return static_cast<size_t>(s % static_cast<long long>(n));
                        };
                        const size_t xs = wrap(x, dir * L::cx[d], nx), ys = wrap(y, dir * L::cy[d], ny),
                                     zs = wrap(z, dir * L::cz[d], nz);
                        out[d * cells + (z * ny + y) * nx + x] = in[d * cells + (zs * ny + ys) * nx + xs];
                    }
    }
    void shift_streams(const float* in, float* out, int dir) const {
        if (q == 9) shift_streams<9>(in, out, dir);
        else if (q == 19) shift_streams<19>(in, out, dir);
        else shift_streams<27>(in, out, dir);
    }

    void save_checkpoint(const char* path, bool raw) {
        if (!path) throw std::invalid_argument("save_checkpoint needs a path");
        use_device();
        std::vector<float> logical(static_cast<size_t>(q) * cells);
        download_populations(logical.data());
        const bool periodic = host_flags.empty();
        std::vector<float> body;
        if (periodic) {
            body.resize(logical.size());
            shift_streams(logical.data(), body.data(), +1);
        }
        std::ofstream o(path, std::ios::binary | std::ios::trunc);
        if (!o) throw std::runtime_error(std::string("cannot open checkpoint for writing: ") + path);
        o.write("CFDCHK3\0", 8);
        put(o, uint32_t{1});
        put(o, uint32_t{0x01020304U});
        put(o, static_cast<uint32_t>(q));
        const uint64_t ext[3] = {nx, ny, nz}, zero[3] = {0, 0, 0};
        for (uint64_t v : ext) put(o, v);  // global
        for (uint64_t v : zero) put(o, v); // begin
        for (uint64_t v : ext) put(o, v);  // extent
        put(o, static_cast<uint64_t>(t));
        put(o, tau);
        put(o, ax);
        put(o, ay);
        put(o, az);
        const std::vector<float>& out = periodic ? body : logical;
        o.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size() * sizeof(float)));
        o.write("CFDESO1\0", 8);
        put(o, static_cast<uint32_t>(t & 1U));
        put(o, static_cast<uint32_t>(periodic ? 0U : 1U));
        put(o, static_cast<uint32_t>(storage));
        put(o, static_cast<uint32_t>(raw ? 1U : 0U));
        if (raw) {
            put(o, geo.row_shift);
            put(o, geo.block_elems);
            put(o, geo.stride);
            put(o, static_cast<uint64_t>(f.bytes));
            std::vector<unsigned char> host(f.bytes);
            d2h(host.data(), f.ptr, f.bytes);
            std::vector<unsigned char> walls(solid_rho ? solid_rho.bytes : 0);
            if (solid_rho) d2h(walls.data(), solid_rho.ptr, solid_rho.bytes);
            sync();
            o.write(reinterpret_cast<const char*>(host.data()), static_cast<std::streamsize>(host.size()));
            put(o, static_cast<uint64_t>(walls.size()));
            o.write(reinterpret_cast<const char*>(walls.data()), static_cast<std::streamsize>(walls.size()));
        }
        if (!o) throw std::runtime_error(std::string("checkpoint write failed: ") + path);
    }

    // load_checkpoint (distributed_pull.hpp:205-245) for this solver: the same checks of magic,
    // format, lattice, extents and configuration; then the raw device state when the trailer has one
    // for this solver's storage and layout (exact for every storage, walls included), otherwise
    // the float populations (a two-grid file restores periodic lattices only).
    void load_checkpoint(const char* path) {
        if (!path) throw std::invalid_argument("load_checkpoint needs a path");
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error(std::string("cannot open checkpoint for reading: ") + path);
        char magic[8] = {};
        in.read(magic, 8);
        if (!in || std::memcmp(magic, "CFDCHK3\0", 8) != 0) throw std::runtime_error("invalid checkpoint magic");
        const auto version = get<uint32_t>(in), endian = get<uint32_t>(in), stored_q = get<uint32_t>(in);
        if (version != 1 || endian != 0x01020304U || stored_q != static_cast<uint32_t>(q))
            throw std::runtime_error("unsupported checkpoint format/descriptor");
        uint64_t global[3], begin[3], extent[3];
        for (auto& v : global) v = get<uint64_t>(in);
        for (auto& v : begin) v = get<uint64_t>(in);
        for (auto& v : extent) v = get<uint64_t>(in);
        const uint64_t mine[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a)
            if (global[a] != mine[a] || extent[a] != mine[a] || begin[a] != 0)
                throw std::runtime_error("checkpoint decomposition does not match block");
        const auto step = get<uint64_t>(in);
        const float ctau = get<float>(in), cax = get<float>(in), cay = get<float>(in), caz = get<float>(in);
        if (ctau != tau || cax != ax || cay != ay || caz != az)
            throw std::runtime_error("checkpoint LBM configuration mismatch");
        std::vector<float> body(static_cast<size_t>(q) * cells);
        in.read(reinterpret_cast<char*>(body.data()), static_cast<std::streamsize>(body.size() * sizeof(float)));
        if (!in) throw std::runtime_error("checkpoint read failed");
        uint32_t convention = 0, stored_storage = 0, raw = 0;
        in.read(magic, 8);
        const bool trailer = in && std::memcmp(magic, "CFDESO1\0", 8) == 0;
        if (trailer) {
            (void)get<uint32_t>(in);  // parity, implied by the time step
            convention = get<uint32_t>(in);
            stored_storage = get<uint32_t>(in);
            raw = get<uint32_t>(in);
            if (!in) throw std::runtime_error("checkpoint read failed");
        }
        use_device();
        if (raw) {
            const auto rs = get<uint32_t>(in), be = get<uint32_t>(in);
            const auto st = get<uint64_t>(in), fb = get<uint64_t>(in);
            if (stored_storage == static_cast<uint32_t>(storage) && rs == geo.row_shift && be == geo.block_elems &&
                st == geo.stride && fb == f.bytes) {
                std::vector<unsigned char> host(fb);
                in.read(reinterpret_cast<char*>(host.data()), static_cast<std::streamsize>(fb));
                const auto wb = get<uint64_t>(in);
                if (!in) throw std::runtime_error("checkpoint read failed");
                // Checked before the buffer is sized from the file.
                if (wb != (solid_rho ? solid_rho.bytes : 0))
                    throw std::runtime_error("checkpoint wall state does not match the solver's flags");
                std::vector<unsigned char> walls(wb);
                in.read(reinterpret_cast<char*>(walls.data()), static_cast<std::streamsize>(wb));
                if (!in) throw std::runtime_error("checkpoint read failed");
                t = step;
                h2d(f.ptr, host.data(), fb);
                if (wb) h2d(solid_rho.ptr, walls.data(), wb);
                sync();
                return;
            }
        }
        if (convention == 0 && !host_flags.empty())
            throw std::runtime_error("a two-grid checkpoint restores a periodic lattice only");
        std::vector<float> logical;
        if (convention == 0) {
            logical.resize(body.size());
            shift_streams(body.data(), logical.data(), -1);
        }
        t = step;
        upload_populations(convention == 0 ? logical.data() : body.data());
    }

    // set_solid_mask(voxelize_surface(surface, grid)) (examples/real_lbm3d_model.cpp:146-151) on the
    // device: the mask is voxelised straight into the new flag array, solid cells become walls at
    // rest, the others keep their flags (keep) or become fluid.
    void set_solid_surface(const double* triangles, size_t count, const double* minimum, const double* maximum,
                           bool keep) {
        if (!minimum || !maximum) throw std::invalid_argument("voxelization needs the grid's minimum and maximum");
        use_device();
        cfd_hip::VoxelGridArgs g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        for (int a = 0; a < 3; ++a) g.minimum[a] = minimum[a], g.maximum[a] = maximum[a];
        DeviceBuffer<unsigned char> fresh;
        fresh.allocate(cells);
        cfd_hip::voxelize_surface(triangles, count, g, fresh.ptr, stream);
        cfd_hip::solid_to_flags(fresh.ptr, keep && flags ? flags.ptr : nullptr, cells, stream);
        std::vector<uint8_t> mirror(cells);
        d2h(mirror.data(), fresh.ptr, cells);
        sync();
        host_flags.swap(mirror);
        install_flags(fresh);
    }

    // Pressure drop between fluid cells of planes x_in and x_out, as a mean over `window` states.
    cfd_hip::AdjointArgs pressure_args(int x_in, int x_out, uint64_t window) const {
        if (x_in < 0 || x_out < 0 || static_cast<size_t>(x_in) >= nx || static_cast<size_t>(x_out) >= nx || x_in == x_out)
            throw std::invalid_argument("objective planes must be two different x indices inside the lattice");
        size_t a = 0, b = 0;
        for (size_t z = 0; z < nz; ++z)
            for (size_t y = 0; y < ny; ++y) {
                const size_t row = (z * ny + y) * nx;
                const bool fa = host_flags.empty() || host_flags[row + static_cast<size_t>(x_in)] == 0;
                const bool fb = host_flags.empty() || host_flags[row + static_cast<size_t>(x_out)] == 0;
                a += fa;
                b += fb;
            }
        if (a == 0 || b == 0) throw std::invalid_argument("objective planes contain no fluid cells");
        cfd_hip::AdjointArgs args{};
        args.kind = cfd_hip::kPressureDrop;
        args.x_in = x_in;
        args.x_out = x_out;
        args.c_in = 1.0 / (3.0 * static_cast<double>(a) * static_cast<double>(window));
        args.c_out = -1.0 / (3.0 * static_cast<double>(b) * static_cast<double>(window));
        return args;
    }

    // Drag objective: the fluid cells next to the registered body and, per cell, the populations
    // that arrive from a body cell (bit d-1 for population d).
    DeviceBuffer<unsigned char> drag_cells, drag_links, objective_scratch;
    uint32_t drag_count = 0;
    int drag_axis = -1;  // -1: no body registered

    void clear_drag_body() {
        drag_cells.reset();
        drag_links.reset();
        drag_count = 0;
        drag_axis = -1;
    }

    template <int Q> void collect_drag_links(const uint8_t* body, std::vector<uint64_t>& pairs) const {
        using L = cfd_hip::Lattice<Q>;
        const auto wrap = [](size_t v, int c, size_t n) { return (v + n + static_cast<size_t>(c + 0)) % n; };
        for (size_t cell = 0; cell < cells; ++cell) {
            if (!body[cell]) continue;
            if ((host_flags[cell] & 0x80U) == 0U) throw std::invalid_argument("drag body cells must be wall cells");
            const size_t x = cell % nx, y = (cell / nx) % ny, z = cell / (nx * ny);
            for (int d = 1; d < Q; ++d) {
                // Population d arrives at n = body + c_d from the body cell.
                const size_t xn = wrap(x, L::cx[d], nx), yn = wrap(y, L::cy[d], ny), zn = wrap(z, L::cz[d], nz);
                const size_t n = (zn * ny + yn) * nx + xn;
                if (host_flags[n] != 0) continue;  // fluid neighbours only
                pairs.push_back((static_cast<uint64_t>(n) << 8) | static_cast<uint64_t>(d));
            }
        }
    }

    void set_drag_body(const uint8_t* body, int axis) {
        use_device();
        if (!body) {
            sync();
            clear_drag_body();
            return;
        }
        if (axis < 0 || axis > 2 || (q == 9 && axis == 2)) throw std::invalid_argument("drag axis must be 0 (x), 1 (y) or 2 (z, 3-D)");
        if (host_flags.empty()) throw std::invalid_argument("the drag objective needs wall flags (set_flags first)");
        std::vector<uint64_t> pairs;
        if (q == 9) collect_drag_links<9>(body, pairs);
        else if (q == 19) collect_drag_links<19>(body, pairs);
        else collect_drag_links<27>(body, pairs);
        std::sort(pairs.begin(), pairs.end());
        std::vector<uint32_t> list, bits;
        for (uint64_t pr : pairs) {
            const uint32_t n = static_cast<uint32_t>(pr >> 8);
            const int d = static_cast<int>(pr & 0xFFU);
            if (list.empty() || list.back() != n) {
                list.push_back(n);
                bits.push_back(0U);
            }
            bits.back() |= 1U << (d - 1);
        }
        if (list.empty()) throw std::invalid_argument("the drag body has no fluid neighbours");
        sync();
        clear_drag_body();
        drag_cells.allocate(list.size() * 4);
        drag_links.allocate(bits.size() * 4);
        h2d(drag_cells.ptr, list.data(), list.size() * 4);
        h2d(drag_links.ptr, bits.data(), bits.size() * 4);
        sync();
        drag_count = static_cast<uint32_t>(list.size());
        drag_axis = axis;
    }

    cfd_hip::AdjointArgs drag_args(uint64_t window) const {
        if (drag_axis < 0) throw std::invalid_argument("no drag body registered (set_drag_body first)");
        cfd_hip::AdjointArgs args{};
        args.kind = cfd_hip::kDrag;
        args.drag_cells = reinterpret_cast<const uint32_t*>(drag_cells.ptr);
        args.drag_links = reinterpret_cast<const uint32_t*>(drag_links.ptr);
        args.drag_count = drag_count;
        args.drag_axis = drag_axis;
        args.c_drag = 1.0 / static_cast<double>(window);
        return args;
    }

    cfd_hip::AdjointArgs objective_args(int kind, int x_in, int x_out, uint64_t window) {
        cfd_hip::AdjointArgs args{};
        if (kind == cfd_hip::kPressureDrop) args = pressure_args(x_in, x_out, window);
        else if (kind == cfd_hip::kDrag) args = drag_args(window);
        else throw std::invalid_argument("objective must be 0 (pressure drop) or 1 (drag)");
        const size_t need = cfd_hip::objective_scratch_doubles(geo, drag_count) * sizeof(double);
        if (objective_scratch.bytes < need) objective_scratch.allocate(need);
        args.scratch = reinterpret_cast<double*>(objective_scratch.ptr);
        return args;
    }

    // j of the current state (FP64).
    double objective(int kind, int x_in, int x_out) {
        use_device();
        const cfd_hip::AdjointOps* adj = cfd_hip::adjoint_ops(q, storage);
        const cfd_hip::AdjointArgs args = objective_args(kind, x_in, x_out, 1);
        DeviceBuffer<unsigned char> acc;
        acc.allocate(sizeof(double));
        const LaunchState st = state();
        adj->zero(st, reinterpret_cast<double*>(acc.ptr), 1);
        adj->objective(st, odd(), bounded(), f.ptr, args, 1.0, reinterpret_cast<double*>(acc.ptr));
        double v = 0.0;
        d2h(&v, acc.ptr, sizeof(double));
        sync();
        return v;
    }

    // J of adjoint_run without the adjoint: `steps` steps in place from the current state, with j
    // accumulated in FP64 over the last `window` states by the adjoint run's own objective kernel in
    // its order, so that J equals adjoint_run's bit for bit.  The state is left at time + steps.
    double objective_run(int kind, uint64_t steps, uint64_t window, int x_in, int x_out) {
        use_device();
        if (steps < 1 || window < 1 || window > steps) throw std::invalid_argument("need 1 <= window <= steps");
        const cfd_hip::AdjointOps* adj = cfd_hip::adjoint_ops(q, storage);
        const cfd_hip::AdjointArgs args = objective_args(kind, x_in, x_out, window);
        DeviceBuffer<unsigned char> acc;
        acc.allocate(sizeof(double));
        double* sum = reinterpret_cast<double*>(acc.ptr);
        const LaunchState st = state();
        adj->zero(st, sum, 1);
        const bool b = bounded();
        const bool forced = ax != 0.0F || ay != 0.0F || az != 0.0F || static_cast<bool>(lacc) || les;
        for (uint64_t k = 0; k < steps; ++k) {
            ops->step(st, odd(), b, forced, f.ptr, f.ptr);
            ++t;
            if (k + 1 >= steps - window + 1) adj->objective(st, odd(), b, f.ptr, args, 1.0, sum);
        }
        double J = 0.0;
        d2h(&J, acc.ptr, sizeof(double));
        sync();
        return J;
    }

    void adjoint_checks(uint64_t steps, uint64_t window) const {
        if (steps < 1 || window < 1 || window > steps) throw std::invalid_argument("need 1 <= window <= steps");
        if (ax != 0.0F || ay != 0.0F || az != 0.0F || lacc) throw std::invalid_argument("the adjoint supports unforced lattices only");
        if (les) throw std::invalid_argument("the adjoint does not differentiate the Smagorinsky closure");
        for (uint8_t v : host_flags)
            if ((v & 0x80U) && (v & 7U) && (wall_u[v & 7U][0] != 0.0F || wall_u[v & 7U][1] != 0.0F || wall_u[v & 7U][2] != 0.0F))
                throw std::invalid_argument("the adjoint supports walls at rest only");
    }

    // Forward `steps` steps from the current state (J = mean of j over the last `window` states),
    // then the discrete adjoint back to the current state.  Checkpointing:
    //  * every > 0: two-level -- a snapshot every `every` steps on the way forward, and per segment
    //    an out-of-place recomputation of its states into replay buffers before the transposed
    //    steps run through them (each step recomputed once; memory steps/every + every - 1 states);
    //  * every == 0: binomial -- `snapshots` state buffers (0: as many as fit, keeping 1/16 of the
    //    free memory in reserve) plus the start state, optimal recomputation (checkpoint_schedule.hpp).
    // grad_on_device: grad_gamma_out is device memory (cells doubles) and receives the gradient
    // directly, with no host copy.
    double adjoint_run(int kind, uint64_t steps, uint64_t window, uint64_t every, uint64_t snapshots, int x_in,
                       int x_out, double* grad_gamma_out, double* grad_peak_out, uint64_t* forward_steps_out,
                       bool grad_on_device = false) {
        use_device();
        if (grad_on_device && (!grad_gamma_out || reinterpret_cast<uintptr_t>(grad_gamma_out) % 8 != 0))
            throw std::invalid_argument("the device gradient needs an 8-byte aligned device pointer");
        adjoint_checks(steps, window);
        const cfd_hip::AdjointOps* adj = cfd_hip::adjoint_ops(q, storage);
        cfd_hip::AdjointArgs args = objective_args(kind, x_in, x_out, window);
        if (!offsets_fit(real))
            throw std::invalid_argument("lattice too large for the adjoint: two FP32 lambda planes must fit in 4 GiB");
        const size_t state_bytes = pop_elems * elem;
        // Checkpoint buffers are carved from one allocation at a 4 KiB pitch, so every buffer
        // starts as aligned as the allocation (hipMalloc: at least 256 B) whatever nx is; the
        // unpadded state size is odd for odd nx.
        const size_t pitch = (state_bytes + 4095) & ~static_cast<size_t>(4095);
        const size_t fixed = pop_elems * real + 2 * cells * 8 + 64;
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        size_t buffers = 0;  // state buffers besides f
        if (every > 0) {
            buffers = static_cast<size_t>((steps + every - 1) / every + std::min<uint64_t>(every, steps) - 1);
        } else {
            if (snapshots == 0) {
                const size_t usable = free_b - free_b / 16;
                snapshots = usable > fixed + 2 * pitch ? (usable - fixed) / pitch - 1 : 0;
                if (snapshots == 0) throw std::invalid_argument("no room for a checkpoint buffer in device memory");
            }
            snapshots = std::min<uint64_t>(snapshots, std::max<uint64_t>(steps - 1, 1));
            buffers = static_cast<size_t>(snapshots) + 1;
        }
        const size_t need = buffers * pitch + fixed;
        if (need > free_b)
            throw std::invalid_argument("checkpoints do not fit in device memory (" + std::to_string(need >> 20) +
                                        " MiB needed, " + std::to_string(free_b >> 20) +
                                        " MiB free): raise the checkpoint interval or lower the snapshot count");
        // One allocation for all checkpoint states, one buffer every `pitch` bytes.
        DeviceBuffer<unsigned char> pool_mem;
        pool_mem.allocate(buffers * pitch);
        std::vector<void*> pool(buffers);
        for (size_t k = 0; k < buffers; ++k) {
            pool[k] = pool_mem.ptr + k * pitch;
            if (reinterpret_cast<uintptr_t>(pool[k]) % 256 != 0)
                throw std::runtime_error("checkpoint buffer is not 256-byte aligned");
        }
        DeviceBuffer<unsigned char> lam, gg, gp, acc;
        lam.allocate(pop_elems * real);
        if (!grad_on_device) gg.allocate(cells * sizeof(double));
        gp.allocate(cells * sizeof(double));
        acc.allocate(sizeof(double));
        args.grad_gamma = grad_on_device ? grad_gamma_out : reinterpret_cast<double*>(gg.ptr);
        args.grad_peak = reinterpret_cast<double*>(gp.ptr);
        const LaunchState st = state();
        adj->zero(st, args.grad_gamma, cells);
        adj->zero(st, args.grad_peak, cells);
        adj->zero(st, reinterpret_cast<double*>(acc.ptr), 1);
        const bool b = bounded();
        const uint64_t t0 = t;
        const auto par = [t0](uint64_t k) { return ((t0 + k) & 1U) != 0U; };
        uint64_t forward_steps = 0;
        // Progress lines on stderr every CFD_LBM_PROGRESS_STEPS forward or transposed steps (0 or
        // unset: none), each after the stream has finished those steps, for runs of many minutes.
        uint64_t progress = 0, transposed_done = 0;
        if (const char* v = std::getenv("CFD_LBM_PROGRESS_STEPS")) progress = std::strtoull(v, nullptr, 10);
        const auto report = [&](const char* phase, uint64_t done, uint64_t total) {
            sync();
            std::fprintf(stderr, "[adjoint] %s: %llu of %llu steps, %llu forward steps so far\n", phase,
                         static_cast<unsigned long long>(done), static_cast<unsigned long long>(total),
                         static_cast<unsigned long long>(forward_steps));
            std::fflush(stderr);
        };
        // Snapshots taken on the way forward: (step index, buffer).
        std::vector<std::pair<uint64_t, void*>> shots;
        std::vector<void*> free_list;
        if (every > 0) {
            for (uint64_t c = 0; c * every < steps; ++c) shots.emplace_back(c * every, pool[c]);
        } else {
            shots.emplace_back(0, pool[0]);
            for (size_t k = 1; k < buffers; ++k) free_list.push_back(pool[k]);
            // The first descent of the binomial plan: its advances are the forward sweep itself.
            std::vector<void*> fl = free_list;
            for (uint64_t a0 = 0; steps - a0 > 1;) {
                const uint64_t m = a0 + cfd_hip::BinomialSchedule::split(steps - a0, fl.size());
                shots.emplace_back(m, fl.back());
                fl.pop_back();
                a0 = m;
            }
        }
        TimingEvent ev0, ev1, ev2;
        hipEvent_t e0 = ev0.e, e1 = ev1.e, e2 = ev2.e;
        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        // Forward sweep.  A step whose result is a checkpoint writes it out of place into the
        // checkpoint's buffer, and the step after it reads from there: storing costs no extra
        // traffic.  Only the start state is copied.  The sweep ends in f (no checkpoint is at T).
        hip_check(hipMemcpyAsync(shots[0].second, f.ptr, state_bytes, hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync start state");
        const void* cur = f.ptr;
        size_t next = 1;
        for (uint64_t k = 0; k < steps; ++k) {
            void* dst = f.ptr;
            if (next < shots.size() && shots[next].first == k + 1) dst = shots[next++].second;
            ops->step(st, par(k), b, false, cur, dst);
            cur = dst;
            ++forward_steps;
            if (k + 1 >= steps - window + 1) adj->objective(st, par(k + 1), b, cur, args, 1.0, reinterpret_cast<double*>(acc.ptr));
            if (progress > 0 && (k + 1) % progress == 0) report("forward sweep", k + 1, steps);
        }
        t = t0 + steps;
        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        const auto transposed = [&](uint64_t tt, const void* state) {
            args.add_objective = tt + window >= steps + 1 ? 1 : 0;
            adj->step(st, par(tt), b, state, lam.ptr, args);
            ++transposed_done;
            if (progress > 0 && transposed_done % progress == 0) report("reverse", transposed_done, steps);
        };
        // States a+1 .. m of the chain into dst, starting from state a in src (out of place first).
        const auto advance = [&](const void* src, void* dst, uint64_t a0, uint64_t m) {
            ops->step(st, par(a0), b, false, src, dst);
            for (uint64_t k = a0 + 1; k < m; ++k) ops->step(st, par(k), b, false, dst, dst);
            forward_steps += m - a0;
        };
        args.add_objective = 0;
        adj->seed(st, par(steps), b, lam.ptr, args);
        if (every > 0) {
            const uint64_t segments = shots.size();
            for (uint64_t c = segments; c-- > 0;) {
                const uint64_t s0 = c * every, e = std::min(steps, s0 + every);
                std::vector<const void*> states(e - s0);
                states[0] = shots[c].second;
                for (uint64_t k = 1; k < e - s0; ++k) {
                    void* dst = pool[segments + k - 1];
                    advance(states[k - 1], dst, s0 + k - 1, s0 + k);
                    states[k] = dst;
                }
                for (uint64_t k = e - s0; k-- > 0;) transposed(s0 + k, states[k]);
            }
        } else {
            // Reverse [a0, b0) with state a0 in src: advance to the split into a free buffer, reverse
            // the right part with one buffer less, then the left part.  On the first descent the
            // split states are already in place from the forward sweep.
            std::function<void(uint64_t, uint64_t, const void*, bool)> reverse = [&](uint64_t a0, uint64_t b0,
                                                                                   const void* src, bool first) {
                while (b0 - a0 > 1) {
                    const uint64_t m = a0 + cfd_hip::BinomialSchedule::split(b0 - a0, free_list.size());
                    void* dst = free_list.back();
                    free_list.pop_back();
                    if (!first) advance(src, dst, a0, m);
                    reverse(m, b0, dst, first);
                    free_list.push_back(dst);
                    b0 = m;
                    first = false;
                }
                transposed(a0, src);
            };
            reverse(0, steps, shots[0].second, true);
        }
        hip_check(hipEventRecord(e2, stream), "hipEventRecord");
        // dJ/du_peak: the inlet cells' contributions summed on the device (fixed order).
        DeviceBuffer<unsigned char> peak_sums;
        peak_sums.allocate((cfd_hip::kSumBlocks + 1) * sizeof(double));
        double* peak_out = reinterpret_cast<double*>(peak_sums.ptr) + cfd_hip::kSumBlocks;
        adj->sum(st, args.grad_peak, cells, reinterpret_cast<double*>(peak_sums.ptr), peak_out);
        double J = 0.0, du = 0.0;
        d2h(&J, acc.ptr, sizeof(double));
        d2h(&du, peak_out, sizeof(double));
        if (grad_gamma_out && !grad_on_device) d2h(grad_gamma_out, gg.ptr, cells * sizeof(double));
        sync();
        (void)hipEventElapsedTime(&last_forward_ms, e0, e1);
        (void)hipEventElapsedTime(&last_reverse_ms, e1, e2);
        last_forward_steps = forward_steps;
        if (grad_peak_out) *grad_peak_out = du;
        if (forward_steps_out) *forward_steps_out = forward_steps;
        return J;
    }

    void set_smagorinsky(bool enabled) {
        if (enabled && (storage == 1 || storage == 3))
            throw std::invalid_argument("Smagorinsky LES runs on FP32 or FP64 storage (the reference's arithmetic)");
        les = enabled;
    }

    // The segment words of a bounded lattice: spans from the flags, then the porous marks of gamma.
    void mark_segments() {
        if (!segments) return;
        const LaunchState st = state();
        ops->segments(st, flags.ptr, seg_x(), reinterpret_cast<uint32_t*>(segments.ptr));
        if (gamma) ops->porous_segments(st, gamma.ptr, seg_x(), reinterpret_cast<uint32_t*>(segments.ptr));
    }

    void set_porosity(const float* g) {
        use_device();
        if (!g) {
            sync();
            gamma.reset();
            mark_segments();
            return;
        }
        for (size_t i = 0; i < cells; ++i)
            if (!(g[i] >= 0.0F && g[i] <= 1.0F)) throw std::invalid_argument("porosity gamma must lie in [0, 1]");
        const float* parts[1] = {g};
        upload_real(gamma, parts, 1);
        mark_segments();
    }

    // set_porosity from device memory (cells floats on this solver's device): validated on the
    // device; the field in use changes only when every value lies in [0, 1].
    void set_porosity_device(const float* g) {
        use_device();
        if (!g) throw std::invalid_argument("set_porosity_device needs a device pointer (set_porosity(NULL) clears)");
        const uint64_t blocks = (cells + 255) / 256;
        DeviceBuffer<unsigned char> fresh, bad;
        fresh.allocate(cells * real);
        bad.allocate(sizeof(uint32_t));
        zero_words(bad.ptr, 1);
        porosity_from_device<<<static_cast<uint32_t>(blocks), 256, 0, stream>>>(
            g, fresh.ptr, static_cast<uint32_t>(real), cells, reinterpret_cast<uint32_t*>(bad.ptr));
        hip_check(hipGetLastError(), "porosity_from_device");
        uint32_t n_bad = 0;
        d2h(&n_bad, bad.ptr, sizeof(n_bad));
        sync();
        if (n_bad != 0) throw std::invalid_argument("porosity gamma must lie in [0, 1]");
        std::swap(gamma.ptr, fresh.ptr);
        std::swap(gamma.bytes, fresh.bytes);
        mark_segments();
    }

    // z-slab halo messages (the slab is this lattice, periodic in z; its neighbours are the slabs
    // below and above).  Elements of one message, in population storage units.
    size_t halo_message_bytes() {
        if (q == 9) throw std::invalid_argument("halo messages cut a 3-D lattice in z (D3Q19 or D3Q27)");
        return ops->halo_elems(state()) * elem;
    }
    void halo_checks() {
        (void)halo_message_bytes();
        if (bounded()) throw std::invalid_argument("halo messages need a lattice without flags (periodic slab)");
        if (t == 0) throw std::invalid_argument("halo messages follow a step: no step has run");
    }
    // After a step: the up and down messages into device buffers of halo_message_bytes() each.
    void halo_pack(void* up, void* down) {
        use_device();
        halo_checks();
        if (!up || !down) throw std::invalid_argument("halo_pack needs two device buffers");
        const LaunchState st = state();
        const bool last_odd = ((t - 1) & 1U) != 0U;
        ops->halo(st, last_odd, false, up, down);
    }
    // The up message of the slab below and the down message of the slab above, into this slab.
    void halo_unpack(const void* from_below, const void* from_above) {
        use_device();
        halo_checks();
        if (!from_below || !from_above) throw std::invalid_argument("halo_unpack needs two device buffers");
        const LaunchState st = state();
        const bool last_odd = ((t - 1) & 1U) != 0U;
        ops->halo(st, last_odd, true, const_cast<void*>(from_below), const_cast<void*>(from_above));
    }

    void set_local_acceleration(const float* lx, const float* ly, const float* lz) {
        use_device();
        if (!lx && !ly && !lz) {
            sync();
            lacc.reset();
            return;
        }
        if (!lx || !ly || !lz) throw std::invalid_argument("local LBM acceleration field size mismatch");
        const float* parts[3] = {lx, ly, lz};
        upload_real(lacc, parts, 3);
    }

    // device_local_acceleration (esoteric_pull_sycl.hpp:349-358): the per-cell acceleration planes
    // (ax, ay, az; cells values each) on the device, allocated zeroed on first use and zeroed again
    // on request.  From then on every step uses the forced kernel.
    unsigned char* device_local_acceleration(bool clear) {
        use_device();
        const bool fresh = !lacc;
        if (fresh) lacc.allocate(3 * cells * real);
        if (fresh || clear) zero_words(lacc.ptr, 3 * cells * real / 4);
        return lacc.ptr;
    }

    void zero_words(unsigned char* p, uint64_t words) {
        const uint64_t blocks = (words + 255) / 256;
        if (blocks > 0x7FFFFFFFull) throw std::invalid_argument("buffer too large to clear");
        if (blocks) zero_words_kernel<<<static_cast<uint32_t>(blocks), 256, 0, stream>>>(reinterpret_cast<uint32_t*>(p), words);
        hip_check(hipGetLastError(), "zero_words");
    }

    void step(uint64_t count) {
        use_device();
        const bool b = bounded();
        const bool forced = ax != 0.0F || ay != 0.0F || az != 0.0F || static_cast<bool>(lacc) || les;
        LaunchState st = state();
        for (uint64_t k = 0; k < count; ++k) {
            ops->step(st, odd(), b, forced, f.ptr, f.ptr);
            ++t;
        }
    }

    float step_timed(uint64_t count) {
        use_device();
        TimingEvent e0, e1;
        float ms = 0.0F;
        hip_check(hipEventRecord(e0.e, stream), "hipEventRecord");
        step(count);
        hip_check(hipEventRecord(e1.e, stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1.e), "hipEventSynchronize");
        ++stats.syncs;
        hip_check(hipEventElapsedTime(&ms, e0.e, e1.e), "hipEventElapsedTime");
        return ms;
    }

    void refresh_macroscopic() {
        use_device();
        if (!macro) macro.allocate(4 * cells * real);
        const LaunchState st = state();
        ops->macroscopic(st, odd(), bounded(), macro.ptr);
    }

    template <class Out> void download_macroscopic(Out* rho, Out* ux, Out* uy, Out* uz) {
        if (!rho || !ux || !uy || !uz) throw std::invalid_argument("download_macroscopic requires four arrays");
        refresh_macroscopic();
        Out* outs[4] = {rho, ux, uy, uz};
        if (real == sizeof(Out)) {
            for (int k = 0; k < 4; ++k) d2h(outs[k], macro.ptr + k * cells * real, cells * real);
            sync();
            return;
        }
        if (real == 4) {
            std::vector<float> tmp(4 * cells);
            d2h(tmp.data(), macro.ptr, 4 * cells * 4);
            sync();
            for (int k = 0; k < 4; ++k)
                for (size_t i = 0; i < cells; ++i) outs[k][i] = static_cast<Out>(tmp[k * cells + i]);
        } else {
            std::vector<double> tmp(4 * cells);
            d2h(tmp.data(), macro.ptr, 4 * cells * 8);
            sync();
            for (int k = 0; k < 4; ++k)
                for (size_t i = 0; i < cells; ++i) outs[k][i] = static_cast<Out>(tmp[k * cells + i]);
        }
    }

    void diagnostics(double* out3) {
        use_device();
        if (!partial) partial.allocate(static_cast<size_t>(geo.tiles) * 3 * sizeof(double));
        if (!reduce_out) reduce_out.allocate(3 * sizeof(double));
        const LaunchState st = state();
        ops->diagnostics(st, odd(), bounded(), reinterpret_cast<double*>(partial.ptr),
                         reinterpret_cast<double*>(reduce_out.ptr));
        d2h(out3, reduce_out.ptr, 3 * sizeof(double));
        sync();
    }

    void index_map(int parity, int64_t* load, int64_t* store) {
        use_device();
        const size_t n = static_cast<size_t>(q) * cells;
        DeviceBuffer<unsigned char> ids, loads;
        ids.allocate(pop_elems * sizeof(uint32_t));
        loads.allocate(n * sizeof(int64_t));
        const LaunchState st = state();
        // Wall-velocity indices do not move slots; the probe sees plain walls.
        DeviceBuffer<unsigned char> probe_flags, probe_links, probe_vel;
        if (bounded()) {
            std::vector<uint8_t> host(cells);
            hip_check(hipMemcpy(host.data(), flags.ptr, cells, hipMemcpyDeviceToHost), "hipMemcpy flags");
            for (auto& v : host)
                if (v & 0x80U) v = 0x80U;
            probe_flags.allocate(cells);
            hip_check(hipMemcpy(probe_flags.ptr, host.data(), cells, hipMemcpyHostToDevice), "hipMemcpy flags");
            probe_links.allocate(cells * sizeof(uint32_t));
            probe_vel.allocate(cells * vel_words() * sizeof(uint32_t));
            ops->links(st, probe_flags.ptr, reinterpret_cast<uint32_t*>(probe_links.ptr),
                       reinterpret_cast<uint32_t*>(probe_vel.ptr));
        }
        ops->index_map(st, (parity & 1) != 0, bounded(), probe_flags.ptr,
                       reinterpret_cast<const uint32_t*>(probe_links.ptr), reinterpret_cast<uint32_t*>(ids.ptr),
                       reinterpret_cast<int64_t*>(loads.ptr));
        std::vector<uint32_t> slots(pop_elems);
        d2h(load, loads.ptr, n * sizeof(int64_t));
        d2h(slots.data(), ids.ptr, slots.size() * sizeof(uint32_t));
        sync();
        // Convert identifiers to slot numbers d*cells + n and invert the store map.
        std::fill(store, store + n, int64_t{-1});
        for (size_t d = 0; d < static_cast<size_t>(q); ++d)
            for (size_t c = 0; c < cells; ++c) {
                const size_t pc = cfd_hip::row_base(geo, static_cast<uint32_t>(c / nx)) + c % nx;
                const uint32_t id = slots[d * geo.stride + pc];
                if (id != 0xFFFFFFFFU) store[id] = static_cast<int64_t>(d * cells + c);
            }
    }

    size_t population_bytes() const { return static_cast<size_t>(q) * cells * elem; }
    size_t resident_bytes() const {
        return f.bytes + flags.bytes + links.bytes + link_vel.bytes + segments.bytes + solid_rho.bytes + gamma.bytes + lacc.bytes + macro.bytes + partial.bytes +
               reduce_out.bytes;
    }
};

namespace cfd_hip {
FlowInfo flow_info(cfd_lbm_solver* s) {
    if (!s) throw std::invalid_argument("null solver handle");
    FlowInfo f;
    f.device = s->device;
    f.stream = s->stream;
    f.nx = s->nx;
    f.ny = s->ny;
    f.nz = s->nz;
    f.cells = s->cells;
    f.real = s->real;
    return f;
}
}  // namespace cfd_hip

namespace {

void copy_error(char* out, size_t cap, const char* text) {
    if (!out || cap == 0U) return;
    const size_t n = std::min(cap - 1U, std::strlen(text));
    std::memcpy(out, text, n);
    out[n] = '\0';
}

template <class F> int guarded(char* error, size_t cap, F&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        copy_error(error, cap, e.what());
        return 1;
    } catch (...) {
        copy_error(error, cap, "unknown C++ exception");
        return 2;
    }
}

cfd_lbm_solver& need(cfd_lbm_solver* s) {
    if (!s) throw std::invalid_argument("null solver handle");
    return *s;
}

}  // namespace

extern "C" {

const char* cfd_lbm_version(void) { return "cfd_lbm_hip 0.1.0 (gfx950)"; }

int cfd_lbm_create(int lattice, size_t nx, size_t ny, size_t nz, float tau, float ax, float ay, float az, int storage,
                   cfd_lbm_solver** out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!out) throw std::invalid_argument("null output handle");
        *out = nullptr;
        *out = new cfd_lbm_solver(lattice, nx, ny, nz, tau, ax, ay, az, storage);
    });
}

int cfd_lbm_destroy(cfd_lbm_solver* s, char* error, size_t cap) {
    return guarded(error, cap, [&] { delete s; });
}

int cfd_lbm_initialize_uniform(cfd_lbm_solver* s, float rho, float ux, float uy, float uz, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need(s).initialize_uniform(rho, ux, uy, uz);
        s->sync();
    });
}

int cfd_lbm_initialize_taylor_green(cfd_lbm_solver* s, float amplitude, float noise, uint64_t seed, char* error,
                                    size_t cap) {
    return guarded(error, cap, [&] { need(s).initialize_taylor_green(amplitude, noise, seed); });
}

int cfd_lbm_initialize_fields(cfd_lbm_solver* s, const float* rho, const float* ux, const float* uy, const float* uz,
                              char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).initialize_fields(rho, ux, uy, uz); });
}

int cfd_lbm_upload_populations(cfd_lbm_solver* s, const float* logical, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).upload_populations(logical); });
}

int cfd_lbm_download_populations(cfd_lbm_solver* s, float* logical, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).download_populations(logical); });
}

int cfd_lbm_download_raw_slots(cfd_lbm_solver* s, float* slots, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).download_raw_slots(slots); });
}

int cfd_lbm_set_flags(cfd_lbm_solver* s, const uint8_t* flags, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).set_flags(flags); });
}

int cfd_lbm_get_flags(cfd_lbm_solver* s, uint8_t* flags, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!flags) throw std::invalid_argument("get_flags requires cells bytes");
        if (need(s).host_flags.empty()) std::memset(flags, 0, s->cells);
        else std::memcpy(flags, s->host_flags.data(), s->cells);
    });
}

int cfd_lbm_set_wall_velocity(cfd_lbm_solver* s, int index, float ux, float uy, float uz, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (index < 1 || index > 7) throw std::invalid_argument("wall velocity index must be 1..7");
        need(s).wall_u[index][0] = ux;
        s->wall_u[index][1] = uy;
        s->wall_u[index][2] = uz;
    });
}

int cfd_lbm_set_porosity(cfd_lbm_solver* s, const float* gamma, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).set_porosity(gamma); });
}

int cfd_lbm_set_porosity_device(cfd_lbm_solver* s, const float* gamma, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).set_porosity_device(gamma); });
}

int cfd_lbm_halo_message_bytes(cfd_lbm_solver* s, size_t* bytes, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!bytes) throw std::invalid_argument("halo_message_bytes needs an output");
        *bytes = need(s).halo_message_bytes();
    });
}

int cfd_lbm_halo_pack(cfd_lbm_solver* s, void* up, void* down, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).halo_pack(up, down); });
}

int cfd_lbm_halo_unpack(cfd_lbm_solver* s, const void* from_below, const void* from_above, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).halo_unpack(from_below, from_above); });
}

int cfd_lbm_set_inlet_outlet(cfd_lbm_solver* s, float u_peak, float rho_out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!std::isfinite(u_peak) || !(rho_out > 0.0F)) throw std::invalid_argument("need finite u_peak and rho_out > 0");
        if (need(s).ny < 3 || (s->q != 9 && s->nz < 3))
            throw std::invalid_argument("inlet profile needs at least one interior cell between walls");
        s->u_peak = u_peak;
        s->rho_out = rho_out;
    });
}

int cfd_lbm_set_local_acceleration(cfd_lbm_solver* s, const float* ax, const float* ay, const float* az, char* error,
                                   size_t cap) {
    return guarded(error, cap, [&] { need(s).set_local_acceleration(ax, ay, az); });
}

int cfd_lbm_step(cfd_lbm_solver* s, uint64_t count, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).step(count); });
}

int cfd_lbm_step_timed(cfd_lbm_solver* s, uint64_t count, float* ms, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const float v = need(s).step_timed(count);
        if (ms) *ms = v;
    });
}

int cfd_lbm_synchronize(cfd_lbm_solver* s, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).sync(); });
}

int cfd_lbm_download_macroscopic(cfd_lbm_solver* s, float* rho, float* ux, float* uy, float* uz, char* error,
                                 size_t cap) {
    return guarded(error, cap, [&] { need(s).download_macroscopic<float>(rho, ux, uy, uz); });
}

int cfd_lbm_download_macroscopic_f64(cfd_lbm_solver* s, double* rho, double* ux, double* uy, double* uz, char* error,
                                     size_t cap) {
    return guarded(error, cap, [&] { need(s).download_macroscopic<double>(rho, ux, uy, uz); });
}

int cfd_lbm_device_local_acceleration(cfd_lbm_solver* s, int clear, void** ax, void** ay, void** az, char* error,
                                      size_t cap) {
    return guarded(error, cap, [&] {
        unsigned char* base = need(s).device_local_acceleration(clear != 0);
        const size_t plane = s->cells * s->real;
        if (ax) *ax = base;
        if (ay) *ay = base + plane;
        if (az) *az = base + 2 * plane;
    });
}

int cfd_lbm_refresh_device_macroscopic(cfd_lbm_solver* s, void** rho, void** ux, void** uy, void** uz, char* error,
                                       size_t cap) {
    return guarded(error, cap, [&] {
        need(s).refresh_macroscopic();
        unsigned char* base = s->macro.ptr;
        const size_t plane = s->cells * s->real;
        if (rho) *rho = base;
        if (ux) *ux = base + plane;
        if (uy) *uy = base + 2 * plane;
        if (uz) *uz = base + 3 * plane;
    });
}

int cfd_lbm_total_mass(cfd_lbm_solver* s, double* mass, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        double out[3];
        need(s).diagnostics(out);
        if (mass) *mass = out[0];
    });
}

int cfd_lbm_diagnostics(cfd_lbm_solver* s, double* mass, double* ke, double* vmax, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        double out[3];
        need(s).diagnostics(out);
        if (mass) *mass = out[0];
        if (ke) *ke = out[1];
        if (vmax) *vmax = out[2];
    });
}

int cfd_lbm_time_step(cfd_lbm_solver* s, uint64_t* out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (out) *out = need(s).t;
    });
}

int cfd_lbm_cells(cfd_lbm_solver* s, size_t* out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (out) *out = need(s).cells;
    });
}

int cfd_lbm_population_bytes(cfd_lbm_solver* s, size_t* out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (out) *out = need(s).population_bytes();
    });
}

int cfd_lbm_resident_bytes(cfd_lbm_solver* s, size_t* out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (out) *out = need(s).resident_bytes();
    });
}

int cfd_lbm_device_name(cfd_lbm_solver* s, char* name, size_t name_cap, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, need(s).device), "hipGetDeviceProperties");
        std::string text = std::string(prop.name) + " (" + prop.gcnArchName + ")";
        copy_error(name, name_cap, text.c_str());
    });
}

int cfd_lbm_stream(cfd_lbm_solver* s, void** stream, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (stream) *stream = need(s).stream;
    });
}

int cfd_lbm_transfer_stats(cfd_lbm_solver* s, uint64_t* stats4, int reset, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        auto& st = need(s).stats;
        if (stats4) {
            stats4[0] = st.h2d;
            stats4[1] = st.d2h;
            stats4[2] = st.d2d;
            stats4[3] = st.syncs;
        }
        if (reset) st = TransferStats{};
    });
}

int cfd_lbm_objective_pressure_drop(cfd_lbm_solver* s, int x_in, int x_out, double* value, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const double v = need(s).objective(cfd_hip::kPressureDrop, x_in, x_out);
        if (value) *value = v;
    });
}

int cfd_lbm_set_smagorinsky(cfd_lbm_solver* s, int enabled, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).set_smagorinsky(enabled != 0); });
}

int cfd_lbm_set_drag_body(cfd_lbm_solver* s, const uint8_t* body, int axis, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).set_drag_body(body, axis); });
}

int cfd_lbm_objective_drag(cfd_lbm_solver* s, double* value, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const double v = need(s).objective(cfd_hip::kDrag, 0, 0);
        if (value) *value = v;
    });
}

int cfd_lbm_objective_run(cfd_lbm_solver* s, int objective_kind, uint64_t steps, uint64_t window, int x_in, int x_out,
                          double* objective, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const double J = need(s).objective_run(objective_kind, steps, window, x_in, x_out);
        if (objective) *objective = J;
    });
}

int cfd_lbm_adjoint_run(cfd_lbm_solver* s, uint64_t steps, uint64_t window, uint64_t checkpoint_every, int x_in,
                        int x_out, double* objective, double* grad_gamma, double* grad_u_peak, char* error,
                        size_t cap) {
    return guarded(error, cap, [&] {
        if (checkpoint_every < 1) throw std::invalid_argument("checkpoint interval must be >= 1");
        const double J = need(s).adjoint_run(cfd_hip::kPressureDrop, steps, window, checkpoint_every, 0, x_in, x_out,
                                             grad_gamma, grad_u_peak, nullptr);
        if (objective) *objective = J;
    });
}

int cfd_lbm_adjoint_run_binomial_device(cfd_lbm_solver* s, int objective_kind, uint64_t steps, uint64_t window,
                                        uint64_t snapshots, int x_in, int x_out, double* objective,
                                        double* grad_gamma, double* grad_u_peak, uint64_t* forward_steps, char* error,
                                        size_t cap) {
    return guarded(error, cap, [&] {
        const double J = need(s).adjoint_run(objective_kind, steps, window, 0, snapshots, x_in, x_out, grad_gamma,
                                             grad_u_peak, forward_steps, true);
        if (objective) *objective = J;
    });
}

int cfd_lbm_adjoint_run_binomial(cfd_lbm_solver* s, int objective_kind, uint64_t steps, uint64_t window,
                                 uint64_t snapshots, int x_in, int x_out, double* objective, double* grad_gamma,
                                 double* grad_u_peak, uint64_t* forward_steps, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const double J = need(s).adjoint_run(objective_kind, steps, window, 0, snapshots, x_in, x_out, grad_gamma,
                                             grad_u_peak, forward_steps);
        if (objective) *objective = J;
    });
}

int cfd_lbm_checkpoint_plan(uint64_t steps, uint64_t snapshots, uint64_t* forward_steps, uint64_t* repetitions,
                            uint64_t* walked_steps, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (steps < 1 || snapshots < 1) throw std::invalid_argument("need steps >= 1 and snapshots >= 1");
        using B = cfd_hip::BinomialSchedule;
        if (forward_steps) *forward_steps = B::forward_steps(steps, snapshots) + 1;  // + the step to the last state
        if (repetitions) *repetitions = B::repetitions(steps, snapshots);
        if (walked_steps) *walked_steps = B::walk(steps, snapshots) + 1;
    });
}

int cfd_lbm_adjoint_timing(cfd_lbm_solver* s, float* forward_ms, float* reverse_ms, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (forward_ms) *forward_ms = need(s).last_forward_ms;
        if (reverse_ms) *reverse_ms = s->last_reverse_ms;
    });
}

int cfd_lbm_save_checkpoint(cfd_lbm_solver* s, const char* path, int raw, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).save_checkpoint(path, raw != 0); });
}

int cfd_lbm_load_checkpoint(cfd_lbm_solver* s, const char* path, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).load_checkpoint(path); });
}

int cfd_lbm_voxelize_surface(const double* triangles, size_t count, size_t nx, size_t ny, size_t nz,
                             const double* minimum, const double* maximum, uint8_t* mask, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!minimum || !maximum || !mask) throw std::invalid_argument("voxelization needs minimum, maximum and a mask");
        cfd_hip::VoxelGridArgs g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        for (int a = 0; a < 3; ++a) g.minimum[a] = minimum[a], g.maximum[a] = maximum[a];
        cfd_hip::validate_voxelization(triangles, count, g);
        if (nx > SIZE_MAX / ny || nx * ny > SIZE_MAX / nz) throw std::invalid_argument("voxelization grid too large");
        const size_t cells = nx * ny * nz;
        hipStream_t stream = nullptr;
        hip_check(hipStreamCreate(&stream), "hipStreamCreate");
        std::unique_ptr<std::remove_pointer<hipStream_t>::type, void (*)(hipStream_t)> own(
            stream, [](hipStream_t st) { (void)hipStreamDestroy(st); });
        DeviceBuffer<unsigned char> d_mask;
        d_mask.allocate(std::max<size_t>(cells, 1));
        cfd_hip::voxelize_surface(triangles, count, g, d_mask.ptr, stream);
        hip_check(hipMemcpyAsync(mask, d_mask.ptr, cells, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    });
}

int cfd_lbm_voxelize_surface_no_alloc(const double* triangles, size_t count, size_t nx, size_t ny, size_t nz,
                                      const double* minimum, const double* maximum, uint8_t* mask, void* stream,
                                      char* error, size_t cap) {
    static_assert(CFD_LBM_VOXELIZE_MAX_TILES == cfd_hip::kVoxelizeMaxTiles, "tile limit");
    static_assert(CFD_LBM_VOXELIZE_MAX_PAIRS == cfd_hip::kVoxelizeMaxPairs, "pair limit");
    return guarded(error, cap, [&] {
        if (!minimum || !maximum) throw std::invalid_argument("voxelization needs minimum and maximum");
        cfd_hip::VoxelGridArgs g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        for (int a = 0; a < 3; ++a) g.minimum[a] = minimum[a], g.maximum[a] = maximum[a];
        cfd_hip::voxelize_surface_no_alloc(triangles, count, g, mask, static_cast<hipStream_t>(stream));
    });
}

int cfd_lbm_voxelize_timing(double* bin_ms, double* lines_ms, uint64_t* pairs, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const cfd_hip::VoxelizeTiming t = cfd_hip::last_voxelize_timing();
        if (bin_ms) *bin_ms = t.bin_ms;
        if (lines_ms) *lines_ms = t.lines_ms;
        if (pairs) *pairs = t.pairs;
    });
}

int cfd_lbm_transform_surface(const double* triangles, size_t count, const double* translation, const double* axis,
                              double angle, double* out, char* error, size_t cap) {
    return guarded(error, cap, [&] { cfd_hip::transform_surface(triangles, count, translation, axis, angle, out); });
}

int cfd_lbm_set_solid_surface(cfd_lbm_solver* s, const double* triangles, size_t count, const double* minimum,
                              const double* maximum, int keep_flags, char* error, size_t cap) {
    return guarded(error, cap, [&] { need(s).set_solid_surface(triangles, count, minimum, maximum, keep_flags != 0); });
}

int cfd_lbm_index_map(cfd_lbm_solver* s, int parity, int64_t* load, int64_t* store, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!load || !store) throw std::invalid_argument("index_map requires two q*cells arrays");
        need(s).index_map(parity, load, store);
    });
}

}  // extern "C"
