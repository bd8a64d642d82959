// esoteric_pull_hip.hpp -- C++ front end of the gfx950 esoteric-pull backend.
//
// EsotericPullHipSolver<Descriptor> is a sibling of the reference's
//   cfd::lbm::EsotericPullSolver<D>      (include/cfd/solvers/lbm/esoteric_pull.hpp:104-278)
//   cfd::lbm::EsotericPullSyclSolver<D>  (include/cfd/solvers/lbm/esoteric_pull_sycl.hpp:18-475)
// with the same member names and meanings, implemented over the C ABI in c_api.h (no HIP or
// torch types in this header).  Descriptor only needs ::q (9, 19 or 27) and ::dimensions, so the
// reference's descriptor types can be used directly; the constructor accepts any config struct
// with the fields of InPlaceLbmConfig (esoteric_pull.hpp:20-29).
#pragma once

#include "cfd_lbm/c_api.h"

#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace cfd::lbm {

struct HipLbmConfig {
    std::size_t nx{128};
    std::size_t ny{128};
    std::size_t nz{128};
    float tau{0.6F};
    float acceleration_x{0.0F};
    float acceleration_y{0.0F};
    float acceleration_z{0.0F};
    bool smagorinsky_les{false};
};

enum class HipPopulationStorage : int {
    fp32 = CFD_LBM_STORE_FP32,
    fp16_shifted = CFD_LBM_STORE_FP16S,
    fp64 = CFD_LBM_STORE_FP64,
    fp32_shifted = CFD_LBM_STORE_FP32S,
    fp64_exact_weights = CFD_LBM_STORE_FP64X,
};

struct HipMacroscopicFields {
    std::vector<float> rho;
    std::vector<float> ux;
    std::vector<float> uy;
    std::vector<float> uz;
};

// include/cfd/core/device_residency.hpp:11-42
struct HipTransferStats {
    std::uint64_t host_to_device_bytes{};
    std::uint64_t device_to_host_bytes{};
    std::uint64_t device_to_device_bytes{};
    std::uint64_t synchronization_points{};
    [[nodiscard]] std::uint64_t explicit_transfer_bytes() const noexcept {
        return host_to_device_bytes + device_to_host_bytes + device_to_device_bytes;
    }
    [[nodiscard]] std::uint64_t host_transfer_bytes() const noexcept {
        return host_to_device_bytes + device_to_host_bytes;
    }
};

// Triangles of any surface type with the shape of cfd::io::TriangleSurface (a `triangles` range of
// {a, b, c} points with x, y, z; include/cfd/io/mesh_io.hpp:9-10) as 9 doubles each.
template <class Surface> std::vector<double> packed_triangles(const Surface& surface) {
    std::vector<double> out;
    out.reserve(surface.triangles.size() * 9);
    for (const auto& t : surface.triangles)
        for (const auto* p : {&t.a, &t.b, &t.c}) {
            out.push_back(static_cast<double>(p->x));
            out.push_back(static_cast<double>(p->y));
            out.push_back(static_cast<double>(p->z));
        }
    return out;
}

// voxelize_surface (src/lbm/voxelize.cpp:5) on the GPU.  Grid is any struct with the fields of
// VoxelGrid3D (nx, ny, nz, minimum, maximum; include/cfd/solvers/lbm/voxelize.hpp:8).  The mask is
// the reference's bit for bit.
template <class Surface, class Grid>
[[nodiscard]] std::vector<unsigned char> voxelize_surface_hip(const Surface& surface, const Grid& grid) {
    const std::vector<double> t = packed_triangles(surface);
    const double lo[3] = {grid.minimum.x, grid.minimum.y, grid.minimum.z};
    const double hi[3] = {grid.maximum.x, grid.maximum.y, grid.maximum.z};
    std::vector<unsigned char> mask(grid.nx * grid.ny * grid.nz);
    char err[512] = {};
    if (cfd_lbm_voxelize_surface(t.data(), t.size() / 9, grid.nx, grid.ny, grid.nz, lo, hi, mask.data(), err,
                                 sizeof(err)) != 0)
        throw std::invalid_argument(err);
    return mask;
}

template <class Descriptor> class EsotericPullHipSolver {
public:
    static constexpr int q = Descriptor::q;

    template <class Config>
    explicit EsotericPullHipSolver(const Config& config,
                                   HipPopulationStorage storage = HipPopulationStorage::fp32)
        : storage_(storage) {
        config_.nx = config.nx;
        config_.ny = config.ny;
        config_.nz = Descriptor::dimensions == 2 ? 1 : config.nz;
        config_.tau = config.tau;
        config_.acceleration_x = config.acceleration_x;
        config_.acceleration_y = config.acceleration_y;
        config_.acceleration_z = config.acceleration_z;
        config_.smagorinsky_les = config.smagorinsky_les;
        call(cfd_lbm_create(q, config_.nx, config_.ny, config_.nz, config_.tau, config_.acceleration_x,
                            config_.acceleration_y, config_.acceleration_z, static_cast<int>(storage), &h_, err_,
                            sizeof(err_)));
        cells_ = config_.nx * config_.ny * config_.nz;
        if (config_.smagorinsky_les) {
            const int rc = cfd_lbm_set_smagorinsky(h_, 1, err_, sizeof(err_));
            if (rc != 0) {
                (void)cfd_lbm_destroy(h_, nullptr, 0);
                h_ = nullptr;
                throw std::invalid_argument(err_);
            }
        }
    }

    ~EsotericPullHipSolver() noexcept {
        if (h_) (void)cfd_lbm_destroy(h_, nullptr, 0);
    }
    EsotericPullHipSolver(const EsotericPullHipSolver&) = delete;
    EsotericPullHipSolver& operator=(const EsotericPullHipSolver&) = delete;

    void initialize_uniform(float rho = 1.0F, float ux = 0.0F, float uy = 0.0F, float uz = 0.0F) {
        call(cfd_lbm_initialize_uniform(h_, rho, ux, uy, uz, err_, sizeof(err_)));
    }
    void initialize_taylor_green(float amplitude = 0.03F, float noise_amplitude = 0.0F, std::uint64_t seed = 0) {
        call(cfd_lbm_initialize_taylor_green(h_, amplitude, noise_amplitude, seed, err_, sizeof(err_)));
    }
    void initialize_fields(std::span<const float> rho, std::span<const float> ux, std::span<const float> uy,
                           std::span<const float> uz) {
        require_cells(rho.size() == cells_ && ux.size() == cells_ && uy.size() == cells_ && uz.size() == cells_,
                      "LBM field size mismatch");
        call(cfd_lbm_initialize_fields(h_, rho.data(), ux.data(), uy.data(), uz.data(), err_, sizeof(err_)));
    }

    void step(std::size_t count = 1) { call(cfd_lbm_step(h_, count, err_, sizeof(err_))); }
    void wait() { call(cfd_lbm_synchronize(h_, err_, sizeof(err_))); }

    void set_solid_mask(std::span<const std::uint8_t> mask) {
        require_cells(mask.size() == cells_, "LBM solid mask size mismatch");
        flags_.resize(cells_);
        for (std::size_t i = 0; i < cells_; ++i) flags_[i] = mask[i] != 0U ? CFD_LBM_WALL : CFD_LBM_FLUID;
        mask_.assign(mask.begin(), mask.end());
        call(cfd_lbm_set_flags(h_, flags_.data(), err_, sizeof(err_)));
    }
    void set_flags(std::span<const std::uint8_t> flags) {
        require_cells(flags.size() == cells_, "LBM flag field size mismatch");
        flags_.assign(flags.begin(), flags.end());
        mask_.resize(cells_);
        for (std::size_t i = 0; i < cells_; ++i) mask_[i] = (flags_[i] & CFD_LBM_WALL) != 0U ? 1U : 0U;
        call(cfd_lbm_set_flags(h_, flags_.data(), err_, sizeof(err_)));
    }
    // set_solid_mask(voxelize_surface(surface, grid)) (examples/real_lbm3d_model.cpp:146-151) on the
    // device; the grid's cell counts are this solver's, minimum/maximum any points with x, y, z.
    // keep_flags keeps the current flags of the cells outside the surface.
    template <class Surface, class Point>
    void set_solid_surface(const Surface& surface, const Point& minimum, const Point& maximum,
                           bool keep_flags = false) {
        const std::vector<double> t = packed_triangles(surface);
        const double lo[3] = {minimum.x, minimum.y, minimum.z};
        const double hi[3] = {maximum.x, maximum.y, maximum.z};
        call(cfd_lbm_set_solid_surface(h_, t.data(), t.size() / 9, lo, hi, keep_flags ? 1 : 0, err_, sizeof(err_)));
        flags_.resize(cells_);
        call(cfd_lbm_get_flags(h_, flags_.data(), err_, sizeof(err_)));
        mask_.resize(cells_);
        for (std::size_t i = 0; i < cells_; ++i) mask_[i] = (flags_[i] & CFD_LBM_WALL) != 0U ? 1U : 0U;
    }
    // Checkpoint / restart in the CFDCHK3 layout (distributed_pull.hpp:176-245) plus the parity.
    void save_checkpoint(const std::string& path, bool raw = true) {
        call(cfd_lbm_save_checkpoint(h_, path.c_str(), raw ? 1 : 0, err_, sizeof(err_)));
    }
    void load_checkpoint(const std::string& path) { call(cfd_lbm_load_checkpoint(h_, path.c_str(), err_, sizeof(err_))); }
    void clear_solid_mask() {
        flags_.clear();
        mask_.clear();
        call(cfd_lbm_set_flags(h_, nullptr, err_, sizeof(err_)));
    }
    [[nodiscard]] bool has_solid_mask() const noexcept { return !mask_.empty(); }
    [[nodiscard]] const std::vector<std::uint8_t>& solid_mask() const noexcept { return mask_; }

    void set_wall_velocity(int index, float ux, float uy, float uz) {
        call(cfd_lbm_set_wall_velocity(h_, index, ux, uy, uz, err_, sizeof(err_)));
    }
    void set_porosity(std::span<const float> gamma) {
        require_cells(gamma.size() == cells_, "porosity field size mismatch");
        call(cfd_lbm_set_porosity(h_, gamma.data(), err_, sizeof(err_)));
    }
    void clear_porosity() { call(cfd_lbm_set_porosity(h_, nullptr, err_, sizeof(err_))); }
    void set_inlet_outlet(float u_peak, float rho_out = 1.0F) {
        call(cfd_lbm_set_inlet_outlet(h_, u_peak, rho_out, err_, sizeof(err_)));
    }

    void set_local_body_acceleration(std::span<const float> ax, std::span<const float> ay,
                                     std::span<const float> az) {
        require_cells(ax.size() == cells_ && ay.size() == cells_ && az.size() == cells_,
                      "local LBM acceleration field size mismatch");
        call(cfd_lbm_set_local_acceleration(h_, ax.data(), ay.data(), az.data(), err_, sizeof(err_)));
        has_local_ = true;
    }
    void clear_local_body_acceleration() {
        call(cfd_lbm_set_local_acceleration(h_, nullptr, nullptr, nullptr, err_, sizeof(err_)));
        has_local_ = false;
    }
    [[nodiscard]] bool has_local_body_acceleration() const noexcept { return has_local_; }

    [[nodiscard]] HipMacroscopicFields compute_macroscopic() const {
        HipMacroscopicFields m{std::vector<float>(cells_), std::vector<float>(cells_), std::vector<float>(cells_),
                               std::vector<float>(cells_)};
        call(cfd_lbm_download_macroscopic(h_, m.rho.data(), m.ux.data(), m.uy.data(), m.uz.data(), err_,
                                          sizeof(err_)));
        return m;
    }
    [[nodiscard]] HipMacroscopicFields download_macroscopic() const { return compute_macroscopic(); }

    struct DeviceMacroscopicView {
        void* rho{};
        void* ux{};
        void* uy{};
        void* uz{};
        std::size_t cells{};
    };
    [[nodiscard]] DeviceMacroscopicView refresh_device_macroscopic() const {
        DeviceMacroscopicView v;
        v.cells = cells_;
        call(cfd_lbm_refresh_device_macroscopic(h_, &v.rho, &v.ux, &v.uy, &v.uz, err_, sizeof(err_)));
        return v;
    }

    [[nodiscard]] double total_mass() const {
        double m = 0.0;
        call(cfd_lbm_total_mass(h_, &m, err_, sizeof(err_)));
        return m;
    }
    [[nodiscard]] double mass() const { return total_mass(); }
    [[nodiscard]] double kinetic_energy() const {
        double m, ke, v;
        call(cfd_lbm_diagnostics(h_, &m, &ke, &v, err_, sizeof(err_)));
        return ke;
    }
    [[nodiscard]] float max_speed() const {
        double m, ke, v;
        call(cfd_lbm_diagnostics(h_, &m, &ke, &v, err_, sizeof(err_)));
        return static_cast<float>(v);
    }
    [[nodiscard]] std::vector<float> logical_populations() const {
        std::vector<float> out(static_cast<std::size_t>(q) * cells_);
        call(cfd_lbm_download_populations(h_, out.data(), err_, sizeof(err_)));
        return out;
    }
    // raw_storage (esoteric_pull.hpp:276-278): the in-place slot array in the order d*cells + n,
    // copied to the host, each slot as a full float population.
    [[nodiscard]] std::vector<float> raw_storage() const {
        std::vector<float> out(static_cast<std::size_t>(q) * cells_);
        call(cfd_lbm_download_raw_slots(h_, out.data(), err_, sizeof(err_)));
        return out;
    }
    void upload_populations(std::span<const float> logical) {
        require_cells(logical.size() == static_cast<std::size_t>(q) * cells_, "population array size mismatch");
        call(cfd_lbm_upload_populations(h_, logical.data(), err_, sizeof(err_)));
    }

    [[nodiscard]] const HipLbmConfig& config() const noexcept { return config_; }
    [[nodiscard]] std::size_t cells() const noexcept { return cells_; }
    [[nodiscard]] std::uint64_t time_step() const {
        std::uint64_t t = 0;
        call(cfd_lbm_time_step(h_, &t, err_, sizeof(err_)));
        return t;
    }
    [[nodiscard]] std::size_t population_bytes() const {
        std::size_t b = 0;
        call(cfd_lbm_population_bytes(h_, &b, err_, sizeof(err_)));
        return b;
    }
    [[nodiscard]] std::size_t two_lattice_population_bytes() const { return 2U * population_bytes(); }
    [[nodiscard]] std::size_t resident_bytes() const {
        std::size_t b = 0;
        call(cfd_lbm_resident_bytes(h_, &b, err_, sizeof(err_)));
        return b;
    }
    [[nodiscard]] std::string device_name() const {
        char name[256];
        call(cfd_lbm_device_name(h_, name, sizeof(name), err_, sizeof(err_)));
        return name;
    }
    [[nodiscard]] void* stream() const {
        void* s = nullptr;
        call(cfd_lbm_stream(h_, &s, err_, sizeof(err_)));
        return s;
    }
    // queue() of the device sibling (esoteric_pull_sycl.hpp:375-376): here the HIP stream every
    // launch of this solver is ordered on (a hipStream_t).
    [[nodiscard]] void* queue() const { return stream(); }

    // device_local_acceleration (esoteric_pull_sycl.hpp:342-353): the per-cell acceleration planes
    // on the device, allocated (zero) on first use and optionally cleared; the next steps apply them.
    struct DeviceAccelerationView {
        float* ax{};
        float* ay{};
        float* az{};
        std::size_t cells{};
    };
    [[nodiscard]] DeviceAccelerationView device_local_acceleration(bool clear = false) {
        void* p[3] = {};
        call(cfd_lbm_device_local_acceleration(h_, clear ? 1 : 0, &p[0], &p[1], &p[2], err_, sizeof(err_)));
        has_local_ = true;
        return {static_cast<float*>(p[0]), static_cast<float*>(p[1]), static_cast<float*>(p[2]), cells_};
    }
    [[nodiscard]] HipTransferStats transfer_stats() const { return stats(false); }
    void reset_transfer_stats() const { (void)stats(true); }
    [[nodiscard]] HipPopulationStorage storage() const noexcept { return storage_; }

private:
    cfd_lbm_solver* h_{nullptr};
    HipLbmConfig config_{};
    HipPopulationStorage storage_;
    std::size_t cells_{};
    std::vector<std::uint8_t> flags_;
    std::vector<std::uint8_t> mask_;
    bool has_local_{false};
    mutable char err_[512]{};

    void call(int rc) const {
        if (rc == 1) throw std::runtime_error(err_);
        if (rc != 0) throw std::runtime_error(std::string("cfd_lbm: ") + err_);
    }
    void require_cells(bool ok, const char* what) const {
        if (!ok) throw std::invalid_argument(what);
    }
    HipTransferStats stats(bool reset) const {
        std::uint64_t v[4] = {};
        call(cfd_lbm_transfer_stats(h_, v, reset ? 1 : 0, err_, sizeof(err_)));
        return HipTransferStats{v[0], v[1], v[2], v[3]};
    }
};

}  // namespace cfd::lbm
