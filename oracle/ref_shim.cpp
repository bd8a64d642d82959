// ref_shim.cpp -- C entry points over the reference's own LBM classes, compiled from the
// reference sources where they lie (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: it lets
// tests/golden/make_golden.py and bench.py's cpu_baseline leg call
//   cfd::lbm::EsotericPullSolver<D>   include/cfd/solvers/lbm/esoteric_pull.hpp:104-515
//   cfd::lbm::OneStepPullSolver<D>    include/cfd/solvers/lbm/one_step_pull.hpp
//   cfd::lbm::PrecisionPullSolver<D,double> include/cfd/solvers/lbm/precision_pull.hpp
//   cfd::lbm::D2Q9Solver              src/lbm/d2q9.cpp
//   cfd::lbm::voxelize_surface, transform_surface   src/lbm/voxelize.cpp:5-6
//   cfd::lbm::DistributedPullBlock<D3Q19>  include/cfd/solvers/lbm/distributed_pull.hpp (one brick;
//                                          CFDCHK3 save/load :176-245, virtual_step)
// through ctypes.  Nothing here re-implements the reference; it only forwards.
#include "cfd/solvers/lbm/d2q9.hpp"
#include "cfd/solvers/lbm/distributed_pull.hpp"
#include "cfd/solvers/lbm/esoteric_pull.hpp"
#include "cfd/solvers/lbm/one_step_pull.hpp"
#include "cfd/solvers/lbm/precision_pull.hpp"
#include "cfd/solvers/lbm/voxelize.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <span>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

using namespace cfd::lbm;

struct EsoBase {
    virtual ~EsoBase() = default;
    virtual void init_tg(float a) = 0;
    virtual void init_uniform(float r, float ux, float uy, float uz) = 0;
    virtual void step(std::size_t n) = 0;
    virtual void macroscopic(float* r, float* ux, float* uy, float* uz) const = 0;
    virtual double mass() const = 0;
    virtual double kinetic_energy() const = 0;
    virtual float max_speed() const = 0;
    virtual void logical(float* out) const = 0;
    virtual void raw(float* out) const = 0;
    virtual void set_raw(const float* in) = 0;
    virtual void set_mask(const std::uint8_t* m) = 0;
    virtual void set_local_acc(const float* ax, const float* ay, const float* az) = 0;
    virtual std::uint64_t time_step() const = 0;
    virtual std::size_t cells() const = 0;
};

template <class D> struct Eso final : EsoBase {
    EsotericPullSolver<D> s;
    explicit Eso(InPlaceLbmConfig c) : s(c) {}
    void init_tg(float a) override { s.initialize_taylor_green(a); }
    void init_uniform(float r, float ux, float uy, float uz) override { s.initialize_uniform(r, ux, uy, uz); }
    void step(std::size_t n) override { s.step(n); }
    void macroscopic(float* r, float* ux, float* uy, float* uz) const override {
        const auto m = s.compute_macroscopic();
        std::memcpy(r, m.rho.data(), m.rho.size() * 4);
        std::memcpy(ux, m.ux.data(), m.ux.size() * 4);
        std::memcpy(uy, m.uy.data(), m.uy.size() * 4);
        std::memcpy(uz, m.uz.data(), m.uz.size() * 4);
    }
    double mass() const override { return s.mass(); }
    double kinetic_energy() const override { return s.kinetic_energy(); }
    float max_speed() const override { return s.max_speed(); }
    void logical(float* out) const override {
        const auto v = s.logical_populations();
        std::memcpy(out, v.data(), v.size() * 4);
    }
    void raw(float* out) const override {
        const auto& r = s.raw_storage().raw();
        std::memcpy(out, r.data(), r.size() * 4);
    }
    // The reference exposes its storage read-only; the shim writes through it to plant a
    // known state (slot identifiers for the index map, or a shared initial state).
    void set_raw(const float* in) override {
        auto& r = const_cast<cfd::core::StaticSoA<float, static_cast<std::size_t>(D::q)>&>(s.raw_storage()).raw();
        std::memcpy(r.data(), in, r.size() * 4);
    }
    void set_mask(const std::uint8_t* m) override {
        if (m) s.set_solid_mask(std::span<const std::uint8_t>(m, s.cells()));
        else s.clear_solid_mask();
    }
    // set_local_body_acceleration / clear_local_body_acceleration (esoteric_pull.hpp:144-157).
    void set_local_acc(const float* ax, const float* ay, const float* az) override {
        if (!ax) {
            s.clear_local_body_acceleration();
            return;
        }
        const std::size_t n = s.cells();
        s.set_local_body_acceleration(std::span<const float>(ax, n), std::span<const float>(ay, n),
                                      std::span<const float>(az, n));
    }
    std::uint64_t time_step() const override { return s.time_step(); }
    std::size_t cells() const override { return s.cells(); }
};

struct PullBase {
    virtual ~PullBase() = default;
    virtual void init_tg(float a) = 0;
    virtual void init_uniform() = 0;
    virtual void step(std::size_t n) = 0;
    virtual void macroscopic(float* r, float* ux, float* uy, float* uz) const = 0;
};
template <class D> struct Pull final : PullBase {
    OneStepPullSolver<D> s;
    explicit Pull(InPlaceLbmConfig c) : s(c) {}
    void init_tg(float a) override { s.initialize_taylor_green(a); }
    void init_uniform() override { s.initialize_uniform(); }
    void step(std::size_t n) override { s.step(n); }
    void macroscopic(float* r, float* ux, float* uy, float* uz) const override {
        const auto m = s.compute_macroscopic();
        std::memcpy(r, m.rho.data(), m.rho.size() * 4);
        std::memcpy(ux, m.ux.data(), m.ux.size() * 4);
        std::memcpy(uy, m.uy.data(), m.uy.size() * 4);
        std::memcpy(uz, m.uz.data(), m.uz.size() * 4);
    }
};

struct PrecBase {
    virtual ~PrecBase() = default;
    virtual void init_tg(double a) = 0;
    virtual void step(std::size_t n) = 0;
    virtual void macroscopic(double* r, double* ux, double* uy, double* uz) const = 0;
};
template <class D> struct Prec final : PrecBase {
    PrecisionPullSolver<D, double> s;
    explicit Prec(TypedLbmConfig<double> c) : s(c) {}
    void init_tg(double a) override { s.initialize_taylor_green(a); }
    void step(std::size_t n) override { s.step(n); }
    void macroscopic(double* r, double* ux, double* uy, double* uz) const override {
        const auto m = s.compute_macroscopic();
        std::memcpy(r, m.rho.data(), m.rho.size() * 8);
        std::memcpy(ux, m.ux.data(), m.ux.size() * 8);
        std::memcpy(uy, m.uy.data(), m.uy.size() * 8);
        std::memcpy(uz, m.uz.data(), m.uz.size() * 8);
    }
};

thread_local char g_error[512];
template <class F> int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        std::snprintf(g_error, sizeof(g_error), "%s", e.what());
        return 1;
    } catch (...) {
        std::snprintf(g_error, sizeof(g_error), "unknown exception");
        return 2;
    }
}

}  // namespace

REF_API const char* ref_last_error(void) { return g_error; }

REF_API int ref_eso_create(int q, std::size_t nx, std::size_t ny, std::size_t nz, float tau, float ax, float ay,
                           float az, void** out) {
    return guarded([&] {
        const InPlaceLbmConfig c{nx, ny, nz, tau, ax, ay, az, false};
        if (q == 9) *out = static_cast<EsoBase*>(new Eso<D2Q9InPlaceDescriptor>(c));
        else if (q == 19) *out = static_cast<EsoBase*>(new Eso<D3Q19Descriptor>(c));
        else if (q == 27) *out = static_cast<EsoBase*>(new Eso<D3Q27Descriptor>(c));
        else throw std::invalid_argument("q must be 9, 19 or 27");
    });
}
// The same with the reference's Smagorinsky-LES switch (InPlaceLbmConfig::smagorinsky_les).
REF_API int ref_eso_create_les(int q, std::size_t nx, std::size_t ny, std::size_t nz, float tau, float ax, float ay,
                               float az, int les, void** out) {
    return guarded([&] {
        const InPlaceLbmConfig c{nx, ny, nz, tau, ax, ay, az, les != 0};
        if (q == 9) *out = static_cast<EsoBase*>(new Eso<D2Q9InPlaceDescriptor>(c));
        else if (q == 19) *out = static_cast<EsoBase*>(new Eso<D3Q19Descriptor>(c));
        else if (q == 27) *out = static_cast<EsoBase*>(new Eso<D3Q27Descriptor>(c));
        else throw std::invalid_argument("q must be 9, 19 or 27");
    });
}
REF_API void ref_eso_destroy(void* h) { delete static_cast<EsoBase*>(h); }
REF_API int ref_eso_init_tg(void* h, float a) { return guarded([&] { static_cast<EsoBase*>(h)->init_tg(a); }); }
REF_API int ref_eso_init_uniform(void* h, float r, float ux, float uy, float uz) {
    return guarded([&] { static_cast<EsoBase*>(h)->init_uniform(r, ux, uy, uz); });
}
REF_API int ref_eso_step(void* h, std::size_t n) { return guarded([&] { static_cast<EsoBase*>(h)->step(n); }); }
REF_API int ref_eso_macroscopic(void* h, float* r, float* ux, float* uy, float* uz) {
    return guarded([&] { static_cast<EsoBase*>(h)->macroscopic(r, ux, uy, uz); });
}
REF_API double ref_eso_mass(void* h) { return static_cast<EsoBase*>(h)->mass(); }
REF_API double ref_eso_kinetic_energy(void* h) { return static_cast<EsoBase*>(h)->kinetic_energy(); }
REF_API float ref_eso_max_speed(void* h) { return static_cast<EsoBase*>(h)->max_speed(); }
REF_API int ref_eso_logical(void* h, float* out) { return guarded([&] { static_cast<EsoBase*>(h)->logical(out); }); }
REF_API int ref_eso_raw(void* h, float* out) { return guarded([&] { static_cast<EsoBase*>(h)->raw(out); }); }
REF_API int ref_eso_set_raw(void* h, const float* in) { return guarded([&] { static_cast<EsoBase*>(h)->set_raw(in); }); }
REF_API int ref_eso_set_mask(void* h, const std::uint8_t* m) {
    return guarded([&] { static_cast<EsoBase*>(h)->set_mask(m); });
}
REF_API int ref_eso_set_local_acceleration(void* h, const float* ax, const float* ay, const float* az) {
    return guarded([&] { static_cast<EsoBase*>(h)->set_local_acc(ax, ay, az); });
}
REF_API std::uint64_t ref_eso_time_step(void* h) { return static_cast<EsoBase*>(h)->time_step(); }

REF_API int ref_pull_create(int q, std::size_t nx, std::size_t ny, std::size_t nz, float tau, float ax, float ay,
                            float az, void** out) {
    return guarded([&] {
        const InPlaceLbmConfig c{nx, ny, nz, tau, ax, ay, az, false};
        if (q == 9) *out = static_cast<PullBase*>(new Pull<D2Q9InPlaceDescriptor>(c));
        else if (q == 19) *out = static_cast<PullBase*>(new Pull<D3Q19Descriptor>(c));
        else if (q == 27) *out = static_cast<PullBase*>(new Pull<D3Q27Descriptor>(c));
        else throw std::invalid_argument("q must be 9, 19 or 27");
    });
}
REF_API void ref_pull_destroy(void* h) { delete static_cast<PullBase*>(h); }
REF_API int ref_pull_init_tg(void* h, float a) { return guarded([&] { static_cast<PullBase*>(h)->init_tg(a); }); }
REF_API int ref_pull_init_uniform(void* h) { return guarded([&] { static_cast<PullBase*>(h)->init_uniform(); }); }
REF_API int ref_pull_step(void* h, std::size_t n) { return guarded([&] { static_cast<PullBase*>(h)->step(n); }); }
REF_API int ref_pull_macroscopic(void* h, float* r, float* ux, float* uy, float* uz) {
    return guarded([&] { static_cast<PullBase*>(h)->macroscopic(r, ux, uy, uz); });
}

REF_API int ref_prec_create(int q, std::size_t nx, std::size_t ny, std::size_t nz, double tau, void** out) {
    return guarded([&] {
        TypedLbmConfig<double> c{};
        c.nx = nx;
        c.ny = ny;
        c.nz = nz;
        c.tau = tau;
        if (q == 9) *out = static_cast<PrecBase*>(new Prec<D2Q9InPlaceDescriptor>(c));
        else if (q == 19) *out = static_cast<PrecBase*>(new Prec<D3Q19Descriptor>(c));
        else if (q == 27) *out = static_cast<PrecBase*>(new Prec<D3Q27Descriptor>(c));
        else throw std::invalid_argument("q must be 9, 19 or 27");
    });
}
REF_API void ref_prec_destroy(void* h) { delete static_cast<PrecBase*>(h); }
REF_API int ref_prec_init_tg(void* h, double a) { return guarded([&] { static_cast<PrecBase*>(h)->init_tg(a); }); }
REF_API int ref_prec_step(void* h, std::size_t n) { return guarded([&] { static_cast<PrecBase*>(h)->step(n); }); }
REF_API int ref_prec_macroscopic(void* h, double* r, double* ux, double* uy, double* uz) {
    return guarded([&] { static_cast<PrecBase*>(h)->macroscopic(r, ux, uy, uz); });
}

REF_API int ref_d2q9_create(std::size_t nx, std::size_t ny, float tau, float ax, float ay, void** out) {
    return guarded([&] { *out = new D2Q9Solver(D2Q9Config{nx, ny, tau, ax, ay}); });
}
REF_API void ref_d2q9_destroy(void* h) { delete static_cast<D2Q9Solver*>(h); }
REF_API int ref_d2q9_set_solid(void* h, std::size_t x, std::size_t y) {
    return guarded([&] { static_cast<D2Q9Solver*>(h)->set_solid(x, y); });
}
REF_API int ref_d2q9_set_wall_velocity(void* h, std::size_t x, std::size_t y, float ux, float uy) {
    return guarded([&] { static_cast<D2Q9Solver*>(h)->set_wall_velocity(x, y, ux, uy); });
}
REF_API int ref_d2q9_init_uniform(void* h, float r, float ux, float uy) {
    return guarded([&] { static_cast<D2Q9Solver*>(h)->initialize_uniform(r, ux, uy); });
}
REF_API int ref_d2q9_init_tg(void* h, float a) {
    return guarded([&] { static_cast<D2Q9Solver*>(h)->initialize_taylor_green(a); });
}
REF_API int ref_d2q9_step(void* h, std::size_t n) { return guarded([&] { static_cast<D2Q9Solver*>(h)->step(n); }); }
REF_API int ref_d2q9_fields(void* h, float* r, float* ux, float* uy) {
    return guarded([&] {
        auto* s = static_cast<D2Q9Solver*>(h);
        const std::size_t n = s->grid().cells();
        std::memcpy(r, s->density().data(), n * 4);
        std::memcpy(ux, s->velocity_x().data(), n * 4);
        std::memcpy(uy, s->velocity_y().data(), n * 4);
    });
}
REF_API double ref_d2q9_mass(void* h) { return static_cast<D2Q9Solver*>(h)->mass(); }

namespace {
cfd::io::TriangleSurface to_surface(const double* tris, std::size_t count) {
    cfd::io::TriangleSurface s;
    s.triangles.resize(count);
    for (std::size_t t = 0; t < count; ++t) {
        const double* v = tris + 9 * t;
        s.triangles[t].a = {v[0], v[1], v[2]};
        s.triangles[t].b = {v[3], v[4], v[5]};
        s.triangles[t].c = {v[6], v[7], v[8]};
    }
    return s;
}
}  // namespace

REF_API int ref_voxelize(const double* tris, std::size_t count, std::size_t nx, std::size_t ny, std::size_t nz,
                         const double* lo, const double* hi, std::uint8_t* mask) {
    return guarded([&] {
        const VoxelGrid3D g{nx, ny, nz, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
        const auto m = voxelize_surface(to_surface(tris, count), g);
        std::memcpy(mask, m.data(), m.size());
    });
}
REF_API int ref_transform_surface(const double* tris, std::size_t count, const double* tr, const double* axis,
                                  double angle, double* out) {
    return guarded([&] {
        const auto s = transform_surface(to_surface(tris, count), {tr[0], tr[1], tr[2]}, {axis[0], axis[1], axis[2]},
                                         angle);
        for (std::size_t t = 0; t < count; ++t) {
            const auto& x = s.triangles[t];
            const double v[9] = {x.a.x, x.a.y, x.a.z, x.b.x, x.b.y, x.b.z, x.c.x, x.c.y, x.c.z};
            std::memcpy(out + 9 * t, v, sizeof(v));
        }
    });
}

// One D3Q19 brick covering the lattice (tests/test_main.cpp:225-266 uses 12x10x8, tau 0.71):
// Taylor-Green start, `steps` virtual steps, CFDCHK3 file at `path`.
REF_API int ref_dist_tg_checkpoint(std::size_t nx, std::size_t ny, std::size_t nz, float tau, float amplitude,
                                   std::size_t steps, const char* path) {
    return guarded([&] {
        const cfd::distributed::Extent3 global{nx, ny, nz};
        const cfd::lbm::DistributedLbmConfig config{global, tau};
        std::vector<cfd::lbm::DistributedPullBlock<cfd::lbm::D3Q19Descriptor>> blocks;
        blocks.emplace_back(config, cfd::distributed::make_brick(global, cfd::distributed::ProcessGrid{1, 1, 1}, 0));
        blocks[0].initialize_taylor_green(amplitude);
        for (std::size_t i = 0; i < steps; ++i) cfd::lbm::virtual_step(blocks);
        blocks[0].save_checkpoint(path);
    });
}
// Restores `path_in` into a fresh brick, runs `steps` virtual steps and saves to `path_out`.
REF_API int ref_dist_restart(std::size_t nx, std::size_t ny, std::size_t nz, float tau, const char* path_in,
                             std::size_t steps, const char* path_out) {
    return guarded([&] {
        const cfd::distributed::Extent3 global{nx, ny, nz};
        const cfd::lbm::DistributedLbmConfig config{global, tau};
        std::vector<cfd::lbm::DistributedPullBlock<cfd::lbm::D3Q19Descriptor>> blocks;
        blocks.emplace_back(config, cfd::distributed::make_brick(global, cfd::distributed::ProcessGrid{1, 1, 1}, 0));
        blocks[0].load_checkpoint(path_in);
        for (std::size_t i = 0; i < steps; ++i) cfd::lbm::virtual_step(blocks);
        blocks[0].save_checkpoint(path_out);
    });
}
