// lbm_oracle.cpp -- CPU restatement of the reference's esoteric-pull LBM path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP kernels in
// cfd_solvers_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.  The product path never links or calls it.
//
// What it restates (reference paths relative to the xsession/cfd_solvers root):
//   * velocity sets and weights ........ include/cfd/solvers/lbm/descriptors.hpp:31-146
//   * periodic_shift / neighbor_index .. include/cfd/solvers/lbm/esoteric_pull.hpp:40-46,304-310
//   * equilibrium, Guo forcing ......... esoteric_pull.hpp:48-70
//   * initialize (even layout) ......... esoteric_pull.hpp:117-137,312-333
//   * load_cell / store_cell ........... esoteric_pull.hpp:349-382
//   * accumulate_macroscopic, collide .. esoteric_pull.hpp:384-432
//   * periodic step_once ............... esoteric_pull.hpp:494-514
//   * bounded step_once_with_solids .... esoteric_pull.hpp:434-492, with the odd-parity repack at
//                                        :483 corrected (read solid_next_[(i+1), n], not [.., forward])
//   * moments / mass / logical pops .... esoteric_pull.hpp:180-267
//   * moving-wall correction ........... src/lbm/d2q9.cpp:217-223 (+6 w rho_w c.u_w, rho_w = max(rho,1e-8))
//   * local body acceleration .......... esoteric_pull.hpp:144-159,414-420
//   * STL ray-parity voxelisation ...... src/lbm/voxelize.cpp:4-5 (every cell against every
//                                        triangle, double), transform_surface voxelize.cpp:4,6
//   * resident coupling modules ........ include/cfd/solvers/lbm/resident_multiphysics_sycl.hpp:
//                                        D3Q7 thermal + Boussinesq :54-198, volume of fluid :200-324,
//                                        particles :326-380, Q-criterion :382-408 (float, periodic)
//
// Extensions the reference does not have (spec restated here, pinned by closed forms in tests/):
//   * seeded noise on the Taylor-Green initial velocity (SplitMix64 counter hash),
//   * 3-D velocity inlet / pressure outlet with Zou-He-style densities and equilibrium populations,
//   * partial bounce-back porosity gamma (Walsh, Burwinkle & Saar 2009 form),
//   * FP16 shifted-population storage h = half_rne((f - w_d) * 4096),
//   * a switch (off by default) for FP64 arithmetic with the weights' exact values.
//
// Arithmetic follows the reference operation by operation (same association order) and the
// library is built with -ffp-contract=off so that it is reproducible bit for bit.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- velocity sets
// Same direction order and float weights as descriptors.hpp (pairs (i, i+1) are opposite).
template <int Q> struct Lat;
template <> struct Lat<9> {
    static constexpr int dims = 2;
    static constexpr int cx[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
    static constexpr int cy[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr float w[9] = {4.0F / 9.0F,  1.0F / 9.0F,  1.0F / 9.0F,  1.0F / 9.0F, 1.0F / 9.0F,
                                   1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F};
    static constexpr double w_exact[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};
template <> struct Lat<19> {
    static constexpr int dims = 3;
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    static constexpr float w[19] = {1.0F / 3.0F,  1.0F / 18.0F, 1.0F / 18.0F, 1.0F / 18.0F, 1.0F / 18.0F,
                                    1.0F / 18.0F, 1.0F / 18.0F, 1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F,
                                    1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F,
                                    1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F};
    static constexpr double w_exact[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                           1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                           1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                           1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};
template <> struct Lat<27> {
    static constexpr int dims = 3;
    static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                                   1, -1, 1, -1, 1, -1, 1, -1};
    static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1,
                                   1, -1, 1, -1, -1, 1, -1, 1};
    static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1,
                                   1, -1, -1, 1, 1, -1, -1, 1};
    static constexpr float w[27] = {
        8.0F / 27.0F,  2.0F / 27.0F,  2.0F / 27.0F,  2.0F / 27.0F,  2.0F / 27.0F,  2.0F / 27.0F,  2.0F / 27.0F,
        1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,
        1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 54.0F,  1.0F / 216.0F, 1.0F / 216.0F,
        1.0F / 216.0F, 1.0F / 216.0F, 1.0F / 216.0F, 1.0F / 216.0F, 1.0F / 216.0F, 1.0F / 216.0F};
    static constexpr double w_exact[27] = {
        8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
        1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
        1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 216.0, 1.0 / 216.0,
        1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
};

// FP64 arithmetic with the weights' exact values instead of the reference's FP32 constants, which
// sum to 1 + 1.49e-8 (D3Q19) and 1 + 7.45e-9 (D2Q9, D3Q27): with those, every BGK collision of an
// FP64 run adds omega times the excess to the cell's density.  No reference counterpart: a switch
// for precision studies, off by default and process-wide (orc_set_exact_fp64_weights).
inline bool g_exact_fp64_weights = false;

inline int opposite(int i) { return i == 0 ? 0 : ((i & 1) != 0 ? i + 1 : i - 1); }

// ---------------------------------------------------------------- flags
// uint8 cell flags shared with the HIP path.
constexpr uint8_t kFluid = 0x00;
constexpr uint8_t kInlet = 0x01;
constexpr uint8_t kOutlet = 0x02;
constexpr uint8_t kWall = 0x80;  // bit 7: halfway bounce-back wall; bits 0-2: wall-velocity index
inline bool is_wall(uint8_t f) { return (f & kWall) != 0; }

// ---------------------------------------------------------------- FP16 codec (IEEE binary16, RNE)
inline uint16_t f32_to_f16_rne(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000U;
    const uint32_t ax = x & 0x7fffffffU;
    if (ax > 0x7f800000U) return static_cast<uint16_t>(sign | 0x7e00U);
    if (ax >= 0x477ff000U) return static_cast<uint16_t>(sign | 0x7c00U);  // >= 65520 rounds to inf
    if (ax < 0x38800000U) {                                                  // below 2^-14: subnormal
        if (ax <= 0x33000000U) return static_cast<uint16_t>(sign);           // <= 2^-25 rounds to 0
        const int e = static_cast<int>(ax >> 23);
        const uint32_t m = (ax & 0x7fffffU) | 0x800000U;
        const int shift = 126 - e;
        uint32_t q = m >> shift;
        const uint32_t rem = m & ((1U << shift) - 1U);
        const uint32_t half = 1U << (shift - 1);
        if (rem > half || (rem == half && (q & 1U))) ++q;
        return static_cast<uint16_t>(sign | q);
    }
    const uint32_t e16 = (ax >> 23) - 127U + 15U;
    const uint32_t m = ax & 0x7fffffU;
    uint32_t q = (e16 << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffU;
    if (rem > 0x1000U || (rem == 0x1000U && (q & 1U))) ++q;
    return static_cast<uint16_t>(sign | q);
}

inline float f16_to_f32(uint16_t h) {
    const uint32_t sign = (static_cast<uint32_t>(h) & 0x8000U) << 16;
    const uint32_t e = (h >> 10) & 0x1fU;
    const uint32_t m = h & 0x3ffU;
    if (e == 0U) {
        const float v = static_cast<float>(m) * 0x1p-24F;
        return sign ? -v : v;
    }
    uint32_t bits;
    if (e == 31U) bits = sign | 0x7f800000U | (m << 13);
    else bits = sign | ((e - 15U + 127U) << 23) | (m << 13);
    float out;
    std::memcpy(&out, &bits, 4);
    return out;
}

constexpr float kHalfScale = 4096.0F;          // fixed per-lattice scale (power of two: exact)
constexpr float kHalfInvScale = 1.0F / 4096.0F;

struct Half {
    uint16_t bits;
};

template <class Store, class Real, int Q> struct Codec;
template <class Real, int Q> struct Codec<Real, Real, Q> {
    static Real load(Real v, int) { return v; }
    static Real store(Real v, int) { return v; }
};
// FP32 storage of the deviation f - w_d (shifted, like FP16S, at 4 B per population).
struct F32Dev {
    float v;
};
template <int Q> struct Codec<F32Dev, float, Q> {
    static float load(F32Dev h, int) { return h.v; }
    static F32Dev store(float v, int) { return F32Dev{v}; }
};
// FP16 shifted storage holds binary16((f - w_d) * 4096); its state is the deviation f - w_d and
// the collision runs on deviations (Shifted below), so the codec only scales.
template <int Q> struct Codec<Half, float, Q> {
    static float load(Half h, int) { return f16_to_f32(h.bits) * kHalfInvScale; }
    static Half store(float v, int) { return Half{f32_to_f16_rne(v * kHalfScale)}; }
};

// ---------------------------------------------------------------- seeded noise (SplitMix64)
// Steele, Lea & Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014.
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// Uniform in [0,1) on a 2^-24 grid: exact in float and double.
template <class Real> inline Real uniform01(uint64_t seed, uint64_t cell, uint64_t component) {
    const uint64_t h = splitmix64(splitmix64(seed) + 4ULL * cell + component);
    return static_cast<Real>(h >> 40) * static_cast<Real>(0x1p-24);
}

// ---------------------------------------------------------------- oracle
class OracleBase {
public:
    virtual ~OracleBase() = default;
    virtual void set_flags(const uint8_t* flags) = 0;
    virtual void set_wall_velocity(int index, double ux, double uy, double uz) = 0;
    virtual void set_gamma(const double* gamma) = 0;
    virtual void set_local_acceleration(const double* ax, const double* ay, const double* az) = 0;
    virtual void set_inlet_outlet(double u_peak, double rho_out) = 0;
    virtual void init_fields(const double* rho, const double* ux, const double* uy, const double* uz) = 0;
    virtual void init_uniform(double rho, double ux, double uy, double uz) = 0;
    virtual void init_taylor_green(double amplitude, double noise, uint64_t seed) = 0;
    virtual void step(uint64_t count) = 0;
    virtual void macroscopic(double* rho, double* ux, double* uy, double* uz) const = 0;
    virtual double mass() const = 0;
    virtual double kinetic_energy() const = 0;
    virtual double max_speed() const = 0;
    virtual void logical(double* out) const = 0;
    virtual void set_logical(const double* in) = 0;
    virtual void raw(double* out) const = 0;
    virtual void load_map(int parity, int64_t* out) const = 0;
    virtual void store_map(int parity, int64_t* out) const = 0;
    virtual void bb_load_map(int parity, int64_t* out) const = 0;
    virtual uint64_t time_step() const = 0;
    virtual void set_time_step(uint64_t t) = 0;
    virtual size_t cells() const = 0;
    virtual double objective(int kind, int x_in, int x_out) const = 0;
    virtual double adjoint_run(int kind, uint64_t steps, uint64_t window, int x_in, int x_out, double* grad_gamma,
                               double* grad_u_peak) = 0;
    virtual void set_drag_body(const uint8_t* body, int axis) = 0;
    virtual void set_smagorinsky(bool enabled) = 0;
};

template <int Q, class Real, class Store> class Oracle final : public OracleBase {
    using L = Lat<Q>;
    using C = Codec<Store, Real, Q>;
    // State = deviation f - w_d from the rest equilibrium (FP16 shifted storage).  The collision
    // then works on deviations: rho = 1 + sum f', f_eq - w = w (rho' + rho (cu + cu^2/2 - uu)),
    // which keeps the non-equilibrium part at FP32 relative precision at low velocities.
    static constexpr bool kShifted = std::is_same<Store, Half>::value || std::is_same<Store, F32Dev>::value;
    static Real to_full(Real v, int d) { return kShifted ? v + w(d) : v; }
    static Real to_state(Real v, int d) { return kShifted ? v - w(d) : v; }

public:
    Oracle(size_t nx, size_t ny, size_t nz, double tau, double ax, double ay, double az)
        : nx_(nx), ny_(ny), nz_(L::dims == 2 ? 1 : nz), tau_(static_cast<Real>(tau)),
          ax_(static_cast<Real>(ax)), ay_(static_cast<Real>(ay)), az_(static_cast<Real>(az)) {
        if (nx_ < 2 || ny_ < 2) throw std::invalid_argument("LBM grid requires nx, ny >= 2");
        if (L::dims == 3 && nz_ < 2) throw std::invalid_argument("3-D LBM grid requires nz >= 2");
        if (!(tau_ > static_cast<Real>(0.5))) throw std::invalid_argument("LBM tau must be > 0.5");
        N_ = nx_ * ny_ * nz_;
        F_.resize(static_cast<size_t>(Q) * N_);
        for (auto& v : wall_u_) v[0] = v[1] = v[2] = Real(0);
        init_uniform(1.0, 0.0, 0.0, 0.0);
    }

    size_t cells() const override { return N_; }
    uint64_t time_step() const override { return t_; }
    void set_time_step(uint64_t t) override { t_ = t; }

    void set_flags(const uint8_t* flags) override {
        if (!flags) {
            flags_.clear();
            return;
        }
        flags_.assign(flags, flags + N_);
    }
    void set_wall_velocity(int index, double ux, double uy, double uz) override {
        if (index < 1 || index > 7) throw std::invalid_argument("wall velocity index must be 1..7");
        wall_u_[index][0] = static_cast<Real>(ux);
        wall_u_[index][1] = static_cast<Real>(uy);
        wall_u_[index][2] = static_cast<Real>(uz);
    }
    void set_gamma(const double* gamma) override {
        if (!gamma) {
            gamma_.clear();
            return;
        }
        gamma_.resize(N_);
        for (size_t n = 0; n < N_; ++n) gamma_[n] = static_cast<Real>(gamma[n]);
    }
    void set_local_acceleration(const double* ax, const double* ay, const double* az) override {
        if (!ax) {
            lax_.clear();
            lay_.clear();
            laz_.clear();
            return;
        }
        lax_.resize(N_);
        lay_.resize(N_);
        laz_.resize(N_);
        for (size_t n = 0; n < N_; ++n) {
            lax_[n] = static_cast<Real>(ax[n]);
            lay_[n] = static_cast<Real>(ay[n]);
            laz_[n] = static_cast<Real>(az[n]);
        }
    }
    void set_smagorinsky(bool enabled) override {
        if (enabled && kShifted) throw std::invalid_argument("Smagorinsky LES runs on FP32 or FP64 storage");
        les_ = enabled;
    }
    void set_inlet_outlet(double u_peak, double rho_out) override {
        u_peak_ = static_cast<Real>(u_peak);
        rho_out_ = static_cast<Real>(rho_out);
    }

    // esoteric_pull.hpp:312-333 -- equilibrium written in the even layout; solids at rest.
    template <class Init> void initialize(Init&& init) {
        t_ = 0;
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < static_cast<int64_t>(N_); ++nn) {
            const size_t n = static_cast<size_t>(nn);
            size_t x, y, z;
            coords(n, x, y, z);
            Real s[4];
            init(n, x, y, z, s);
            const bool solid = !flags_.empty() && is_wall(flags_[n]);
            const Real ux = solid ? Real(0) : s[1];
            const Real uy = solid ? Real(0) : s[2];
            const Real uz = solid ? Real(0) : s[3];
            const Real rp = s[0] - Real(1);
            at(0, n) = C::store(feq_state(0, rp, s[0], ux, uy, uz), 0);
            for (int i = 1; i < Q; i += 2) {
                const size_t nb = neighbor(x, y, z, i + 1);
                at(i, nb) = C::store(feq_state(i, rp, s[0], ux, uy, uz), i);
                at(i + 1, n) = C::store(feq_state(i + 1, rp, s[0], ux, uy, uz), i + 1);
            }
        }
    }

    void init_fields(const double* rho, const double* ux, const double* uy, const double* uz) override {
        initialize([&](size_t n, size_t, size_t, size_t, Real* s) {
            s[0] = static_cast<Real>(rho[n]);
            s[1] = static_cast<Real>(ux[n]);
            s[2] = static_cast<Real>(uy[n]);
            s[3] = static_cast<Real>(uz[n]);
        });
    }
    void init_uniform(double rho, double ux, double uy, double uz) override {
        const Real r = static_cast<Real>(rho), a = static_cast<Real>(ux), b = static_cast<Real>(uy),
                   c = static_cast<Real>(uz);
        initialize([=](size_t, size_t, size_t, size_t, Real* s) {
            s[0] = r;
            s[1] = a;
            s[2] = b;
            s[3] = c;
        });
    }
    // esoteric_pull.hpp:121-137, plus optional seeded uniform noise of amplitude `noise` per component.
    void init_taylor_green(double amplitude_d, double noise_d, uint64_t seed) override {
        const Real amplitude = static_cast<Real>(amplitude_d);
        const Real noise = static_cast<Real>(noise_d);
        const Real two_pi = Real(2) * static_cast<Real>(3.14159265358979323846);
        initialize([=](size_t n, size_t x, size_t y, size_t z, Real* s) {
            const Real xf = (static_cast<Real>(x) + Real(0.5)) / static_cast<Real>(nx_);
            const Real yf = (static_cast<Real>(y) + Real(0.5)) / static_cast<Real>(ny_);
            const Real sx = std::sin(two_pi * xf);
            const Real cx = std::cos(two_pi * xf);
            const Real sy = std::sin(two_pi * yf);
            const Real cy = std::cos(two_pi * yf);
            Real z_factor = Real(1);
            if (L::dims == 3) {
                const Real zf = (static_cast<Real>(z) + Real(0.5)) / static_cast<Real>(nz_);
                z_factor = std::cos(two_pi * zf);
            }
            s[0] = Real(1);
            s[1] = amplitude * sx * cy * z_factor;
            s[2] = -amplitude * cx * sy * z_factor;
            s[3] = Real(0);
            if (noise != Real(0)) {
                s[1] = s[1] + noise * (Real(2) * uniform01<Real>(seed, n, 0) - Real(1));
                s[2] = s[2] + noise * (Real(2) * uniform01<Real>(seed, n, 1) - Real(1));
                if (L::dims == 3) s[3] = s[3] + noise * (Real(2) * uniform01<Real>(seed, n, 2) - Real(1));
            }
        });
    }

    void step(uint64_t count) override {
        for (uint64_t k = 0; k < count; ++k) {
            if (flags_.empty()) step_periodic();
            else step_bounded();
        }
    }

    void macroscopic(double* rho_out, double* ux_out, double* uy_out, double* uz_out) const override {
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < static_cast<int64_t>(N_); ++nn) {
            const size_t n = static_cast<size_t>(nn);
            Real m[4];
            cell_moments(n, m);
            rho_out[n] = static_cast<double>(m[0]);
            ux_out[n] = static_cast<double>(m[1]);
            uy_out[n] = static_cast<double>(m[2]);
            uz_out[n] = static_cast<double>(m[3]);
        }
    }

    // esoteric_pull.hpp:211-220: per-cell sum of loaded populations, accumulated in double.
    double mass() const override {
        double total = 0.0;
        for (size_t n = 0; n < N_; ++n) {
            Real fin[Q];
            load(n, fin);
            Real rho = Real(0);
            for (int d = 0; d < Q; ++d) rho += fin[d];
            total += kShifted ? 1.0 + static_cast<double>(rho) : static_cast<double>(rho);
        }
        return total;
    }

    // 0.5 * sum rho |u|^2 over non-wall cells (wall cells excluded as in src/lbm/d2q9.cpp mass/energy).
    double kinetic_energy() const override {
        double total = 0.0;
        for (size_t n = 0; n < N_; ++n) {
            if (!flags_.empty() && is_wall(flags_[n])) continue;
            Real m[4];
            cell_moments(n, m);
            total += static_cast<double>(m[0]) * (static_cast<double>(m[1]) * m[1] +
                                                  static_cast<double>(m[2]) * m[2] + static_cast<double>(m[3]) * m[3]);
        }
        return 0.5 * total;
    }

    double max_speed() const override {
        double best = 0.0;
        for (size_t n = 0; n < N_; ++n) {
            if (!flags_.empty() && is_wall(flags_[n])) continue;
            Real m[4];
            cell_moments(n, m);
            const Real s = std::sqrt(m[1] * m[1] + m[2] * m[2] + m[3] * m[3]);
            best = std::max(best, static_cast<double>(s));
        }
        return best;
    }

    void logical(double* out) const override {
        for (size_t n = 0; n < N_; ++n) {
            Real fin[Q];
            load(n, fin);
            for (int d = 0; d < Q; ++d) out[static_cast<size_t>(d) * N_ + n] = static_cast<double>(to_full(fin[d], d));
        }
    }
    // Writes logical populations (the inverse of load) at the current parity.
    void set_logical(const double* in) override {
        const bool odd = (t_ & 1U) != 0U;
        for (size_t n = 0; n < N_; ++n) {
            size_t x, y, z;
            coords(n, x, y, z);
            at(0, n) = C::store(to_state(static_cast<Real>(in[n]), 0), 0);
            for (int i = 1; i < Q; i += 2) {
                const size_t nb = neighbor(x, y, z, i + 1);
                const Real a = to_state(static_cast<Real>(in[static_cast<size_t>(i) * N_ + n]), i);
                const Real b = to_state(static_cast<Real>(in[static_cast<size_t>(i + 1) * N_ + n]), i + 1);
                if (odd) {
                    at(i + 1, nb) = C::store(a, i + 1);
                    at(i, n) = C::store(b, i);
                } else {
                    at(i, nb) = C::store(a, i);
                    at(i + 1, n) = C::store(b, i + 1);
                }
            }
        }
    }
    void raw(double* out) const override {
        for (size_t k = 0; k < F_.size(); ++k) {
            const int d = static_cast<int>(k / N_);
            out[k] = static_cast<double>(to_full(C::load(F_[k], d), d));
        }
    }

    // Slot (d' * N + n') from which fin[d] of cell n is read at the given parity (esoteric_pull.hpp:349-365).
    void load_map(int parity, int64_t* out) const override {
        for (size_t n = 0; n < N_; ++n) {
            size_t x, y, z;
            coords(n, x, y, z);
            out[n] = slot(0, n);
            for (int i = 1; i < Q; i += 2) {
                const size_t nb = neighbor(x, y, z, i + 1);
                if (parity & 1) {
                    out[static_cast<size_t>(i) * N_ + n] = slot(i + 1, nb);
                    out[static_cast<size_t>(i + 1) * N_ + n] = slot(i, n);
                } else {
                    out[static_cast<size_t>(i) * N_ + n] = slot(i, nb);
                    out[static_cast<size_t>(i + 1) * N_ + n] = slot(i + 1, n);
                }
            }
        }
    }
    // Slot into which fout[d] of cell n is written at the given parity (esoteric_pull.hpp:367-382).
    void store_map(int parity, int64_t* out) const override {
        for (size_t n = 0; n < N_; ++n) {
            size_t x, y, z;
            coords(n, x, y, z);
            out[n] = slot(0, n);
            for (int i = 1; i < Q; i += 2) {
                const size_t nb = neighbor(x, y, z, i + 1);
                if (parity & 1) {
                    out[static_cast<size_t>(i) * N_ + n] = slot(i, n);
                    out[static_cast<size_t>(i + 1) * N_ + n] = slot(i + 1, nb);
                } else {
                    out[static_cast<size_t>(i) * N_ + n] = slot(i + 1, n);
                    out[static_cast<size_t>(i + 1) * N_ + n] = slot(i, nb);
                }
            }
        }
    }
    // Physical slot read by the single-pass in-place bounded kernel: a population whose upstream
    // cell is a wall is read from the slot of the opposite parity, which holds the value this cell
    // stored for the opposite direction one step earlier (halfway bounce-back).  -1 for wall cells.
    void bb_load_map(int parity, int64_t* out) const override {
        std::vector<int64_t> normal(static_cast<size_t>(Q) * N_), other(static_cast<size_t>(Q) * N_);
        load_map(parity, normal.data());
        load_map(parity ^ 1, other.data());
        for (size_t n = 0; n < N_; ++n) {
            size_t x, y, z;
            coords(n, x, y, z);
            const bool wall = !flags_.empty() && is_wall(flags_[n]);
            for (int d = 0; d < Q; ++d) {
                const size_t k = static_cast<size_t>(d) * N_ + n;
                if (wall) {
                    out[k] = -1;
                    continue;
                }
                const size_t up = neighbor(x, y, z, opposite(d));
                const bool bb = d != 0 && !flags_.empty() && is_wall(flags_[up]);
                out[k] = bb ? other[k] : normal[k];
            }
        }
    }

private:
    size_t nx_, ny_, nz_, N_{};
    Real tau_, ax_, ay_, az_;
    std::vector<Store> F_;
    std::vector<uint8_t> flags_;
    std::vector<uint8_t> body_;  // drag body (wall cells), empty: none
    std::vector<Real> scratch_cur_, scratch_nxt_;  // bounded-step scratch lattices
    bool les_{false};                              // Smagorinsky relaxation time
    int drag_axis_{0};
    std::vector<Real> gamma_, lax_, lay_, laz_;
    Real wall_u_[8][3];
    Real u_peak_{0}, rho_out_{1};
    uint64_t t_{0};

    Store& at(int d, size_t n) { return F_[static_cast<size_t>(d) * N_ + n]; }
    const Store& at(int d, size_t n) const { return F_[static_cast<size_t>(d) * N_ + n]; }
    int64_t slot(int d, size_t n) const { return static_cast<int64_t>(static_cast<size_t>(d) * N_ + n); }

    void coords(size_t n, size_t& x, size_t& y, size_t& z) const {
        x = n % nx_;
        const size_t yz = n / nx_;
        y = yz % ny_;
        z = yz / ny_;
    }
    static size_t shift(size_t c, int delta, size_t extent) {
        if (delta > 0) return c + 1U == extent ? 0U : c + 1U;
        if (delta < 0) return c == 0U ? extent - 1U : c - 1U;
        return c;
    }
    size_t neighbor(size_t x, size_t y, size_t z, int d) const {
        return (shift(z, L::cz[d], nz_) * ny_ + shift(y, L::cy[d], ny_)) * nx_ + shift(x, L::cx[d], nx_);
    }

    static Real w(int d) {
        if constexpr (std::is_same_v<Real, double>) {
            if (g_exact_fp64_weights) return L::w_exact[d];
        }
        return static_cast<Real>(L::w[d]);
    }
    static Real feq(int d, Real rho, Real ux, Real uy, Real uz) {
        const Real cu = Real(3) * (static_cast<Real>(L::cx[d]) * ux + static_cast<Real>(L::cy[d]) * uy +
                                   static_cast<Real>(L::cz[d]) * uz);
        const Real uu = Real(1.5) * (ux * ux + uy * uy + uz * uz);
        return w(d) * rho * (Real(1) + cu + Real(0.5) * cu * cu - uu);
    }
    // Shifted arithmetic (the GPU's FP16S/FP32S kernels, operation for operation; no reference
    // counterpart): uu = 3/2 fma(uz, uz, fma(uy, uy, ux ux)), cu = 3 c.u and
    //   f_eq - w_d = w_d fma(rho, fma(cu/2, cu, -uu) + cu, rho'),   rho' = rho - 1.
    static Real shifted_uu(Real ux, Real uy, Real uz) {
        return Real(1.5) * std::fma(uz, uz, std::fma(uy, uy, ux * ux));
    }
    static Real feq_dev(int d, Real rp, Real rho, Real ux, Real uy, Real uz) {
        const Real cu = Real(3) * (static_cast<Real>(L::cx[d]) * ux + static_cast<Real>(L::cy[d]) * uy +
                                   static_cast<Real>(L::cz[d]) * uz);
        const Real uu = shifted_uu(ux, uy, uz);
        return w(d) * std::fma(rho, std::fma(Real(0.5) * cu, cu, -uu) + cu, rp);
    }
    static Real feq_state(int d, Real rp, Real rho, Real ux, Real uy, Real uz) {
        return kShifted ? feq_dev(d, rp, rho, ux, uy, uz) : feq(d, rho, ux, uy, uz);
    }
    // esoteric_pull.hpp:72-97 (Smagorinsky-Lilly closure on the non-equilibrium momentum flux):
    // Pi_ab = sum_d c_da c_db (f_d - f_eq_d), Q = sum_ab Pi_ab^2 (row-major), and
    // tau_local = (tau + sqrt(tau^2 + C sqrt(Q) / rho)) / 2 with C = 16 sqrt(2) / (3 pi^2) in float.
    Real smagorinsky_tau(const Real* fin, Real rho, Real ux, Real uy, Real uz) const {
        if (!(rho > Real(0))) return tau_;
        Real pi[3][3] = {};
        for (int d = 0; d < Q; ++d) {
            const Real neq = fin[d] - feq(d, rho, ux, uy, uz);
            const Real e[3] = {static_cast<Real>(L::cx[d]), static_cast<Real>(L::cy[d]), static_cast<Real>(L::cz[d])};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) pi[a][b] += e[a] * e[b] * neq;
        }
        Real q = Real(0);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) q += pi[a][b] * pi[a][b];
        const float pi_f = 3.14159265358979323846F;
        const Real c = static_cast<Real>(16.0F * 1.41421356237309504880F / (3.0F * pi_f * pi_f));
        return Real(0.5) * (tau_ + std::sqrt(tau_ * tau_ + c * std::sqrt(q) / rho));
    }
    static Real guo(int d, Real rho, Real ux, Real uy, Real uz, Real ax, Real ay, Real az, Real omega) {
        if (ax == Real(0) && ay == Real(0) && az == Real(0)) return Real(0);
        const Real cx = static_cast<Real>(L::cx[d]), cy = static_cast<Real>(L::cy[d]), cz = static_cast<Real>(L::cz[d]);
        const Real fx = rho * ax, fy = rho * ay, fz = rho * az;
        const Real cu = cx * ux + cy * uy + cz * uz;
        const Real cf = cx * fx + cy * fy + cz * fz;
        const Real uf = ux * fx + uy * fy + uz * fz;
        return w(d) * (Real(1) - Real(0.5) * omega) * (Real(3) * (cf - uf) + Real(9) * cu * cf);
    }

    void load(size_t n, Real* fin) const {
        size_t x, y, z;
        coords(n, x, y, z);
        const bool odd = (t_ & 1U) != 0U;
        fin[0] = C::load(at(0, n), 0);
        for (int i = 1; i < Q; i += 2) {
            const size_t nb = neighbor(x, y, z, i + 1);
            if (odd) {
                fin[i] = C::load(at(i + 1, nb), i + 1);
                fin[i + 1] = C::load(at(i, n), i);
            } else {
                fin[i] = C::load(at(i, nb), i);
                fin[i + 1] = C::load(at(i + 1, n), i + 1);
            }
        }
    }
    void store(size_t n, size_t x, size_t y, size_t z, bool odd, const Real* fout) {
        at(0, n) = C::store(fout[0], 0);
        for (int i = 1; i < Q; i += 2) {
            const size_t nb = neighbor(x, y, z, i + 1);
            if (odd) {
                at(i, n) = C::store(fout[i], i);
                at(i + 1, nb) = C::store(fout[i + 1], i + 1);
            } else {
                at(i + 1, n) = C::store(fout[i], i + 1);
                at(i, nb) = C::store(fout[i + 1], i);
            }
        }
    }

    static void accumulate(const Real* fin, Real& rho, Real& ux, Real& uy, Real& uz) {
        Real rp;
        accumulate(fin, rp, rho, ux, uy, uz);
    }
    // rho' = rho - 1 as well; shifted states sum to rho'.
    static void accumulate(const Real* fin, Real& rp, Real& rho, Real& ux, Real& uy, Real& uz) {
        Real s = Real(0);
        ux = uy = uz = Real(0);
        for (int d = 0; d < Q; ++d) {
            const Real v = fin[d];
            s += v;
            ux += v * static_cast<Real>(L::cx[d]);
            uy += v * static_cast<Real>(L::cy[d]);
            uz += v * static_cast<Real>(L::cz[d]);
        }
        if (kShifted) {
            rp = s;
            rho = Real(1) + s;
            if (rho > Real(0)) {  // one division and a residual correction: u0 = m c, u = fma(fma(-u0, rho, m), c, u0)
                const Real c = Real(1) / rho;
                const auto q = [&](Real m) {
                    const Real u0 = m * c;
                    return std::fma(std::fma(-u0, rho, m), c, u0);
                };
                ux = q(ux);
                uy = q(uy);
                uz = q(uz);
            } else {
                ux = uy = uz = Real(0);
            }
            return;
        }
        rho = s;
        rp = s - Real(1);
        if (rho > Real(0)) {
            ux /= rho;
            uy /= rho;
            uz /= rho;
        } else {
            ux = uy = uz = Real(0);
        }
    }

    void acceleration(size_t n, Real& ax, Real& ay, Real& az) const {
        const bool local = !lax_.empty();
        ax = ax_ + (local ? lax_[n] : Real(0));
        ay = ay_ + (local ? lay_[n] : Real(0));
        az = az_ + (local ? laz_[n] : Real(0));
    }

    Real inlet_velocity(size_t y, size_t z) const {
        const Real eta = (static_cast<Real>(y) - Real(0.5)) / static_cast<Real>(ny_ - 2);
        Real u = u_peak_ * (Real(4) * eta * (Real(1) - eta));
        if (L::dims == 3) {
            const Real zeta = (static_cast<Real>(z) - Real(0.5)) / static_cast<Real>(nz_ - 2);
            u = u * (Real(4) * zeta * (Real(1) - zeta));
        }
        return u;
    }
    // Zou-He density of a velocity-inlet cell at x = 0 from the populations that come from the interior.
    // Shifted: the weights' share of S0 + 2 S- is 1, so rho' = (S0' + 2 S-' + u) / (1 - u).
    Real inlet_density(const Real* fin, Real uin) const {
        Real rp;
        return inlet_density(fin, uin, rp);
    }
    Real inlet_density(const Real* fin, Real uin, Real& rp) const {
        Real s0 = Real(0), sn = Real(0);
        for (int d = 0; d < Q; ++d) {
            if (L::cx[d] == 0) s0 += fin[d];
            else if (L::cx[d] < 0) sn += fin[d];
        }
        const Real s = s0 + Real(2) * sn;
        if (kShifted) {
            rp = (s + uin) / (Real(1) - uin);
            return Real(1) + rp;
        }
        const Real rho = s / (Real(1) - uin);
        rp = rho - Real(1);
        return rho;
    }
    // Zou-He normal velocity of a pressure-outlet cell at x = nx-1.
    Real outlet_velocity(const Real* fin) const {
        Real s0 = Real(0), sp = Real(0);
        for (int d = 0; d < Q; ++d) {
            if (L::cx[d] == 0) s0 += fin[d];
            else if (L::cx[d] > 0) sp += fin[d];
        }
        if (kShifted) return ((Real(1) - rho_out_) + (s0 + Real(2) * sp)) / rho_out_;
        return Real(-1) + (s0 + Real(2) * sp) / rho_out_;
    }

    // Collision of one non-wall cell; returns the density used for moving-wall corrections.
    Real collide(size_t n, size_t y, size_t z, const Real* fin, Real* fout) const {
        const uint8_t flag = flags_.empty() ? kFluid : flags_[n];
        if (flag == kInlet) {
            const Real uin = inlet_velocity(y, z);
            Real rp;
            const Real rho = inlet_density(fin, uin, rp);
            for (int d = 0; d < Q; ++d) fout[d] = feq_state(d, rp, rho, uin, Real(0), Real(0));
            return rho;
        }
        if (flag == kOutlet) {
            const Real ux = outlet_velocity(fin);
            const Real rp = rho_out_ - Real(1);
            for (int d = 0; d < Q; ++d) fout[d] = feq_state(d, rp, rho_out_, ux, Real(0), Real(0));
            return rho_out_;
        }
        Real rp, rho, ux, uy, uz, ax, ay, az;
        accumulate(fin, rp, rho, ux, uy, uz);
        acceleration(n, ax, ay, az);
        ux += Real(0.5) * ax;
        uy += Real(0.5) * ay;
        uz += Real(0.5) * az;
        const Real omega = Real(1) / (les_ ? smagorinsky_tau(fin, rho, ux, uy, uz) : tau_);
        if (kShifted) {
            // Relaxation fma(-omega, f' - f_eq', f'); opposite directions share fma(cu/2, cu, -uu).
            const Real uu = shifted_uu(ux, uy, uz);
            // f' - f_eq' = fma(-w_d, y_d, f') with y_d = fma(rho, x_d, rho'), one rounding.
            fout[0] = std::fma(-omega, std::fma(-w(0), std::fma(rho, -uu, rp), fin[0]), fin[0]);
            for (int i = 1; i < Q; i += 2) {
                const Real cu = Real(3) * (static_cast<Real>(L::cx[i]) * ux + static_cast<Real>(L::cy[i]) * uy +
                                           static_cast<Real>(L::cz[i]) * uz);
                const Real common = std::fma(Real(0.5) * cu, cu, -uu);
                const Real da = std::fma(-w(i), std::fma(rho, common + cu, rp), fin[i]);
                const Real db = std::fma(-w(i), std::fma(rho, common - cu, rp), fin[i + 1]);
                fout[i] = std::fma(-omega, da, fin[i]);
                fout[i + 1] = std::fma(-omega, db, fin[i + 1]);
            }
            for (int d = 0; d < Q; ++d) fout[d] += guo(d, rho, ux, uy, uz, ax, ay, az, omega);
        } else {
            for (int d = 0; d < Q; ++d) {
                Real p = fin[d];
                p -= omega * (p - feq(d, rho, ux, uy, uz));
                p += guo(d, rho, ux, uy, uz, ax, ay, az, omega);
                fout[d] = p;
            }
        }
        if (!gamma_.empty()) {
            const Real g = gamma_[n];
            Real mixed[Q];
            for (int d = 0; d < Q; ++d)
                mixed[d] = kShifted ? std::fma(Real(1) - g, fout[d], g * fin[opposite(d)])
                                    : (Real(1) - g) * fout[d] + g * fin[opposite(d)];
            for (int d = 0; d < Q; ++d) fout[d] = mixed[d];
        }
        return rho;
    }

    // esoteric_pull.hpp:494-514
    void step_periodic() {
        const bool odd = (t_ & 1U) != 0U;
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < static_cast<int64_t>(N_); ++nn) {
            const size_t n = static_cast<size_t>(nn);
            size_t x, y, z;
            coords(n, x, y, z);
            Real fin[Q], fout[Q];
            load(n, fin);
            collide(n, y, z, fin, fout);
            store(n, x, y, z, odd, fout);
        }
        ++t_;
    }

    // esoteric_pull.hpp:434-492 with the :483 repack corrected; moving walls per src/lbm/d2q9.cpp:217-223.
    void step_bounded() {
        const size_t total = static_cast<size_t>(Q) * N_;
        // The reference's two scratch lattices, kept between steps (every entry is rewritten).
        if (scratch_cur_.size() != total) {
            scratch_cur_.assign(total, Real(0));
            scratch_nxt_.assign(total, Real(0));
        }
        std::vector<Real>& cur = scratch_cur_;
        std::vector<Real>& nxt = scratch_nxt_;
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < static_cast<int64_t>(N_); ++nn) {
            const size_t n = static_cast<size_t>(nn);
            Real fin[Q];
            load(n, fin);
            for (int d = 0; d < Q; ++d) cur[static_cast<size_t>(d) * N_ + n] = fin[d];
        }
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < static_cast<int64_t>(N_); ++nn) {
            const size_t n = static_cast<size_t>(nn);
            if (is_wall(flags_[n])) {
                for (int d = 0; d < Q; ++d) nxt[static_cast<size_t>(d) * N_ + n] = cur[static_cast<size_t>(d) * N_ + n];
                continue;
            }
            size_t x, y, z;
            coords(n, x, y, z);
            Real fin[Q], fout[Q];
            for (int d = 0; d < Q; ++d) fin[d] = cur[static_cast<size_t>(d) * N_ + n];
            const Real rho = collide(n, y, z, fin, fout);
            const Real rho_wall = std::max(rho, static_cast<Real>(1.0e-8));
            for (int d = 0; d < Q; ++d) {
                const size_t target = neighbor(x, y, z, d);
                const uint8_t tf = flags_[target];
                if (is_wall(tf)) {
                    Real value = fout[d];
                    const int k = tf & 7;
                    if (k != 0) {
                        const Real cu = static_cast<Real>(L::cx[d]) * wall_u_[k][0] +
                                        static_cast<Real>(L::cy[d]) * wall_u_[k][1] +
                                        static_cast<Real>(L::cz[d]) * wall_u_[k][2];
                        value -= Real(6) * w(d) * rho_wall * cu;
                    }
                    nxt[static_cast<size_t>(opposite(d)) * N_ + n] = value;
                } else {
                    nxt[static_cast<size_t>(d) * N_ + target] = fout[d];
                }
            }
        }
        const bool new_odd = ((t_ + 1U) & 1U) != 0U;
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < static_cast<int64_t>(N_); ++nn) {
            const size_t n = static_cast<size_t>(nn);
            size_t x, y, z;
            coords(n, x, y, z);
            at(0, n) = C::store(nxt[n], 0);
            for (int i = 1; i < Q; i += 2) {
                const size_t backward = neighbor(x, y, z, i + 1);
                if (new_odd) {
                    at(i, n) = C::store(nxt[static_cast<size_t>(i + 1) * N_ + n], i);  // corrected :483
                    at(i + 1, backward) = C::store(nxt[static_cast<size_t>(i) * N_ + n], i + 1);
                } else {
                    at(i, backward) = C::store(nxt[static_cast<size_t>(i) * N_ + n], i);
                    at(i + 1, n) = C::store(nxt[static_cast<size_t>(i + 1) * N_ + n], i + 1);
                }
            }
        }
        ++t_;
    }


public:
    // ------------------------------------------------------------ objective and adjoint
    // kind 0, pressure drop between the fluid cells of plane x_in and plane x_out:
    //   j(f) = mean_{x = x_in} rho / 3 - mean_{x = x_out} rho / 3,   rho = sum_d f_d (logical).
    // kind 1, drag on the registered body (momentum exchange of halfway bounce-back): the population
    // f*_opp(d) that left fluid cell n towards a body cell at n - c_d returns as f_d(n), so the body
    // receives 2 c_opp(d) f_d(n):
    //   j(f) = sum over fluid n and d with n - c_d in the body of -2 c_d[axis] f_d(n) (logical).
    double objective(int kind, int x_in, int x_out) const override {
        std::vector<Real> fin(static_cast<size_t>(Q) * N_);
        logical_state(fin.data());
        return kind == 1 ? drag_objective(fin.data()) : plane_objective(fin.data(), x_in, x_out);
    }

    void set_drag_body(const uint8_t* body, int axis) override {
        if (!body) {
            body_.clear();
            return;
        }
        if (axis < 0 || axis > 2) throw std::invalid_argument("drag axis must be 0, 1 or 2");
        if (flags_.empty()) throw std::invalid_argument("the drag objective needs wall flags");
        body_.assign(body, body + N_);
        for (size_t n = 0; n < N_; ++n)
            if (body_[n] && !is_wall(flags_[n])) throw std::invalid_argument("drag body cells must be wall cells");
        drag_axis_ = axis;
    }

    // Runs `steps` forward steps from the current state and returns
    //   J = (1/window) * sum_{t = steps-window+1}^{steps} j(f_t),
    // with dJ/dgamma per cell and dJ/du_peak from the textbook discrete adjoint in logical space:
    // reverse streaming (the transpose of the pull/bounce-back permutation) and the transposed
    // collision Jacobian, evaluated in double precision on the stored forward states.  Supports
    // unforced BGK with partial bounce-back, stationary walls, velocity inlet and pressure outlet.
    double adjoint_run(int kind, uint64_t steps, uint64_t window, int x_in, int x_out, double* grad_gamma,
                       double* grad_u_peak) override {
        if (window < 1 || window > steps) throw std::invalid_argument("objective window must be in [1, steps]");
        if (kind == 1 && body_.empty()) throw std::invalid_argument("no drag body registered");
        if (ax_ != Real(0) || ay_ != Real(0) || az_ != Real(0) || !lax_.empty())
            throw std::invalid_argument("the adjoint supports unforced lattices only");
        for (int k = 1; k < 8; ++k)
            if (wall_u_[k][0] != Real(0) || wall_u_[k][1] != Real(0) || wall_u_[k][2] != Real(0))
                if (!flags_.empty())
                    for (uint8_t f : flags_)
                        if (is_wall(f) && (f & 7) == k) throw std::invalid_argument("the adjoint supports walls at rest only");
        const size_t QN = static_cast<size_t>(Q) * N_;
        std::vector<std::vector<Real>> states(steps + 1, std::vector<Real>(QN));
        logical_state(states[0].data());
        double J = 0.0;
        for (uint64_t t = 0; t < steps; ++t) {
            step(1);
            logical_state(states[t + 1].data());
            if (t + 1 >= steps - window + 1)
                J += kind == 1 ? drag_objective(states[t + 1].data()) : plane_objective(states[t + 1].data(), x_in, x_out);
        }
        J /= static_cast<double>(window);

        // Adjoint of the objective's density sum on each plane.
        size_t count_in = 0, count_out = 0;
        for (size_t n = 0; n < N_; ++n) {
            if (!flags_.empty() && flags_[n] != kFluid) continue;
            const int x = static_cast<int>(n % nx_);
            count_in += x == x_in;
            count_out += x == x_out;
        }
        const double c_in = count_in ? 1.0 / (3.0 * static_cast<double>(count_in) * static_cast<double>(window)) : 0.0;
        const double c_out = count_out ? -1.0 / (3.0 * static_cast<double>(count_out) * static_cast<double>(window)) : 0.0;
        auto add_source = [&](std::vector<double>& lam) {
            if (kind == 1) {
                for (size_t n = 0; n < N_; ++n)
                    for (int d = 1; d < Q; ++d)
                        if (body_link(n, d))
                            lam[static_cast<size_t>(d) * N_ + n] += -2.0 * c_axis(d) / static_cast<double>(window);
                return;
            }
            for (size_t n = 0; n < N_; ++n) {
                if (!flags_.empty() && flags_[n] != kFluid) continue;
                const int x = static_cast<int>(n % nx_);
                const double c = x == x_in ? c_in : (x == x_out ? c_out : 0.0);
                if (c != 0.0)
                    for (int d = 0; d < Q; ++d) lam[static_cast<size_t>(d) * N_ + n] += c;
            }
        };
        std::vector<double> lam(QN, 0.0), lam_out(QN, 0.0);
        add_source(lam);  // lambda_T = dj(f_T)/df_T
        std::fill(grad_gamma, grad_gamma + N_, 0.0);
        double du = 0.0;
        const double omega = 1.0 / static_cast<double>(tau_);
        for (uint64_t t = steps; t-- > 0;) {
            // Reverse streaming: f_{t+1}[d][n + c_d] = fout_t[d][n], or f_{t+1}[opp d][n] = fout_t[d][n]
            // when n + c_d is a wall (halfway bounce-back).  Wall cells keep their populations.
            for (size_t n = 0; n < N_; ++n) {
                size_t x, y, z;
                coords(n, x, y, z);
                const bool wall = !flags_.empty() && is_wall(flags_[n]);
                for (int d = 0; d < Q; ++d) {
                    const size_t k = static_cast<size_t>(d) * N_ + n;
                    if (wall) {
                        lam_out[k] = lam[k];
                        continue;
                    }
                    const size_t target = neighbor(x, y, z, d);
                    if (!flags_.empty() && is_wall(flags_[target])) lam_out[k] = lam[static_cast<size_t>(opposite(d)) * N_ + n];
                    else lam_out[k] = lam[static_cast<size_t>(d) * N_ + target];
                }
            }
            // Transposed collision Jacobian at every non-wall cell.
            const std::vector<Real>& fs = states[t];
            for (size_t n = 0; n < N_; ++n) {
                const uint8_t flag = flags_.empty() ? kFluid : flags_[n];
                if (is_wall(flag)) {
                    for (int d = 0; d < Q; ++d) lam[static_cast<size_t>(d) * N_ + n] = lam_out[static_cast<size_t>(d) * N_ + n];
                    continue;
                }
                size_t x, y, z;
                coords(n, x, y, z);
                double f[Q], lo[Q], mu[Q];
                for (int d = 0; d < Q; ++d) {
                    f[d] = static_cast<double>(to_full(fs[static_cast<size_t>(d) * N_ + n], d));
                    lo[d] = lam_out[static_cast<size_t>(d) * N_ + n];
                }
                double dgamma = 0.0, dpeak = 0.0;
                collision_vjp(flag, n, y, z, f, lo, omega, mu, dgamma, dpeak);
                for (int d = 0; d < Q; ++d) lam[static_cast<size_t>(d) * N_ + n] = mu[d];
                grad_gamma[n] += dgamma;
                du += dpeak;
            }
            if (t >= steps - window + 1) add_source(lam);  // dj(f_t)/df_t for t in the window
        }
        if (grad_u_peak) *grad_u_peak = du;
        return J;
    }

private:
    // Loaded populations in the state representation (deviations for shifted storage).
    void logical_state(Real* out) const {
        for (size_t n = 0; n < N_; ++n) {
            Real fin[Q];
            load(n, fin);
            for (int d = 0; d < Q; ++d) out[static_cast<size_t>(d) * N_ + n] = fin[d];
        }
    }

    double c_axis(int d) const {
        return static_cast<double>(drag_axis_ == 0 ? L::cx[d] : (drag_axis_ == 1 ? L::cy[d] : L::cz[d]));
    }

    // Population d of fluid cell n arrives from a body cell (the upstream cell n - c_d).
    bool body_link(size_t n, int d) const {
        if (flags_[n] != kFluid) return false;
        size_t x, y, z;
        coords(n, x, y, z);
        return body_[neighbor(x, y, z, opposite(d))] != 0;
    }

    double drag_objective(const Real* fin) const {
        if (body_.empty()) throw std::invalid_argument("no drag body registered");
        double s = 0.0;
        for (size_t n = 0; n < N_; ++n)
            for (int d = 1; d < Q; ++d)
                if (body_link(n, d))
                    s += -2.0 * c_axis(d) * static_cast<double>(to_full(fin[static_cast<size_t>(d) * N_ + n], d));
        return s;
    }

    double plane_objective(const Real* fin, int x_in, int x_out) const {
        double s_in = 0.0, s_out = 0.0;
        size_t c_in = 0, c_out = 0;
        for (size_t n = 0; n < N_; ++n) {
            if (!flags_.empty() && flags_[n] != kFluid) continue;
            const int x = static_cast<int>(n % nx_);
            if (x != x_in && x != x_out) continue;
            Real rho = Real(0);
            for (int d = 0; d < Q; ++d) rho += fin[static_cast<size_t>(d) * N_ + n];
            const double r = kShifted ? 1.0 + static_cast<double>(rho) : static_cast<double>(rho);
            if (x == x_in) {
                s_in += r;
                ++c_in;
            } else {
                s_out += r;
                ++c_out;
            }
        }
        if (c_in == 0 || c_out == 0) throw std::invalid_argument("objective planes contain no fluid cells");
        return s_in / (3.0 * static_cast<double>(c_in)) - s_out / (3.0 * static_cast<double>(c_out));
    }

    // d feq_d / d(rho, u) for the second-order equilibrium, in double.
    static void feq_grad(int d, double rho, double ux, double uy, double uz, double& e, double (&gu)[3]) {
        const double cx = L::cx[d], cy = L::cy[d], cz = L::cz[d];
        const double w = g_exact_fp64_weights ? L::w_exact[d] : static_cast<double>(L::w[d]);
        const double cu = cx * ux + cy * uy + cz * uz;
        const double uu = ux * ux + uy * uy + uz * uz;
        e = w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        gu[0] = w * rho * (3.0 * cx + 9.0 * cu * cx - 3.0 * ux);
        gu[1] = w * rho * (3.0 * cy + 9.0 * cu * cy - 3.0 * uy);
        gu[2] = w * rho * (3.0 * cz + 9.0 * cu * cz - 3.0 * uz);
    }

    // mu = (d fout / d fin)^T lambda_out for one cell, plus d/dgamma and d/du_peak.
    void collision_vjp(uint8_t flag, size_t n, size_t y, size_t z, const double* f, const double* lo, double omega,
                       double* mu, double& dgamma, double& dpeak) const {
        if (flag == kInlet) {
            // u_in = u_peak * P(y, z) with P the parabolic profile of inlet_velocity().
            const double eta = (static_cast<double>(y) - 0.5) / static_cast<double>(ny_ - 2);
            double P = 4.0 * eta * (1.0 - eta);
            if (L::dims == 3) {
                const double zeta = (static_cast<double>(z) - 0.5) / static_cast<double>(nz_ - 2);
                P *= 4.0 * zeta * (1.0 - zeta);
            }
            const double uin = static_cast<double>(inlet_velocity(y, z));
            double s0 = 0.0, sn = 0.0;
            for (int d = 0; d < Q; ++d) {
                if (L::cx[d] == 0) s0 += f[d];
                else if (L::cx[d] < 0) sn += f[d];
            }
            const double rho = (s0 + 2.0 * sn) / (1.0 - uin);
            double g_rho = 0.0, g_ux = 0.0;
            for (int d = 0; d < Q; ++d) {
                double e, gu[3];
                feq_grad(d, rho, uin, 0.0, 0.0, e, gu);
                g_rho += lo[d] * e;
                g_ux += lo[d] * gu[0];
            }
            for (int d = 0; d < Q; ++d) {
                const double coef = L::cx[d] == 0 ? 1.0 : (L::cx[d] < 0 ? 2.0 : 0.0);
                mu[d] = g_rho * coef / (1.0 - uin);
            }
            // rho depends on u_in through 1 / (1 - u_in); u_in = u_peak * P.
            dpeak = (g_ux + g_rho * rho / (1.0 - uin)) * P;
            dgamma = 0.0;
            return;
        }
        if (flag == kOutlet) {
            const double rho = static_cast<double>(rho_out_);
            double s0 = 0.0, sp = 0.0;
            for (int d = 0; d < Q; ++d) {
                if (L::cx[d] == 0) s0 += f[d];
                else if (L::cx[d] > 0) sp += f[d];
            }
            const double ux = -1.0 + (s0 + 2.0 * sp) / rho;
            double g_ux = 0.0;
            for (int d = 0; d < Q; ++d) {
                double e, gu[3];
                feq_grad(d, rho, ux, 0.0, 0.0, e, gu);
                g_ux += lo[d] * gu[0];
            }
            for (int d = 0; d < Q; ++d) {
                const double coef = L::cx[d] == 0 ? 1.0 : (L::cx[d] > 0 ? 2.0 : 0.0);
                mu[d] = g_ux * coef / rho;
            }
            dgamma = dpeak = 0.0;
            return;
        }
        double rho = 0.0, m[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < Q; ++d) {
            rho += f[d];
            m[0] += f[d] * L::cx[d];
            m[1] += f[d] * L::cy[d];
            m[2] += f[d] * L::cz[d];
        }
        const double u[3] = {m[0] / rho, m[1] / rho, m[2] / rho};
        const double g = gamma_.empty() ? 0.0 : static_cast<double>(gamma_[n]);
        double a[Q], fstar[Q];
        double g_rho = 0.0, G[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < Q; ++d) {
            double e, gu[3];
            feq_grad(d, rho, u[0], u[1], u[2], e, gu);
            fstar[d] = f[d] - omega * (f[d] - e * rho);
            a[d] = (1.0 - g) * lo[d];
            g_rho += omega * a[d] * e;
            for (int c = 0; c < 3; ++c) G[c] += omega * a[d] * gu[c] / rho;
        }
        const double Gu = G[0] * u[0] + G[1] * u[1] + G[2] * u[2];
        dgamma = 0.0;
        for (int k = 0; k < Q; ++k) {
            const double Gc = G[0] * L::cx[k] + G[1] * L::cy[k] + G[2] * L::cz[k];
            mu[k] = (1.0 - omega) * a[k] + g * lo[opposite(k)] + g_rho + Gc - Gu;
            dgamma += lo[k] * (f[opposite(k)] - fstar[k]);
        }
        dpeak = 0.0;
    }

private:
    // Macroscopic state reported per cell (esoteric_pull.hpp:180-209): fluid u = m/rho + a/2,
    // wall cells rho = sum of frozen populations and u = 0, inlet/outlet cells report the state
    // their boundary rule imposes.
    void cell_moments(size_t n, Real* m) const {
        Real fin[Q];
        load(n, fin);
        const uint8_t flag = flags_.empty() ? kFluid : flags_[n];
        if (is_wall(flag)) {
            Real rho = Real(0);
            for (int d = 0; d < Q; ++d) rho += fin[d];
            m[0] = kShifted ? Real(1) + rho : rho;
            m[1] = m[2] = m[3] = Real(0);
            return;
        }
        size_t x, y, z;
        coords(n, x, y, z);
        if (flag == kInlet) {
            const Real uin = inlet_velocity(y, z);
            m[0] = inlet_density(fin, uin);
            m[1] = uin;
            m[2] = m[3] = Real(0);
            return;
        }
        if (flag == kOutlet) {
            m[0] = rho_out_;
            m[1] = outlet_velocity(fin);
            m[2] = m[3] = Real(0);
            return;
        }
        Real rho, ux, uy, uz, ax, ay, az;
        accumulate(fin, rho, ux, uy, uz);
        acceleration(n, ax, ay, az);
        m[0] = rho;
        m[1] = ux + Real(0.5) * ax;
        m[2] = uy + Real(0.5) * ay;
        m[3] = uz + Real(0.5) * az;
    }
};

template <int Q> OracleBase* make_q(int storage, size_t nx, size_t ny, size_t nz, double tau, double ax, double ay,
                                    double az) {
    switch (storage) {
    case 0: return new Oracle<Q, float, float>(nx, ny, nz, tau, ax, ay, az);
    case 1: return new Oracle<Q, float, Half>(nx, ny, nz, tau, ax, ay, az);
    case 2: return new Oracle<Q, double, double>(nx, ny, nz, tau, ax, ay, az);
    case 3: return new Oracle<Q, float, F32Dev>(nx, ny, nz, tau, ax, ay, az);
    default: throw std::invalid_argument("storage must be 0 (fp32), 1 (fp16 shifted), 2 (fp64) or 3 (fp32 shifted)");
    }
}

// ---------------------------------------------------------------- voxelisation (voxelize.cpp:4-6)
// A cell is solid when the ray from its centre (nudged by 1e-10 dx in y and 2e-10 dx in z) along
// +x crosses the surface an odd number of times.  Brute force over all cells and triangles, as the
// reference does, in double with its operation order (-ffp-contract=off).
struct V3 {
    double x, y, z;
};
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 vcross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Moller-Trumbore along d = (1,0,0) with the thresholds of voxelize.cpp:4.
inline bool ray_x_hits(V3 o, const double* t) {
    const V3 a{t[0], t[1], t[2]}, b{t[3], t[4], t[5]}, c{t[6], t[7], t[8]};
    const V3 d{1, 0, 0};
    const V3 e1 = vsub(b, a), e2 = vsub(c, a);
    const V3 h = vcross(d, e2);
    const double det = vdot(e1, h);
    if (std::fabs(det) < 1e-12) return false;
    const double inv = 1 / det;
    const V3 s = vsub(o, a);
    const double u = inv * vdot(s, h);
    if (u < 0 || u > 1) return false;
    const V3 q = vcross(s, e1);
    const double v = inv * vdot(d, q);
    if (v < 0 || u + v > 1) return false;
    return inv * vdot(e2, q) > 1e-10;
}

inline void voxelize(const double* tris, size_t count, size_t nx, size_t ny, size_t nz, const double* lo,
                     const double* hi, uint8_t* mask) {
    if (count == 0 || nx == 0 || ny == 0 || nz == 0 || !(hi[0] > lo[0]) || !(hi[1] > lo[1]) || !(hi[2] > lo[2]))
        throw std::invalid_argument("invalid voxelization inputs");
    const double dx = (hi[0] - lo[0]) / static_cast<double>(nx);
    const double dy = (hi[1] - lo[1]) / static_cast<double>(ny);
    const double dz = (hi[2] - lo[2]) / static_cast<double>(nz);
#pragma omp parallel for schedule(dynamic)
    for (long long kk = 0; kk < static_cast<long long>(nz); ++kk) {
        const size_t k = static_cast<size_t>(kk);
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const V3 o{lo[0] + (static_cast<double>(i) + .5) * dx, lo[1] + (static_cast<double>(j) + .5) * dy + 1e-10 * dx,
                           lo[2] + (static_cast<double>(k) + .5) * dz + 2e-10 * dx};
                unsigned parity = 0;
                for (size_t t = 0; t < count; ++t) parity ^= ray_x_hits(o, tris + 9 * t) ? 1U : 0U;
                mask[(k * ny + j) * nx + i] = static_cast<uint8_t>(parity);
            }
    }
}

// Rodrigues rotation about a unit axis through the origin, then a translation (voxelize.cpp:4,6).
inline void transform(const double* in, size_t count, const double* tr, const double* axis, double angle,
                      double* out) {
    const V3 ax0{axis[0], axis[1], axis[2]};
    const double n = std::sqrt(vdot(ax0, ax0));
    for (size_t v = 0; v < 3 * count; ++v) {
        const V3 p{in[3 * v], in[3 * v + 1], in[3 * v + 2]};
        V3 r = p;
        if (!(n < 1e-30)) {
            const double m = 1 / n;
            const V3 k{ax0.x * m, ax0.y * m, ax0.z * m};
            const double c = std::cos(angle), sn = std::sin(angle);
            const V3 kp = vcross(k, p);
            const double along = vdot(k, p) * (1 - c);
            r = {(p.x * c + kp.x * sn) + k.x * along, (p.y * c + kp.y * sn) + k.y * along,
                 (p.z * c + kp.z * sn) + k.z * along};
        }
        out[3 * v] = r.x + tr[0];
        out[3 * v + 1] = r.y + tr[1];
        out[3 * v + 2] = r.z + tr[2];
    }
}

// ---------------------------------------------------------------- resident coupling modules
// Periodic lattices of nx*ny*nz cells, cell n = (z*ny + y)*nx + x, float arithmetic in the
// reference's order (resident_multiphysics_sycl.hpp:21-48 for the index helpers).
namespace mp {
struct Grid {
    size_t nx, ny, nz;
    size_t cells() const { return nx * ny * nz; }
    size_t at(size_t x, size_t y, size_t z) const { return (z * ny + y) * nx + x; }
    static size_t shift(size_t v, int delta, size_t n) {
        if (delta > 0) return v + 1 == n ? 0 : v + 1;
        if (delta < 0) return v == 0 ? n - 1 : v - 1;
        return v;
    }
    void split(size_t n, size_t& x, size_t& y, size_t& z) const {
        x = n % nx;
        y = (n / nx) % ny;
        z = n / (nx * ny);
    }
};

// D3Q7: 0, +x, -x, +y, -y, +z, -z; weights 1/4, 1/8 (:188-191).
inline int c7x(int d) { return d == 1 ? 1 : d == 2 ? -1 : 0; }
inline int c7y(int d) { return d == 3 ? 1 : d == 4 ? -1 : 0; }
inline int c7z(int d) { return d == 5 ? 1 : d == 6 ? -1 : 0; }
inline float w7(int d) { return d == 0 ? 0.25F : 0.125F; }

inline void thermal_fill(float* g, const float* t, float value, size_t cells) {
    for (size_t n = 0; n < cells; ++n)
        for (int d = 0; d < 7; ++d) g[d * cells + n] = w7(d) * (t ? t[n] : value);
}

// step (:93-125): pull from x - c, BGK towards w T (1 + 4 c.u).
inline void thermal_step(const float* in, float* out, const float* ux, const float* uy, const float* uz, Grid g,
                         float omega) {
    const size_t cells = g.cells();
    for (long long nn = 0; nn < static_cast<long long>(cells); ++nn) {
        const size_t n = static_cast<size_t>(nn);
        size_t x, y, z;
        g.split(n, x, y, z);
        float local[7];
        float temperature = 0.0F;
        for (int d = 0; d < 7; ++d) {
            const size_t src = g.at(Grid::shift(x, -c7x(d), g.nx), Grid::shift(y, -c7y(d), g.ny),
                                    Grid::shift(z, -c7z(d), g.nz));
            local[d] = in[d * cells + src];
            temperature += local[d];
        }
        for (int d = 0; d < 7; ++d) {
            const float cu = static_cast<float>(c7x(d)) * ux[n] + static_cast<float>(c7y(d)) * uy[n] +
                             static_cast<float>(c7z(d)) * uz[n];
            const float geq = w7(d) * temperature * (1.0F + 4.0F * cu);
            out[d * cells + n] = local[d] - omega * (local[d] - geq);
        }
    }
}

inline void thermal_scalar(const float* g, float* t, size_t cells) {
    for (size_t n = 0; n < cells; ++n) {
        float v = 0.0F;
        for (int d = 0; d < 7; ++d) v += g[d * cells + n];
        t[n] = v;
    }
}

// add_boussinesq_acceleration (:140-155)
inline void boussinesq(const float* t, float* ax, float* ay, float* az, size_t cells, float t_ref, float beta,
                       float gx, float gy, float gz) {
    for (size_t n = 0; n < cells; ++n) {
        const float scale = -beta * (t[n] - t_ref);
        ax[n] += scale * gx;
        ay[n] += scale * gy;
        az[n] += scale * gz;
    }
}

// initialize_half_space_x (:225-236)
inline void vof_half_space(float* phi, Grid g, float interface_x) {
    for (size_t n = 0; n < g.cells(); ++n) {
        const float left = static_cast<float>(n % g.nx), right = left + 1.0F;
        phi[n] = interface_x <= left ? 0.0F : interface_x >= right ? 1.0F : interface_x - left;
    }
}

// advect (:238-263): upwind face fluxes, clamped to [0, 1].
inline void vof_advect(const float* phi, float* out, const float* ux, const float* uy, const float* uz, Grid g,
                       float dt) {
    for (size_t n = 0; n < g.cells(); ++n) {
        size_t x, y, z;
        g.split(n, x, y, z);
        const size_t xp = g.at(Grid::shift(x, 1, g.nx), y, z), xm = g.at(Grid::shift(x, -1, g.nx), y, z);
        const size_t yp = g.at(x, Grid::shift(y, 1, g.ny), z), ym = g.at(x, Grid::shift(y, -1, g.ny), z);
        const size_t zp = g.at(x, y, Grid::shift(z, 1, g.nz)), zm = g.at(x, y, Grid::shift(z, -1, g.nz));
        const float uxp = 0.5F * (ux[n] + ux[xp]), uxm = 0.5F * (ux[xm] + ux[n]);
        const float uyp = 0.5F * (uy[n] + uy[yp]), uym = 0.5F * (uy[ym] + uy[n]);
        const float uzp = 0.5F * (uz[n] + uz[zp]), uzm = 0.5F * (uz[zm] + uz[n]);
        const float fxp = uxp >= 0.0F ? uxp * phi[n] : uxp * phi[xp];
        const float fxm = uxm >= 0.0F ? uxm * phi[xm] : uxm * phi[n];
        const float fyp = uyp >= 0.0F ? uyp * phi[n] : uyp * phi[yp];
        const float fym = uym >= 0.0F ? uym * phi[ym] : uym * phi[n];
        const float fzp = uzp >= 0.0F ? uzp * phi[n] : uzp * phi[zp];
        const float fzm = uzm >= 0.0F ? uzm * phi[zm] : uzm * phi[n];
        float value = phi[n] - dt * ((fxp - fxm) + (fyp - fym) + (fzp - fzm));
        value = value < 0.0F ? 0.0F : value > 1.0F ? 1.0F : value;
        out[n] = value;
    }
}

inline void grad(const float* phi, Grid g, size_t n, float& gx, float& gy, float& gz) {
    size_t x, y, z;
    g.split(n, x, y, z);
    gx = 0.5F * (phi[g.at(Grid::shift(x, 1, g.nx), y, z)] - phi[g.at(Grid::shift(x, -1, g.nx), y, z)]);
    gy = 0.5F * (phi[g.at(x, Grid::shift(y, 1, g.ny), z)] - phi[g.at(x, Grid::shift(y, -1, g.ny), z)]);
    gz = 0.5F * (phi[g.at(x, y, Grid::shift(z, 1, g.nz))] - phi[g.at(x, y, Grid::shift(z, -1, g.nz))]);
}

// refresh_geometry (:268-297): geo = [normal x | normal y | normal z | curvature].
inline void vof_geometry(const float* phi, float* geo, Grid g) {
    const size_t c = g.cells();
    for (size_t n = 0; n < c; ++n) {
        float gx, gy, gz;
        grad(phi, g, n, gx, gy, gz);
        const float mag = std::sqrt(gx * gx + gy * gy + gz * gz);
        if (mag > 1.0e-12F) {
            geo[n] = gx / mag;
            geo[c + n] = gy / mag;
            geo[2 * c + n] = gz / mag;
        } else {
            geo[n] = geo[c + n] = geo[2 * c + n] = 0.0F;
        }
    }
    for (size_t n = 0; n < c; ++n) {
        size_t x, y, z;
        g.split(n, x, y, z);
        const float dxn = geo[g.at(Grid::shift(x, 1, g.nx), y, z)] - geo[g.at(Grid::shift(x, -1, g.nx), y, z)];
        const float dyn = geo[c + g.at(x, Grid::shift(y, 1, g.ny), z)] - geo[c + g.at(x, Grid::shift(y, -1, g.ny), z)];
        const float dzn =
            geo[2 * c + g.at(x, y, Grid::shift(z, 1, g.nz))] - geo[2 * c + g.at(x, y, Grid::shift(z, -1, g.nz))];
        geo[3 * c + n] = -0.5F * ((dxn + dyn) + dzn);
    }
}

// add_surface_tension_acceleration (:299-317), with the curvature of vof_geometry.
inline void vof_tension(const float* phi, const float* curvature, float* ax, float* ay, float* az, Grid g,
                        float sigma, float density) {
    for (size_t n = 0; n < g.cells(); ++n) {
        float gx, gy, gz;
        grad(phi, g, n, gx, gy, gz);
        const float c = sigma * curvature[n] / density;
        ax[n] += c * gx;
        ay[n] += c * gy;
        az[n] += c * gz;
    }
}

inline long long floor_to_ll(float v) {
    long long i = static_cast<long long>(v);
    if (static_cast<float>(i) > v) --i;
    return i;
}
inline float periodic_coordinate(float value, size_t extent) {
    const float e = static_cast<float>(extent);
    const long long turns = floor_to_ll(value / e);
    value -= static_cast<float>(turns) * e;
    if (value >= e) value -= e;
    if (value < 0.0F) value += e;
    return value;
}
inline size_t wrap_cell(long long v, size_t extent) {
    const long long e = static_cast<long long>(extent);
    v %= e;
    if (v < 0) v += e;
    return static_cast<size_t>(v);
}

// advance (:346-366), particles in order (the two-way sums in that order).
inline void particles_advance(float* s, size_t count, const float* ux, const float* uy, const float* uz, float* ax,
                              float* ay, float* az, Grid g, float dt, bool two_way, float gx, float gy, float gz) {
    for (size_t p = 0; p < count; ++p) {
        float px = periodic_coordinate(s[p], g.nx), py = periodic_coordinate(s[count + p], g.ny),
              pz = periodic_coordinate(s[2 * count + p], g.nz);
        float vx = s[3 * count + p], vy = s[4 * count + p], vz = s[5 * count + p];
        const float mass = s[6 * count + p], response = s[7 * count + p];
        const long long x0 = floor_to_ll(px), y0 = floor_to_ll(py), z0 = floor_to_ll(pz);
        const float tx = px - static_cast<float>(x0), ty = py - static_cast<float>(y0), tz = pz - static_cast<float>(z0);
        float fu = 0.0F, fv = 0.0F, fw = 0.0F;
        size_t ids[8];
        float weights[8];
        int k = 0;
        for (int dz = 0; dz <= 1; ++dz)
            for (int dy = 0; dy <= 1; ++dy)
                for (int dx = 0; dx <= 1; ++dx) {
                    const float wx = dx ? tx : 1.0F - tx, wy = dy ? ty : 1.0F - ty, wz = dz ? tz : 1.0F - tz;
                    const size_t n = g.at(wrap_cell(x0 + dx, g.nx), wrap_cell(y0 + dy, g.ny), wrap_cell(z0 + dz, g.nz));
                    const float w = wx * wy * wz;
                    ids[k] = n;
                    weights[k] = w;
                    ++k;
                    fu += w * ux[n];
                    fv += w * uy[n];
                    fw += w * uz[n];
                }
        const float dragx = mass * (fu - vx) / response, dragy = mass * (fv - vy) / response,
                    dragz = mass * (fw - vz) / response;
        if (two_way)
            for (int j = 0; j < 8; ++j) {
                ax[ids[j]] += -dragx * weights[j];
                ay[ids[j]] += -dragy * weights[j];
                az[ids[j]] += -dragz * weights[j];
            }
        vx += dt * (dragx / mass + gx);
        vy += dt * (dragy / mass + gy);
        vz += dt * (dragz / mass + gz);
        px = periodic_coordinate(px + dt * vx, g.nx);
        py = periodic_coordinate(py + dt * vy, g.ny);
        pz = periodic_coordinate(pz + dt * vz, g.nz);
        s[p] = px;
        s[count + p] = py;
        s[2 * count + p] = pz;
        s[3 * count + p] = vx;
        s[4 * count + p] = vy;
        s[5 * count + p] = vz;
    }
}

// q_criterion (:394-404)
inline void q_criterion(const float* ux, const float* uy, const float* uz, float* q, Grid g) {
    for (size_t n = 0; n < g.cells(); ++n) {
        size_t x, y, z;
        g.split(n, x, y, z);
        float gr[3][3] = {};
        for (int axis = 0; axis < 3; ++axis) {
            size_t xp = x, xm = x, yp = y, ym = y, zp = z, zm = z;
            if (axis == 0) {
                xp = Grid::shift(x, 1, g.nx);
                xm = Grid::shift(x, -1, g.nx);
            } else if (axis == 1) {
                yp = Grid::shift(y, 1, g.ny);
                ym = Grid::shift(y, -1, g.ny);
            } else {
                zp = Grid::shift(z, 1, g.nz);
                zm = Grid::shift(z, -1, g.nz);
            }
            const size_t a = g.at(xp, yp, zp), b = g.at(xm, ym, zm);
            gr[0][axis] = 0.5F * (ux[a] - ux[b]);
            gr[1][axis] = 0.5F * (uy[a] - uy[b]);
            gr[2][axis] = 0.5F * (uz[a] - uz[b]);
        }
        float s2 = 0.0F, o2 = 0.0F;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const float sij = 0.5F * (gr[i][j] + gr[j][i]), oij = 0.5F * (gr[i][j] - gr[j][i]);
                s2 += sij * sij;
                o2 += oij * oij;
            }
        q[n] = 0.5F * (o2 - s2);
    }
}
}  // namespace mp

}  // namespace orc

// ---------------------------------------------------------------- C ABI for the tests (ctypes)
#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {
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
orc::OracleBase* H(void* h) { return static_cast<orc::OracleBase*>(h); }
}  // namespace

ORC_API const char* orc_last_error(void) { return g_error; }
ORC_API int orc_create(int q, int storage, size_t nx, size_t ny, size_t nz, double tau, double ax, double ay,
                       double az, void** out) {
    return guarded([&] {
        switch (q) {
        case 9: *out = orc::make_q<9>(storage, nx, ny, nz, tau, ax, ay, az); break;
        case 19: *out = orc::make_q<19>(storage, nx, ny, nz, tau, ax, ay, az); break;
        case 27: *out = orc::make_q<27>(storage, nx, ny, nz, tau, ax, ay, az); break;
        default: throw std::invalid_argument("lattice must have q = 9, 19 or 27");
        }
    });
}
ORC_API void orc_destroy(void* h) { delete H(h); }
ORC_API int orc_set_flags(void* h, const uint8_t* f) { return guarded([&] { H(h)->set_flags(f); }); }
ORC_API int orc_set_wall_velocity(void* h, int k, double ux, double uy, double uz) {
    return guarded([&] { H(h)->set_wall_velocity(k, ux, uy, uz); });
}
ORC_API int orc_set_gamma(void* h, const double* g) { return guarded([&] { H(h)->set_gamma(g); }); }
ORC_API int orc_set_local_acceleration(void* h, const double* ax, const double* ay, const double* az) {
    return guarded([&] { H(h)->set_local_acceleration(ax, ay, az); });
}
ORC_API int orc_set_inlet_outlet(void* h, double u_peak, double rho_out) {
    return guarded([&] { H(h)->set_inlet_outlet(u_peak, rho_out); });
}
ORC_API int orc_init_fields(void* h, const double* r, const double* ux, const double* uy, const double* uz) {
    return guarded([&] { H(h)->init_fields(r, ux, uy, uz); });
}
ORC_API int orc_init_uniform(void* h, double r, double ux, double uy, double uz) {
    return guarded([&] { H(h)->init_uniform(r, ux, uy, uz); });
}
ORC_API int orc_init_taylor_green(void* h, double a, double noise, uint64_t seed) {
    return guarded([&] { H(h)->init_taylor_green(a, noise, seed); });
}
ORC_API int orc_step(void* h, uint64_t count) { return guarded([&] { H(h)->step(count); }); }
ORC_API int orc_macroscopic(void* h, double* r, double* ux, double* uy, double* uz) {
    return guarded([&] { H(h)->macroscopic(r, ux, uy, uz); });
}
ORC_API double orc_mass(void* h) { return H(h)->mass(); }
ORC_API double orc_kinetic_energy(void* h) { return H(h)->kinetic_energy(); }
ORC_API double orc_max_speed(void* h) { return H(h)->max_speed(); }
ORC_API int orc_logical(void* h, double* out) { return guarded([&] { H(h)->logical(out); }); }
ORC_API int orc_set_logical(void* h, const double* in) { return guarded([&] { H(h)->set_logical(in); }); }
ORC_API int orc_raw(void* h, double* out) { return guarded([&] { H(h)->raw(out); }); }
ORC_API int orc_load_map(void* h, int p, int64_t* out) { return guarded([&] { H(h)->load_map(p, out); }); }
ORC_API int orc_store_map(void* h, int p, int64_t* out) { return guarded([&] { H(h)->store_map(p, out); }); }
ORC_API int orc_bb_load_map(void* h, int p, int64_t* out) { return guarded([&] { H(h)->bb_load_map(p, out); }); }
ORC_API uint64_t orc_time_step(void* h) { return H(h)->time_step(); }
ORC_API void orc_set_time_step(void* h, uint64_t t) { H(h)->set_time_step(t); }
ORC_API size_t orc_cells(void* h) { return H(h)->cells(); }
ORC_API int orc_objective(void* h, int kind, int x_in, int x_out, double* out) {
    return guarded([&] { *out = H(h)->objective(kind, x_in, x_out); });
}
ORC_API void orc_set_exact_fp64_weights(int enabled) { orc::g_exact_fp64_weights = enabled != 0; }

ORC_API int orc_set_smagorinsky(void* h, int enabled) {
    return guarded([&] { H(h)->set_smagorinsky(enabled != 0); });
}
ORC_API int orc_set_drag_body(void* h, const uint8_t* body, int axis) {
    return guarded([&] { H(h)->set_drag_body(body, axis); });
}
ORC_API int orc_adjoint_run(void* h, int kind, uint64_t steps, uint64_t window, int x_in, int x_out, double* J,
                            double* grad_gamma, double* grad_u_peak) {
    return guarded([&] { *J = H(h)->adjoint_run(kind, steps, window, x_in, x_out, grad_gamma, grad_u_peak); });
}
ORC_API uint16_t orc_f32_to_f16(float f) { return orc::f32_to_f16_rne(f); }
ORC_API float orc_f16_to_f32(uint16_t h) { return orc::f16_to_f32(h); }
// The SplitMix64 output function: the generator seeded with s returns splitmix64(s + k * gamma)
// as its k-th output (k = 0, 1, ...).  Vectorised for the tests that pin it.
ORC_API void orc_splitmix64(const uint64_t* in, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = orc::splitmix64(in[i]);
}
// uniform01(seed, cell, component) for cells [0, n) of one component (the noise of init_taylor_green).
ORC_API void orc_uniform01_cells(uint64_t seed, uint64_t comp, size_t n, double* out) {
    for (size_t i = 0; i < n; ++i) out[i] = static_cast<double>(orc::uniform01<float>(seed, i, comp));
}
ORC_API double orc_uniform01(uint64_t seed, uint64_t cell, uint64_t comp) {
    return static_cast<double>(orc::uniform01<float>(seed, cell, comp));
}
ORC_API int orc_voxelize(const double* tris, size_t count, size_t nx, size_t ny, size_t nz, const double* lo,
                         const double* hi, uint8_t* mask) {
    return guarded([&] { orc::voxelize(tris, count, nx, ny, nz, lo, hi, mask); });
}
ORC_API int orc_transform_surface(const double* tris, size_t count, const double* tr, const double* axis,
                                  double angle, double* out) {
    return guarded([&] { orc::transform(tris, count, tr, axis, angle, out); });
}
// resident coupling modules (periodic float lattices)
ORC_API void orc_thermal_fill(float* g, const float* t, float value, size_t cells) { orc::mp::thermal_fill(g, t, value, cells); }
ORC_API void orc_thermal_step(const float* in, float* out, const float* ux, const float* uy, const float* uz,
                              size_t nx, size_t ny, size_t nz, float omega) {
    orc::mp::thermal_step(in, out, ux, uy, uz, {nx, ny, nz}, omega);
}
ORC_API void orc_thermal_scalar(const float* g, float* t, size_t cells) { orc::mp::thermal_scalar(g, t, cells); }
ORC_API void orc_boussinesq(const float* t, float* ax, float* ay, float* az, size_t cells, float t_ref, float beta,
                            float gx, float gy, float gz) {
    orc::mp::boussinesq(t, ax, ay, az, cells, t_ref, beta, gx, gy, gz);
}
ORC_API void orc_vof_half_space(float* phi, size_t nx, size_t ny, size_t nz, float x) {
    orc::mp::vof_half_space(phi, {nx, ny, nz}, x);
}
ORC_API void orc_vof_advect(const float* phi, float* out, const float* ux, const float* uy, const float* uz, size_t nx,
                            size_t ny, size_t nz, float dt) {
    orc::mp::vof_advect(phi, out, ux, uy, uz, {nx, ny, nz}, dt);
}
ORC_API void orc_vof_geometry(const float* phi, float* geo, size_t nx, size_t ny, size_t nz) {
    orc::mp::vof_geometry(phi, geo, {nx, ny, nz});
}
ORC_API void orc_vof_tension(const float* phi, const float* curvature, float* ax, float* ay, float* az, size_t nx,
                             size_t ny, size_t nz, float sigma, float density) {
    orc::mp::vof_tension(phi, curvature, ax, ay, az, {nx, ny, nz}, sigma, density);
}
ORC_API void orc_particles_advance(float* s, size_t count, const float* ux, const float* uy, const float* uz, float* ax,
                                   float* ay, float* az, size_t nx, size_t ny, size_t nz, float dt, int two_way,
                                   float gx, float gy, float gz) {
    orc::mp::particles_advance(s, count, ux, uy, uz, ax, ay, az, {nx, ny, nz}, dt, two_way != 0, gx, gy, gz);
}
ORC_API void orc_q_criterion(const float* ux, const float* uy, const float* uz, float* q, size_t nx, size_t ny,
                             size_t nz) {
    orc::mp::q_criterion(ux, uy, uz, q, {nx, ny, nz});
}
