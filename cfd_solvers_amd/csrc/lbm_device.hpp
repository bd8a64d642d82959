// lbm_device.hpp -- device building blocks of the gfx950 esoteric-pull stream-collide.
//
// Everything here is written for one thread per cell, x fastest across the 64-lane wavefront,
// populations in structure-of-arrays planes F[d][n] (d-major, one padded plane per direction).
// Per-cell state lives in registers: the q loaded populations, the 9 neighbour row bases and 3
// wrapped x coordinates; no LDS, no division by grid sizes in the cell body.
//
// Reference behaviour followed (paths relative to the xsession/cfd_solvers root):
//   velocity sets ............. include/cfd/solvers/lbm/descriptors.hpp:31-146
//   index map (load/store) .... include/cfd/solvers/lbm/esoteric_pull.hpp:349-382
//   equilibrium / Guo ......... esoteric_pull.hpp:48-70
//   macroscopic accumulation .. esoteric_pull.hpp:384-406
//   BGK collision ............. esoteric_pull.hpp:408-432
// Arithmetic keeps the reference's association order; the library is compiled with
// -ffp-contract=off, so an FP32 step reproduces the reference's FP32 step bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace cfd_hip {

// ------------------------------------------------------------------ velocity sets
template <int Q> struct Lattice;
template <> struct Lattice<9> {
    static constexpr int dims = 2;
    static constexpr int cx[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
    static constexpr int cy[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr float w[9] = {4.0F / 9.0F,  1.0F / 9.0F,  1.0F / 9.0F,  1.0F / 9.0F, 1.0F / 9.0F,
                                   1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F, 1.0F / 36.0F};
    static constexpr double w_exact[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};
template <> struct Lattice<19> {
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
template <> struct Lattice<27> {
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

constexpr int opposite(int i) { return i == 0 ? 0 : ((i & 1) != 0 ? i + 1 : i - 1); }

// ------------------------------------------------------------------ compile-time loops
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// ------------------------------------------------------------------ flags
constexpr uint8_t kFluid = 0x00;
constexpr uint8_t kInlet = 0x01;
constexpr uint8_t kOutlet = 0x02;
constexpr uint8_t kWall = 0x80;
__device__ __forceinline__ bool is_wall(uint8_t f) { return (f & kWall) != 0; }
__device__ __forceinline__ bool is_wall(uint32_t f) { return (f & kWall) != 0; }

// ------------------------------------------------------------------ population storage
// fp32: the value itself, collision in the reference's arithmetic.  fp64: verification build.
// fp16s: binary16 of (f - w_d) * 4096 (round to nearest even, the hardware conversion).  The
// state of an fp16s lattice is the deviation f' = f - w_d from the rest equilibrium, and the
// collision works on deviations throughout (rho = 1 + sum f'; see collide): at low velocities the
// non-equilibrium part f - f_eq is many orders below f itself, and forming it from full
// populations in FP32 loses it to cancellation (a 1 % bias in a porous channel at u ~ 3e-4),
// while the deviations keep it to FP32 relative precision.
constexpr float kHalfScale = 4096.0F;
constexpr float kHalfInvScale = 1.0F / 4096.0F;

template <class T> struct Storage;
template <> struct Storage<float> {
    using Real = float;
    static constexpr bool kShifted = false;
    static constexpr bool kExactWeights = false;
    static constexpr float kScale = 1.0F, kInvScale = 1.0F;
    template <int Q, int D> __device__ __forceinline__ static float decode(float v) { return v; }
    template <int Q, int D> __device__ __forceinline__ static float encode(float v) { return v; }
};
// The state of an fp16s lattice is kept scaled by 4096 in registers as well (a power of two: every
// operation on the scaled values is exact scaling of the same operation on the deviations), so
// decoding and encoding are plain conversions.
template <> struct Storage<_Float16> {
    using Real = float;
    static constexpr bool kShifted = true;
    static constexpr bool kExactWeights = false;
    static constexpr float kScale = kHalfScale, kInvScale = kHalfInvScale;
    template <int Q, int D> __device__ __forceinline__ static float decode(_Float16 h) { return static_cast<float>(h); }
    // The barrier keeps the compiler from fusing the producing multiply or fma into the conversion
    // (v_fma_mix*_f16 rounds once, straight to binary16): the stored value is the binary16 rounding
    // of the FP32 result, on every path.
    template <int Q, int D> __device__ __forceinline__ static _Float16 encode(float v) {
        asm volatile("" : "+v"(v));
        return static_cast<_Float16>(v);
    }
};
// fp32s: FP32 storage of the deviation f - w_d, collision on deviations (the accuracy of the
// shifted arithmetic at the FP32 layout's cost; fp32 keeps the reference's arithmetic bit for bit).
struct F32Dev {
    float v;
};
template <> struct Storage<F32Dev> {
    using Real = float;
    static constexpr bool kShifted = true;
    static constexpr bool kExactWeights = false;
    static constexpr float kScale = 1.0F, kInvScale = 1.0F;
    template <int Q, int D> __device__ __forceinline__ static float decode(F32Dev h) { return h.v; }
    template <int Q, int D> __device__ __forceinline__ static F32Dev encode(float v) { return F32Dev{v}; }
};
// Slot identifiers, used only by the index-map probe.
template <> struct Storage<uint32_t> {
    using Real = uint32_t;
    static constexpr bool kShifted = false;
    static constexpr bool kExactWeights = false;
    static constexpr float kScale = 1.0F, kInvScale = 1.0F;
    template <int Q, int D> __device__ __forceinline__ static uint32_t decode(uint32_t v) { return v; }
    template <int Q, int D> __device__ __forceinline__ static uint32_t encode(uint32_t v) { return v; }
};
template <> struct Storage<double> {
    using Real = double;
    static constexpr bool kShifted = false;
    static constexpr bool kExactWeights = false;
    static constexpr double kScale = 1.0, kInvScale = 1.0;
    template <int Q, int D> __device__ __forceinline__ static double decode(double v) { return v; }
    template <int Q, int D> __device__ __forceinline__ static double encode(double v) { return v; }
};
// fp64x: the FP64 build with the lattice weights' exact values (4/9, 1/3, 1/18, ...) instead of the
// reference's FP32 constants, which sum to 1 + 1.49e-8 (D3Q19) or 1 + 7.45e-9 (D2Q9, D3Q27) and so
// add omega times that excess to a cell's density at every collision.  A reference for precision
// studies; fp64 keeps the reference's constants.
struct F64Exact {
    double v;
};
template <> struct Storage<F64Exact> {
    using Real = double;
    static constexpr bool kShifted = false;
    static constexpr bool kExactWeights = true;
    static constexpr double kScale = 1.0, kInvScale = 1.0;
    template <int Q, int D> __device__ __forceinline__ static double decode(F64Exact v) { return v.v; }
    template <int Q, int D> __device__ __forceinline__ static F64Exact encode(double v) { return F64Exact{v}; }
};

// Lattice weight w_D in the arithmetic of storage S.
template <class S, int Q, int D, class Real> __host__ __device__ constexpr Real lattice_weight() {
    if constexpr (S::kExactWeights) return static_cast<Real>(Lattice<Q>::w_exact[D]);
    else return static_cast<Real>(Lattice<Q>::w[D]);
}

// Population d of the state representation -> the population itself (deviation / scale + w_d
// when shifted).
template <class S, int Q, int D, class Real> __device__ __forceinline__ Real full_population(Real v) {
    if constexpr (S::kShifted) return v * static_cast<Real>(S::kInvScale) + static_cast<Real>(Lattice<Q>::w[D]);
    else return v;
}
// The population itself -> the state representation.
template <class S, int Q, int D, class Real> __device__ __forceinline__ Real state_population(Real v) {
    if constexpr (S::kShifted) return (v - static_cast<Real>(Lattice<Q>::w[D])) * static_cast<Real>(S::kScale);
    else return v;
}

// ------------------------------------------------------------------ launch geometry
// A block is a bx*by = 256-thread tile of one z plane (bx spans x: 64, 128 or 256, so that the 64
// lanes of a wavefront lie in one row and one 64-cell row segment).  Tiles are
// numbered x-fastest.  The hardware deals consecutive blocks round-robin to the 8 XCDs, so the
// chip streams through neighbouring rows together; xcd_remap (off by default: it measured slower
// on MI355X) instead gives each XCD one contiguous slab of tiles.
// Population layout.  The lattice's x-rows (row r = z*ny + y) are grouped into blocks of 2^row_shift
// consecutive rows; a block holds the q direction planes of its cells one after another, `stride`
// elements apart, and blocks are `block_elems` elements apart.  Slot (d, cell (r, x)) is element
//   (r >> row_shift) * block_elems + d * stride + (r & (2^row_shift - 1)) * nx + x.
// row_shift = 31 puts every row in one block: plain structure-of-arrays planes, the reference's
// StaticSoA order (soa_field.hpp:34-36), with `stride` >= cells.  Blocks of ~4096 cells keep the 19
// slot streams of a wavefront inside one 300 KB region, which the in-place pattern streams faster.
struct Geometry {
    uint32_t nx, ny, nz;
    uint32_t bx, by;
    uint32_t tiles_x, tiles_y, tiles;
    uint32_t xcd_remap;
    uint32_t row_shift;    // log2 rows per block; 31: one block (plain planes)
    uint32_t block_elems;  // elements per block (q * stride); unused when row_shift = 31
    uint64_t stride;       // elements between consecutive direction planes of a block
};

// Element of direction plane 0 holding the first cell of x-row r.
__host__ __device__ __forceinline__ uint32_t row_base(const Geometry& g, uint32_t r) {
    return (r >> g.row_shift) * g.block_elems + (r & ((1U << g.row_shift) - 1U)) * g.nx;
}

__device__ __forceinline__ bool cell_coords(const Geometry& g, uint32_t& x, uint32_t& y, uint32_t& z) {
    uint32_t b = blockIdx.x;
    if (g.xcd_remap) b = (b & 7U) * (g.tiles >> 3) + (b >> 3);
    const uint32_t tx = b % g.tiles_x;
    const uint32_t r = b / g.tiles_x;
    const uint32_t ty = r % g.tiles_y;
    z = r / g.tiles_y;
    x = tx * g.bx + threadIdx.x;
    y = ty * g.by + threadIdx.y;
    return x < g.nx && y < g.ny;
}

// Periodic neighbourhood of (x, y, z): 3 wrapped x coordinates and, for the 9 wrapped rows, their
// plane-0 element (population slots) and their linear cell index (flags, descriptors, fields).
struct Neighbourhood {
    uint32_t xs[3];
    uint32_t rows[3][3];   // [dy + 1][dz + 1]: row_base
    uint32_t lrows[3][3];  // [dy + 1][dz + 1]: row * nx
    // Element of direction plane 0 for the neighbour (dx, dy, dz).
    __device__ __forceinline__ uint32_t cell(int dx, int dy, int dz) const { return rows[dy + 1][dz + 1] + xs[dx + 1]; }
    // Linear cell index (z*ny + y)*nx + x of the neighbour.
    __device__ __forceinline__ uint32_t lin(int dx, int dy, int dz) const { return lrows[dy + 1][dz + 1] + xs[dx + 1]; }
};

__device__ __forceinline__ Neighbourhood neighbourhood(const Geometry& g, uint32_t x, uint32_t y, uint32_t z) {
    Neighbourhood h;
    h.xs[0] = x == 0U ? g.nx - 1U : x - 1U;
    h.xs[1] = x;
    h.xs[2] = x + 1U == g.nx ? 0U : x + 1U;
    const uint32_t ys[3] = {y == 0U ? g.ny - 1U : y - 1U, y, y + 1U == g.ny ? 0U : y + 1U};
    const uint32_t zs[3] = {z == 0U ? g.nz - 1U : z - 1U, z, z + 1U == g.nz ? 0U : z + 1U};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const uint32_t r = zs[b] * g.ny + ys[a];
            h.rows[a][b] = row_base(g, r);
            h.lrows[a][b] = r * g.nx;
        }
    return h;
}

// ------------------------------------------------------------------ esoteric slots
// For the opposite pair (i, i+1) of cell n, with up = n - c_i and dn = n + c_i:
//   even parity: fin[i] <- F[i][up],   fin[i+1] <- F[i+1][n];  fout[i] -> F[i+1][n], fout[i+1] -> F[i][up]
//   odd parity:  fin[i] <- F[i+1][up], fin[i+1] <- F[i][n];    fout[i] -> F[i][n],   fout[i+1] -> F[i+1][up]
// (esoteric_pull.hpp:349-382).  Halfway bounce-back in place: when the upstream cell of a
// population is a wall, the population is read from the slot of the opposite parity, where this
// cell stored the opposite population one step earlier.  Wall cells do nothing, and the two slots
// of a pair at a wall cell are touched only by the one fluid cell facing it, so the update stays
// race-free with a single pass over the lattice.
// Per-cell link descriptor of a bounded lattice, built from the flags whenever they change:
//   bits 0..q-2   link l is a wall link: link 2p is population i = 2p+1 (upstream cell n - c_i),
//                 link 2p+1 is population i+1 (upstream cell n + c_i)
//   bits 26..27   cell type: 0 fluid, 1 velocity inlet, 2 pressure outlet, 3 wall
//   bit 31        at least one wall link has a non-zero wall velocity
// One 32-bit load replaces the cell's own flag and its q-1 neighbour flags in the hot kernel.
constexpr uint32_t kTypeShift = 26;
constexpr uint32_t kTypeWall = 3;
constexpr uint32_t kMovingBit = 31;

// Per-cell slot state carried from the loads to the stores, kept small so the bounded kernel
// stays close to the periodic kernel's register budget.
template <int Q> struct PairSlots {
    static constexpr int kVelWords = (Q - 1 + 9) / 10;  // 3-bit wall-velocity index per link, 10 per word
    uint32_t up[Q];  // byte offset of cell up = n - c_i in a plane, index by i (odd)
    uint32_t self;   // byte offset of the cell itself in a plane
    uint32_t desc;   // link descriptor (bounded lattices only)
    __device__ __forceinline__ bool is_wall_link(int link) const { return ((desc >> link) & 1U) != 0U; }
};

// Link bits (descriptor layout above) of a fluid cell whose only wall neighbours are the cells of
// the x column at x + side (side -1 or +1).
template <int Q> constexpr uint32_t x_wall_links(int side) {
    uint32_t m = 0U;
    for (int pair = 0; pair < (Q - 1) / 2; ++pair) {
        const int cx = Lattice<Q>::cx[2 * pair + 1];
        if (cx == -side) m |= 1U << (2 * pair);     // population i from n - c_i, at x + side
        if (cx == side) m |= 1U << (2 * pair + 1);  // population i + 1 from n + c_i, at x + side
    }
    return m;
}

// Direction plane holding fin[i] (a) and fin[i+1] (b) of pair i.
template <bool Odd> __device__ __forceinline__ int dir_a(int i, bool wall_up) { return (Odd != wall_up) ? i + 1 : i; }
template <bool Odd> __device__ __forceinline__ int dir_b(int i, bool wall_dn) { return (Odd != wall_dn) ? i : i + 1; }

// ------------------------------------------------------------------ collision arithmetic
// c_d . (ux, uy, uz) with the reference's (a + b) + c association; zero terms are skipped and
// unit products folded, which leaves every result bit-identical.
template <int Q, int D, class Real> __device__ __forceinline__ Real cdot(Real ux, Real uy, Real uz) {
    using L = Lattice<Q>;
    Real s = Real(0);
    bool any = false;
    if constexpr (L::cx[D] != 0) {
        s = L::cx[D] > 0 ? ux : -ux;
        any = true;
    }
    if constexpr (L::cy[D] != 0) {
        if (any) s = L::cy[D] > 0 ? s + uy : s - uy;
        else s = L::cy[D] > 0 ? uy : -uy;
        any = true;
    }
    if constexpr (L::cz[D] != 0) {
        if (any) s = L::cz[D] > 0 ? s + uz : s - uz;
        else s = L::cz[D] > 0 ? uz : -uz;
    }
    return s;
}

template <class S, int Q, int D, class Real>
__device__ __forceinline__ Real equilibrium(Real rho, Real ux, Real uy, Real uz, Real uu) {
    const Real cu = Real(3) * cdot<Q, D>(ux, uy, uz);
    return lattice_weight<S, Q, D, Real>() * rho * (Real(1) + cu + Real(0.5) * cu * cu - uu);
}

// Shifted arithmetic (fp16s, fp32s; no reference counterpart).  With cu = 3 c.u and
// uu = 3/2 |u|^2:
//   f_eq - w_d = w_d fma(rho, fma(cu/2, cu, -uu) + cu, rho')      rho' = rho - 1
// scaled by the storage scale K (exact).  Opposite directions share fma(cu/2, cu, -uu).
template <class S, int Q, int D, class Real> __device__ __forceinline__ Real scaled_weight() {
    return static_cast<Real>(Lattice<Q>::w[D]) * static_cast<Real>(S::kScale);
}
template <class S, int Q, int D, class Real>
__device__ __forceinline__ Real equilibrium_dev(Real rp, Real rho, Real ux, Real uy, Real uz, Real uu) {
    const Real cu = Real(3) * cdot<Q, D>(ux, uy, uz);
    const Real x = __builtin_fmaf(Real(0.5) * cu, cu, -uu) + cu;
    return scaled_weight<S, Q, D, Real>() * __builtin_fmaf(rho, x, rp);
}
// 3/2 |u|^2 as the shifted arithmetic forms it.
template <class Real> __device__ __forceinline__ Real shifted_uu(Real ux, Real uy, Real uz) {
    return Real(1.5) * __builtin_fmaf(uz, uz, __builtin_fmaf(uy, uy, ux * ux));
}

// Equilibrium in the state representation of storage S (rp = rho - 1, used when shifted).
template <class S, int Q, int D, class Real>
__device__ __forceinline__ Real equilibrium_state(Real rp, Real rho, Real ux, Real uy, Real uz, Real uu) {
    if constexpr (S::kShifted) return equilibrium_dev<S, Q, D>(rp, rho, ux, uy, uz, uu);
    else return equilibrium<S, Q, D>(rho, ux, uy, uz, uu);
}
// 3/2 |u|^2 in the arithmetic of storage S.
template <class S, class Real> __device__ __forceinline__ Real kinetic_term(Real ux, Real uy, Real uz) {
    if constexpr (S::kShifted) return shifted_uu(ux, uy, uz);
    else return Real(1.5) * (ux * ux + uy * uy + uz * uz);
}

template <class S, int Q, int D, class Real>
__device__ __forceinline__ Real guo(Real ux, Real uy, Real uz, Real fx, Real fy, Real fz, Real uf, Real wfac) {
    const Real cu = cdot<Q, D>(ux, uy, uz);
    const Real cf = cdot<Q, D>(fx, fy, fz);
    return lattice_weight<S, Q, D, Real>() * wfac * (Real(3) * (cf - uf) + Real(9) * cu * cf);
}

// Density and momentum in direction order 0..q-1 (esoteric_pull.hpp:384-406).
template <int Q, class Real>
__device__ __forceinline__ void moments(const Real (&f)[Q], Real& rho, Real& mx, Real& my, Real& mz) {
    using L = Lattice<Q>;
    rho = mx = my = mz = Real(0);
    static_for<Q>([&](auto DD) {
        constexpr int d = decltype(DD)::value;
        const Real v = f[d];
        rho += v;
        if constexpr (L::cx[d] > 0) mx += v;
        if constexpr (L::cx[d] < 0) mx -= v;
        if constexpr (L::cy[d] > 0) my += v;
        if constexpr (L::cy[d] < 0) my -= v;
        if constexpr (L::cz[d] > 0) mz += v;
        if constexpr (L::cz[d] < 0) mz -= v;
    });
}

constexpr uint32_t kSpanBits = 0xFFFFU;        // lo | hi << 8 of a segment word
constexpr uint32_t kPorousSegment = 1U << 16;  // the segment holds a cell with gamma != 0
// Bits 17..18 of a segment word: the span is a resting x wall and its one fluid neighbour, whose
// 3x3x3 neighbourhood is plain fluid but for the wall's x column.  1: wall at lane lo, fluid at
// lo + 1; 2: fluid at lo, wall at lo + 1.  The fluid cell's wall links then follow from the
// lattice alone (x_wall_links), without its link descriptor.
constexpr uint32_t kXWallShift = 17;

template <class Real> struct Physics {
    Real omega;
    Real tau;
    uint32_t les;               // Smagorinsky relaxation time (general kernel variant only)
    Real ax, ay, az;
    Real u_peak, rho_out;
    Real wall_u[8][3];
    const uint8_t* flags;       // null: fully periodic lattice
    const Real* gamma;          // null: no partial bounce-back
    const Real* lax;            // null: no local body acceleration (lay = lax + cells, laz = lay + cells)
    const Real* solid_rho;      // density frozen in wall cells
    const uint32_t* links;      // per-cell link descriptor (bounded lattices)
    const uint32_t* link_vel;   // per cell, (q-1+9)/10 words of 3-bit wall-velocity indices per link
    // Per 64-cell row segment of a bounded lattice: the span lo | hi << 8 of the cells [lo, hi) that
    // are not plain fluid with a plain-fluid 3x3x3 neighbourhood, 0 if none, kPorousSegment when
    // a cell of the segment has a nonzero partial bounce-back fraction, and the x-wall pattern in
    // bits kXWallShift..+1.  A 32-bit word, read through the scalar cache by the wavefront that
    // updates the segment.
    const uint32_t* segments;
    uint32_t seg_x;             // segments per row, ceil(nx / 64)
    uint32_t cells;
};

// Sum of the populations that reach an x-boundary cell from inside the domain.
template <int Q, int SignX, class Real> __device__ __forceinline__ Real sum_cx(const Real (&f)[Q]) {
    Real s = Real(0);
    static_for<Q>([&](auto DD) {
        constexpr int d = decltype(DD)::value;
        if constexpr (Lattice<Q>::cx[d] == SignX) s += f[d];
    });
    return s;
}

template <class Real>
__device__ __forceinline__ Real inlet_velocity(const Physics<Real>& p, const Geometry& g, uint32_t y, uint32_t z,
                                               bool three_d) {
    const Real eta = (static_cast<Real>(y) - Real(0.5)) / static_cast<Real>(g.ny - 2U);
    Real u = p.u_peak * (Real(4) * eta * (Real(1) - eta));
    if (three_d) {
        const Real zeta = (static_cast<Real>(z) - Real(0.5)) / static_cast<Real>(g.nz - 2U);
        u = u * (Real(4) * zeta * (Real(1) - zeta));
    }
    return u;
}

// Density of a velocity-inlet cell (Zou-He, from the populations arriving from inside) as
// rho' = rho - 1 and rho.  Unshifted: rho = (S0 + 2 S-) / (1 - u) as in src/lbm/d2q9.cpp:152-168.
// Shifted (S0 + 2 S- of the weights is 1): rho' = (S0' + 2 S-' + u) / (1 - u).
template <class S, int Q, class Real>
__device__ __forceinline__ void inlet_density(const Real (&f)[Q], Real uin, Real& rp, Real& rho) {
    const Real s = sum_cx<Q, 0>(f) + Real(2) * sum_cx<Q, -1>(f);
    if constexpr (S::kShifted) {
        rp = (s * static_cast<Real>(S::kInvScale) + uin) / (Real(1) - uin);
        rho = Real(1) + rp;
    } else {
        rho = s / (Real(1) - uin);
        rp = rho - Real(1);
    }
}
// Normal velocity of a pressure-outlet cell: -1 + (S0 + 2 S+) / rho_out (shifted: the weights'
// part of the sum is 1, so u = (1 - rho_out + S0' + 2 S+') / rho_out).
template <class S, int Q, class Real>
__device__ __forceinline__ Real outlet_velocity(const Physics<Real>& p, const Real (&f)[Q]) {
    const Real s = sum_cx<Q, 0>(f) + Real(2) * sum_cx<Q, 1>(f);
    if constexpr (S::kShifted) return ((Real(1) - p.rho_out) + s * static_cast<Real>(S::kInvScale)) / p.rho_out;
    else return Real(-1) + s / p.rho_out;
}

// m / rho from c = 1 / rho with one residual correction: u0 = m c, u = u0 + (m - u0 rho) c, the
// residual exact by fma.  A bare m * (1 / rho) is biased by its double rounding (a 2e-6 drift of the
// mean velocity in a forced porous channel); the corrected quotient is not.
template <class Real> __device__ __forceinline__ Real corrected_quotient(Real m, Real rho, Real c) {
    const Real u0 = m * c;
    return __builtin_fmaf(__builtin_fmaf(-u0, rho, m), c, u0);
}

// Density (and rho' = rho - 1) and velocity m / rho of a fluid cell from its state.  The shifted
// arithmetic divides once (corrected_quotient); scaled states are unscaled first (exact).
template <class S, int Q, class Real>
__device__ __forceinline__ void fluid_moments(const Real (&f)[Q], Real& rp, Real& rho, Real& ux, Real& uy, Real& uz) {
    Real mx, my, mz, s;
    moments<Q>(f, s, mx, my, mz);
    if constexpr (S::kShifted) {
        const Real k = static_cast<Real>(S::kInvScale);
        rp = s * k;
        rho = Real(1) + rp;
        if (rho > Real(0)) {
            const Real c = Real(1) / rho;
            ux = corrected_quotient(mx * k, rho, c);
            uy = corrected_quotient(my * k, rho, c);
            uz = corrected_quotient(mz * k, rho, c);
        } else {
            ux = uy = uz = Real(0);
        }
        return;
    } else {
        rho = s;
        rp = s - Real(1);
    }
    if (rho > Real(0)) {
        ux = mx / rho;
        uy = my / rho;
        uz = mz / rho;
    } else {
        ux = uy = uz = Real(0);
    }
}

// Macroscopic state a non-wall cell reports (esoteric_pull.hpp:180-209 for fluid cells).
template <int Q, class S, class Real>
__device__ __forceinline__ void cell_state(const Physics<Real>& p, const Geometry& g, uint8_t flag, uint32_t n,
                                           uint32_t y, uint32_t z, const Real (&f)[Q], Real& rho, Real& ux, Real& uy,
                                           Real& uz) {
    if (flag == kInlet) {
        const Real uin = inlet_velocity(p, g, y, z, Lattice<Q>::dims == 3);
        Real rp;
        inlet_density<S, Q>(f, uin, rp, rho);
        ux = uin;
        uy = uz = Real(0);
        return;
    }
    if (flag == kOutlet) {
        rho = p.rho_out;
        ux = outlet_velocity<S, Q>(p, f);
        uy = uz = Real(0);
        return;
    }
    Real rp;
    fluid_moments<S, Q>(f, rp, rho, ux, uy, uz);
    Real ax = p.ax, ay = p.ay, az = p.az;
    if (p.lax) {
        ax = ax + p.lax[n];
        ay = ay + p.lax[p.cells + n];
        az = az + p.lax[2U * p.cells + n];
    }
    ux = ux + Real(0.5) * ax;
    uy = uy + Real(0.5) * ay;
    uz = uz + Real(0.5) * az;
}

// Smagorinsky-Lilly relaxation time (esoteric_pull.hpp:72-97): Pi_ab = sum_d c_da c_db (f_d - f_eq_d),
// Q = sum_ab Pi_ab^2 in row-major order, tau_local = (tau + sqrt(tau^2 + C sqrt(Q) / rho)) / 2 with
// C = 16 sqrt(2) / (3 pi^2) evaluated in float.  The reference's additions of zero terms are
// skipped (exact: they add +-0).  Unshifted storage only.
template <int Q, class S, class Real>
__device__ __forceinline__ Real smagorinsky_tau(const Real (&f)[Q], Real rho, Real ux, Real uy, Real uz, Real uu,
                                                Real tau) {
    using L = Lattice<Q>;
    if (!(rho > Real(0))) return tau;
    Real pxx = Real(0), pxy = Real(0), pxz = Real(0), pyy = Real(0), pyz = Real(0), pzz = Real(0);
    static_for<Q>([&](auto DD) {
        constexpr int d = decltype(DD)::value;
        constexpr int cx = L::cx[d], cy = L::cy[d], cz = L::cz[d];
        const Real neq = f[d] - equilibrium<S, Q, d>(rho, ux, uy, uz, uu);
        if constexpr (cx != 0) pxx = pxx + neq;
        if constexpr (cy != 0) pyy = pyy + neq;
        if constexpr (cz != 0) pzz = pzz + neq;
        if constexpr (cx * cy > 0) pxy = pxy + neq;
        if constexpr (cx * cy < 0) pxy = pxy - neq;
        if constexpr (cx * cz > 0) pxz = pxz + neq;
        if constexpr (cx * cz < 0) pxz = pxz - neq;
        if constexpr (cy * cz > 0) pyz = pyz + neq;
        if constexpr (cy * cz < 0) pyz = pyz - neq;
    });
    Real q = pxx * pxx;
    q = q + pxy * pxy;
    q = q + pxz * pxz;
    q = q + pxy * pxy;
    q = q + pyy * pyy;
    q = q + pyz * pyz;
    q = q + pxz * pxz;
    q = q + pyz * pyz;
    q = q + pzz * pzz;
    constexpr float kPi = 3.14159265358979323846F;
    const Real c = static_cast<Real>(16.0F * 1.41421356237309504880F / (3.0F * kPi * kPi));
    return Real(0.5) * (tau + sqrt(tau * tau + c * sqrt(q) / rho));
}

// BGK relaxation of the state f (populations, or scaled deviations for shifted storage S) towards
// the equilibrium of (rho, u): the reference's f - omega (f - f_eq) for unshifted storage, and
// fma(-omega, fma(-w_d, y_d, f'), f') with the pair-shared equilibrium terms for shifted storage.
template <int Q, class S, class Real>
__device__ __forceinline__ void relax(Real (&f)[Q], Real rp, Real rho, Real ux, Real uy, Real uz, Real uu,
                                      Real omega) {
    if constexpr (S::kShifted) {
        // f' - f_eq' = fma(-w_d K, y_d, f') with y_d = fma(rho, x_d, rho'), one rounding.
        const Real d0 = __builtin_fmaf(-scaled_weight<S, Q, 0, Real>(), __builtin_fmaf(rho, -uu, rp), f[0]);
        f[0] = __builtin_fmaf(-omega, d0, f[0]);
        static_for<(Q - 1) / 2>([&](auto P) {
            constexpr int i = 2 * decltype(P)::value + 1;
            const Real cu = Real(3) * cdot<Q, i>(ux, uy, uz);
            const Real common = __builtin_fmaf(Real(0.5) * cu, cu, -uu);
            const Real wk = scaled_weight<S, Q, i, Real>();  // w_i = w_(i+1)
            const Real da = __builtin_fmaf(-wk, __builtin_fmaf(rho, common + cu, rp), f[i]);
            const Real db = __builtin_fmaf(-wk, __builtin_fmaf(rho, common - cu, rp), f[i + 1]);
            f[i] = __builtin_fmaf(-omega, da, f[i]);
            f[i + 1] = __builtin_fmaf(-omega, db, f[i + 1]);
        });
    } else {
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            f[d] = f[d] - omega * (f[d] - equilibrium<S, Q, d>(rho, ux, uy, uz, uu));
        });
    }
}

// Collision of one non-wall cell in registers, on the state representation of storage S
// (populations, or scaled deviations when shifted).  Returns the density used for moving-wall
// momentum corrections (src/lbm/d2q9.cpp:217-223 uses the fluid cell's density).
// Forced selects the Guo-forcing variant at compile time: the unforced variant carries no force
// temporaries and fits twice the waves per SIMD.  With Forced, a zero acceleration still takes the
// unforced arithmetic (the reference's guo_force returns 0 then), so both variants agree bitwise.
// gamma: the partial bounce-back fractions, or null where the cell's segment holds none (the
// blend with gamma = 0 returns the BGK result unchanged, up to the sign of an exact zero).
template <int Q, bool Forced, class S, class Real>
__device__ __forceinline__ Real collide(const Physics<Real>& p, const Real* gamma, const Geometry& g, uint8_t flag,
                                        uint32_t n, uint32_t y, uint32_t z, Real (&f)[Q]) {
    if (flag == kInlet || flag == kOutlet) {
        Real rp, rho, ux;
        if (flag == kInlet) {
            ux = inlet_velocity(p, g, y, z, Lattice<Q>::dims == 3);
            inlet_density<S, Q>(f, ux, rp, rho);
        } else {
            rho = p.rho_out;
            rp = rho - Real(1);
            ux = outlet_velocity<S, Q>(p, f);
        }
        const Real uu = kinetic_term<S>(ux, Real(0), Real(0));
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            f[d] = equilibrium_state<S, Q, d>(rp, rho, ux, Real(0), Real(0), uu);
        });
        return rho;
    }
    Real rp, rho, ux, uy, uz;
    fluid_moments<S, Q>(f, rp, rho, ux, uy, uz);
    Real ax = Real(0), ay = Real(0), az = Real(0);
    if constexpr (Forced) {
        ax = p.ax;
        ay = p.ay;
        az = p.az;
        if (p.lax) {
            ax = ax + p.lax[n];
            ay = ay + p.lax[p.cells + n];
            az = az + p.lax[2U * p.cells + n];
        }
    }
    // The reference always adds a/2 (esoteric_pull.hpp:418-420); u + 0 is u, so the unforced
    // variant skips it.
    if constexpr (Forced) {
        ux = ux + Real(0.5) * ax;
        uy = uy + Real(0.5) * ay;
        uz = uz + Real(0.5) * az;
    }
    const Real uu = kinetic_term<S>(ux, uy, uz);
    Real omega = p.omega;
    if constexpr (Forced && !S::kShifted)
        if (p.les) omega = Real(1) / smagorinsky_tau<Q, S>(f, rho, ux, uy, uz, uu, p.tau);
    Real fin_copy[Q];
    if (gamma) {
        static_for<Q>([&](auto DD) { fin_copy[decltype(DD)::value] = f[decltype(DD)::value]; });
    }
    if (!Forced || (ax == Real(0) && ay == Real(0) && az == Real(0))) {
        relax<Q, S>(f, rp, rho, ux, uy, uz, uu, omega);
    } else {
        const Real fx = rho * ax, fy = rho * ay, fz = rho * az;
        const Real uf = ux * fx + uy * fy + uz * fz;
        const Real wfac = Real(1) - Real(0.5) * omega;
        relax<Q, S>(f, rp, rho, ux, uy, uz, uu, omega);
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            Real src = guo<S, Q, d>(ux, uy, uz, fx, fy, fz, uf, wfac);
            if constexpr (S::kShifted) src = src * static_cast<Real>(S::kScale);
            f[d] = f[d] + src;
        });
    }
    if (gamma) {
        // Partial bounce-back (Walsh, Burwinkle & Saar 2009): blend the BGK result with the
        // pre-collision population of the opposite direction (w_d = w_opp(d): the blend is the
        // same on deviations).
        const Real gm = gamma[n];
        const Real keep = Real(1) - gm;
        Real mixed[Q];
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            if constexpr (S::kShifted) mixed[d] = __builtin_fmaf(keep, f[d], gm * fin_copy[opposite(d)]);
            else mixed[d] = keep * f[d] + gm * fin_copy[opposite(d)];
        });
        static_for<Q>([&](auto DD) { f[decltype(DD)::value] = mixed[decltype(DD)::value]; });
    }
    return rho;
}

// ------------------------------------------------------------------ seeded noise (SplitMix64)
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
template <class Real>
__host__ __device__ __forceinline__ Real uniform01(uint64_t seed_hash, uint64_t cell, uint64_t component) {
    const uint64_t h = splitmix64(seed_hash + 4ULL * cell + component);
    return static_cast<Real>(h >> 40) * static_cast<Real>(0x1p-24);
}

}  // namespace cfd_hip
