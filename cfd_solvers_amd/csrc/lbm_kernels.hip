// lbm_kernels.hip -- gfx950 kernels of the esoteric-pull LBM backend and their launchers.
//
// Hot kernel: stream_collide<Q, T, Odd, Bounded>, one fused in-place pass per time step
// (the single-grid scheme of include/cfd/solvers/lbm/esoteric_pull.hpp:494-514; the reference's
// device sibling is the SYCL kernel at esoteric_pull_sycl.hpp:87-211).  Everything else here runs
// once per setup or per diagnostic: initialisation, bounce-back slot preparation, macroscopic
// fields, FP64 reductions, the index-map probe.
#include "lbm_kernels.hpp"
#include "lbm_cell.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>

namespace cfd_hip {

namespace {

// ------------------------------------------------------------------ hot kernel
template <int Q, class T, bool Odd, bool Bounded, bool Forced>
__device__ __forceinline__ void update_cell(const void* src, void* dst, const Geometry& g,
                                            const Physics<typename Storage<T>::Real>& p,
                                            const typename Storage<T>::Real* gamma, const Neighbourhood& h,
                                            uint8_t flag, uint32_t desc, uint32_t n, uint32_t x, uint32_t y,
                                            uint32_t z) {
    using Real = typename Storage<T>::Real;
    Real fin[Q];
    PairSlots<Q> s;
    load_cell<Q, T, Odd, Bounded>(src, g, desc, h, fin, s);
    const Real rho = collide<Q, Forced, Storage<T>>(p, gamma, g, flag, n, y, z, fin);
    store_cell<Q, T, Odd, Bounded>(dst, g, p, n, x, y, z, rho, fin, s);
}

#if defined(CFD_LBM_WAVES_PER_EU)
#define CFD_LBM_OCCUPANCY __attribute__((amdgpu_waves_per_eu(CFD_LBM_WAVES_PER_EU)))
#else
#define CFD_LBM_OCCUPANCY
#endif

// In place when src == dst (the time step); src != dst writes the next state into another buffer
// (recomputation of stored states for the adjoint).  Every slot the next step reads is written.
template <int Q, class T, bool Odd, bool Bounded, bool Forced>
__global__ __launch_bounds__(256) CFD_LBM_OCCUPANCY void stream_collide(const void* src, void* dst, const Geometry g,
                                                                        const Physics<typename Storage<T>::Real> p) {
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    if constexpr (Bounded) {
        // The 64 lanes of a wavefront lie in one 64-cell row segment (bx >= 64), so its span is
        // wavefront-uniform: one scalar load, and a branch on it costs no lane masking.  A segment
        // whose cells are all plain fluid (span 0) takes the periodic path: no link loads, no
        // bounce-back selects.  Otherwise only the cells inside the span [lo, hi) can have a nonzero
        // link descriptor; the others skip its load.  Segments without a porous cell skip the
        // partial bounce-back (its loads of gamma and its blend) as a whole.
        const uint32_t k = __builtin_amdgcn_readfirstlane((z * g.ny + y) * p.seg_x + (x >> 6));
        using ConstantWords = const __attribute__((address_space(4))) uint32_t*;  // scalar loads
        const uint32_t word = ((ConstantWords)p.segments)[k];
        const auto* gamma = (word & kPorousSegment) != 0U ? p.gamma : nullptr;
        const uint32_t span = word & kSpanBits;
        if (span == 0U) {
            update_cell<Q, T, Odd, false, Forced>(src, dst, g, p, gamma, h, kFluid, 0U, n, x, y, z);
            return;
        }
        const uint32_t xi = x & 63U;
        const uint32_t side = (word >> kXWallShift) & 3U;
        if (!Forced && side != 0U) {  // the forced kernels keep their register budget
            // A resting x wall and its one fluid neighbour (the segments at the ends of a row in a
            // walled box): the neighbour's wall links follow from the lattice, with no descriptor
            // load ahead of the population loads.
            const uint32_t lo = span & 0xFFU;
            if (xi == (side == 1U ? lo : lo + 1U)) return;  // the wall cell
            const uint32_t links = side == 1U ? x_wall_links<Q>(-1) : x_wall_links<Q>(1);
            const uint32_t desc = xi == (side == 1U ? lo + 1U : lo) ? links : 0U;
            update_cell<Q, T, Odd, true, Forced>(src, dst, g, p, gamma, h, kFluid, desc, n, x, y, z);
            return;
        }
        const uint32_t desc = (xi >= (span & 0xFFU) && xi < (span >> 8)) ? p.links[n] : 0U;
        const uint32_t type = (desc >> kTypeShift) & 3U;
        if (type == kTypeWall) return;
        update_cell<Q, T, Odd, true, Forced>(src, dst, g, p, gamma, h, static_cast<uint8_t>(type), desc, n, x, y, z);
    } else {
        update_cell<Q, T, Odd, false, Forced>(src, dst, g, p, p.gamma, h, kFluid, 0U, n, x, y, z);
    }
}

// The span of each 64-cell row segment: cells [lo, hi) of the segment that are not plain fluid
// with a plain-fluid 3x3x3 neighbourhood (periodic), stored as lo | hi << 8; 0 when every cell is
// (the interior segments).  Cells outside the span have link descriptor 0.  A two-cell span that
// is a resting x wall and its fluid neighbour also records that pattern (kXWallShift).
__global__ __launch_bounds__(256) void mark_segments(const uint8_t* __restrict__ flags, const Geometry g,
                                                     uint32_t seg_x, uint32_t* __restrict__ segments) {
    const uint32_t total = seg_x * g.ny * g.nz;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const uint32_t sx = k % seg_x;
    const uint32_t r = k / seg_x;
    const uint32_t y = r % g.ny;
    const uint32_t z = r / g.ny;
    const uint32_t xs = sx * 64U, xe = min(xs + 64U, g.nx);
    uint32_t rows[9];
    for (int dz = -1, i = 0; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy, ++i) {
            const uint32_t zz = (z + g.nz + static_cast<uint32_t>(dz)) % g.nz;
            const uint32_t yy = (y + g.ny + static_cast<uint32_t>(dy)) % g.ny;
            rows[i] = (zz * g.ny + yy) * g.nx;
        }
    // column(x): the 3x3 flags at x in the rows above are all plain fluid.
    auto column = [&](uint32_t xx) {
        for (int i = 0; i < 9; ++i)
            if (flags[rows[i] + xx] != kFluid) return false;
        return true;
    };
    bool left = column(xs == 0U ? g.nx - 1U : xs - 1U), mid = column(xs);
    uint32_t lo = 64U, hi = 0U;
    for (uint32_t x = xs; x < xe; ++x) {
        const bool right = column(x + 1U == g.nx ? 0U : x + 1U);
        if (!(left && mid && right)) {
            lo = min(lo, x - xs);
            hi = x - xs + 1U;
        }
        left = mid;
        mid = right;
    }
    // The x-wall pattern (kXWallShift): a resting wall column and a plain-fluid column on the
    // other side of the span's fluid cell.
    uint32_t side = 0U;
    if (hi == lo + 2U) {
        auto wall_column = [&](uint32_t xx) {
            for (int i = 0; i < 9; ++i)
                if (flags[rows[i] + xx] != kWall) return false;
            return true;
        };
        const uint32_t xa = xs + lo, xb = xa + 1U;
        if (wall_column(xa) && column(xb) && column(xb + 1U == g.nx ? 0U : xb + 1U)) side = 1U;
        else if (wall_column(xb) && column(xa) && column(xa == 0U ? g.nx - 1U : xa - 1U)) side = 2U;
    }
    segments[k] = hi == 0U ? 0U : lo | (hi << 8) | (side << kXWallShift);
}

// Sets kPorousSegment in the word of every row segment that holds a cell with gamma != 0 (after
// mark_segments has written the spans).  Launched on the step's grid: each wavefront covers one
// segment (bx >= 64), and its first lane, x = 64 * segment, owns the word.
template <class Real>
__global__ __launch_bounds__(256) void mark_porous_segments(const Real* __restrict__ gamma, const Geometry g,
                                                            uint32_t seg_x, uint32_t* __restrict__ segments) {
    uint32_t x, y, z;
    const bool inside = cell_coords(g, x, y, z);
    const bool porous = inside && gamma[(z * g.ny + y) * g.nx + x] != Real(0);
    const uint64_t any = __ballot(porous);
    if (inside && (x & 63U) == 0U && any != 0U) segments[(z * g.ny + y) * seg_x + (x >> 6)] |= kPorousSegment;
}

// Link descriptor of every cell from the flags (layout in lbm_device.hpp).
template <int Q>
__global__ __launch_bounds__(256) void build_links(const uint8_t* __restrict__ flags, const Geometry g,
                                                   uint32_t* __restrict__ links, uint32_t* __restrict__ link_vel) {
    using L = Lattice<Q>;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    const uint8_t own = flags[n];
    const uint32_t type = is_wall(own) ? kTypeWall : (own == kInlet ? 1U : (own == kOutlet ? 2U : 0U));
    constexpr int kW = PairSlots<Q>::kVelWords;
    uint32_t desc = type << kTypeShift;
    uint32_t vel[kW] = {};
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint8_t fu = flags[h.lin(-L::cx[i], -L::cy[i], -L::cz[i])];
        const uint8_t fd = flags[h.lin(L::cx[i], L::cy[i], L::cz[i])];
        if (is_wall(fu)) {
            desc |= 1U << (i - 1);
            vel[(i - 1) / 10] |= static_cast<uint32_t>(fu & 7U) << (3 * ((i - 1) % 10));
        }
        if (is_wall(fd)) {
            desc |= 1U << i;
            vel[i / 10] |= static_cast<uint32_t>(fd & 7U) << (3 * (i % 10));
        }
    });
    bool moving = false;
    for (int w = 0; w < kW; ++w) {
        moving = moving || vel[w] != 0U;
        link_vel[static_cast<uint64_t>(n) * kW + w] = vel[w];
    }
    if (moving) desc |= 1U << kMovingBit;
    links[n] = desc;
}

// ------------------------------------------------------------------ initialisation
template <class Real> struct InitArgs {
    int mode;  // 0 uniform, 1 Taylor-Green tables, 2 per-cell fields
    Real rho, ux, uy, uz;
    Real amplitude, noise;
    uint64_t seed_hash;
    const Real* tables;  // sx[nx], cx[nx], sy[ny], cy[ny], zf[nz]
    const Real* fields;  // rho, ux, uy, uz planes of `cells`
};

// esoteric_pull.hpp:312-333: f = f_eq(rho, u) written in the even layout; walls at rest.
template <int Q, class T>
__global__ __launch_bounds__(256) void initialize(void* __restrict__ f, const Geometry g,
                                                  const InitArgs<typename Storage<T>::Real> a,
                                                  const uint8_t* __restrict__ flags, uint32_t cells) {
    using L = Lattice<Q>;
    using S = Storage<T>;
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);  // flags, fields, noise
    const uint32_t np = h.cell(0, 0, 0);  // population slots
    Real rho = a.rho, ux = a.ux, uy = a.uy, uz = a.uz;
    if (a.mode == 1) {
        const Real* sx = a.tables;
        const Real* cx = sx + g.nx;
        const Real* sy = cx + g.nx;
        const Real* cy = sy + g.ny;
        const Real* zf = cy + g.ny;
        rho = Real(1);
        ux = a.amplitude * sx[x] * cy[y] * zf[z];
        uy = -a.amplitude * cx[x] * sy[y] * zf[z];
        uz = Real(0);
        if (a.noise != Real(0)) {
            ux = ux + a.noise * (Real(2) * uniform01<Real>(a.seed_hash, n, 0) - Real(1));
            uy = uy + a.noise * (Real(2) * uniform01<Real>(a.seed_hash, n, 1) - Real(1));
            if constexpr (L::dims == 3) uz = uz + a.noise * (Real(2) * uniform01<Real>(a.seed_hash, n, 2) - Real(1));
        }
    } else if (a.mode == 2) {
        rho = a.fields[n];
        ux = a.fields[cells + n];
        uy = a.fields[2U * cells + n];
        uz = a.fields[3U * cells + n];
    }
    if (flags && is_wall(flags[n])) ux = uy = uz = Real(0);
    const Real uu = kinetic_term<S>(ux, uy, uz);
    const Real rp = rho - Real(1);
    plane<T>(f, g, 0)[np] = S::template encode<Q, 0>(equilibrium_state<S, Q, 0>(rp, rho, ux, uy, uz, uu));
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t up = h.cell(-L::cx[i], -L::cy[i], -L::cz[i]);
        plane<T>(f, g, i)[up] = S::template encode<Q, i>(equilibrium_state<S, Q, i>(rp, rho, ux, uy, uz, uu));
        plane<T>(f, g, i + 1)[np] =
            S::template encode<Q, i + 1>(equilibrium_state<S, Q, i + 1>(rp, rho, ux, uy, uz, uu));
    });
}

// Writes caller-supplied logical populations (q planes of `cells`) at the given parity.
template <int Q, class T, bool Odd>
__global__ __launch_bounds__(256) void upload_logical(void* __restrict__ f, const Geometry g,
                                                      const typename Storage<T>::Real* __restrict__ in,
                                                      uint32_t cells) {
    using L = Lattice<Q>;
    using S = Storage<T>;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0), np = h.cell(0, 0, 0);
    plane<T>(f, g, 0)[np] = S::template encode<Q, 0>(state_population<S, Q, 0>(in[n]));
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t up = h.cell(-L::cx[i], -L::cy[i], -L::cz[i]);
        const auto va = state_population<S, Q, i>(in[static_cast<uint64_t>(i) * cells + n]);
        const auto vb = state_population<S, Q, i + 1>(in[static_cast<uint64_t>(i + 1) * cells + n]);
        if constexpr (Odd) {
            plane<T>(f, g, i + 1)[up] = S::template encode<Q, i + 1>(va);
            plane<T>(f, g, i)[np] = S::template encode<Q, i>(vb);
        } else {
            plane<T>(f, g, i)[up] = S::template encode<Q, i>(va);
            plane<T>(f, g, i + 1)[np] = S::template encode<Q, i + 1>(vb);
        }
    });
}

// Bounce-back slot preparation for a change of flags at a given parity.
//   ToBounceBack = true : copy normal slot -> opposite-parity slot for every wall link of a
//                         non-wall cell (after initialisation or upload under `flags`).
//   ToBounceBack = false: copy opposite-parity slot -> normal slot (before `flags` are replaced),
//                         which restores the plain esoteric layout of the fluid populations.
template <int Q, class T, bool Odd, bool ToBounceBack>
__global__ __launch_bounds__(256) void bounce_back_slots(void* __restrict__ f, const Geometry g,
                                                         const uint8_t* __restrict__ flags) {
    using L = Lattice<Q>;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.cell(0, 0, 0);
    if (is_wall(flags[h.lin(0, 0, 0)])) return;
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t up = h.cell(-L::cx[i], -L::cy[i], -L::cz[i]);
        if (is_wall(flags[h.lin(-L::cx[i], -L::cy[i], -L::cz[i])])) {
            T* normal = Odd ? plane<T>(f, g, i + 1) : plane<T>(f, g, i);
            T* other = Odd ? plane<T>(f, g, i) : plane<T>(f, g, i + 1);
            if (ToBounceBack) other[up] = normal[up];
            else normal[up] = other[up];
        }
        if (is_wall(flags[h.lin(L::cx[i], L::cy[i], L::cz[i])])) {
            T* normal = Odd ? plane<T>(f, g, i) : plane<T>(f, g, i + 1);
            T* other = Odd ? plane<T>(f, g, i + 1) : plane<T>(f, g, i);
            if (ToBounceBack) other[n] = normal[n];
            else normal[n] = other[n];
        }
    });
}

// Density of the cells that become walls, from their plain (normal-slot) populations; cells
// that already were walls keep their frozen density.
template <int Q, class T, bool Odd>
__global__ __launch_bounds__(256) void freeze_wall_density(const void* __restrict__ f, const Geometry g,
                                                           const uint8_t* __restrict__ flags,
                                                           const uint8_t* __restrict__ old_flags,
                                                           typename Storage<T>::Real* __restrict__ solid_rho) {
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    if (!is_wall(flags[n])) return;
    if (old_flags && is_wall(old_flags[n])) return;
    Real fin[Q];
    PairSlots<Q> s;
    load_cell<Q, T, Odd, false>(f, g, 0U, h, fin, s);
    Real rho = Real(0);
    static_for<Q>([&](auto DD) { rho += fin[decltype(DD)::value]; });
    if constexpr (Storage<T>::kShifted) rho = Real(1) + rho * static_cast<Real>(Storage<T>::kInvScale);
    solid_rho[n] = rho;
}

// Cells that stop being walls restart from rest equilibrium at their frozen density, written in
// the plain layout (wall cells keep no populations of their own while they are walls).
template <int Q, class T, bool Odd>
__global__ __launch_bounds__(256) void thaw_walls(void* __restrict__ f, const Geometry g,
                                                  const uint8_t* __restrict__ old_flags,
                                                  const uint8_t* __restrict__ flags,
                                                  const typename Storage<T>::Real* __restrict__ solid_rho) {
    using L = Lattice<Q>;
    using S = Storage<T>;
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0), np = h.cell(0, 0, 0);
    if (!is_wall(old_flags[n]) || (flags && is_wall(flags[n]))) return;
    const Real rho = solid_rho[n];
    const Real rp = rho - Real(1);
    const Real zero = Real(0);
    plane<T>(f, g, 0)[np] = S::template encode<Q, 0>(equilibrium_state<S, Q, 0>(rp, rho, zero, zero, zero, zero));
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t up = h.cell(-L::cx[i], -L::cy[i], -L::cz[i]);
        const Real va = equilibrium_state<S, Q, i>(rp, rho, zero, zero, zero, zero);
        const Real vb = equilibrium_state<S, Q, i + 1>(rp, rho, zero, zero, zero, zero);
        if constexpr (Odd) {
            plane<T>(f, g, i + 1)[up] = S::template encode<Q, i + 1>(va);
            plane<T>(f, g, i)[np] = S::template encode<Q, i>(vb);
        } else {
            plane<T>(f, g, i)[up] = S::template encode<Q, i>(va);
            plane<T>(f, g, i + 1)[np] = S::template encode<Q, i + 1>(vb);
        }
    });
}

// ------------------------------------------------------------------ macroscopic fields
template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void macroscopic(const void* __restrict__ f, const Geometry g,
                                                   const Physics<typename Storage<T>::Real> p,
                                                   typename Storage<T>::Real* __restrict__ out) {
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    const uint32_t cells = p.cells;
    uint8_t flag = kFluid;
    if constexpr (Bounded) {
        flag = p.flags[n];
        if (is_wall(flag)) {
            out[n] = p.solid_rho[n];
            out[cells + n] = Real(0);
            out[2U * cells + n] = Real(0);
            out[3U * cells + n] = Real(0);
            return;
        }
    }
    Real fin[Q];
    PairSlots<Q> s;
    load_cell<Q, T, Odd, Bounded>(f, g, link_desc<Bounded>(p.links, n), h, fin, s);
    Real rho, ux, uy, uz;
    cell_state<Q, Storage<T>>(p, g, flag, n, y, z, fin, rho, ux, uy, uz);
    out[n] = rho;
    out[cells + n] = ux;
    out[2U * cells + n] = uy;
    out[3U * cells + n] = uz;
}

template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void download_logical(const void* __restrict__ f, const Geometry g,
                                                        const Physics<typename Storage<T>::Real> p,
                                                        typename Storage<T>::Real* __restrict__ out) {
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    if constexpr (Bounded) {
        if (is_wall(p.flags[n])) {
            const Real rho = p.solid_rho[n];
            static_for<Q>([&](auto DD) {
                constexpr int d = decltype(DD)::value;
                out[static_cast<uint64_t>(d) * p.cells + n] = lattice_weight<Storage<T>, Q, d, Real>() * rho;
            });
            return;
        }
    }
    Real fin[Q];
    PairSlots<Q> s;
    load_cell<Q, T, Odd, Bounded>(f, g, link_desc<Bounded>(p.links, n), h, fin, s);
    static_for<Q>([&](auto DD) {
        constexpr int d = decltype(DD)::value;
        out[static_cast<uint64_t>(d) * p.cells + n] = full_population<Storage<T>, Q, d>(fin[d]);
    });
}

// Every slot of the in-place array, decoded to a full population, in the order d * cells + n of
// the reference's raw storage (esoteric_pull.hpp:276-278): plane d of cell n holds whichever
// population the esoteric scheme left there at this parity.
template <int Q, class T>
__global__ __launch_bounds__(256) void download_slots(const void* __restrict__ f, const Geometry g, uint32_t cells,
                                                      typename Storage<T>::Real* __restrict__ out) {
    using S = Storage<T>;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const uint32_t row = z * g.ny + y;
    const uint32_t n = row * g.nx + x;
    const uint32_t e = row_base(g, row) + x;
    static_for<Q>([&](auto DD) {
        constexpr int d = decltype(DD)::value;
        // The slot of plane d holds a population of direction d or of its opposite, which has the
        // same weight (hence the same shift).
        out[static_cast<uint64_t>(d) * cells + n] =
            full_population<S, Q, d>(S::template decode<Q, d>(load_at<T>(plane<T>(f, g, d), e * sizeof(T))));
    });
}

// ------------------------------------------------------------------ FP64 reductions
// Per block: mass (sum of per-cell population sums; wall cells contribute their frozen density),
// kinetic energy 0.5 rho |u|^2 and max |u| over non-wall cells.  Fixed tree order: deterministic.
template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void diagnostics_partial(const void* __restrict__ f, const Geometry g,
                                                           const Physics<typename Storage<T>::Real> p,
                                                           double* __restrict__ partial) {
    using Real = typename Storage<T>::Real;
    __shared__ double sm[3][256];
    const uint32_t tid = threadIdx.y * blockDim.x + threadIdx.x;
    double mass = 0.0, ke = 0.0, vmax = 0.0;
    uint32_t x, y, z;
    if (cell_coords(g, x, y, z)) {
        const Neighbourhood h = neighbourhood(g, x, y, z);
        const uint32_t n = h.lin(0, 0, 0);
        uint8_t flag = kFluid;
        bool wall = false;
        if constexpr (Bounded) {
            flag = p.flags[n];
            wall = is_wall(flag);
        }
        if (wall) {
            mass = static_cast<double>(p.solid_rho[n]);
        } else {
            Real fin[Q];
            PairSlots<Q> s;
            load_cell<Q, T, Odd, Bounded>(f, g, link_desc<Bounded>(p.links, n), h, fin, s);
            Real sum = Real(0);
            static_for<Q>([&](auto DD) { sum += fin[decltype(DD)::value]; });
            // Shifted storage: the sum of deviations is rho - 1 (added in FP64).
            mass = Storage<T>::kShifted ? 1.0 + static_cast<double>(sum * static_cast<Real>(Storage<T>::kInvScale))
                                        : static_cast<double>(sum);
            Real rho, ux, uy, uz;
            cell_state<Q, Storage<T>>(p, g, flag, n, y, z, fin, rho, ux, uy, uz);
            ke = static_cast<double>(rho) * (static_cast<double>(ux) * ux + static_cast<double>(uy) * uy +
                                             static_cast<double>(uz) * uz);
            vmax = static_cast<double>(sqrt(ux * ux + uy * uy + uz * uz));
        }
    }
    sm[0][tid] = mass;
    sm[1][tid] = ke;
    sm[2][tid] = vmax;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            sm[0][tid] += sm[0][tid + w];
            sm[1][tid] += sm[1][tid + w];
            sm[2][tid] = fmax(sm[2][tid], sm[2][tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        partial[3U * blockIdx.x + 0] = sm[0][0];
        partial[3U * blockIdx.x + 1] = sm[1][0];
        partial[3U * blockIdx.x + 2] = sm[2][0];
    }
}

__global__ __launch_bounds__(256) void diagnostics_final(const double* __restrict__ partial, uint32_t count,
                                                         double* __restrict__ out) {
    __shared__ double sm[3][256];
    const uint32_t tid = threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    for (uint32_t k = tid; k < count; k += 256U) {
        a += partial[3U * k + 0];
        b += partial[3U * k + 1];
        c = fmax(c, partial[3U * k + 2]);
    }
    sm[0][tid] = a;
    sm[1][tid] = b;
    sm[2][tid] = c;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            sm[0][tid] += sm[0][tid + w];
            sm[1][tid] += sm[1][tid + w];
            sm[2][tid] = fmax(sm[2][tid], sm[2][tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = sm[0][0];
        out[1] = 0.5 * sm[1][0];
        out[2] = sm[2][0];
    }
}

// ------------------------------------------------------------------ index-map probe
// The slots are read off the hot kernel's own load_cell / store_cell: every slot of a plane array
// holds its identifier d*cells + n, load_cell returns the identifiers it read, and store_cell
// writes the identifier of each (direction, cell) it stores into the slot it stores to.
__global__ __launch_bounds__(256) void fill_ids(uint32_t* __restrict__ ids, const Geometry g, uint32_t cells,
                                                int q, uint32_t empty) {
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const uint32_t n = (z * g.ny + y) * g.nx + x;
    const uint32_t np = row_base(g, z * g.ny + y) + x;
    for (int d = 0; d < q; ++d) ids[static_cast<uint64_t>(d) * g.stride + np] = empty ? empty : d * cells + n;
}

template <int Q, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void probe_load(const uint32_t* __restrict__ ids, const Geometry g,
                                                  const uint8_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ links, uint32_t cells,
                                                  int64_t* __restrict__ load) {
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    if (Bounded && is_wall(flags[n])) {
        for (int d = 0; d < Q; ++d) load[static_cast<uint64_t>(d) * cells + n] = -1;
        return;
    }
    uint32_t fin[Q];
    PairSlots<Q> s;
    load_cell<Q, uint32_t, Odd, Bounded>(ids, g, link_desc<Bounded>(links, n), h, fin, s);
    for (int d = 0; d < Q; ++d) load[static_cast<uint64_t>(d) * cells + n] = fin[d];
}

template <int Q, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void probe_store(uint32_t* __restrict__ ids, const Geometry g,
                                                   const Physics<uint32_t> p) {
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    if (Bounded && is_wall(p.flags[n])) return;
    uint32_t fin[Q];
    PairSlots<Q> s;
    load_cell<Q, uint32_t, Odd, Bounded>(ids, g, link_desc<Bounded>(p.links, n), h, fin, s);  // records the slots
    uint32_t fout[Q];
    for (int d = 0; d < Q; ++d) fout[d] = d * p.cells + n;
    store_cell<Q, uint32_t, Odd, Bounded>(ids, g, p, n, x, y, z, 1U, fout, s);
}

// ------------------------------------------------------------------ z-slab halo messages
// A lattice split into z-slabs runs the periodic kernel on each slab; the z wrap then puts the slots
// a boundary cell shares with the neighbouring slab into the slab's own opposite face layer.  After
// every step each slab sends 5 (D3Q19) / 9 (D3Q27) slot planes of those layers up and as many down,
// and writes what it receives into the same slots (the receiver's slots sit where the sender's do):
//   pair (i, i+1) with c_i,z = +1, top layer z = nz-1:    up: plane i+1 after an even step, i after an odd one
//                                                        down: the other plane of the pair
//   pair (i, i+1) with c_i,z = -1, bottom layer z = 0:     up: plane i after an even step, i+1 after an odd one
//                                                        down: the other plane
// (the slot ownership of esoteric_pull.hpp:349-382 across the cut; the crossing populations of
// include/cfd/distributed/selective_halo.hpp:15-21).  Message element k*nx*ny + y*nx + x is crossing
// pair k (in pair order) at (x, y).  With one slab that is its own neighbour, unpacking its own
// messages changes nothing.
template <int Q> __host__ __device__ constexpr int crossing_pairs() {
    int n = 0;
    for (int i = 1; i < Q; i += 2)
        if (Lattice<Q>::cz[i] != 0) ++n;
    return n;
}

// One launch covers both messages: elements [0, E) of the up message, then [E, 2E) of the down one.
// Each thread moves V consecutive elements of a row (V = 16 bytes' worth when nx allows 16-byte
// aligned chunks, else 1): rows are contiguous in the population planes and in the message alike.
template <int Q, class T, int V>
__global__ __launch_bounds__(256) void halo_planes(void* __restrict__ f, const Geometry g, uint32_t odd,
                                                   uint32_t unpack, T* __restrict__ up_msg, T* __restrict__ down_msg) {
    using L = Lattice<Q>;
    constexpr uint32_t kPairs = static_cast<uint32_t>(crossing_pairs<Q>());
    if constexpr (kPairs == 0) return;
    const uint32_t layer = g.nx * g.ny;
    const uint64_t per = static_cast<uint64_t>(kPairs) * layer / V;  // chunks per message
    const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= 2U * per) return;
    const uint32_t down = gid >= per ? 1U : 0U;
    const uint64_t id = (down ? gid - per : gid) * V;  // first element of the chunk
    T* __restrict__ msg = down ? down_msg : up_msg;
    const uint32_t k = static_cast<uint32_t>(id / layer), c = static_cast<uint32_t>(id % layer);
    const uint32_t x = c % g.nx, y = c / g.nx;
    int i = 1;
    for (uint32_t seen = 0;; i += 2)
        if (L::cz[i] != 0 && seen++ == k) break;
    const bool top = L::cz[i] > 0;
    // up, top layer: i+1 even / i odd; up, bottom layer: i even / i+1 odd; down: the other plane.
    const bool second = top ? (odd == 0U) : (odd != 0U);
    const int d = (second != (down != 0U)) ? i + 1 : i;
    const uint32_t z = top ? g.nz - 1U : 0U;
    T* slot = plane<T>(f, g, d) + row_base(g, z * g.ny + y) + x;
    if constexpr (V == 1) {
        if (unpack) *slot = msg[id];
        else msg[id] = *slot;
    } else {
        using Chunk = unsigned int __attribute__((ext_vector_type(4)));
        static_assert(V * sizeof(T) == sizeof(Chunk), "16-byte chunks");
        Chunk* a = reinterpret_cast<Chunk*>(slot);
        Chunk* b = reinterpret_cast<Chunk*>(msg + id);
        if (unpack) *a = *b;
        else *b = *a;
    }
}

// ------------------------------------------------------------------ launchers
inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline dim3 block_of(const Geometry& g) { return dim3(g.bx, g.by, 1); }
inline dim3 grid_of(const Geometry& g) { return dim3(g.tiles, 1, 1); }

template <int Q, class T> struct Ops final : KernelOps {
    using Real = typename Storage<T>::Real;

    template <bool Odd, bool Bounded, bool Forced>
    static void step_t(const LaunchState& st, const void* src, void* dst) {
        hipLaunchKernelGGL((stream_collide<Q, T, Odd, Bounded, Forced>), grid_of(st.geo), block_of(st.geo),
                           Bounded ? st.step_lds : st.step_lds_periodic, st.stream, src, dst, st.geo,
                           *static_cast<const Physics<Real>*>(st.physics));
    }
    void step(const LaunchState& st, bool odd, bool bounded, bool forced, const void* src, void* dst) const override {
        per_parity(st, odd, bounded, [&](auto O, auto B) {
            if (forced) step_t<decltype(O)::value, decltype(B)::value, true>(st, src, dst);
            else step_t<decltype(O)::value, decltype(B)::value, false>(st, src, dst);
        });
        check(hipGetLastError(), "stream_collide launch");
    }
    void initialize(const LaunchState& st, const void* init_args, const uint8_t* flags) const override {
        hipLaunchKernelGGL((cfd_hip::initialize<Q, T>), grid_of(st.geo), block_of(st.geo), 0, st.stream, st.f, st.geo,
                           *static_cast<const InitArgs<Real>*>(init_args), flags, st.cells);
        check(hipGetLastError(), "initialize launch");
    }
    void upload(const LaunchState& st, bool odd, const void* logical) const override {
        if (odd)
            hipLaunchKernelGGL((upload_logical<Q, T, true>), grid_of(st.geo), block_of(st.geo), 0, st.stream, st.f,
                               st.geo, static_cast<const Real*>(logical), st.cells);
        else
            hipLaunchKernelGGL((upload_logical<Q, T, false>), grid_of(st.geo), block_of(st.geo), 0, st.stream, st.f,
                               st.geo, static_cast<const Real*>(logical), st.cells);
        check(hipGetLastError(), "upload launch");
    }
    void bounce_back(const LaunchState& st, bool odd, bool to_bounce_back, const uint8_t* flags) const override {
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid_of(st.geo), block_of(st.geo), 0, st.stream, st.f, st.geo, flags);
        };
        if (odd) to_bounce_back ? go(bounce_back_slots<Q, T, true, true>) : go(bounce_back_slots<Q, T, true, false>);
        else to_bounce_back ? go(bounce_back_slots<Q, T, false, true>) : go(bounce_back_slots<Q, T, false, false>);
        check(hipGetLastError(), "bounce_back_slots launch");
    }
    void freeze(const LaunchState& st, bool odd, const uint8_t* flags, const uint8_t* old_flags,
                void* solid_rho) const override {
        if (odd)
            hipLaunchKernelGGL((freeze_wall_density<Q, T, true>), grid_of(st.geo), block_of(st.geo), 0, st.stream,
                               st.f, st.geo, flags, old_flags, static_cast<Real*>(solid_rho));
        else
            hipLaunchKernelGGL((freeze_wall_density<Q, T, false>), grid_of(st.geo), block_of(st.geo), 0, st.stream,
                               st.f, st.geo, flags, old_flags, static_cast<Real*>(solid_rho));
        check(hipGetLastError(), "freeze_wall_density launch");
    }
    void segments(const LaunchState& st, const uint8_t* flags, uint32_t seg_x, uint32_t* out) const override {
        const uint32_t total = seg_x * st.geo.ny * st.geo.nz;
        hipLaunchKernelGGL(mark_segments, dim3((total + 255U) / 256U), dim3(256), 0, st.stream, flags, st.geo, seg_x,
                           out);
        check(hipGetLastError(), "mark_segments launch");
    }
    void porous_segments(const LaunchState& st, const void* gamma, uint32_t seg_x, uint32_t* segments) const override {
        hipLaunchKernelGGL(mark_porous_segments<Real>, grid_of(st.geo), block_of(st.geo), 0, st.stream,
                           static_cast<const Real*>(gamma), st.geo, seg_x, segments);
        check(hipGetLastError(), "mark_porous_segments launch");
    }
    void thaw(const LaunchState& st, bool odd, const uint8_t* old_flags, const uint8_t* flags,
              const void* solid_rho) const override {
        if (odd)
            hipLaunchKernelGGL((thaw_walls<Q, T, true>), grid_of(st.geo), block_of(st.geo), 0, st.stream, st.f, st.geo,
                               old_flags, flags, static_cast<const Real*>(solid_rho));
        else
            hipLaunchKernelGGL((thaw_walls<Q, T, false>), grid_of(st.geo), block_of(st.geo), 0, st.stream, st.f,
                               st.geo, old_flags, flags, static_cast<const Real*>(solid_rho));
        check(hipGetLastError(), "thaw_walls launch");
    }
    template <class K> static void per_parity(const LaunchState& st, bool odd, bool bounded, K k) {
        if (odd) bounded ? k(std::true_type{}, std::true_type{}) : k(std::true_type{}, std::false_type{});
        else bounded ? k(std::false_type{}, std::true_type{}) : k(std::false_type{}, std::false_type{});
    }
    void macroscopic(const LaunchState& st, bool odd, bool bounded, void* out) const override {
        per_parity(st, odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((cfd_hip::macroscopic<Q, T, decltype(O)::value, decltype(B)::value>), grid_of(st.geo),
                               block_of(st.geo), 0, st.stream, st.f, st.geo,
                               *static_cast<const Physics<Real>*>(st.physics), static_cast<Real*>(out));
        });
        check(hipGetLastError(), "macroscopic launch");
    }
    void download(const LaunchState& st, bool odd, bool bounded, void* out) const override {
        per_parity(st, odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((download_logical<Q, T, decltype(O)::value, decltype(B)::value>), grid_of(st.geo),
                               block_of(st.geo), 0, st.stream, st.f, st.geo,
                               *static_cast<const Physics<Real>*>(st.physics), static_cast<Real*>(out));
        });
        check(hipGetLastError(), "download_logical launch");
    }
    void slots(const LaunchState& st, void* out) const override {
        hipLaunchKernelGGL((download_slots<Q, T>), grid_of(st.geo), block_of(st.geo), 0, st.stream, st.f, st.geo,
                           st.cells, static_cast<Real*>(out));
        check(hipGetLastError(), "download_slots launch");
    }
    void diagnostics(const LaunchState& st, bool odd, bool bounded, double* partial, double* out) const override {
        per_parity(st, odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((diagnostics_partial<Q, T, decltype(O)::value, decltype(B)::value>), grid_of(st.geo),
                               block_of(st.geo), 0, st.stream, st.f, st.geo,
                               *static_cast<const Physics<Real>*>(st.physics), partial);
        });
        check(hipGetLastError(), "diagnostics_partial launch");
        hipLaunchKernelGGL(diagnostics_final, dim3(1), dim3(256), 0, st.stream, partial, st.geo.tiles, out);
        check(hipGetLastError(), "diagnostics_final launch");
    }
    size_t halo_elems(const LaunchState& st) const override {
        return static_cast<size_t>(crossing_pairs<Q>()) * st.geo.nx * st.geo.ny;
    }
    void halo(const LaunchState& st, bool odd, bool unpack, void* up, void* down) const override {
        const uint64_t n = 2U * halo_elems(st);
        if (n == 0) return;
        // 16-byte chunks when every row starts 16-byte aligned in the planes (the plane bases are
        // 256-byte aligned and a row starts nx elements after the previous one) and in the
        // messages (both 16-byte aligned).
        constexpr uint32_t kV = 16U / static_cast<uint32_t>(sizeof(T));
        const bool wide = kV > 1 && (st.geo.nx % kV) == 0 && (reinterpret_cast<uintptr_t>(st.f) % 16U) == 0 &&
                          (reinterpret_cast<uintptr_t>(up) % 16U) == 0 && (reinterpret_cast<uintptr_t>(down) % 16U) == 0 &&
                          (st.geo.stride % kV) == 0 && (st.geo.row_shift == 31 || (st.geo.block_elems % kV) == 0);
        const uint64_t threads = wide ? n / kV : n;
        const uint64_t blocks = (threads + 255) / 256;
        const dim3 grid(static_cast<uint32_t>(blocks));
        if (wide)
            hipLaunchKernelGGL((halo_planes<Q, T, kV>), grid, dim3(256), 0, st.stream, st.f, st.geo, odd ? 1U : 0U,
                               unpack ? 1U : 0U, static_cast<T*>(up), static_cast<T*>(down));
        else
            hipLaunchKernelGGL((halo_planes<Q, T, 1>), grid, dim3(256), 0, st.stream, st.f, st.geo, odd ? 1U : 0U,
                               unpack ? 1U : 0U, static_cast<T*>(up), static_cast<T*>(down));
        check(hipGetLastError(), "halo_planes launch");
    }
    void links(const LaunchState& st, const uint8_t* flags, uint32_t* out, uint32_t* vel) const override {
        hipLaunchKernelGGL((build_links<Q>), grid_of(st.geo), block_of(st.geo), 0, st.stream, flags, st.geo, out,
                           vel);
        check(hipGetLastError(), "build_links launch");
    }
    void index_map(const LaunchState& st, bool odd, bool bounded, const uint8_t* flags, const uint32_t* links,
                   uint32_t* ids, int64_t* load) const override {
        hipLaunchKernelGGL(fill_ids, grid_of(st.geo), block_of(st.geo), 0, st.stream, ids, st.geo, st.cells, Q, 0U);
        per_parity(st, odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((probe_load<Q, decltype(O)::value, decltype(B)::value>), grid_of(st.geo),
                               block_of(st.geo), 0, st.stream, ids, st.geo, flags, links, st.cells, load);
        });
        check(hipGetLastError(), "probe_load launch");
        hipLaunchKernelGGL(fill_ids, grid_of(st.geo), block_of(st.geo), 0, st.stream, ids, st.geo, st.cells, Q,
                           0xFFFFFFFFU);
        Physics<uint32_t> p{};
        p.flags = flags;
        p.links = links;
        p.cells = st.cells;  // probe flags carry no wall velocities: link_vel is never read
        per_parity(st, odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((probe_store<Q, decltype(O)::value, decltype(B)::value>), grid_of(st.geo),
                               block_of(st.geo), 0, st.stream, ids, st.geo, p);
        });
        check(hipGetLastError(), "probe_store launch");
    }
};

template <int Q> const KernelOps* ops_for_storage(int storage) {
    static const Ops<Q, float> f32;
    static const Ops<Q, _Float16> f16;
    static const Ops<Q, double> f64;
    static const Ops<Q, F32Dev> f32s;
    static const Ops<Q, F64Exact> f64x;
    switch (storage) {
    case 0: return &f32;
    case 1: return &f16;
    case 2: return &f64;
    case 3: return &f32s;
    case 4: return &f64x;
    default: throw std::invalid_argument("storage must be FP32 (0), FP16S (1), FP64 (2), FP32S (3) or FP64X (4)");
    }
}

}  // namespace

const KernelOps* kernel_ops(int q, int storage) {
    switch (q) {
    case 9: return ops_for_storage<9>(storage);
    case 19: return ops_for_storage<19>(storage);
    case 27: return ops_for_storage<27>(storage);
    default: throw std::invalid_argument("lattice must be D2Q9 (9), D3Q19 (19) or D3Q27 (27)");
    }
}

size_t init_args_size(int storage) {
    return storage == 2 || storage == 4 ? sizeof(InitArgs<double>) : sizeof(InitArgs<float>);
}

void fill_init_args(int storage, void* out, int mode, double rho, double ux, double uy, double uz, double amplitude,
                    double noise, uint64_t seed, const void* tables, const void* fields) {
    auto fill = [&](auto* a) {
        using Real = std::remove_reference_t<decltype(a->rho)>;
        a->mode = mode;
        a->rho = static_cast<Real>(rho);
        a->ux = static_cast<Real>(ux);
        a->uy = static_cast<Real>(uy);
        a->uz = static_cast<Real>(uz);
        a->amplitude = static_cast<Real>(amplitude);
        a->noise = static_cast<Real>(noise);
        a->seed_hash = splitmix64(seed);
        a->tables = static_cast<const Real*>(tables);
        a->fields = static_cast<const Real*>(fields);
    };
    if (storage == 2 || storage == 4) fill(static_cast<InitArgs<double>*>(out));
    else fill(static_cast<InitArgs<float>*>(out));
}

}  // namespace cfd_hip
