// lbm_cell.hpp -- per-cell population access shared by the forward and adjoint kernels.
//
// load_cell / store_cell implement the esoteric-pull slot map of one cell (the reference's
// load_cell / store_cell, include/cfd/solvers/lbm/esoteric_pull.hpp:349-382) with in-place halfway
// bounce-back, on a uniform plane base plus 32-bit byte offsets.
#pragma once

#include "lbm_device.hpp"

namespace cfd_hip {

template <class T> __device__ __forceinline__ T* plane(void* base, const Geometry& g, int d) {
    return static_cast<T*>(base) + static_cast<uint64_t>(d) * g.stride;
}
template <class T> __device__ __forceinline__ const T* plane(const void* base, const Geometry& g, int d) {
    return static_cast<const T*>(base) + static_cast<uint64_t>(d) * g.stride;
}

// Byte-offset access: a uniform 64-bit plane base (SGPRs) plus a 32-bit per-lane byte offset
// (one VGPR), which maps onto the global_load/store "saddr" form.  Two consecutive planes must
// fit in 4 GiB (checked at solver creation).
template <class T> __device__ __forceinline__ T load_at(const void* base, uint32_t byte_off) {
    const T* p = reinterpret_cast<const T*>(static_cast<const unsigned char*>(base) + byte_off);
#if defined(CFD_LBM_NT_LOAD)
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
// byte_off should be opaque to the optimiser at the store (store_cell passes it through
// opaque_offset): otherwise it materialises all 64-bit store addresses before the collision and
// keeps them live (2 VGPRs each), which halves the occupancy.
template <class T> __device__ __forceinline__ void store_at(void* base, uint32_t byte_off, T v) {
    T* p = reinterpret_cast<T*>(static_cast<unsigned char*>(base) + byte_off);
    // Streaming stores: no slot written in a step is read again within that step, and the whole
    // lattice passes through the caches every step (measured ~1 % faster than plain stores).
    if constexpr (__is_same(T, F32Dev)) __builtin_nontemporal_store(v.v, reinterpret_cast<float*>(p));
    else if constexpr (__is_same(T, F64Exact)) __builtin_nontemporal_store(v.v, reinterpret_cast<double*>(p));
    else __builtin_nontemporal_store(v, p);
}

// An offset the optimiser cannot look through, once per distinct offset: a barrier at every store
// would copy an offset that several stores share.
__device__ __forceinline__ uint32_t opaque_offset(uint32_t byte_off) {
    asm volatile("" : "+v"(byte_off));
    return byte_off;
}

// The link descriptor of cell n on a bounded lattice (0 on a periodic one).
template <bool Bounded> __device__ __forceinline__ uint32_t link_desc(const uint32_t* links, uint32_t n) {
    if constexpr (Bounded) return links[n];
    else return 0U;
}

// Loads the logical populations of a non-wall cell (bounce-back aware) and records the slots.
// desc is the cell's link descriptor (used only when Bounded); slots come from the neighbourhood.
template <int Q, class T, bool Odd, bool Bounded>
__device__ __forceinline__ void load_cell(const void* f, const Geometry& g, uint32_t desc, const Neighbourhood& h,
                                          typename Storage<T>::Real (&fin)[Q], PairSlots<Q>& s) {
    using L = Lattice<Q>;
    using S = Storage<T>;
    const uint32_t off_n = h.cell(0, 0, 0) * static_cast<uint32_t>(sizeof(T));
    s.self = off_n;
    const uint32_t plane_bytes = static_cast<uint32_t>(g.stride * sizeof(T));
    if constexpr (Bounded) s.desc = desc;
    fin[0] = S::template decode<Q, 0>(load_at<T>(plane<T>(f, g, 0), off_n));
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t up = h.cell(-L::cx[i], -L::cy[i], -L::cz[i]);
        const uint32_t off_up = up * static_cast<uint32_t>(sizeof(T));
        s.up[i] = off_up;
        if constexpr (Bounded) {
            const bool wup = s.is_wall_link(i - 1);  // population i comes from a wall at up
            const bool wdn = s.is_wall_link(i);      // population i+1 comes from a wall at dn
            // Plane i holds the pair; a wall link selects the partner plane (one plane further).
            const uint32_t sel_a = (Odd != wup) ? plane_bytes : 0U;
            const uint32_t sel_b = (Odd != wdn) ? 0U : plane_bytes;
            fin[i] = S::template decode<Q, i>(load_at<T>(plane<T>(f, g, i), off_up + sel_a));
            fin[i + 1] = S::template decode<Q, i + 1>(load_at<T>(plane<T>(f, g, i), off_n + sel_b));
        } else {
            // The slot of either plane of a pair holds a population of the direction it is
            // loaded as (opposite directions share one weight, hence one FP16 shift).
            fin[i] = S::template decode<Q, i>(load_at<T>(plane<T>(f, g, Odd ? i + 1 : i), off_up));
            fin[i + 1] = S::template decode<Q, i + 1>(load_at<T>(plane<T>(f, g, Odd ? i : i + 1), off_n));
        }
    });
}

// The moving-wall momentum term in the state representation (scaled for shifted storage).
template <class S, class Real> __device__ __forceinline__ Real wall_term(Real v) {
    if constexpr (S::kShifted) return v * static_cast<Real>(S::kScale);
    else return v;
}

template <int Q, class T, bool Odd, bool Bounded>
__device__ __forceinline__ void store_cell(void* f, const Geometry& g, const Physics<typename Storage<T>::Real>& p,
                                           uint32_t n, uint32_t x, uint32_t y, uint32_t z,
                                           typename Storage<T>::Real rho, typename Storage<T>::Real (&fout)[Q],
                                           const PairSlots<Q>& s) {
    using L = Lattice<Q>;
    using S = Storage<T>;
    using Real = typename Storage<T>::Real;
    if constexpr (Bounded) {
        if ((s.desc >> kMovingBit) & 1U) {
            // Moving wall: a population bounced off a wall of velocity u_w is stored as
            // f - 6 w rho_w (c . u_w), the push form of src/lbm/d2q9.cpp:217-223.  Only cells next
            // to a moving wall come here; they read their per-link wall-velocity indices.
            constexpr int kW = PairSlots<Q>::kVelWords;
            uint32_t vel[kW];
            for (int w = 0; w < kW; ++w) vel[w] = p.link_vel[static_cast<uint64_t>(n) * kW + w];
            const Real rho_wall = rho > Real(1.0e-8) ? rho : Real(1.0e-8);
            static_for<(Q - 1) / 2>([&](auto P) {
                constexpr int i = 2 * decltype(P)::value + 1;
                // fout[i] travels to dn = n + c_i (link i); fout[i+1] to up = n - c_i (link i-1).
                const uint32_t ka = (vel[i / 10] >> (3 * (i % 10))) & 7U;
                if (ka != 0U) {
                    const Real cu = static_cast<Real>(L::cx[i]) * p.wall_u[ka][0] +
                                    static_cast<Real>(L::cy[i]) * p.wall_u[ka][1] +
                                    static_cast<Real>(L::cz[i]) * p.wall_u[ka][2];
                    fout[i] = fout[i] - wall_term<S>(Real(6) * lattice_weight<S, Q, i, Real>() * rho_wall * cu);
                }
                const uint32_t kb = (vel[(i - 1) / 10] >> (3 * ((i - 1) % 10))) & 7U;
                if (kb != 0U) {
                    const Real cu = static_cast<Real>(L::cx[i + 1]) * p.wall_u[kb][0] +
                                    static_cast<Real>(L::cy[i + 1]) * p.wall_u[kb][1] +
                                    static_cast<Real>(L::cz[i + 1]) * p.wall_u[kb][2];
                    fout[i + 1] = fout[i + 1] - wall_term<S>(Real(6) * lattice_weight<S, Q, i + 1, Real>() * rho_wall * cu);
                }
            });
        }
    }
    const uint32_t off_n = opaque_offset(s.self);
    store_at<T>(plane<T>(f, g, 0), off_n, S::template encode<Q, 0>(fout[0]));
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t off_up = opaque_offset(s.up[i]);
        if constexpr (Odd) {
            store_at<T>(plane<T>(f, g, i), off_n, S::template encode<Q, i>(fout[i]));
            store_at<T>(plane<T>(f, g, i + 1), off_up, S::template encode<Q, i + 1>(fout[i + 1]));
        } else {
            store_at<T>(plane<T>(f, g, i + 1), off_n, S::template encode<Q, i>(fout[i]));
            store_at<T>(plane<T>(f, g, i), off_up, S::template encode<Q, i + 1>(fout[i + 1]));
        }
    });
}

}  // namespace cfd_hip
