// lbm_adjoint.hip -- discrete adjoint of the in-place esoteric-pull step, and the pressure-drop
// objective it differentiates.
//
// In slot space one time step is cell-local: cell n reads its slots L(n) (esoteric map with
// bounce-back) and writes its slots W(n), and no other cell touches them.  The transposed step is
// therefore cell-local and in place as well: cell n reads lambda_{t+1} from W(n), applies the
// transposed collision Jacobian at the forward state f_t (read from L(n) of a stored state), and
// writes lambda_t to L(n).  A bounce-back link reads slot X (opposite parity) and writes slot Y;
// X is written one step later and never read again, Y is overwritten without being read, so the
// adjoint writes the Jacobian term to X and zero to Y -- no pass-through terms are needed.
//
// Differentiated physics: unforced BGK with partial bounce-back gamma, stationary halfway walls,
// the velocity inlet (parabolic, Zou-He density) and the pressure outlet (Zou-He velocity).
// Gradients: dJ/dgamma per cell and dJ/du_peak, accumulated in FP64.  With FP16 shifted storage
// the forward states are the decoded FP16 states and lambda stays FP32: the adjoint is the one of
// the FP32 collision evaluated at the stored states (the rounding of the store is not
// differentiated).  The reference has no adjoint; this is checked against a logical-space adjoint
// (reverse streaming + transposed collision) and central finite differences in the tests.
#include "lbm_adjoint.hpp"
#include "lbm_cell.hpp"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace cfd_hip {

namespace {

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// d feq_d / d rho = e and d feq_d / d u = gu for the second-order equilibrium.
template <class S, int Q, int D, class Real>
__device__ __forceinline__ void feq_grad(Real rho, Real ux, Real uy, Real uz, Real uu_plain, Real& e, Real (&gu)[3]) {
    using L = Lattice<Q>;
    const Real w = lattice_weight<S, Q, D, Real>();
    const Real cu = cdot<Q, D>(ux, uy, uz);
    e = w * (Real(1) + Real(3) * cu + Real(4.5) * cu * cu - Real(1.5) * uu_plain);
    const Real wr = w * rho;
    gu[0] = wr * (Real(3) * static_cast<Real>(L::cx[D]) + Real(9) * cu * static_cast<Real>(L::cx[D]) - Real(3) * ux);
    gu[1] = wr * (Real(3) * static_cast<Real>(L::cy[D]) + Real(9) * cu * static_cast<Real>(L::cy[D]) - Real(3) * uy);
    gu[2] = wr * (Real(3) * static_cast<Real>(L::cz[D]) + Real(9) * cu * static_cast<Real>(L::cz[D]) - Real(3) * uz);
}

template <class Real> __device__ __forceinline__ Real fma_real(Real a, Real b, Real c) {
    if constexpr (sizeof(Real) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// s += c_D v, component by component (c_D in {-1, 0, 1}^3).
template <int Q, int D, class Real> __device__ __forceinline__ void add_c(Real (&s)[3], Real v) {
    using L = Lattice<Q>;
    if constexpr (L::cx[D] > 0) s[0] += v;
    if constexpr (L::cx[D] < 0) s[0] -= v;
    if constexpr (L::cy[D] > 0) s[1] += v;
    if constexpr (L::cy[D] < 0) s[1] -= v;
    if constexpr (L::cz[D] > 0) s[2] += v;
    if constexpr (L::cz[D] < 0) s[2] -= v;
}

// mu = (d fout / d fin)^T lam for one cell; also d/dgamma and d/du_peak of lam . fout.
template <int Q, class S, class Real>
__device__ __forceinline__ void collision_vjp(const Physics<Real>& p, const Geometry& g, uint8_t flag, uint32_t n,
                                              uint32_t y, uint32_t z, const Real (&f)[Q], const Real (&lam)[Q],
                                              Real (&mu)[Q], double& dgamma, double& dpeak) {
    using L = Lattice<Q>;
    dgamma = 0.0;
    dpeak = 0.0;
    if (flag == kInlet) {
        const Real uin = inlet_velocity(p, g, y, z, L::dims == 3);
        const Real eta = (static_cast<Real>(y) - Real(0.5)) / static_cast<Real>(g.ny - 2U);
        Real P = Real(4) * eta * (Real(1) - eta);
        if (L::dims == 3) {
            const Real zeta = (static_cast<Real>(z) - Real(0.5)) / static_cast<Real>(g.nz - 2U);
            P = P * (Real(4) * zeta * (Real(1) - zeta));
        }
        const Real inv = Real(1) / (Real(1) - uin);
        const Real rho = (sum_cx<Q, 0>(f) + Real(2) * sum_cx<Q, -1>(f)) * inv;
        Real g_rho = Real(0), g_ux = Real(0);
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            Real e, gu[3];
            feq_grad<S, Q, d>(rho, uin, Real(0), Real(0), uin * uin, e, gu);
            g_rho += lam[d] * e;
            g_ux += lam[d] * gu[0];
        });
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            constexpr int c = L::cx[d];
            mu[d] = c == 0 ? g_rho * inv : (c < 0 ? Real(2) * g_rho * inv : Real(0));
        });
        dpeak = static_cast<double>((g_ux + g_rho * rho * inv) * P);
        return;
    }
    if (flag == kOutlet) {
        const Real rho = p.rho_out;
        const Real ux = Real(-1) + (sum_cx<Q, 0>(f) + Real(2) * sum_cx<Q, 1>(f)) / rho;
        Real g_ux = Real(0);
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            Real e, gu[3];
            feq_grad<S, Q, d>(rho, ux, Real(0), Real(0), ux * ux, e, gu);
            g_ux += lam[d] * gu[0];
        });
        const Real s = g_ux / rho;
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            constexpr int c = L::cx[d];
            mu[d] = c == 0 ? s : (c > 0 ? Real(2) * s : Real(0));
        });
        return;
    }
    // BGK with partial bounce-back: fout_d = (1 - gamma) (f_d - omega (f_d - feq_d)) + gamma f_opp(d).
    // With a_d = (1 - gamma) lam_d, b_d = omega w_d a_d and cu_d = c_d . u, the contraction of lam
    // with the gradients of feq_d(rho, u) needs only three sums, S0 = sum_d b_d, S1 = sum_d b_d c_d
    // and S2 = sum_d b_d cu_d c_d:
    //   g_rho = sum_d omega a_d dfeq_d/drho      = S0 (1 - 1.5 |u|^2) + 3 u.S1 + 4.5 u.S2
    //   G     = sum_d omega a_d dfeq_d/du / rho  = 3 S1 + 9 S2 - 3 u S0
    // and mu_j = (1 - omega) a_j + gamma lam_opp(j) + g_rho - G.u + c_j.G, since drho/df_j = 1 and
    // du/df_j = (c_j - u) / rho.  Opposite directions share w_d and flip c_d and cu_d, so each pair
    // adds the sum and the difference of its two multipliers.  The sums associate differently from
    // the per-direction form (feq_grad, kept for the inlet and outlet), which changes results only
    // by rounding.
    Real rho, mx, my, mz;
    moments<Q>(f, rho, mx, my, mz);
    const Real ux = mx / rho, uy = my / rho, uz = mz / rho;
    const Real uu = ux * ux + uy * uy + uz * uz;
    const Real omega = p.omega;
    const bool porous = p.gamma != nullptr;
    const Real gm = porous ? p.gamma[n] : Real(0);
    const Real keep = Real(1) - gm;
    const Real ok = omega * keep;                  // b_d = ok w_d lam_d
    const Real om1k = (Real(1) - omega) * keep;    // (1 - omega) a_d = om1k lam_d
    const Real quad = Real(1) - Real(1.5) * uu;
    const Real w0 = lattice_weight<S, Q, 0, Real>();
    Real S0 = ok * w0 * lam[0];
    Real S1[3] = {Real(0), Real(0), Real(0)}, S2[3] = {Real(0), Real(0), Real(0)};
    // dJ/dgamma = sum_d lam_d (f_opp(d) - fstar_d), fstar_d = f_d - omega (f_d - feq_d), for porous
    // lattices only.
    Real dg = Real(0);
    if (porous) {
        const Real fs = f[0] - omega * (f[0] - rho * w0 * quad);
        dg = lam[0] * (f[0] - fs);
    }
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const Real w = lattice_weight<S, Q, i, Real>();
        const Real cu = cdot<Q, i>(ux, uy, uz);
        const Real kw = ok * w;
        const Real bs = kw * (lam[i] + lam[i + 1]);
        const Real bd = kw * (lam[i] - lam[i + 1]);
        S0 += bs;
        const Real bscu = bs * cu;
        add_c<Q, i>(S1, bd);
        add_c<Q, i>(S2, bscu);
        if (porous) {
            const Real rw = rho * w;
            const Real even = rw * fma_real(Real(4.5) * cu, cu, quad);  // feq of the pair without the odd term
            const Real odd = rw * (Real(3) * cu);
            const Real fs_a = f[i] - omega * (f[i] - (even + odd));
            const Real fs_b = f[i + 1] - omega * (f[i + 1] - (even - odd));
            dg += lam[i] * (f[i + 1] - fs_a) + lam[i + 1] * (f[i] - fs_b);
        }
    });
    const Real G0 = Real(3) * S1[0] + Real(9) * S2[0] - Real(3) * ux * S0;
    const Real G1 = Real(3) * S1[1] + Real(9) * S2[1] - Real(3) * uy * S0;
    const Real G2 = Real(3) * S1[2] + Real(9) * S2[2] - Real(3) * uz * S0;
    const Real g_rho = S0 * quad + Real(3) * (ux * S1[0] + uy * S1[1] + uz * S1[2]) +
                       Real(4.5) * (ux * S2[0] + uy * S2[1] + uz * S2[2]);
    const Real base = g_rho - (G0 * ux + G1 * uy + G2 * uz);
    mu[0] = om1k * lam[0] + gm * lam[0] + base;
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const Real Gc = cdot<Q, i>(G0, G1, G2);
        mu[i] = om1k * lam[i] + gm * lam[i + 1] + (base + Gc);
        mu[i + 1] = om1k * lam[i + 1] + gm * lam[i] + (base - Gc);
    });
    dgamma = porous ? static_cast<double>(dg) : 0.0;
}

// Objective plane weight of a cell: c_in on plane x_in, c_out on plane x_out (fluid cells only).
template <class Real>
__device__ __forceinline__ Real plane_weight(const AdjointArgs& a, uint32_t x, bool fluid) {
    if (!fluid) return Real(0);
    if (static_cast<int>(x) == a.x_in) return static_cast<Real>(a.c_in);
    if (static_cast<int>(x) == a.x_out) return static_cast<Real>(a.c_out);
    return Real(0);
}

// One transposed step at the given parity, in place on lam.  fstate holds the forward state f_t.
template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void adjoint_step(const void* __restrict__ fstate, void* __restrict__ lam_buf,
                                                    const Geometry g, const Physics<typename Storage<T>::Real> p,
                                                    const AdjointArgs a) {
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0);
    uint8_t flag = kFluid;
    if constexpr (Bounded) {
        const uint32_t type = (p.links[n] >> kTypeShift) & 3U;
        if (type == kTypeWall) return;
        flag = static_cast<uint8_t>(type);
    }
    Real f[Q];
    PairSlots<Q> s;
    load_cell<Q, T, Odd, Bounded>(fstate, g, link_desc<Bounded>(p.links, n), h, f, s);
    if constexpr (Storage<T>::kShifted)  // the transposed collision takes the populations themselves
        static_for<Q>([&](auto DD) {
            constexpr int d = decltype(DD)::value;
            f[d] = full_population<Storage<T>, Q, d>(f[d]);
        });
    // lambda_{t+1} at the slots this cell wrote (its store slots).
    const uint32_t off_n = h.cell(0, 0, 0) * static_cast<uint32_t>(sizeof(Real));
    const uint32_t rb = static_cast<uint32_t>(sizeof(Real));
    const uint32_t plane_bytes = static_cast<uint32_t>(g.stride * sizeof(Real));
    Real lam[Q];
    lam[0] = load_at<Real>(plane<Real>(lam_buf, g, 0), off_n);
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t off_up = s.up[i] / static_cast<uint32_t>(sizeof(T)) * rb;
        lam[i] = load_at<Real>(plane<Real>(lam_buf, g, Odd ? i : i + 1), off_n);
        lam[i + 1] = load_at<Real>(plane<Real>(lam_buf, g, Odd ? i + 1 : i), off_up);
    });
    Real mu[Q];
    double dgamma, dpeak;
    collision_vjp<Q, Storage<T>>(p, g, flag, n, y, z, f, lam, mu, dgamma, dpeak);
    if (a.add_objective && a.kind == kPressureDrop) {
        const Real c = plane_weight<Real>(a, x, flag == kFluid);
        if (c != Real(0)) static_for<Q>([&](auto DD) { mu[decltype(DD)::value] += c; });
    }
    // lambda_t at the slots this cell read; zero on bounce-back store slots that were not read.
    // (Each store offset passes its own barrier: one barrier per distinct offset measured one
    // VGPR more here.)
    store_at<Real>(plane<Real>(lam_buf, g, 0), opaque_offset(off_n), mu[0]);
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t off_up = s.up[i] / static_cast<uint32_t>(sizeof(T)) * rb;
        bool wup = false, wdn = false;
        if constexpr (Bounded) {
            wup = s.is_wall_link(i - 1);
            wdn = s.is_wall_link(i);
        }
        const uint32_t sel_a = (Odd != wup) ? plane_bytes : 0U;
        const uint32_t sel_b = (Odd != wdn) ? 0U : plane_bytes;
        store_at<Real>(plane<Real>(lam_buf, g, i), opaque_offset(off_up + sel_a), mu[i]);
        store_at<Real>(plane<Real>(lam_buf, g, i), opaque_offset(off_n + sel_b), mu[i + 1]);
        if (wup) store_at<Real>(plane<Real>(lam_buf, g, Odd ? i + 1 : i), opaque_offset(off_up), Real(0));
        if (wdn) store_at<Real>(plane<Real>(lam_buf, g, Odd ? i : i + 1), opaque_offset(off_n), Real(0));
    });
    if (p.gamma && flag == kFluid) a.grad_gamma[n] += dgamma;
    if (flag == kInlet) a.grad_peak[n] += dpeak;
}

// lambda_T = dJ/df_T: zero everywhere plus the objective's weight at the slots read at parity T.
template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void adjoint_seed(void* __restrict__ lam_buf, const Geometry g,
                                                    const Physics<typename Storage<T>::Real> p, const AdjointArgs a) {
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.cell(0, 0, 0);
    for (int d = 0; d < Q; ++d) static_cast<Real*>(lam_buf)[static_cast<uint64_t>(d) * g.stride + n] = Real(0);
    (void)h;
}

template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void plane_source(void* __restrict__ lam_buf, const Geometry g,
                                                           const Physics<typename Storage<T>::Real> p,
                                                           const AdjointArgs a) {
    using L = Lattice<Q>;
    using Real = typename Storage<T>::Real;
    uint32_t x, y, z;
    if (!cell_coords(g, x, y, z)) return;
    const Neighbourhood h = neighbourhood(g, x, y, z);
    const uint32_t n = h.lin(0, 0, 0), np = h.cell(0, 0, 0);
    bool fluid = true;
    uint32_t desc = 0;
    if constexpr (Bounded) {
        desc = p.links[n];
        fluid = ((desc >> kTypeShift) & 3U) == 0U;
    }
    const Real c = plane_weight<Real>(a, x, fluid);
    if (c == Real(0)) return;
    Real* lam = static_cast<Real*>(lam_buf);
    lam[np] += c;
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const uint32_t up = h.cell(-L::cx[i], -L::cy[i], -L::cz[i]);
        const bool wup = Bounded && ((desc >> (i - 1)) & 1U);
        const bool wdn = Bounded && ((desc >> i) & 1U);
        lam[static_cast<uint64_t>(dir_a<Odd>(i, wup)) * g.stride + up] += c;
        lam[static_cast<uint64_t>(dir_b<Odd>(i, wdn)) * g.stride + np] += c;
    });
}

// Drag source: lam at the slots the listed cells read at this parity += c_drag * (-2 c_d[axis]) for
// every population d that arrives from the body.  Each slot belongs to one cell: no races.
template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void drag_source(void* __restrict__ lam_buf, const Geometry g, const AdjointArgs a) {
    using L = Lattice<Q>;
    using Real = typename Storage<T>::Real;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.drag_count) return;
    const uint32_t n = a.drag_cells[k];
    const uint32_t bits = a.drag_links[k];
    const uint32_t x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
    const Neighbourhood h = neighbourhood(g, x, y, z);
    Real* lam = static_cast<Real*>(lam_buf);
    const Real c = static_cast<Real>(a.c_drag);
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1;
        const Real ci = static_cast<Real>(a.drag_axis == 0 ? L::cx[i] : (a.drag_axis == 1 ? L::cy[i] : L::cz[i]));
        if ((bits >> (i - 1)) & 1U) {
            const uint32_t up = h.cell(-L::cx[i], -L::cy[i], -L::cz[i]);
            lam[static_cast<uint64_t>(dir_a<Odd>(i, true)) * g.stride + up] += Real(-2) * ci * c;
        }
        if ((bits >> i) & 1U)  // population i+1 has c = -c_i
            lam[static_cast<uint64_t>(dir_b<Odd>(i, true)) * g.stride + h.cell(0, 0, 0)] += Real(2) * ci * c;
    });
}

// Per-block partial sums of the drag over the listed cells (FP64, fixed tree order).
template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void drag_partial(const void* __restrict__ f, const Geometry g,
                                                    const Physics<typename Storage<T>::Real> p, const AdjointArgs a) {
    using L = Lattice<Q>;
    using Real = typename Storage<T>::Real;
    __shared__ double sm[256];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (k < a.drag_count) {
        const uint32_t n = a.drag_cells[k];
        const uint32_t bits = a.drag_links[k];
        const uint32_t x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
        const Neighbourhood h = neighbourhood(g, x, y, z);
        Real fin[Q];
        PairSlots<Q> s;
        load_cell<Q, T, Odd, Bounded>(f, g, link_desc<Bounded>(p.links, n), h, fin, s);
        static_for<Q - 1>([&](auto DD) {
            constexpr int d = decltype(DD)::value + 1;
            const int cd = a.drag_axis == 0 ? L::cx[d] : (a.drag_axis == 1 ? L::cy[d] : L::cz[d]);
            if ((bits >> (d - 1)) & 1U)
                v += -2.0 * static_cast<double>(cd) *
                     static_cast<double>(full_population<Storage<T>, Q, d>(fin[d]));
        });
    }
    sm[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.scratch[blockIdx.x] = sm[0];
}

// acc += weight * sum(partial[0..count)), one block in a fixed order.
__global__ __launch_bounds__(1024) void sum_partials(const double* __restrict__ partial, uint32_t count, double weight,
                                                     double* acc) {
    __shared__ double sm[1024];
    double v = 0.0;
    for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) v += partial[k];
    sm[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += weight * sm[0];
}

// Per-block partial sums of j(f) = sum_{x = x_in} rho * c_in + sum_{x = x_out} rho * c_out over fluid
// cells, in FP64 (one thread per cell of the two planes; fixed tree order).
template <int Q, class T, bool Odd, bool Bounded>
__global__ __launch_bounds__(256) void plane_partial(const void* __restrict__ f, const Geometry g,
                                                     const Physics<typename Storage<T>::Real> p,
                                                     const AdjointArgs a) {
    using Real = typename Storage<T>::Real;
    __shared__ double sm[256];
    const uint32_t per_plane = g.ny * g.nz;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (k < 2U * per_plane) {
        const uint32_t which = k / per_plane;
        const uint32_t r = k % per_plane;
        const uint32_t y = r % g.ny, z = r / g.ny;
        const uint32_t x = static_cast<uint32_t>(which == 0 ? a.x_in : a.x_out);
        const Neighbourhood h = neighbourhood(g, x, y, z);
        const uint32_t n = h.lin(0, 0, 0);
        bool fluid = true;
        if constexpr (Bounded) fluid = ((p.links[n] >> kTypeShift) & 3U) == 0U;
        if (fluid) {
            Real fin[Q];
            PairSlots<Q> s;
            load_cell<Q, T, Odd, Bounded>(f, g, link_desc<Bounded>(p.links, n), h, fin, s);
            Real rho = Real(0);
            static_for<Q>([&](auto DD) { rho += fin[decltype(DD)::value]; });
            const double rho_d = Storage<T>::kShifted
                                     ? 1.0 + static_cast<double>(rho * static_cast<Real>(Storage<T>::kInvScale))
                                     : static_cast<double>(rho);
            v = rho_d * (which == 0 ? a.c_in : a.c_out);
        }
    }
    sm[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.scratch[blockIdx.x] = sm[0];
}

// partial[b] = sum of block b's contiguous share of v[0..n), in a fixed tree order.
__global__ __launch_bounds__(256) void sum_chunks_f64(const double* __restrict__ v, uint64_t n,
                                                      double* __restrict__ partial) {
    __shared__ double sm[256];
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t begin = chunk * blockIdx.x, end = begin + chunk < n ? begin + chunk : n;
    double s = 0.0;
    for (uint64_t i = begin + threadIdx.x; i < end; i += blockDim.x) s += v[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void fill_zero_f64(double* __restrict__ v, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = 0.0;
}

template <int Q, class T> struct AdjOps final : AdjointOps {
    using Real = typename Storage<T>::Real;
    template <class K> static void per_parity(bool odd, bool bounded, K k) {
        if (odd) bounded ? k(std::true_type{}, std::true_type{}) : k(std::true_type{}, std::false_type{});
        else bounded ? k(std::false_type{}, std::true_type{}) : k(std::false_type{}, std::false_type{});
    }
    static dim3 block(const Geometry& g) { return dim3(g.bx, g.by, 1); }
    static dim3 grid(const Geometry& g) { return dim3(g.tiles, 1, 1); }

    static unsigned list_blocks(const AdjointArgs& a) { return (a.drag_count + 255U) / 256U; }
    static unsigned plane_blocks(const Geometry& g) { return (2U * g.ny * g.nz + 255U) / 256U; }

    // The objective's dj/df source added onto lam at parity odd (the drag part; the pressure-drop
    // part of adjoint_step is fused into the transposed step itself).
    void drag(const LaunchState& st, bool odd, bool bounded, void* lam, const AdjointArgs& a) const {
        if (a.drag_count == 0) return;
        per_parity(odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((drag_source<Q, T, decltype(O)::value, decltype(B)::value>), dim3(list_blocks(a)),
                               dim3(256), 0, st.stream, lam, st.geo, a);
        });
        check(hipGetLastError(), "drag_source launch");
    }

    void step(const LaunchState& st, bool odd, bool bounded, const void* fstate, void* lam,
              const AdjointArgs& a) const override {
        per_parity(odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((adjoint_step<Q, T, decltype(O)::value, decltype(B)::value>), grid(st.geo),
                               block(st.geo), st.adjoint_lds, st.stream, fstate, lam, st.geo,
                               *static_cast<const Physics<Real>*>(st.physics), a);
        });
        check(hipGetLastError(), "adjoint_step launch");
        if (a.add_objective && a.kind == kDrag) drag(st, odd, bounded, lam, a);
    }
    void seed(const LaunchState& st, bool odd, bool bounded, void* lam, const AdjointArgs& a) const override {
        per_parity(odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((adjoint_seed<Q, T, decltype(O)::value, decltype(B)::value>), grid(st.geo),
                               block(st.geo), 0, st.stream, lam, st.geo,
                               *static_cast<const Physics<Real>*>(st.physics), a);
            if (a.kind == kPressureDrop)
                hipLaunchKernelGGL((plane_source<Q, T, decltype(O)::value, decltype(B)::value>), grid(st.geo),
                                   block(st.geo), 0, st.stream, lam, st.geo,
                                   *static_cast<const Physics<Real>*>(st.physics), a);
        });
        check(hipGetLastError(), "adjoint_seed launch");
        if (a.kind == kDrag) drag(st, odd, bounded, lam, a);
    }
    void objective(const LaunchState& st, bool odd, bool bounded, const void* f, const AdjointArgs& a, double weight,
                   double* acc) const override {
        if (a.kind == kDrag) {
            if (a.drag_count == 0) return;
            per_parity(odd, bounded, [&](auto O, auto B) {
                hipLaunchKernelGGL((drag_partial<Q, T, decltype(O)::value, decltype(B)::value>), dim3(list_blocks(a)),
                                   dim3(256), 0, st.stream, f, st.geo, *static_cast<const Physics<Real>*>(st.physics),
                                   a);
            });
            hipLaunchKernelGGL(sum_partials, dim3(1), dim3(1024), 0, st.stream, a.scratch, list_blocks(a),
                               weight * a.c_drag, acc);
            check(hipGetLastError(), "drag objective launch");
            return;
        }
        const unsigned blocks = plane_blocks(st.geo);
        per_parity(odd, bounded, [&](auto O, auto B) {
            hipLaunchKernelGGL((plane_partial<Q, T, decltype(O)::value, decltype(B)::value>), dim3(blocks), dim3(256),
                               0, st.stream, f, st.geo, *static_cast<const Physics<Real>*>(st.physics), a);
        });
        hipLaunchKernelGGL(sum_partials, dim3(1), dim3(1024), 0, st.stream, a.scratch, blocks, weight, acc);
        check(hipGetLastError(), "pressure-drop objective launch");
    }
    void zero(const LaunchState& st, double* v, uint64_t n) const override {
        hipLaunchKernelGGL(fill_zero_f64, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st.stream, v,
                           n);
        check(hipGetLastError(), "fill_zero launch");
    }
    void sum(const LaunchState& st, const double* v, uint64_t n, double* partial, double* out) const override {
        zero(st, out, 1);
        hipLaunchKernelGGL(sum_chunks_f64, dim3(kSumBlocks), dim3(256), 0, st.stream, v, n, partial);
        hipLaunchKernelGGL(sum_partials, dim3(1), dim3(1024), 0, st.stream, partial, kSumBlocks, 1.0, out);
        check(hipGetLastError(), "sum launch");
    }
};

template <int Q> const AdjointOps* adj_for_storage(int storage) {
    static const AdjOps<Q, float> f32;
    static const AdjOps<Q, _Float16> f16;
    static const AdjOps<Q, double> f64;
    static const AdjOps<Q, F32Dev> f32s;
    static const AdjOps<Q, F64Exact> f64x;
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

const AdjointOps* adjoint_ops(int q, int storage) {
    switch (q) {
    case 9: return adj_for_storage<9>(storage);
    case 19: return adj_for_storage<19>(storage);
    case 27: return adj_for_storage<27>(storage);
    default: throw std::invalid_argument("lattice must be D2Q9 (9), D3Q19 (19) or D3Q27 (27)");
    }
}

}  // namespace cfd_hip
