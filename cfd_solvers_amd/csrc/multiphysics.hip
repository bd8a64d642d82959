// multiphysics.hip -- device-resident modules that couple to the flow solver through its device
// views (macroscopic fields and the per-cell local acceleration), on the flow's stream:
//
//   thermal   D3Q7 passive scalar / temperature lattice and Boussinesq buoyancy
//             (include/cfd/solvers/lbm/resident_multiphysics_sycl.hpp:54-198)
//   surface   volume-of-fluid fill fraction: upwind advection, interface normals and curvature,
//             surface-tension acceleration (resident_multiphysics_sycl.hpp:200-324)
//   particles Lagrangian point particles: trilinear fluid velocity, Stokes drag, two-way force
//             spreading (resident_multiphysics_sycl.hpp:326-380)
//   q         Q-criterion of the velocity field (resident_multiphysics_sycl.hpp:382-408)
//
// All lattices are periodic in x, y, z with the flow solver's nx, ny, nz, cell n = (z*ny + y)*nx + x.
// The per-cell arithmetic keeps the reference's float operations in its order (the file is built
// with -ffp-contract=off), so every field except the atomically accumulated two-way particle force
// is reproducible bit for bit; the global sums (total scalar, volume) are FP64 and deterministic.
#include "../../include/cfd_lbm/c_api.h"
#include "solver_view.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace {

using cfd_hip::FlowInfo;

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct Dims {
    uint32_t nx, ny, nz, cells;
};

constexpr int kBlock = 256;
constexpr int kReduceBlocks = 1024;

__device__ __forceinline__ uint32_t at(uint32_t x, uint32_t y, uint32_t z, const Dims& d) {
    return (z * d.ny + y) * d.nx + x;
}

// Element i of a field plane, addressed as the plane's uniform base plus the 32-bit byte offset
// 4i: the global_load/store "saddr" form, one VGPR per address instead of a 64-bit pair.  Planes
// hold at most 2^30 cells (checked when a module is created), so the offset fits.
__device__ __forceinline__ float ld(const float* base, uint32_t i) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const unsigned char*>(base) + (i << 2));
}
__device__ __forceinline__ float& el(float* base, uint32_t i) {
    return *reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(base) + (i << 2));
}

// A cell of the stencil kernels and its periodic neighbours along +-x, +-y, +-z (linear indices).
struct Around {
    uint32_t n, xp, xm, yp, ym, zp, zm;
};

// The stencil kernels run on a 3-D launch shape (stencil_shape): x and y from the block and thread
// indices, one z plane per grid layer, so the coordinates need no divisions.  Each thread takes
// Rows rows, blockDim.y apart, so that a workgroup covers Rows * blockDim.y consecutive rows and
// reads most y neighbours through its own XCD's L2 (round-robin dispatch deals consecutive
// workgroups to the 8 XCDs).  The wrapped neighbour offsets are formed in modular unsigned
// arithmetic; every result lies in [0, cells).
__device__ __forceinline__ Around around(const Dims& d, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t plane = d.nx * d.ny;
    const uint32_t n = z * plane + y * d.nx + x;
    Around a;
    a.n = n;
    a.xp = x + 1U == d.nx ? n + 1U - d.nx : n + 1U;
    a.xm = x == 0U ? n + d.nx - 1U : n - 1U;
    a.yp = y + 1U == d.ny ? n + d.nx - plane : n + d.nx;
    a.ym = y == 0U ? n + plane - d.nx : n - d.nx;
    a.zp = z + 1U == d.nz ? n + plane - d.cells : n + plane;
    a.zm = z == 0U ? n + d.cells - plane : n - plane;
    return a;
}
// With Remap, each XCD takes consecutive row blocks of a plane, so that its blocks share their
// halo rows in its L2 (when the blocks per plane are a multiple of 8).
template <uint32_t Rows, bool Remap, class F>
__device__ __forceinline__ void for_stencil_cells(const Dims& d, F&& f) {
    uint32_t bx = blockIdx.x, by = blockIdx.y;
    const uint32_t per_plane = gridDim.x * gridDim.y;
    if (Remap && (per_plane & 7U) == 0U) {
        // Dispatch deals block p of a plane to XCD p % 8; block p takes the place of block
        // (p % 8) * per_plane / 8 + p / 8, so each XCD holds per_plane / 8 consecutive blocks.
        const uint32_t p = blockIdx.y * gridDim.x + blockIdx.x;
        const uint32_t q = (p & 7U) * (per_plane >> 3) + (p >> 3);
        bx = q % gridDim.x;
        by = q / gridDim.x;
    }
    const uint32_t x = bx * blockDim.x + threadIdx.x;
    const uint32_t y0 = by * Rows * blockDim.y + threadIdx.y;
    if (x >= d.nx) return;
#pragma unroll
    for (uint32_t k = 0; k < Rows; ++k) {
        const uint32_t y = y0 + k * blockDim.y;
        if (y < d.ny) f(around(d, x, y, blockIdx.z));
    }
}

// Rows per thread of the kernels on the 3-D launch, the fastest of 1, 2, 4 and 8 at 256^3 (1 / 2 /
// 4 / 8 rows: thermal step 198 / 207 / 215 / 238 us, interface geometry 125 / 124 / 120 / 134 us,
// surface tension with the geometry 241 / 231 / 232 / 252 us).  The D3Q7 step reads each of its 7
// planes once, from a different neighbour per plane, so it has no neighbour reuse to exploit.
constexpr uint32_t kRowsThermal = 1;
constexpr uint32_t kRowsGeometry = 4;
constexpr uint32_t kRowsTension = 2;
// The per-XCD row-block remap, where it measured faster at 256^3 (off / on: interface geometry
// 123.8 / 118.3 us; thermal step 195.1 / 207.6 us, surface tension 233 / 245 us).
constexpr bool kRemapGeometry = true;

// ------------------------------------------------------------------ z-marching stencil tiles
// Fill advection and the Q-criterion read three or four fields at their neighbours.  They march
// along z through a tile of kTileX x (kTileWaves * R) cells: each thread keeps the values of its R
// cells at z-1, z and z+1 in registers, and the block stages plane z with a one-cell halo in x and
// y in LDS for the x and y neighbours.  Every input value is then read from memory about once
// (the tile's halo and the plane below and above each z chunk come on top) instead of once per
// neighbour that uses it.  The loads of plane z+1+kPrefetch are issued before the arithmetic of
// plane z, into a ring of register sets that is never copied (a copy would wait for the loads in
// flight).  The LDS tile is double-buffered, so one barrier per plane suffices: a thread writes
// buffer b again only after the barrier of the plane in between, which every thread reaches after
// its reads of buffer b.  At 256^3 the tiles take fill advection from 82.2 to 75.9 us and the
// Q-criterion from 59.7 to 50.9 us (rows per thread 1 of 1, 2, 4; 2048 of 512-4096 blocks;
// prefetch 1 of 1, 2).  The normals (one field; three IEEE divisions and a square root per cell),
// the curvature (already at the copy rate of its 18 B per cell) and the surface tension measured
// 5-24 % slower on tiles and keep the 3-D launch.
constexpr uint32_t kTileX = 64;     // one wavefront per tile row
constexpr uint32_t kTileWaves = 4;  // block = kTileX x kTileWaves threads
// Blocks a launch aims for (tiles per plane x z chunks): eight 256-thread blocks on each CU.
constexpr uint32_t kZBlocks = 2048;
// Planes a thread's loads run ahead of the plane it computes (1 or 2).
constexpr uint32_t kPrefetch = 1;
constexpr uint32_t kRowsAdvect = 1;
constexpr uint32_t kRowsQ = 1;

template <int NF> struct Inputs {
    const float* p[NF];
};
// The values a 7-point stencil reads around one cell, per input field.
template <int NF> struct Nbr {
    float c[NF], xm[NF], xp[NF], ym[NF], yp[NF], zm[NF], zp[NF];
};

constexpr int bits_below(uint32_t mask, int f) {
    int k = 0;
    for (int i = 0; i < f; ++i) k += (mask >> i) & 1U;
    return k;
}

// ZM: fields whose z neighbours the stencil reads; XYM: fields whose x or y neighbours it reads
// (staged in LDS).  Every field's own value at the cell is available.  A block also reads the
// planes just below and above its z chunk, so no kernel may write its own inputs.  op(nbr, n) runs once per
// cell n of the tile.  Launch: block (kTileX, kTileWaves), grid (tiles in x, tiles in y, z chunks
// of zchunk planes).  Consecutive blocks of the flattened grid go to one XCD (when the block count
// is a multiple of 8), so a tile's halo rows mostly come from its own XCD's L2.
template <int NF, uint32_t ZM, uint32_t XYM, uint32_t R, class Op>
__device__ __forceinline__ void zmarch(const Inputs<NF>& in, const Dims& d, uint32_t zchunk, Op&& op) {
    constexpr uint32_t TY = kTileWaves * R;
    constexpr uint32_t W = kTileX + 2;
    constexpr int NXY = bits_below(XYM, NF);
    static_assert(2 * kTileX + 2 * TY <= kTileX * kTileWaves, "one halo value per thread");
    __shared__ float tile[2][NXY > 0 ? NXY : 1][(TY + 2) * W];

    uint32_t b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const uint32_t total = gridDim.x * gridDim.y * gridDim.z;
    if ((total & 7U) == 0U) b = (b & 7U) * (total >> 3) + (b >> 3);
    const uint32_t bx = b % gridDim.x;
    b /= gridDim.x;
    const uint32_t by = b % gridDim.y;
    const uint32_t bz = b / gridDim.y;
    const uint32_t z0 = bz * zchunk;
    if (z0 >= d.nz) return;  // the whole block
    const uint32_t z1 = min(z0 + zchunk, d.nz);
    const uint32_t tx = threadIdx.x, ty = threadIdx.y;
    const uint32_t x0 = bx * kTileX, y0 = by * TY;
    const uint32_t w = min(kTileX, d.nx - x0), h = min(TY, d.ny - y0);
    const uint32_t plane = d.nx * d.ny;

    uint32_t own[R];
    bool rin[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        rin[r] = tx < w && ty + kTileWaves * r < h;
        own[r] = (y0 + ty + kTileWaves * r) * d.nx + x0 + tx;
    }
    // Halo duty: threads 0..63 the row below the tile, 64..127 the row above, then the column left
    // and the column right of the tile (all wrapped).  The row above sits right after the tile's
    // last row and the right column right after its last column, also in partial tiles.
    int hslot = -1;
    uint32_t hoff = 0;
    if constexpr (NXY > 0) {
        const uint32_t t = ty * kTileX + tx;
        if (t < kTileX) {
            if (t < w) {
                hoff = (y0 == 0U ? d.ny - 1U : y0 - 1U) * d.nx + x0 + t;
                hslot = static_cast<int>(1U + t);
            }
        } else if (t < 2U * kTileX) {
            const uint32_t u = t - kTileX;
            if (u < w) {
                hoff = (y0 + h == d.ny ? 0U : y0 + h) * d.nx + x0 + u;
                hslot = static_cast<int>((h + 1U) * W + 1U + u);
            }
        } else if (t < 2U * kTileX + TY) {
            const uint32_t u = t - 2U * kTileX;
            if (u < h) {
                hoff = (y0 + u) * d.nx + (x0 == 0U ? d.nx - 1U : x0 - 1U);
                hslot = static_cast<int>((1U + u) * W);
            }
        } else if (t < 2U * kTileX + 2U * TY) {
            const uint32_t u = t - 2U * kTileX - TY;
            if (u < h) {
                hoff = (y0 + u) * d.nx + (x0 + w == d.nx ? 0U : x0 + w);
                hslot = static_cast<int>((1U + u) * W + w + 1U);
            }
        }
    }

    // Ring of S register sets: plane z0 + j lives in set j % S.  Computing plane z reads sets
    // z-1, z, z+1 and loads plane z + D + 1 into the set of plane z - 2; the ring is never copied,
    // so loads stay in flight until the plane they fetch is used, D steps later.
    constexpr uint32_t S = kPrefetch + 3;
    static_assert(S == 4 || S == 5, "the z loop below is unrolled for rings of 4 or 5 sets");
    float v[S][NF][R];
    float hv[S][NXY > 0 ? NXY : 1];
    auto load_plane = [&](uint32_t set, uint32_t zz) {
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (uint32_t r = 0; r < R; ++r)
                if (rin[r]) v[set][f][r] = ld(in.p[f], zz * plane + own[r]);
    };
    auto load_halo = [&](uint32_t set, uint32_t zz) {
        if constexpr (NXY > 0) {
            if (hslot < 0) return;
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if ((XYM >> f) & 1U) hv[set][bits_below(XYM, f)] = ld(in.p[f], zz * plane + hoff);
        }
    };
#pragma unroll
    for (uint32_t s = 0; s < S; ++s) {
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (uint32_t r = 0; r < R; ++r) v[s][f][r] = 0.0F;
#pragma unroll
        for (int k = 0; k < (NXY > 0 ? NXY : 1); ++k) hv[s][k] = 0.0F;
    }
    // Prologue: planes z0 - 1 .. z0 + D and the halos of planes z0 .. z0 + D - 1 (those up to z1).
    load_plane(S - 1, z0 == 0U ? d.nz - 1U : z0 - 1U);
    uint32_t za = z0;  // wrapped plane index of the next load
#pragma unroll
    for (uint32_t j = 0; j <= kPrefetch; ++j) {
        if (z0 + j <= z1) {
            load_plane(j, za);
            if (j < kPrefetch && z0 + j < z1) load_halo(j, za);
        }
        za = za + 1U == d.nz ? 0U : za + 1U;
    }
    // za is now plane z0 + D + 1 (wrapped); zh the plane of the next halo load, z0 + D.
    uint32_t zh = za == 0U ? d.nz - 1U : za - 1U;

    auto step = [&](auto K, uint32_t z) {
        constexpr uint32_t k = decltype(K)::value;
        constexpr uint32_t sp = (k + S - 1) % S, sc = k, sn = (k + 1) % S, sl = (k + kPrefetch + 1) % S;
        constexpr uint32_t hs = (k + kPrefetch) % S;
        const uint32_t lb = (z - z0) & 1U;  // LDS buffer
        if constexpr (NXY > 0) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                if (!((XYM >> f) & 1U)) continue;
                const int kk = bits_below(XYM, f);
                float* t = tile[lb][kk];
#pragma unroll
                for (uint32_t r = 0; r < R; ++r)
                    if (rin[r]) t[(1U + ty + kTileWaves * r) * W + 1U + tx] = v[sc][f][r];
                if (hslot >= 0) t[hslot] = hv[sc][kk];
            }
        }
        if (z + kPrefetch < z1) {
            load_plane(sl, za);
            load_halo(hs, zh);
        }
        za = za + 1U == d.nz ? 0U : za + 1U;
        zh = zh + 1U == d.nz ? 0U : zh + 1U;
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            if (!rin[r]) continue;
            Nbr<NF> s{};
            const uint32_t row = 1U + ty + kTileWaves * r;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                s.c[f] = v[sc][f][r];
                if ((ZM >> f) & 1U) {
                    s.zm[f] = v[sp][f][r];
                    s.zp[f] = v[sn][f][r];
                }
                if constexpr (NXY > 0) {
                    if ((XYM >> f) & 1U) {
                        const float* t = tile[lb][bits_below(XYM, f)];
                        s.xm[f] = t[row * W + tx];
                        s.xp[f] = t[row * W + tx + 2U];
                        s.ym[f] = t[(row - 1U) * W + 1U + tx];
                        s.yp[f] = t[(row + 1U) * W + 1U + tx];
                    }
                }
            }
            op(s, z * plane + own[r]);
        }
    };
    for (uint32_t z = z0;; z += S) {
        step(std::integral_constant<uint32_t, 0>{}, z);
        if (z + 1U >= z1) break;
        step(std::integral_constant<uint32_t, 1>{}, z + 1U);
        if (z + 2U >= z1) break;
        step(std::integral_constant<uint32_t, 2>{}, z + 2U);
        if (z + 3U >= z1) break;
        step(std::integral_constant<uint32_t, 3>{}, z + 3U);
        if (z + 4U >= z1) break;
        if constexpr (S > 4) {
            step(std::integral_constant<uint32_t, 4>{}, z + 4U);
            if (z + 5U >= z1) break;
        }
    }
}

// ------------------------------------------------------------------ D3Q7 thermal lattice
// Directions 0, +x, -x, +y, -y, +z, -z; weights 1/4 and 1/8 (resident_multiphysics_sycl.hpp:188-191).
__device__ __forceinline__ float q7_weight(int d) { return d == 0 ? 0.25F : 0.125F; }
__device__ __forceinline__ float q7_cx(int d) { return d == 1 ? 1.0F : d == 2 ? -1.0F : 0.0F; }
__device__ __forceinline__ float q7_cy(int d) { return d == 3 ? 1.0F : d == 4 ? -1.0F : 0.0F; }
__device__ __forceinline__ float q7_cz(int d) { return d == 5 ? 1.0F : d == 6 ? -1.0F : 0.0F; }

__global__ void __launch_bounds__(kBlock) thermal_fill(float* g, const float* field, float value, Dims d) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.cells) return;
    const float t = field ? field[n] : value;
#pragma unroll
    for (int q = 0; q < 7; ++q) g[static_cast<size_t>(q) * d.cells + n] = q7_weight(q) * t;
}

// Pull streaming from the periodic neighbours, then BGK towards w T (1 + 4 c.u)
// (resident_multiphysics_sycl.hpp:94-124).
__global__ void __launch_bounds__(kBlock) thermal_step(const float* __restrict__ in, float* __restrict__ out,
                                                       const float* __restrict__ ux, const float* __restrict__ uy,
                                                       const float* __restrict__ uz, Dims d, float omega) {
    for_stencil_cells<kRowsThermal, false>(d, [&](const Around& a) {
        const uint32_t n = a.n;
        const uint32_t src[7] = {n, a.xm, a.xp, a.ym, a.yp, a.zm, a.zp};
        float local[7];
        float temperature = 0.0F;
    #pragma unroll
        for (int q = 0; q < 7; ++q) {
            local[q] = ld(in + static_cast<size_t>(q) * d.cells, src[q]);
            temperature += local[q];
        }
        const float u = ld(ux, n), v = ld(uy, n), w = ld(uz, n);
    #pragma unroll
        for (int q = 0; q < 7; ++q) {
            const float cu = q7_cx(q) * u + q7_cy(q) * v + q7_cz(q) * w;
            const float geq = q7_weight(q) * temperature * (1.0F + 4.0F * cu);
            el(out + static_cast<size_t>(q) * d.cells, n) = local[q] - omega * (local[q] - geq);
        }
    });
}

__global__ void __launch_bounds__(kBlock) thermal_scalar(const float* __restrict__ g, float* __restrict__ s, Dims d) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.cells) return;
    float value = 0.0F;
#pragma unroll
    for (int q = 0; q < 7; ++q) value += g[static_cast<size_t>(q) * d.cells + n];
    s[n] = value;
}

// The temperature refresh and the Boussinesq coupling a += -beta (T - T_ref) g
// (resident_multiphysics_sycl.hpp:140-155) in one pass: T = the sum of the 7 populations (the
// thermal_scalar order) is written to the temperature plane and drives the coupling.
__global__ void __launch_bounds__(kBlock) boussinesq_from_populations(const float* __restrict__ g, float* __restrict__ t,
                                                                      float* ax, float* ay, float* az, Dims d,
                                                                      float t_ref, float beta, float gx, float gy,
                                                                      float gz) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.cells) return;
    float value = 0.0F;
#pragma unroll
    for (int q = 0; q < 7; ++q) value += g[static_cast<size_t>(q) * d.cells + n];
    t[n] = value;
    const float scale = -beta * (value - t_ref);
    ax[n] += scale * gx;
    ay[n] += scale * gy;
    az[n] += scale * gz;
}

// ------------------------------------------------------------------ FP64 sums (deterministic)
__global__ void __launch_bounds__(kBlock) sum_partial(const float* __restrict__ v, uint64_t count, double* partial) {
    __shared__ double part[kBlock];
    double s = 0.0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        s += static_cast<double>(v[i]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(w)) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = part[0];
}

__global__ void __launch_bounds__(kBlock) sum_final(const double* __restrict__ partial, int count, double* out) {
    __shared__ double part[kBlock];
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += kBlock) s += partial[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(w)) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// ------------------------------------------------------------------ volume of fluid
__global__ void __launch_bounds__(kBlock) vof_fill(float* phi, const float* field, float value, int half_space,
                                                   float interface_x, Dims d) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.cells) return;
    if (half_space) {  // resident_multiphysics_sycl.hpp:226-236
        const float left = static_cast<float>(n % d.nx), right = left + 1.0F;
        phi[n] = interface_x <= left ? 0.0F : interface_x >= right ? 1.0F : interface_x - left;
    } else {
        phi[n] = field ? field[n] : value;
    }
}

// Upwind fluxes through the six faces with face velocities averaged from the two cells, then
// clamped to [0, 1] (resident_multiphysics_sycl.hpp:238-263).  Fields: 0 fill (all neighbours),
// 1 ux (x neighbours), 2 uy (y neighbours), 3 uz (z neighbours).
__global__ void __launch_bounds__(kBlock) vof_advect(const Inputs<4> in, float* __restrict__ out, Dims d, float dt,
                                                     uint32_t zchunk) {
    zmarch<4, 0b1001U, 0b0111U, kRowsAdvect>(in, d, zchunk, [&](const Nbr<4>& s, uint32_t n) {
        const float uxp = 0.5F * (s.c[1] + s.xp[1]), uxm = 0.5F * (s.xm[1] + s.c[1]);
        const float uyp = 0.5F * (s.c[2] + s.yp[2]), uym = 0.5F * (s.ym[2] + s.c[2]);
        const float uzp = 0.5F * (s.c[3] + s.zp[3]), uzm = 0.5F * (s.zm[3] + s.c[3]);
        const float p = s.c[0];
        const float fxp = uxp >= 0.0F ? uxp * p : uxp * s.xp[0];
        const float fxm = uxm >= 0.0F ? uxm * s.xm[0] : uxm * p;
        const float fyp = uyp >= 0.0F ? uyp * p : uyp * s.yp[0];
        const float fym = uym >= 0.0F ? uym * s.ym[0] : uym * p;
        const float fzp = uzp >= 0.0F ? uzp * p : uzp * s.zp[0];
        const float fzm = uzm >= 0.0F ? uzm * s.zm[0] : uzm * p;
        float value = p - dt * ((fxp - fxm) + (fyp - fym) + (fzp - fzm));
        value = value < 0.0F ? 0.0F : value > 1.0F ? 1.0F : value;
        el(out, n) = value;
    });
}

__device__ __forceinline__ void fill_gradient(const float* phi, const Around& a, float& gx, float& gy, float& gz) {
    gx = 0.5F * (ld(phi, a.xp) - ld(phi, a.xm));
    gy = 0.5F * (ld(phi, a.yp) - ld(phi, a.ym));
    gz = 0.5F * (ld(phi, a.zp) - ld(phi, a.zm));
}

// Unit interface normal grad(phi)/|grad(phi)| (zero where the gradient vanishes).
__global__ void __launch_bounds__(kBlock) vof_normals(const float* __restrict__ phi, float* __restrict__ geo, Dims d) {
    for_stencil_cells<kRowsGeometry, kRemapGeometry>(d, [&](const Around& a) {
        const uint32_t n = a.n;
        float gx, gy, gz;
        fill_gradient(phi, a, gx, gy, gz);
        const float mag = sqrtf(gx * gx + gy * gy + gz * gz);
        float* nxv = geo;
        float* nyv = geo + d.cells;
        float* nzv = geo + 2 * static_cast<size_t>(d.cells);
        if (mag > 1.0e-12F) {
            el(nxv, n) = gx / mag;
            el(nyv, n) = gy / mag;
            el(nzv, n) = gz / mag;
        } else {
            el(nxv, n) = 0.0F;
            el(nyv, n) = 0.0F;
            el(nzv, n) = 0.0F;
        }
    });
}

// curvature = -div(n) by central differences (resident_multiphysics_sycl.hpp:287-296).
__global__ void __launch_bounds__(kBlock) vof_curvature(float* geo, Dims d) {
    for_stencil_cells<kRowsGeometry, kRemapGeometry>(d, [&](const Around& a) {
        const uint32_t n = a.n;
        const float* nxv = geo;
        const float* nyv = geo + d.cells;
        const float* nzv = geo + 2 * static_cast<size_t>(d.cells);
        el(geo + 3 * static_cast<size_t>(d.cells), n) =
            -0.5F * ((ld(nxv, a.xp) - ld(nxv, a.xm)) + (ld(nyv, a.yp) - ld(nyv, a.ym)) + (ld(nzv, a.zp) - ld(nzv, a.zm)));
    });
}

// a += (sigma kappa / rho) grad(phi)  (resident_multiphysics_sycl.hpp:299-317)
__global__ void __launch_bounds__(kBlock) vof_tension(const float* __restrict__ phi, const float* __restrict__ curvature,
                                                      float* ax, float* ay, float* az, Dims d, float sigma,
                                                      float density) {
    for_stencil_cells<kRowsTension, false>(d, [&](const Around& a) {
        const uint32_t n = a.n;
        float gx, gy, gz;
        fill_gradient(phi, a, gx, gy, gz);
        const float c = sigma * ld(curvature, n) / density;
        el(ax, n) += c * gx;
        el(ay, n) += c * gy;
        el(az, n) += c * gz;
    });
}

// ------------------------------------------------------------------ particles
__device__ __forceinline__ long long floor_to_ll(float value) {
    long long i = static_cast<long long>(value);
    if (static_cast<float>(i) > value) --i;
    return i;
}
__device__ __forceinline__ float periodic_coordinate(float value, uint32_t extent) {
    const float e = static_cast<float>(extent);
    const long long turns = floor_to_ll(value / e);
    value -= static_cast<float>(turns) * e;
    if (value >= e) value -= e;
    if (value < 0.0F) value += e;
    return value;
}
__device__ __forceinline__ uint32_t wrap_cell(long long v, uint32_t extent) {
    const long long e = static_cast<long long>(extent);
    // Corners of a wrapped position lie in [0, extent + 1]: one subtraction (the 64-bit remainder
    // below gives the same value there and keeps any other input inside the lattice).
    if (v >= 0 && v < 2 * e) return static_cast<uint32_t>(v >= e ? v - e : v);
    v %= e;
    if (v < 0) v += e;
    return static_cast<uint32_t>(v);
}

// State: 8 planes of count floats (x, y, z, vx, vy, vz, mass, response time).  Trilinear fluid
// velocity, Stokes drag m (u - v) / tau_p, explicit Euler, periodic positions; with two_way the
// reaction -drag is spread to the 8 surrounding cells with the same weights
// (resident_multiphysics_sycl.hpp:346-366).
__global__ void __launch_bounds__(kBlock) particles_advance(float* s, uint32_t count, const float* __restrict__ ux,
                                                            const float* __restrict__ uy, const float* __restrict__ uz,
                                                            float* ax, float* ay, float* az, Dims d, float dt,
                                                            int two_way, float gx, float gy, float gz) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    float px = periodic_coordinate(s[p], d.nx), py = periodic_coordinate(s[count + p], d.ny),
          pz = periodic_coordinate(s[2U * count + p], d.nz);
    float vx = s[3U * count + p], vy = s[4U * count + p], vz = s[5U * count + p];
    const float mass = s[6U * count + p], response = s[7U * count + p];
    const long long x0 = floor_to_ll(px), y0 = floor_to_ll(py), z0 = floor_to_ll(pz);
    const float tx = px - static_cast<float>(x0), ty = py - static_cast<float>(y0), tz = pz - static_cast<float>(z0);
    float fu = 0.0F, fv = 0.0F, fw = 0.0F;
    uint32_t ids[8];
    float weights[8];
    int k = 0;
    for (int dz = 0; dz <= 1; ++dz)
        for (int dy = 0; dy <= 1; ++dy)
            for (int dx = 0; dx <= 1; ++dx) {
                const float wx = dx ? tx : 1.0F - tx, wy = dy ? ty : 1.0F - ty, wz = dz ? tz : 1.0F - tz;
                const uint32_t n = at(wrap_cell(x0 + dx, d.nx), wrap_cell(y0 + dy, d.ny), wrap_cell(z0 + dz, d.nz), d);
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
            atomicAdd(&ax[ids[j]], -dragx * weights[j]);
            atomicAdd(&ay[ids[j]], -dragy * weights[j]);
            atomicAdd(&az[ids[j]], -dragz * weights[j]);
        }
    vx += dt * (dragx / mass + gx);
    vy += dt * (dragy / mass + gy);
    vz += dt * (dragz / mass + gz);
    px = periodic_coordinate(px + dt * vx, d.nx);
    py = periodic_coordinate(py + dt * vy, d.ny);
    pz = periodic_coordinate(pz + dt * vz, d.nz);
    s[p] = px;
    s[count + p] = py;
    s[2U * count + p] = pz;
    s[3U * count + p] = vx;
    s[4U * count + p] = vy;
    s[5U * count + p] = vz;
}

// ------------------------------------------------------------------ Q-criterion
// Q = (|Omega|^2 - |S|^2) / 2 from central differences (resident_multiphysics_sycl.hpp:394-404).
__global__ void __launch_bounds__(kBlock) q_criterion(const Inputs<3> u, float* __restrict__ q, Dims d, uint32_t zchunk) {
    zmarch<3, 0b111U, 0b111U, kRowsQ>(u, d, zchunk, [&](const Nbr<3>& s, uint32_t n) {
        float g[3][3];
    #pragma unroll
        for (int c = 0; c < 3; ++c) {
            g[c][0] = 0.5F * (s.xp[c] - s.xm[c]);
            g[c][1] = 0.5F * (s.yp[c] - s.ym[c]);
            g[c][2] = 0.5F * (s.zp[c] - s.zm[c]);
        }
        float s2 = 0.0F, o2 = 0.0F;
    #pragma unroll
        for (int i = 0; i < 3; ++i)
    #pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float sij = 0.5F * (g[i][j] + g[j][i]), oij = 0.5F * (g[i][j] - g[j][i]);
                s2 += sij * sij;
                o2 += oij * oij;
            }
        el(q, n) = 0.5F * (o2 - s2);
    });
}

// ------------------------------------------------------------------ host side
uint32_t blocks_for(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

// Launch shape of the stencil kernels: bx x by threads over x and y (bx the smallest power of two
// from 64 to 256 that is >= nx, or 256; by = 256 / bx), rows rows per thread, one z plane per grid
// layer.
struct StencilShape {
    dim3 grid, block;
};
StencilShape stencil_shape(const Dims& d, uint32_t rows) {
    uint32_t bx = 64;
    while (bx < static_cast<uint32_t>(kBlock) && bx < d.nx) bx <<= 1;
    const uint32_t by = kBlock / bx;
    return {dim3((d.nx + bx - 1) / bx, (d.ny + by * rows - 1) / (by * rows), d.nz), dim3(bx, by, 1)};
}

// Launch shape of the z-marching kernels: kTileX x (kTileWaves * rows) tiles, and z chunks sized so
// that the launch has about kZBlocks blocks.
struct ZShape {
    dim3 grid, block;
    uint32_t zchunk;
};
ZShape zmarch_shape(const Dims& d, uint32_t rows) {
    const uint32_t ty = kTileWaves * rows;
    const uint32_t tiles_x = (d.nx + kTileX - 1) / kTileX, tiles_y = (d.ny + ty - 1) / ty;
    const uint64_t tiles = static_cast<uint64_t>(tiles_x) * tiles_y;
    uint32_t chunks = tiles >= kZBlocks ? 1U : static_cast<uint32_t>((kZBlocks + tiles - 1) / tiles);
    chunks = std::min(chunks, d.nz);
    const uint32_t zchunk = (d.nz + chunks - 1) / chunks;
    chunks = (d.nz + zchunk - 1) / zchunk;
    return {dim3(tiles_x, tiles_y, chunks), dim3(kTileX, kTileWaves, 1), zchunk};
}

struct Stats {
    uint64_t h2d = 0, d2h = 0, d2d = 0, syncs = 0;
};

struct Module {
    FlowInfo flow{};
    Dims dims{};
    Stats stats{};
    explicit Module(cfd_lbm_solver* s, uint32_t min_extent, const char* what) : flow(cfd_hip::flow_info(s)) {
        if (flow.real != 4)
            throw std::invalid_argument(std::string(what) +
                                        " couples to FP32 device views: the flow needs FP32, FP16S or FP32S storage");
        if (flow.nx < min_extent || flow.ny < min_extent || flow.nz < min_extent)
            throw std::invalid_argument(std::string("invalid resident ") + what + " lattice");
        if (flow.cells > (1ull << 30))  // field planes are addressed with 32-bit byte offsets
            throw std::invalid_argument(std::string("resident ") + what + " lattice too large: at most 2^30 cells");
        dims = {static_cast<uint32_t>(flow.nx), static_cast<uint32_t>(flow.ny), static_cast<uint32_t>(flow.nz),
                static_cast<uint32_t>(flow.cells)};
        const StencilShape widest = stencil_shape(dims, 1);
        if (widest.grid.y > 65535U || widest.grid.z > 65535U)  // grid limits of the y and z launch dimensions
            throw std::invalid_argument(std::string("resident ") + what + " lattice too tall: at most 65535 z planes and " +
                                        std::to_string(65535U * widest.block.y) + " rows per plane");
    }
    StencilShape shape(uint32_t rows) const { return stencil_shape(dims, rows); }
    ZShape zshape(uint32_t rows) const { return zmarch_shape(dims, rows); }
    void use() const { check(hipSetDevice(flow.device), "hipSetDevice"); }
    float* alloc(uint64_t floats) const {
        void* p = nullptr;
        check(hipMalloc(&p, floats * sizeof(float)), "hipMalloc");
        return static_cast<float*>(p);
    }
    void h2d(void* dst, const void* src, size_t bytes) {
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, flow.stream), "hipMemcpyAsync");
        stats.h2d += bytes;
    }
    void d2h(void* dst, const void* src, size_t bytes) {
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, flow.stream), "hipMemcpyAsync");
        stats.d2h += bytes;
    }
    void sync() {
        check(hipStreamSynchronize(flow.stream), "hipStreamSynchronize");
        ++stats.syncs;
    }
    // FP64 sum of count floats on the device; one synchronisation.
    double sum(const float* v, uint64_t count) {
        if (!partial) {
            void* p = nullptr;
            check(hipMalloc(&p, (kReduceBlocks + 1) * sizeof(double)), "hipMalloc");
            partial = static_cast<double*>(p);
        }
        sum_partial<<<kReduceBlocks, kBlock, 0, flow.stream>>>(v, count, partial);
        sum_final<<<1, kBlock, 0, flow.stream>>>(partial, kReduceBlocks, partial + kReduceBlocks);
        check(hipGetLastError(), "sum");
        // The 8-byte result is not a field transfer: the ledger counts it as a synchronisation only,
        // as the reference's does for its shared-memory reduction scalar.
        double out = 0.0;
        check(hipMemcpyAsync(&out, partial + kReduceBlocks, sizeof(double), hipMemcpyDeviceToHost, flow.stream),
              "hipMemcpyAsync");
        sync();
        return out;
    }
    ~Module() {
        if (partial) (void)hipFree(partial);
    }
    Module(const Module&) = delete;
    Module& operator=(const Module&) = delete;
    double* partial = nullptr;
};

void need_views(std::initializer_list<const void*> v, const char* what) {
    for (const void* p : v)
        if (!p) throw std::invalid_argument(std::string(what) + " needs device fields from the flow solver");
}

template <class F> int guarded(char* error, size_t cap, F&& fn) {
    try {
        fn();
        if (error && cap) error[0] = '\0';
        return 0;
    } catch (const std::exception& e) {
        if (error && cap) std::snprintf(error, cap, "%s", e.what());
        return 1;
    } catch (...) {
        if (error && cap) std::snprintf(error, cap, "unknown error");
        return 2;
    }
}

}  // namespace

struct cfd_lbm_thermal : Module {
    float tau = 0.0F;
    float* g = nullptr;
    float* next = nullptr;
    float* scalar = nullptr;
    cfd_lbm_thermal(cfd_lbm_solver* s, float diffusivity) : Module(s, 2, "thermal") {
        if (!(diffusivity > 0.0F) || !std::isfinite(diffusivity)) throw std::invalid_argument("invalid resident thermal lattice");
        tau = 0.5F + 4.0F * diffusivity;  // D = c_s^2 (tau - 1/2) with c_s^2 = 1/4
        use();
        try {
            g = alloc(7ull * dims.cells);
            next = alloc(7ull * dims.cells);
            fill(nullptr, 0.0F);
        } catch (...) {
            release();  // a constructor that throws runs no destructor
            throw;
        }
    }
    ~cfd_lbm_thermal() { release(); }
    void release() {
        (void)hipStreamSynchronize(flow.stream);
        for (float** p : {&g, &next, &scalar})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
    }
    void fill(const float* device_field, float value) {
        thermal_fill<<<blocks_for(dims.cells), kBlock, 0, flow.stream>>>(g, device_field, value, dims);
        check(hipGetLastError(), "thermal_fill");
    }
    float* refresh_scalar() {
        if (!scalar) scalar = alloc(dims.cells);
        thermal_scalar<<<blocks_for(dims.cells), kBlock, 0, flow.stream>>>(g, scalar, dims);
        check(hipGetLastError(), "thermal_scalar");
        return scalar;
    }
};

struct cfd_lbm_surface : Module {
    float* fill = nullptr;
    float* next = nullptr;
    float* geometry = nullptr;  // normal x, y, z and curvature planes
    explicit cfd_lbm_surface(cfd_lbm_solver* s) : Module(s, 2, "free-surface") {
        use();
        try {
            fill = alloc(dims.cells);
            next = alloc(dims.cells);
            geometry = alloc(4ull * dims.cells);
            init(nullptr, 0.0F, false, 0.0F);
        } catch (...) {
            release();  // a constructor that throws runs no destructor
            throw;
        }
    }
    ~cfd_lbm_surface() { release(); }
    void release() {
        (void)hipStreamSynchronize(flow.stream);
        for (float** p : {&fill, &next, &geometry})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
    }
    void init(const float* device_field, float value, bool half, float x) {
        vof_fill<<<blocks_for(dims.cells), kBlock, 0, flow.stream>>>(fill, device_field, value, half ? 1 : 0, x, dims);
        check(hipGetLastError(), "vof_fill");
    }
    void refresh_geometry() {
        const StencilShape sh = shape(kRowsGeometry);
        vof_normals<<<sh.grid, sh.block, 0, flow.stream>>>(fill, geometry, dims);
        vof_curvature<<<sh.grid, sh.block, 0, flow.stream>>>(geometry, dims);
        check(hipGetLastError(), "vof_geometry");
    }
};

struct cfd_lbm_particles : Module {
    uint32_t count = 0;
    float* state = nullptr;
    // Device slot j holds the caller's particle order[j]: upload sorts the particles by the cell
    // they start in, so that neighbouring lanes gather neighbouring cells' velocities.  Each
    // particle's update is independent of the others, so the order changes no particle's result.
    std::vector<uint32_t> order;
    cfd_lbm_particles(cfd_lbm_solver* s, size_t n) : Module(s, 2, "particle") {
        if (n > 0x7FFFFFFFull) throw std::invalid_argument("too many particles");
        count = static_cast<uint32_t>(n);
        use();
        if (count) state = alloc(8ull * count);
        order.resize(count);
        for (uint32_t j = 0; j < count; ++j) order[j] = j;
    }
    ~cfd_lbm_particles() {
        (void)hipStreamSynchronize(flow.stream);
        if (state) (void)hipFree(state);
    }
};

struct cfd_lbm_flow_diagnostics : Module {
    float* q = nullptr;
    explicit cfd_lbm_flow_diagnostics(cfd_lbm_solver* s) : Module(s, 3, "Q-criterion") {
        use();
        q = alloc(dims.cells);
    }
    ~cfd_lbm_flow_diagnostics() {
        (void)hipStreamSynchronize(flow.stream);
        if (q) (void)hipFree(q);
    }
};

namespace {
template <class T> T& need_obj(T* p) {
    if (!p) throw std::invalid_argument("null module handle");
    return *p;
}
template <class T> void fill_stats(T& m, uint64_t* out, int reset) {
    if (out) {
        out[0] = m.stats.h2d;
        out[1] = m.stats.d2h;
        out[2] = m.stats.d2d;
        out[3] = m.stats.syncs;
    }
    if (reset) m.stats = {};
}
}  // namespace

extern "C" {

// ------------------------------------------------------------------ thermal
int cfd_lbm_thermal_create(cfd_lbm_solver* flow, float diffusivity, cfd_lbm_thermal** out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!out) throw std::invalid_argument("thermal_create needs an output handle");
        *out = new cfd_lbm_thermal(flow, diffusivity);
    });
}
int cfd_lbm_thermal_destroy(cfd_lbm_thermal* t, char* error, size_t cap) {
    return guarded(error, cap, [&] { delete t; });
}
int cfd_lbm_thermal_initialize_uniform(cfd_lbm_thermal* t, float temperature, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!std::isfinite(temperature)) throw std::invalid_argument("thermal state must be finite");
        need_obj(t).use();
        t->fill(nullptr, temperature);
    });
}
int cfd_lbm_thermal_initialize_fields(cfd_lbm_thermal* t, const float* temperature, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!temperature) throw std::invalid_argument("thermal_initialize_fields needs cells values");
        need_obj(t).use();
        for (uint32_t i = 0; i < t->dims.cells; ++i)
            if (!std::isfinite(temperature[i])) throw std::invalid_argument("thermal state must be finite");
        float* tmp = t->refresh_scalar();  // scratch plane, rewritten by the next refresh
        t->h2d(tmp, temperature, t->dims.cells * sizeof(float));
        t->fill(tmp, 0.0F);
        t->sync();
    });
}
int cfd_lbm_thermal_step(cfd_lbm_thermal* t, const float* ux, const float* uy, const float* uz, uint64_t count,
                         char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(t).use();
        need_views({ux, uy, uz}, "thermal step");
        const float omega = 1.0F / t->tau;
        const StencilShape sh = t->shape(kRowsThermal);
        for (uint64_t k = 0; k < count; ++k) {
            thermal_step<<<sh.grid, sh.block, 0, t->flow.stream>>>(t->g, t->next, ux, uy, uz, t->dims, omega);
            check(hipGetLastError(), "thermal_step");
            std::swap(t->g, t->next);
        }
    });
}
int cfd_lbm_thermal_refresh_device_scalar(cfd_lbm_thermal* t, void** scalar, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(t).use();
        float* s = t->refresh_scalar();
        if (scalar) *scalar = s;
    });
}
int cfd_lbm_thermal_download_scalar(cfd_lbm_thermal* t, float* out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!out) throw std::invalid_argument("thermal_download_scalar needs cells values");
        need_obj(t).use();
        t->d2h(out, t->refresh_scalar(), t->dims.cells * sizeof(float));
        t->sync();
    });
}
int cfd_lbm_thermal_add_boussinesq(cfd_lbm_thermal* t, float* ax, float* ay, float* az, float t_ref, float beta,
                                   double gx, double gy, double gz, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(t).use();
        need_views({ax, ay, az}, "Boussinesq coupling");
        if (!std::isfinite(t_ref) || !std::isfinite(beta)) throw std::invalid_argument("invalid resident Boussinesq coupling");
        if (!t->scalar) t->scalar = t->alloc(t->dims.cells);
        boussinesq_from_populations<<<blocks_for(t->dims.cells), kBlock, 0, t->flow.stream>>>(
            t->g, t->scalar, ax, ay, az, t->dims, t_ref, beta, static_cast<float>(gx), static_cast<float>(gy),
            static_cast<float>(gz));
        check(hipGetLastError(), "boussinesq");
    });
}
int cfd_lbm_thermal_total_scalar(cfd_lbm_thermal* t, double* total, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(t).use();
        const double v = t->sum(t->g, 7ull * t->dims.cells);
        if (total) *total = v;
    });
}
int cfd_lbm_thermal_resident_bytes(cfd_lbm_thermal* t, size_t* bytes, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (bytes) *bytes = (14ull + (need_obj(t).scalar ? 1ull : 0ull)) * t->dims.cells * sizeof(float);
    });
}
int cfd_lbm_thermal_transfer_stats(cfd_lbm_thermal* t, uint64_t* stats4, int reset, char* error, size_t cap) {
    return guarded(error, cap, [&] { fill_stats(need_obj(t), stats4, reset); });
}

// ------------------------------------------------------------------ free surface
int cfd_lbm_surface_create(cfd_lbm_solver* flow, cfd_lbm_surface** out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!out) throw std::invalid_argument("surface_create needs an output handle");
        *out = new cfd_lbm_surface(flow);
    });
}
int cfd_lbm_surface_destroy(cfd_lbm_surface* s, char* error, size_t cap) {
    return guarded(error, cap, [&] { delete s; });
}
int cfd_lbm_surface_initialize_uniform(cfd_lbm_surface* s, float fill, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!(fill >= 0.0F && fill <= 1.0F)) throw std::invalid_argument("free-surface fill must be in [0,1]");
        need_obj(s).use();
        s->init(nullptr, fill, false, 0.0F);
    });
}
int cfd_lbm_surface_initialize_half_space_x(cfd_lbm_surface* s, float interface_x, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!std::isfinite(interface_x)) throw std::invalid_argument("invalid free-surface interface position");
        need_obj(s).use();
        s->init(nullptr, 0.0F, true, interface_x);
    });
}
int cfd_lbm_surface_upload_fill(cfd_lbm_surface* s, const float* fill, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!fill) throw std::invalid_argument("surface_upload_fill needs cells values");
        need_obj(s).use();
        for (uint32_t i = 0; i < s->dims.cells; ++i)
            if (!(fill[i] >= 0.0F && fill[i] <= 1.0F)) throw std::invalid_argument("free-surface fill must be in [0,1]");
        s->h2d(s->fill, fill, s->dims.cells * sizeof(float));
        s->sync();
    });
}
int cfd_lbm_surface_download_fill(cfd_lbm_surface* s, float* fill, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!fill) throw std::invalid_argument("surface_download_fill needs cells values");
        need_obj(s).use();
        s->d2h(fill, s->fill, s->dims.cells * sizeof(float));
        s->sync();
    });
}
int cfd_lbm_surface_advect(cfd_lbm_surface* s, const float* ux, const float* uy, const float* uz, float dt, char* error,
                           size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(s).use();
        need_views({ux, uy, uz}, "free-surface advection");
        if (!(dt > 0.0F)) throw std::invalid_argument("invalid resident free-surface advection");
        const ZShape sh = s->zshape(kRowsAdvect);
        vof_advect<<<sh.grid, sh.block, 0, s->flow.stream>>>(Inputs<4>{{s->fill, ux, uy, uz}}, s->next, s->dims, dt,
                                                             sh.zchunk);
        check(hipGetLastError(), "vof_advect");
        std::swap(s->fill, s->next);
    });
}
int cfd_lbm_surface_fill_device(cfd_lbm_surface* s, void** fill, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (fill) *fill = need_obj(s).fill;
    });
}
int cfd_lbm_surface_refresh_geometry(cfd_lbm_surface* s, void** normal_x, void** normal_y, void** normal_z,
                                     void** curvature, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(s).use();
        s->refresh_geometry();
        const size_t c = s->dims.cells;
        if (normal_x) *normal_x = s->geometry;
        if (normal_y) *normal_y = s->geometry + c;
        if (normal_z) *normal_z = s->geometry + 2 * c;
        if (curvature) *curvature = s->geometry + 3 * c;
    });
}
int cfd_lbm_surface_add_surface_tension(cfd_lbm_surface* s, float* ax, float* ay, float* az, float surface_tension,
                                        float density, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(s).use();
        need_views({ax, ay, az}, "surface-tension coupling");
        if (!(surface_tension >= 0.0F) || !(density > 0.0F))
            throw std::invalid_argument("invalid resident surface-tension coupling");
        s->refresh_geometry();
        const StencilShape sh = s->shape(kRowsTension);
        vof_tension<<<sh.grid, sh.block, 0, s->flow.stream>>>(
            s->fill, s->geometry + 3ull * s->dims.cells, ax, ay, az, s->dims, surface_tension, density);
        check(hipGetLastError(), "vof_tension");
    });
}
int cfd_lbm_surface_volume(cfd_lbm_surface* s, double* volume, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(s).use();
        const double v = s->sum(s->fill, s->dims.cells);
        if (volume) *volume = v;
    });
}
int cfd_lbm_surface_transfer_stats(cfd_lbm_surface* s, uint64_t* stats4, int reset, char* error, size_t cap) {
    return guarded(error, cap, [&] { fill_stats(need_obj(s), stats4, reset); });
}

// ------------------------------------------------------------------ particles
int cfd_lbm_particles_create(cfd_lbm_solver* flow, size_t count, cfd_lbm_particles** out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!out) throw std::invalid_argument("particles_create needs an output handle");
        *out = new cfd_lbm_particles(flow, count);
    });
}
int cfd_lbm_particles_destroy(cfd_lbm_particles* p, char* error, size_t cap) {
    return guarded(error, cap, [&] { delete p; });
}
int cfd_lbm_particles_upload(cfd_lbm_particles* p, const double* particles, size_t count, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (count != need_obj(p).count) throw std::invalid_argument("resident particle count mismatch");
        if (!count) return;
        if (!particles) throw std::invalid_argument("particles_upload needs count x 8 values");
        p->use();
        std::vector<uint64_t> key(count);
        for (size_t i = 0; i < count; ++i) {
            const double* q = particles + 8 * i;
            if (!(q[6] > 0.0) || !(q[7] > 0.0)) throw std::invalid_argument("particle mass and response time must be positive");
            for (int k = 0; k < 6; ++k)
                if (!std::isfinite(q[k])) throw std::invalid_argument("particle position and velocity must be finite");
            auto cell = [](double v, uint32_t extent) {
                const double e = static_cast<double>(extent);
                double w = std::fmod(v, e);
                if (w < 0.0) w += e;
                const uint64_t c = static_cast<uint64_t>(w);
                return c < extent ? c : uint64_t{0};
            };
            const uint64_t n = (cell(q[2], p->dims.nz) * p->dims.ny + cell(q[1], p->dims.ny)) * p->dims.nx +
                               cell(q[0], p->dims.nx);
            key[i] = (n << 32) | i;  // cell, then the caller's order: a deterministic permutation
        }
        std::sort(key.begin(), key.end());
        p->order.resize(count);
        std::vector<float> host(8ull * count);
        for (size_t j = 0; j < count; ++j) {
            const uint32_t i = static_cast<uint32_t>(key[j] & 0xFFFFFFFFull);
            p->order[j] = i;
            for (int k = 0; k < 8; ++k) host[k * count + j] = static_cast<float>(particles[8 * static_cast<size_t>(i) + k]);
        }
        p->h2d(p->state, host.data(), host.size() * sizeof(float));
        p->sync();
    });
}
int cfd_lbm_particles_download(cfd_lbm_particles* p, double* particles, size_t count, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (count != need_obj(p).count) throw std::invalid_argument("resident particle count mismatch");
        if (!count) return;
        if (!particles) throw std::invalid_argument("particles_download needs count x 8 values");
        p->use();
        std::vector<float> host(8ull * count);
        p->d2h(host.data(), p->state, host.size() * sizeof(float));
        p->sync();
        for (size_t j = 0; j < count; ++j)
            for (int k = 0; k < 8; ++k)
                particles[8 * static_cast<size_t>(p->order[j]) + k] = static_cast<double>(host[k * count + j]);
    });
}
int cfd_lbm_particles_advance(cfd_lbm_particles* p, const float* ux, const float* uy, const float* uz, float* ax,
                              float* ay, float* az, float dt, int two_way, double gx, double gy, double gz, char* error,
                              size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(p).use();
        need_views({ux, uy, uz, ax, ay, az}, "particle coupling");
        if (!(dt > 0.0F)) throw std::invalid_argument("invalid resident particle coupling");
        if (!p->count) return;
        particles_advance<<<blocks_for(p->count), kBlock, 0, p->flow.stream>>>(
            p->state, p->count, ux, uy, uz, ax, ay, az, p->dims, dt, two_way ? 1 : 0, static_cast<float>(gx),
            static_cast<float>(gy), static_cast<float>(gz));
        check(hipGetLastError(), "particles_advance");
    });
}
int cfd_lbm_particles_transfer_stats(cfd_lbm_particles* p, uint64_t* stats4, int reset, char* error, size_t cap) {
    return guarded(error, cap, [&] { fill_stats(need_obj(p), stats4, reset); });
}

// ------------------------------------------------------------------ Q-criterion
int cfd_lbm_flow_diagnostics_create(cfd_lbm_solver* flow, cfd_lbm_flow_diagnostics** out, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        if (!out) throw std::invalid_argument("flow_diagnostics_create needs an output handle");
        *out = new cfd_lbm_flow_diagnostics(flow);
    });
}
int cfd_lbm_flow_diagnostics_destroy(cfd_lbm_flow_diagnostics* d, char* error, size_t cap) {
    return guarded(error, cap, [&] { delete d; });
}
int cfd_lbm_q_criterion(cfd_lbm_flow_diagnostics* d, const float* ux, const float* uy, const float* uz, void** q,
                        char* error, size_t cap) {
    return guarded(error, cap, [&] {
        need_obj(d).use();
        need_views({ux, uy, uz}, "Q-criterion");
        const ZShape sh = d->zshape(kRowsQ);
        q_criterion<<<sh.grid, sh.block, 0, d->flow.stream>>>(Inputs<3>{{ux, uy, uz}}, d->q, d->dims, sh.zchunk);
        check(hipGetLastError(), "q_criterion");
        if (q) *q = d->q;
    });
}

// ------------------------------------------------------------------ device copies for callers
int cfd_lbm_device_download(cfd_lbm_solver* flow, const void* device, void* host, size_t bytes, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const FlowInfo f = cfd_hip::flow_info(flow);
        if (bytes && (!device || !host)) throw std::invalid_argument("device_download needs both pointers");
        check(hipSetDevice(f.device), "hipSetDevice");
        check(hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, f.stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(f.stream), "hipStreamSynchronize");
    });
}
int cfd_lbm_device_upload(cfd_lbm_solver* flow, void* device, const void* host, size_t bytes, char* error, size_t cap) {
    return guarded(error, cap, [&] {
        const FlowInfo f = cfd_hip::flow_info(flow);
        if (bytes && (!device || !host)) throw std::invalid_argument("device_upload needs both pointers");
        check(hipSetDevice(f.device), "hipSetDevice");
        check(hipMemcpyAsync(device, host, bytes, hipMemcpyHostToDevice, f.stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(f.stream), "hipStreamSynchronize");
    });
}

}  // extern "C"
