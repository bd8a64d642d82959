// voxelize.hip -- ray-parity voxelisation of a closed triangle surface on the GPU.
//
// Reference: src/lbm/voxelize.cpp:4-5 (CPU, double) and src/sycl/voxelize_sycl.cpp:13-68 (device,
// float).  A cell is solid when the ray from its centre along +x crosses the surface an odd number
// of times; the centre is nudged by (0, 1e-10 dx, 2e-10 dx) so that rays miss shared edges and
// vertices of axis-aligned meshes, and the crossing test is Moller-Trumbore with the reference's
// thresholds (|det| < 1e-12 parallel, hit when the ray parameter exceeds 1e-10).
//
// The reference tests every triangle against every cell.  For a ray along +x, the barycentric
// coordinates u and v of the crossing do not depend on the ray origin's x (the x terms meet the
// zero components of d = (1,0,0)), so here they are evaluated once per x-line and triangle:
//   1. binning: every triangle whose yz bounding box (plus half a cell) overlaps a tile of 8 x 8
//      x-lines is listed under that tile (count, scan, fill);
//   2. lines: one workgroup per tile.  Its four wavefronts test the tile's triangles against the
//      64 lines (u, v), keep the crossing triangles of each line in LDS, and then sweep the cells
//      of the lines along x (each wavefront its own lines, 8 cells per lane), evaluating the ray
//      parameter of each crossing with the reference's per-cell formula and writing the parity
//      bytes (64-bit stores, coalesced along x).
// All arithmetic is double with the reference's operation order and no contraction, so the mask
// equals voxelize.cpp's bit for bit.  Work: (lines x triangles of a tile) u,v tests + (cells x
// crossings per line) distance tests, instead of cells x triangles.
#include "voxelize.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

// Working memory of voxelize_surface_no_alloc, one copy per GPU for the life of the process:
// header (pair total, non-finite flag), tile counts, tile offsets, tile cursors, pair list.
// Host code never names it after set-up: its device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[cfd_hip::kVoxelizeScratchBytes];

namespace cfd_hip {
namespace {

constexpr int kTileJ = 8, kTileK = 8, kTileLines = kTileJ * kTileK;  // 64 x-lines per tile
constexpr int kThreads = 256;                                          // 4 wavefronts per tile
constexpr int kChunk = 64;  // triangles per LDS round: at most kChunk crossings per line
constexpr int kWaves = kThreads / 64;
constexpr int kLaneCells = 8;  // consecutive cells of a line per lane in the sweep
constexpr int kScanThreads = 1024;

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct Tri {
    double ax, ay, az, bx, by, bz, cx, cy, cz;
};

// Line-constant parts of one crossing triangle (voxelize.cpp:4 names: edges e1, e2, 1/det, the
// vertex a, s = o - a with o on the line, q = s x e1): the terms of the ray parameter that do not
// depend on the cell's x.
struct Crossing {
    double inv, e1y, e1z, e2y, e2z, ax, sz_e1x, sy_e1x, e2x_qx;
};

__device__ __forceinline__ double read_lane(double v, uint32_t lane) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), static_cast<int>(lane)));
    const uint32_t hi =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), static_cast<int>(lane)));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Crossing `lane` of the wavefront, read from that lane's registers (lane uniform).
__device__ __forceinline__ Crossing read_lane(const Crossing& x, uint32_t lane) {
    return {read_lane(x.inv, lane),    read_lane(x.e1y, lane),    read_lane(x.e1z, lane),
            read_lane(x.e2y, lane),    read_lane(x.e2z, lane),    read_lane(x.ax, lane),
            read_lane(x.sz_e1x, lane), read_lane(x.sy_e1x, lane), read_lane(x.e2x_qx, lane)};
}

struct LineGrid {
    uint32_t nx, ny, nz, tiles_j, tiles_k;
    uint32_t packed;  // rows start 8-byte aligned (nx % 8 == 0 and an aligned mask): 64-bit stores
    double minx, miny, minz, dx, dy, dz;
};

// det = dot(e1, d x e2) with d = (1,0,0): d x e2 = (0*e2z - 0*e2y, 0*e2x - 1*e2z, 1*e2y - 0*e2x).
__device__ __forceinline__ double tri_det(const Tri& t) {
    const double e1x = t.bx - t.ax, e1y = t.by - t.ay, e1z = t.bz - t.az;
    const double e2x = t.cx - t.ax, e2y = t.cy - t.ay, e2z = t.cz - t.az;
    const double hx = 0.0 * e2z - 0.0 * e2y, hy = 0.0 * e2x - 1.0 * e2z, hz = 1.0 * e2y - 0.0 * e2x;
    return e1x * hx + e1y * hy + e1z * hz;
}

// Range of line tiles [t0, t1] along one axis whose lines can cross [lo, hi]: lines at least half
// a cell outside the bounding box are left out; the rounding error of u and v is ~1e-16 of the
// coordinates, far below half a cell.  Returns false when no line of the axis is in range.
__device__ __forceinline__ bool tile_range(double lo, double hi, double origin, double step, uint32_t n, uint32_t tile,
                                           uint32_t& t0, uint32_t& t1) {
    const double slack = 1e-9 * (fabs(lo) + fabs(hi) + fabs(origin));
    double a = floor((lo - slack - origin) / step) - 1.0;
    double b = floor((hi + slack - origin) / step) + 1.0;
    if (!(b >= 0.0) || !(a <= static_cast<double>(n) - 1.0)) return false;
    a = fmax(a, 0.0);
    b = fmin(b, static_cast<double>(n) - 1.0);
    t0 = static_cast<uint32_t>(a) / tile;
    t1 = static_cast<uint32_t>(b) / tile;
    return true;
}

__device__ __forceinline__ bool tri_tiles(const Tri& t, const LineGrid& g, uint32_t& j0, uint32_t& j1, uint32_t& k0,
                                          uint32_t& k1) {
    if (fabs(tri_det(t)) < 1e-12) return false;  // voxelize.cpp:4: parallel to the ray, never hit
    const double ylo = fmin(t.ay, fmin(t.by, t.cy)), yhi = fmax(t.ay, fmax(t.by, t.cy));
    const double zlo = fmin(t.az, fmin(t.bz, t.cz)), zhi = fmax(t.az, fmax(t.bz, t.cz));
    return tile_range(ylo, yhi, g.miny, g.dy, g.ny, kTileJ, j0, j1) &&
           tile_range(zlo, zhi, g.minz, g.dz, g.nz, kTileK, k0, k1);
}

__global__ void __launch_bounds__(256) bin_count(const Tri* tris, uint32_t count, LineGrid g, uint32_t* tile_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Tri t = tris[i];
    uint32_t j0, j1, k0, k1;
    if (!tri_tiles(t, g, j0, j1, k0, k1)) return;
    for (uint32_t k = k0; k <= k1; ++k)
        for (uint32_t j = j0; j <= j1; ++j) atomicAdd(&tile_count[k * g.tiles_j + j], 1U);
}

// Exclusive scan of tile_count into tile_offset[0..tiles] (tile_offset[tiles] = total), one
// workgroup; the total is also written as 64 bits so that overflow of 32-bit offsets is caught.
__global__ void __launch_bounds__(kScanThreads) bin_scan(const uint32_t* tile_count, uint32_t tiles,
                                                          uint32_t* tile_offset, unsigned long long* total) {
    __shared__ unsigned long long part[kScanThreads];
    const uint32_t per = (tiles + kScanThreads - 1) / kScanThreads;
    const uint32_t begin = threadIdx.x * per, end = min(begin + per, tiles);
    unsigned long long s = 0;
    for (uint32_t i = begin; i < end; ++i) s += tile_count[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const unsigned long long v = threadIdx.x >= static_cast<unsigned>(d) ? part[threadIdx.x - d] : 0ULL;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = part[threadIdx.x] - s;  // exclusive prefix of this thread's chunk
    for (uint32_t i = begin; i < end; ++i) {
        tile_offset[i] = static_cast<uint32_t>(run);
        run += tile_count[i];
    }
    if (threadIdx.x == kScanThreads - 1) {
        tile_offset[tiles] = static_cast<uint32_t>(part[threadIdx.x]);
        *total = part[threadIdx.x];
    }
}

__global__ void __launch_bounds__(256) bin_fill(const Tri* tris, uint32_t count, LineGrid g, const uint32_t* tile_offset,
                                                uint32_t* cursor, uint32_t* list) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Tri t = tris[i];
    uint32_t j0, j1, k0, k1;
    if (!tri_tiles(t, g, j0, j1, k0, k1)) return;
    for (uint32_t k = k0; k <= k1; ++k)
        for (uint32_t j = j0; j <= j1; ++j) {
            const uint32_t tile = k * g.tiles_j + j;
            const uint32_t p = atomicAdd(&cursor[tile], 1U);
            if (p < tile_offset[tile + 1] - tile_offset[tile]) list[tile_offset[tile] + p] = i;
        }
}

__device__ __forceinline__ double line_y(const LineGrid& g, uint32_t j) {
    return g.miny + (static_cast<double>(j) + .5) * g.dy + 1e-10 * g.dx;
}
__device__ __forceinline__ double line_z(const LineGrid& g, uint32_t k) {
    return g.minz + (static_cast<double>(k) + .5) * g.dz + 2e-10 * g.dx;
}

// One workgroup per tile of 8 x 8 x-lines.  Lines of the tile: j = tj*8 + (l & 7), k = tk*8 + (l >> 3).
__global__ void __launch_bounds__(kThreads) voxelize_lines(const Tri* tris, LineGrid g, const uint32_t* tile_offset,
                                                           const uint32_t* list, uint8_t* mask) {
    __shared__ uint32_t cross[kTileLines][kChunk];
    __shared__ uint32_t ncross[kTileLines];
    const uint32_t tile = blockIdx.x;
    const uint32_t tj = tile % g.tiles_j, tk = tile / g.tiles_j;
    const uint32_t begin = tile_offset[tile], len = tile_offset[tile + 1] - begin;
    const uint32_t lane_line = threadIdx.x % kTileLines, sub = threadIdx.x / kTileLines;
    const uint32_t my_j = tj * kTileJ + (lane_line % kTileJ), my_k = tk * kTileK + (lane_line / kTileJ);
    const bool my_valid = my_j < g.ny && my_k < g.nz;
    const double oy = line_y(g, my_j), oz = line_z(g, my_k);
    const uint32_t rounds = len == 0 ? 1 : (len + kChunk - 1) / kChunk;
    for (uint32_t r = 0; r < rounds; ++r) {
        if (threadIdx.x < kTileLines) ncross[threadIdx.x] = 0;
        __syncthreads();
        // u, v of voxelize.cpp:4 for this line; they do not depend on the origin's x.
        const uint32_t c0 = r * kChunk, cn = min(len - min(len, c0), static_cast<uint32_t>(kChunk));
        if (my_valid)
            for (uint32_t c = sub; c < cn; c += kThreads / kTileLines) {
                const uint32_t id = list[begin + c0 + c];
                const Tri t = tris[id];
                const double e1x = t.bx - t.ax, e1y = t.by - t.ay, e1z = t.bz - t.az;
                const double e2x = t.cx - t.ax, e2y = t.cy - t.ay, e2z = t.cz - t.az;
                const double hx = 0.0 * e2z - 0.0 * e2y, hy = 0.0 * e2x - 1.0 * e2z, hz = 1.0 * e2y - 0.0 * e2x;
                const double det = e1x * hx + e1y * hy + e1z * hz;
                if (fabs(det) < 1e-12) continue;
                const double inv = 1 / det;
                // s = o - a: its x component only meets hx = +-0 in u and 0 in v, so any finite
                // origin x gives the same u and v; the line's first cell centre stands in.
                const double sx = g.minx - t.ax, sy = oy - t.ay, sz = oz - t.az;
                const double u = inv * (sx * hx + sy * hy + sz * hz);
                if (u < 0 || u > 1) continue;
                const double qx = sy * e1z - sz * e1y;
                const double v = inv * (1.0 * qx + 0.0 * (sz * e1x - sx * e1z) + 0.0 * (sx * e1y - sy * e1x));
                if (v < 0 || u + v > 1) continue;
                const uint32_t p = atomicAdd(&ncross[lane_line], 1U);
                cross[lane_line][p] = id;
            }
        __syncthreads();
        // Sweep the lines along x, each wavefront its own lines: the ray parameter of every
        // crossing, per cell (voxelize.cpp:4).  Lane c holds the line-constant parts of crossing c
        // (1/det, edges, a_x, and the products s_z e1_x, s_y e1_x, e2_x q_x that do not depend on
        // the cell); the wavefront reads them back lane by lane (uniform, no LDS, no barrier).
        // Each lane takes 8 consecutive cells and writes their 8 parity bytes with one 64-bit
        // store where the row allows it.  Every rounded operation is the reference's, in its order.
        const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
        for (uint32_t l = wave; l < kTileLines; l += kWaves) {
            const uint32_t j = tj * kTileJ + (l % kTileJ), k = tk * kTileK + (l / kTileJ);
            if (j >= g.ny || k >= g.nz) continue;
            const uint32_t n = ncross[l];
            if (r > 0 && n == 0) continue;
            Crossing mine{};
            if (lane < n) {
                const Tri t = tris[cross[l][lane]];
                const double e1x = t.bx - t.ax, e1y = t.by - t.ay, e1z = t.bz - t.az;
                const double e2x = t.cx - t.ax, e2y = t.cy - t.ay, e2z = t.cz - t.az;
                const double hx = 0.0 * e2z - 0.0 * e2y, hy = 0.0 * e2x - 1.0 * e2z, hz = 1.0 * e2y - 0.0 * e2x;
                const double sy = line_y(g, j) - t.ay, sz = line_z(g, k) - t.az;
                mine.inv = 1 / (e1x * hx + e1y * hy + e1z * hz);
                mine.e1y = e1y, mine.e1z = e1z, mine.e2y = e2y, mine.e2z = e2z, mine.ax = t.ax;
                mine.sz_e1x = sz * e1x;
                mine.sy_e1x = sy * e1x;
                mine.e2x_qx = e2x * (sy * e1z - sz * e1y);
            }
            uint8_t* row = mask + (static_cast<size_t>(k) * g.ny + j) * g.nx;
            for (uint32_t i0 = lane * kLaneCells; i0 < g.nx; i0 += 64 * kLaneCells) {
                uint32_t par = 0;  // bit e: parity of cell i0 + e
                double ox[kLaneCells];
                if (n > 0) {
#pragma unroll
                    for (int e = 0; e < kLaneCells; ++e) ox[e] = g.minx + (static_cast<double>(i0 + e) + .5) * g.dx;
                }
                for (uint32_t c = 0; c < n; ++c) {
                    const Crossing x = read_lane(mine, c);
#pragma unroll
                    for (int e = 0; e < kLaneCells; ++e) {
                        const double sx = ox[e] - x.ax;
                        const double qy = x.sz_e1x - sx * x.e1z, qz = sx * x.e1y - x.sy_e1x;
                        const double dist = x.inv * (x.e2x_qx + x.e2y * qy + x.e2z * qz);
                        par ^= (dist > 1e-10 ? 1U : 0U) << e;
                    }
                }
                if (g.packed && i0 + kLaneCells <= g.nx) {
                    uint64_t bytes = 0;
#pragma unroll
                    for (int e = 0; e < kLaneCells; ++e) bytes |= static_cast<uint64_t>((par >> e) & 1U) << (8 * e);
                    uint64_t* dst = reinterpret_cast<uint64_t*>(row + i0);
                    *dst = r == 0 ? bytes : (*dst ^ bytes);
                } else {
                    for (uint32_t e = 0; e < static_cast<uint32_t>(kLaneCells) && i0 + e < g.nx; ++e) {
                        const uint8_t bit = static_cast<uint8_t>((par >> e) & 1U);
                        row[i0 + e] = r == 0 ? bit : static_cast<uint8_t>(row[i0 + e] ^ bit);
                    }
                }
            }
        }
        __syncthreads();  // ncross and cross are rewritten by the next round
    }
}

__global__ void __launch_bounds__(256) zero_words(uint32_t* p, uint32_t words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) p[i] = 0U;
}

// flag = 1 when any of the `values` doubles is NaN or infinite.
__global__ void __launch_bounds__(256) flag_non_finite(const double* values, uint64_t n, uint32_t* flag) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    bool bad = false;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        bad |= !isfinite(values[i]);
    if (bad) atomicOr(flag, 1U);
}

__global__ void __launch_bounds__(256) solid_flags(uint8_t* mask, const uint8_t* base, size_t cells) {
    const size_t n = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= cells) return;
    mask[n] = mask[n] != 0 ? static_cast<uint8_t>(0x80) : (base ? base[n] : static_cast<uint8_t>(0));
}

template <class T> struct Scratch {
    T* ptr = nullptr;
    explicit Scratch(size_t n) { check(hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc"); }
    ~Scratch() { (void)hipFree(ptr); }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
};

struct Event {
    hipEvent_t e = nullptr;
    Event() { check(hipEventCreate(&e), "hipEventCreate"); }
    ~Event() { (void)hipEventDestroy(e); }
};

// Rodrigues rotation in the operation order of voxelize.cpp:4 (rotate), so that the transformed
// vertices equal the reference's bit for bit (the file is built with -ffp-contract=off).
void rotate_point(const double p[3], const double axis_in[3], double a, double out[3]) {
    const double n = std::sqrt(axis_in[0] * axis_in[0] + axis_in[1] * axis_in[1] + axis_in[2] * axis_in[2]);
    if (n < 1e-30) {
        out[0] = p[0], out[1] = p[1], out[2] = p[2];
        return;
    }
    const double r = 1 / n;
    const double ux = axis_in[0] * r, uy = axis_in[1] * r, uz = axis_in[2] * r;
    const double c = std::cos(a), s = std::sin(a);
    const double kx = uy * p[2] - uz * p[1], ky = uz * p[0] - ux * p[2], kz = ux * p[1] - uy * p[0];
    const double along = (ux * p[0] + uy * p[1] + uz * p[2]) * (1 - c);
    out[0] = (p[0] * c + kx * s) + ux * along;
    out[1] = (p[1] * c + ky * s) + uy * along;
    out[2] = (p[2] * c + kz * s) + uz * along;
}

}  // namespace

thread_local VoxelizeTiming g_last_timing;

VoxelizeTiming last_voxelize_timing() { return g_last_timing; }

void validate_voxelization(const double* triangles, size_t count, const VoxelGridArgs& grid) {
    if (!triangles || count == 0 || grid.nx == 0 || grid.ny == 0 || grid.nz == 0 ||
        !(grid.maximum[0] > grid.minimum[0]) || !(grid.maximum[1] > grid.minimum[1]) ||
        !(grid.maximum[2] > grid.minimum[2]))
        throw std::invalid_argument("invalid voxelization inputs");
    if (count > 0xFFFFFFFFull || grid.nx > 0xFFFFFFFFull || grid.ny > 0xFFFFFFFFull || grid.nz > 0xFFFFFFFFull)
        throw std::invalid_argument("voxelization grid or surface too large");
    for (size_t i = 0; i < count * 9; ++i)
        if (!std::isfinite(triangles[i])) throw std::invalid_argument("surface vertices must be finite");
    const uint64_t tiles = ((grid.ny + kTileJ - 1) / kTileJ) * ((grid.nz + kTileK - 1) / kTileK);
    if (tiles >= 0x7FFFFFFFull) throw std::invalid_argument("voxelization grid has too many lines");
}

void voxelize_surface(const double* triangles, size_t count, const VoxelGridArgs& grid, uint8_t* device_mask,
                      hipStream_t stream, VoxelizeTiming* timing) {
    validate_voxelization(triangles, count, grid);
    if (!device_mask) throw std::invalid_argument("voxelization needs an output mask");
    LineGrid g{};
    g.nx = static_cast<uint32_t>(grid.nx);
    g.ny = static_cast<uint32_t>(grid.ny);
    g.nz = static_cast<uint32_t>(grid.nz);
    g.tiles_j = (g.ny + kTileJ - 1) / kTileJ;
    g.packed = (grid.nx % 8 == 0 && reinterpret_cast<uintptr_t>(device_mask) % 8 == 0) ? 1U : 0U;
    g.tiles_k = (g.nz + kTileK - 1) / kTileK;
    g.minx = grid.minimum[0];
    g.miny = grid.minimum[1];
    g.minz = grid.minimum[2];
    // voxelize.cpp:5: dx = (max - min) / n with n converted to double.
    g.dx = (grid.maximum[0] - grid.minimum[0]) / static_cast<double>(grid.nx);
    g.dy = (grid.maximum[1] - grid.minimum[1]) / static_cast<double>(grid.ny);
    g.dz = (grid.maximum[2] - grid.minimum[2]) / static_cast<double>(grid.nz);
    const uint32_t tiles = g.tiles_j * g.tiles_k;
    const uint32_t n = static_cast<uint32_t>(count);

    Scratch<Tri> d_tris(count);
    Scratch<uint32_t> d_count(tiles), d_offset(tiles + 1), d_cursor(tiles);
    Scratch<unsigned long long> d_total(1);
    Event e0, e1, e2;
    check(hipMemcpyAsync(d_tris.ptr, triangles, count * sizeof(Tri), hipMemcpyHostToDevice, stream), "copy triangles");
    check(hipEventRecord(e0.e, stream), "hipEventRecord");
    check(hipMemsetAsync(d_count.ptr, 0, tiles * sizeof(uint32_t), stream), "hipMemsetAsync");
    check(hipMemsetAsync(d_cursor.ptr, 0, tiles * sizeof(uint32_t), stream), "hipMemsetAsync");
    const uint32_t tri_blocks = (n + 255) / 256;
    bin_count<<<tri_blocks, 256, 0, stream>>>(d_tris.ptr, n, g, d_count.ptr);
    check(hipGetLastError(), "bin_count");
    bin_scan<<<1, kScanThreads, 0, stream>>>(d_count.ptr, tiles, d_offset.ptr, d_total.ptr);
    check(hipGetLastError(), "bin_scan");
    unsigned long long total = 0;
    check(hipMemcpyAsync(&total, d_total.ptr, sizeof(total), hipMemcpyDeviceToHost, stream), "copy total");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (total > 0xFFFFFFFFull) throw std::invalid_argument("surface too large to bin (more than 2^32 triangle-tile pairs)");
    Scratch<uint32_t> d_list(total);
    check(hipMemsetAsync(d_list.ptr, 0, std::max<unsigned long long>(total, 1) * sizeof(uint32_t), stream),
          "hipMemsetAsync");
    bin_fill<<<tri_blocks, 256, 0, stream>>>(d_tris.ptr, n, g, d_offset.ptr, d_cursor.ptr, d_list.ptr);
    check(hipGetLastError(), "bin_fill");
    check(hipEventRecord(e1.e, stream), "hipEventRecord");
    voxelize_lines<<<tiles, kThreads, 0, stream>>>(d_tris.ptr, g, d_offset.ptr, d_list.ptr, device_mask);
    check(hipGetLastError(), "voxelize_lines");
    check(hipEventRecord(e2.e, stream), "hipEventRecord");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    VoxelizeTiming t{};
    float ms = 0;
    check(hipEventElapsedTime(&ms, e0.e, e1.e), "hipEventElapsedTime");
    t.bin_ms = ms;
    check(hipEventElapsedTime(&ms, e1.e, e2.e), "hipEventElapsedTime");
    t.lines_ms = ms;
    t.pairs = total;
    g_last_timing = t;
    if (timing) *timing = t;
}

namespace {
// Device address of `scratch` on each GPU, looked up once per device with that device current.
std::mutex g_scratch_mutex;
std::vector<unsigned char*> g_scratch_ptr;

unsigned char* scratch_on_current_device() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr.size() <= static_cast<size_t>(dev)) g_scratch_ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p || reinterpret_cast<uintptr_t>(p) % 256 != 0) throw std::runtime_error("voxelize scratch is misaligned");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void voxelize_surface_no_alloc(const double* device_triangles, size_t count, const VoxelGridArgs& grid,
                               uint8_t* device_mask, hipStream_t stream) {
    if (!device_triangles || count == 0 || grid.nx == 0 || grid.ny == 0 || grid.nz == 0 ||
        !(grid.maximum[0] > grid.minimum[0]) || !(grid.maximum[1] > grid.minimum[1]) ||
        !(grid.maximum[2] > grid.minimum[2]))
        throw std::invalid_argument("invalid voxelization inputs");
    if (!device_mask) throw std::invalid_argument("voxelization needs an output mask");
    if (reinterpret_cast<uintptr_t>(device_triangles) % alignof(double) != 0)
        throw std::invalid_argument("device triangles must be 8-byte aligned doubles");
    if (count > 0xFFFFFFFFull || grid.nx > 0xFFFFFFFFull || grid.ny > 0xFFFFFFFFull || grid.nz > 0xFFFFFFFFull)
        throw std::invalid_argument("voxelization grid or surface too large");
    const uint64_t tiles64 = ((grid.ny + kTileJ - 1) / kTileJ) * ((grid.nz + kTileK - 1) / kTileK);
    if (tiles64 > kVoxelizeMaxTiles)
        throw std::invalid_argument("voxelization grid has " + std::to_string(tiles64) +
                                    " tiles of 8 x 8 x-lines; the no-allocation voxelizer takes at most " +
                                    std::to_string(kVoxelizeMaxTiles));
    LineGrid g{};
    g.nx = static_cast<uint32_t>(grid.nx);
    g.ny = static_cast<uint32_t>(grid.ny);
    g.nz = static_cast<uint32_t>(grid.nz);
    g.tiles_j = (g.ny + kTileJ - 1) / kTileJ;
    g.tiles_k = (g.nz + kTileK - 1) / kTileK;
    g.packed = (grid.nx % 8 == 0 && reinterpret_cast<uintptr_t>(device_mask) % 8 == 0) ? 1U : 0U;
    g.minx = grid.minimum[0];
    g.miny = grid.minimum[1];
    g.minz = grid.minimum[2];
    g.dx = (grid.maximum[0] - grid.minimum[0]) / static_cast<double>(grid.nx);
    g.dy = (grid.maximum[1] - grid.minimum[1]) / static_cast<double>(grid.ny);
    g.dz = (grid.maximum[2] - grid.minimum[2]) / static_cast<double>(grid.nz);
    const uint32_t tiles = g.tiles_j * g.tiles_k;
    const uint32_t n = static_cast<uint32_t>(count);

    unsigned char* base = scratch_on_current_device();
    auto* total = reinterpret_cast<unsigned long long*>(base);
    auto* bad = reinterpret_cast<uint32_t*>(base + 8);
    auto* counts = reinterpret_cast<uint32_t*>(base + kVoxelizeCountOffset);
    auto* cursor = reinterpret_cast<uint32_t*>(base + kVoxelizeCursorOffset);
    auto* offsets = reinterpret_cast<uint32_t*>(base + kVoxelizeTileOffset);
    auto* list = reinterpret_cast<uint32_t*>(base + kVoxelizeListOffset);
    const Tri* tris = reinterpret_cast<const Tri*>(device_triangles);

    // The previous call's header, counts and cursors are cleared by a kernel (no memset node).
    const uint32_t words = static_cast<uint32_t>(kVoxelizeTileOffset / 4);
    zero_words<<<(words + 255) / 256, 256, 0, stream>>>(reinterpret_cast<uint32_t*>(base), words);
    check(hipGetLastError(), "zero_words");
    const uint64_t values = static_cast<uint64_t>(count) * 9;
    const uint32_t check_blocks = static_cast<uint32_t>(std::min<uint64_t>((values + 255) / 256, 1024));
    flag_non_finite<<<check_blocks, 256, 0, stream>>>(device_triangles, values, bad);
    check(hipGetLastError(), "flag_non_finite");
    const uint32_t tri_blocks = (n + 255) / 256;
    bin_count<<<tri_blocks, 256, 0, stream>>>(tris, n, g, counts);
    check(hipGetLastError(), "bin_count");
    bin_scan<<<1, kScanThreads, 0, stream>>>(counts, tiles, offsets, total);
    check(hipGetLastError(), "bin_scan");
    unsigned long long header[2] = {0, 0};  // pair total, non-finite flag (low word)
    check(hipMemcpyAsync(header, base, sizeof(header), hipMemcpyDeviceToHost, stream), "copy pair total");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if ((header[1] & 0xFFFFFFFFull) != 0) throw std::invalid_argument("surface vertices must be finite");
    if (header[0] > kVoxelizeMaxPairs)
        throw std::invalid_argument("surface binned to " + std::to_string(header[0]) +
                                    " (triangle, tile) pairs; the no-allocation voxelizer takes at most " +
                                    std::to_string(kVoxelizeMaxPairs));
    bin_fill<<<tri_blocks, 256, 0, stream>>>(tris, n, g, offsets, cursor, list);
    check(hipGetLastError(), "bin_fill");
    voxelize_lines<<<tiles, kThreads, 0, stream>>>(tris, g, offsets, list, device_mask);
    check(hipGetLastError(), "voxelize_lines");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void solid_to_flags(uint8_t* mask, const uint8_t* base, size_t cells, hipStream_t stream) {
    const size_t blocks = (cells + 255) / 256;
    if (blocks > 0x7FFFFFFFull) throw std::invalid_argument("too many cells");
    solid_flags<<<static_cast<uint32_t>(blocks), 256, 0, stream>>>(mask, base, cells);
    check(hipGetLastError(), "solid_flags");
}

void transform_surface(const double* in, size_t count, const double translation[3], const double axis[3], double angle,
                       double* out) {
    if ((!in || !out) && count > 0) throw std::invalid_argument("transform_surface needs input and output triangles");
    if (!translation || !axis) throw std::invalid_argument("transform_surface needs a translation and an axis");
    for (size_t v = 0; v < count * 3; ++v) {
        double r[3];
        rotate_point(in + 3 * v, axis, angle, r);
        out[3 * v + 0] = r[0] + translation[0];
        out[3 * v + 1] = r[1] + translation[1];
        out[3 * v + 2] = r[2] + translation[2];
    }
}

}  // namespace cfd_hip
