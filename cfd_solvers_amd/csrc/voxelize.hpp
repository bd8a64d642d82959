// voxelize.hpp -- GPU ray-parity voxelisation of a triangle surface (src/lbm/voxelize.cpp:4-5,
// src/sycl/voxelize_sycl.cpp:13-68) and the rigid transform of a surface (voxelize.cpp:6).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cfd_hip {

// VoxelGrid3D (include/cfd/solvers/lbm/voxelize.hpp:8): nx*ny*nz cells spanning [minimum, maximum].
struct VoxelGridArgs {
    size_t nx = 0, ny = 0, nz = 0;
    double minimum[3] = {0, 0, 0}, maximum[3] = {0, 0, 0};
};

// voxelize_surface_no_alloc's working memory: a static device array per GPU (`scratch` in
// voxelize.hip) holding a 256-byte header (pair total as 64 bits, non-finite flag), the tile
// counts and cursors (kVoxelizeMaxTiles words each), the tile offsets (kVoxelizeMaxTiles + 1
// words) and the pair list (kVoxelizeMaxPairs words): 17 MiB.
constexpr uint32_t kVoxelizeMaxTiles = 1U << 16;           // tiles of 8 x 8 x-lines: (ny+7)/8 * (nz+7)/8
constexpr uint64_t kVoxelizeMaxPairs = uint64_t{1} << 22;  // (triangle, tile) pairs
constexpr size_t kVoxelizeCountOffset = 256;
constexpr size_t kVoxelizeCursorOffset = kVoxelizeCountOffset + 4 * size_t{kVoxelizeMaxTiles};
constexpr size_t kVoxelizeTileOffset = kVoxelizeCursorOffset + 4 * size_t{kVoxelizeMaxTiles};
constexpr size_t kVoxelizeListOffset = kVoxelizeTileOffset + ((4 * (size_t{kVoxelizeMaxTiles} + 1) + 255) / 256) * 256;
constexpr size_t kVoxelizeScratchBytes = kVoxelizeListOffset + 4 * kVoxelizeMaxPairs;

// Per-stage device times of the last voxelize_surface call on this thread (HIP events, ms):
// binning (count + scan + fill) and the line kernel.
struct VoxelizeTiming {
    double bin_ms = 0, lines_ms = 0;
    uint64_t pairs = 0;  // (triangle, line tile) pairs binned
};

// Throws std::invalid_argument for the inputs voxelize.cpp:5 rejects ("invalid voxelization
// inputs"), for non-finite vertex coordinates and for sizes beyond the kernels' 32-bit indices.
void validate_voxelization(const double* triangles, size_t count, const VoxelGridArgs& grid);

// Writes the inside/outside mask of the closed surface `triangles` (count x 9 doubles: a, b, c
// as x,y,z triples; host memory) into `device_mask` (nx*ny*nz bytes, cell (i,j,k) at
// (k*ny + j)*nx + i), on `stream`, and waits for it.  Throws std::invalid_argument for the
// inputs voxelize.cpp:5 rejects and for non-finite vertex coordinates.
void voxelize_surface(const double* triangles, size_t count, const VoxelGridArgs& grid, uint8_t* device_mask,
                      hipStream_t stream, VoxelizeTiming* timing = nullptr);

// voxelize_surface without device allocation: `device_triangles` (count x 9 doubles) and
// `device_mask` are caller-owned device memory, the working memory is the static per-GPU array
// above, and the work runs on `stream` (the call returns when the mask is written).  Throws
// std::invalid_argument for what voxelize_surface rejects (non-finite vertices are found on the
// device) and for grids of more than kVoxelizeMaxTiles tiles or surfaces of more than
// kVoxelizeMaxPairs (triangle, tile) pairs.  Calls that overlap on one GPU race on the array.
void voxelize_surface_no_alloc(const double* device_triangles, size_t count, const VoxelGridArgs& grid,
                               uint8_t* device_mask, hipStream_t stream);

// Timing of the calling thread's last voxelize_surface call.
VoxelizeTiming last_voxelize_timing();

// flags[n] = mask[n] ? 0x80 : (base ? base[n] : 0), in place over `mask` (device pointers).
void solid_to_flags(uint8_t* mask, const uint8_t* base, size_t cells, hipStream_t stream);

// transform_surface (voxelize.cpp:6): every vertex rotated about `axis` (through the origin) by
// `angle` radians with Rodrigues' formula, then translated; host arithmetic, out may equal in.
void transform_surface(const double* in, size_t count, const double translation[3], const double axis[3],
                       double angle, double* out);

}  // namespace cfd_hip
