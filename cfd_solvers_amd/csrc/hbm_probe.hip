// hbm_probe.hip -- bandwidth probes of known traffic, for bench.py.  Not part of the LBM C ABI.
//
//   hbm_probe_copy     a plain device-to-device copy: the card's streaming bandwidth, and the
//                      known-size reference against which rocprofv3's FETCH_SIZE / WRITE_SIZE
//                      counters are scaled;
//   hbm_probe_inplace  the esoteric-pull access pattern alone (D3Q19, periodic, in place, the
//                      row-blocked population layout, streaming stores) with a trivial update:
//                      the bandwidth an in-place stream-collide of that layout can reach.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace {

// One float4 per thread: on MI355X this streams at ~6.1 TB/s, against ~4.6 TB/s for a grid-stride
// loop over 64 blocks per CU.
__global__ __launch_bounds__(256) void copy_f4(const float4* __restrict__ src, float4* __restrict__ dst,
                                               uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

constexpr int kCx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int kCy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int kCz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

template <int I> struct Int {
    static constexpr int value = I;
};
template <class F, int... I> __device__ __forceinline__ void each_pair(F&& f, std::integer_sequence<int, I...>) {
    (f(Int<2 * I + 1>{}), ...);
}

// One 256-cell x-row per workgroup.  Even-parity esoteric pull: population i+1 of the cell is read
// in place, population i from the cell up = n - c_i; the updated values go back to the opposite
// slots.  Layout: blocks of 2^rs rows, 19 planes of block_cells elements each.
template <class T>
__global__ __launch_bounds__(256) void inplace_pattern(T* __restrict__ f, uint32_t ny, uint32_t nz, uint32_t rs) {
    constexpr uint32_t nx = 256;
    const uint32_t y = blockIdx.x % ny, z = blockIdx.x / ny, x = threadIdx.x;
    const uint32_t block_cells = nx << rs, stride = 19U * block_cells, mask = (1U << rs) - 1U;
    const auto row = [&](int dy, int dz) {
        const uint32_t r = ((z + nz + dz) % nz) * ny + (y + ny + dy) % ny;
        return (r >> rs) * stride + (r & mask) * nx;
    };
    const uint32_t n = row(0, 0) + x;
    float s = static_cast<float>(f[n]);
    float a[19], b[19];
    uint32_t up[19];
    each_pair([&](auto I) {
        constexpr int i = decltype(I)::value;
        up[i] = row(-kCy[i], -kCz[i]) + (x + nx - kCx[i]) % nx;
        a[i] = static_cast<float>(f[static_cast<uint64_t>(i) * block_cells + up[i]]);
        b[i] = static_cast<float>(f[static_cast<uint64_t>(i + 1) * block_cells + n]);
        s += a[i] + b[i];
    }, std::make_integer_sequence<int, 9>{});
    __builtin_nontemporal_store(static_cast<T>(s * 1e-3F), f + n);
    each_pair([&](auto I) {
        constexpr int i = decltype(I)::value;
        __builtin_nontemporal_store(static_cast<T>(a[i] * 0.999F + s * 1e-9F),
                                    f + static_cast<uint64_t>(i + 1) * block_cells + n);
        __builtin_nontemporal_store(static_cast<T>(b[i] * 0.999F + s * 1e-9F),
                                    f + static_cast<uint64_t>(i) * block_cells + up[i]);
    }, std::make_integer_sequence<int, 9>{});
}

}  // namespace

// Average time of the in-place pattern on a 256 x ny x nz lattice of elem_bytes (2: FP16, 4: FP32)
// populations in blocks of 2^row_shift rows.  Algorithmic traffic: 19 reads + 19 writes per cell.
extern "C" __attribute__((visibility("default"))) int hbm_probe_inplace(int elem_bytes, uint32_t ny, uint32_t nz,
                                                                        uint32_t row_shift, int repeats,
                                                                        float* ms_per_pass) {
    if ((elem_bytes != 2 && elem_bytes != 4) || ny < 1 || nz < 1 || repeats < 1 || !ms_per_pass || row_shift > 16)
        return 1;
    const uint64_t rows = static_cast<uint64_t>(ny) * nz;
    if ((rows & ((uint64_t{1} << row_shift) - 1U)) != 0 || rows >= (uint64_t{1} << 31)) return 1;
    const uint64_t elems = rows * 256U * 19U;
    if (elems >= (uint64_t{1} << 32)) return 1;
    void* f = nullptr;
    if (hipMalloc(&f, elems * static_cast<uint64_t>(elem_bytes)) != hipSuccess) return 2;
    (void)hipMemset(f, 0, elems * static_cast<uint64_t>(elem_bytes));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const auto launch = [&] {
        if (elem_bytes == 4)
            hipLaunchKernelGGL(inplace_pattern<float>, dim3(static_cast<unsigned>(rows)), dim3(256), 0, nullptr,
                               static_cast<float*>(f), ny, nz, row_shift);
        else
            hipLaunchKernelGGL(inplace_pattern<_Float16>, dim3(static_cast<unsigned>(rows)), dim3(256), 0, nullptr,
                               static_cast<_Float16*>(f), ny, nz, row_shift);
    };
    launch();
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < repeats; ++r) launch();
    (void)hipEventRecord(e1, nullptr);
    const hipError_t st = hipEventSynchronize(e1);
    float ms = 0.0F;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_pass = ms / static_cast<float>(repeats);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(f);
    return st == hipSuccess ? 0 : 3;
}

extern "C" __attribute__((visibility("default"))) int hbm_probe_copy(uint64_t bytes, int repeats, float* ms_per_copy) {
    if (bytes < 16 || repeats < 1 || !ms_per_copy) return 1;
    const uint64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return 2;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return 2;
    }
    (void)hipMemset(a, 0, n * 16);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    if ((n + 255) / 256 >= (uint64_t{1} << 31)) return 1;
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(copy_f4, dim3(grid), dim3(256), 0, nullptr, static_cast<const float4*>(a),
                       static_cast<float4*>(b), n);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(copy_f4, dim3(grid), dim3(256), 0, nullptr, static_cast<const float4*>(a),
                           static_cast<float4*>(b), n);
    (void)hipEventRecord(e1, nullptr);
    const hipError_t st = hipEventSynchronize(e1);
    float ms = 0.0F;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_copy = ms / static_cast<float>(repeats);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return st == hipSuccess ? 0 : 3;
}
