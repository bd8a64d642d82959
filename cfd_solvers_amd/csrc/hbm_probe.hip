// hbm_probe.hip -- a plain device-to-device copy kernel of known size.
//
// Measurement helper for bench.py: it gives the achievable streaming bandwidth of the card and is
// the known-size reference against which rocprofv3's FETCH_SIZE / WRITE_SIZE counters are scaled.
// Not part of the LBM C ABI.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// One float4 per thread: on MI355X this streams at ~6.1 TB/s, against ~4.6 TB/s for a grid-stride
// loop over 64 blocks per CU.
__global__ __launch_bounds__(256) void copy_f4(const float4* __restrict__ src, float4* __restrict__ dst,
                                               uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

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
