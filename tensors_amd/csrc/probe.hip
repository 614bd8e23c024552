// Streaming bandwidth probes: the measured-HBM denominator for the roofline
// and the known-size reference kernel for scaling rocprof FETCH/WRITE counters.
#include "common.hpp"

namespace tg {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kBlock) void copy_kernel(const f4* __restrict__ src, f4* __restrict__ dst,
                                                      int64_t n16) {
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kBlock)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ __launch_bounds__(kBlock) void read_kernel(const f4* __restrict__ src, int64_t n16,
                                                      float* __restrict__ sink) {
    f4 acc = {0, 0, 0, 0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kBlock)
        acc += __builtin_nontemporal_load(src + i);
    const float s = acc[0] + acc[1] + acc[2] + acc[3];
    if (s == 1234.5f) sink[0] = s;  // keeps the loads alive; practically never taken
}

}  // namespace tg

using namespace tg;

extern "C" int tg_probe_copy(const void* src, void* dst, int64_t bytes, void* stream) {
    if (!src || !dst || bytes < 0 || bytes % 16) return fail(TG_ERR_INVALID, "tg_probe_copy: bytes must be a multiple of 16");
    const int64_t n16 = bytes / 16;
    const int grid = (int)std::min<int64_t>((n16 + kBlock - 1) / kBlock, 256 * 32);
    if (grid > 0) hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(kBlock), 0, as_stream(stream), (const f4*)src, (f4*)dst, n16);
    TG_LAUNCH_CHECK();
    return TG_OK;
}

extern "C" int tg_probe_read(const void* src, int64_t bytes, float* sink, void* stream) {
    if (!src || !sink || bytes < 0 || bytes % 16) return fail(TG_ERR_INVALID, "tg_probe_read: bytes must be a multiple of 16");
    const int64_t n16 = bytes / 16;
    const int grid = (int)std::min<int64_t>((n16 + kBlock - 1) / kBlock, 256 * 32);
    if (grid > 0) hipLaunchKernelGGL(read_kernel, dim3(grid), dim3(kBlock), 0, as_stream(stream), (const f4*)src, n16, sink);
    TG_LAUNCH_CHECK();
    return TG_OK;
}
