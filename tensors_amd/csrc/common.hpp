// Shared helpers of the gfx950 library: status/error plumbing, launch
// geometry, XCD-aware block remapping and wave-level reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <algorithm>

#include "tensors_amd.h"

namespace tg {

// Records a message for tg_last_error() (thread-local) and returns `code`.
int fail(int code, const char* fmt, ...);

#define TG_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t tg_e_ = (expr);                                                     \
        if (tg_e_ != hipSuccess)                                                       \
            return ::tg::fail(TG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(tg_e_)); \
    } while (0)

// Checks the launch that was just enqueued.
#define TG_LAUNCH_CHECK() TG_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;      // CDNA wavefront width
constexpr int kBlock = 256;    // threads per workgroup of the streaming kernels
constexpr int kXcds = 8;       // MI355X: 8 XCDs, each with its own L2

// V consecutive elements as one vector load / store (V = 2, 4: 8- or 16-byte
// aligned addresses).
template <typename T, int V>
struct VecOf {
    typedef T type __attribute__((ext_vector_type(V)));
};

template <typename T, int V>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, T (&out)[V]) {
    if constexpr (V == 1) {
        out[0] = *p;
    } else {
        using VT = typename VecOf<T, V>::type;
        const VT v = *reinterpret_cast<const VT*>(p);
#pragma unroll
        for (int i = 0; i < V; ++i) out[i] = v[i];
    }
}

template <typename T, int V>
__device__ __forceinline__ void store_vec(T* __restrict__ p, const T (&in)[V]) {
    if constexpr (V == 1) {
        *p = in[0];
    } else {
        using VT = typename VecOf<T, V>::type;
        VT v;
#pragma unroll
        for (int i = 0; i < V; ++i) v[i] = in[i];
        *reinterpret_cast<VT*>(p) = v;
    }
}

// Workgroups are dispatched round-robin over the XCDs.  Give XCD x the
// contiguous logical range [x*q + min(x,r), ...) so neighbouring row blocks
// (which touch neighbouring X rows for banded matrices) share one L2.
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds, idx = b / kXcds;
    return x * q + (x < r ? x : r) + idx;
}

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
    return __shfl_xor(v, mask, kWave);
}

// Sum over the `width` lanes of an aligned lane group (width a power of 2).
template <int WIDTH, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = WIDTH / 2; off > 0; off >>= 1) v += shfl_xor(v, off);
    return v;
}

}  // namespace tg
