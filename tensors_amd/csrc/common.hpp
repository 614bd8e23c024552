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

// Sums R values per lane over the lanes of a wave that share c = lane % kpad
// (kpad a power of 2 <= 32), in place: a reduce-scatter over lane bits 5, 4 and 3
// as far as they lie above the column bits (each level keeps half of the values and
// adds the partner's copy of that half), then butterflies over the remaining bits
// down to kpad.  Afterwards the lanes with (lane % low) < kpad, low = the lowest
// scattered bit (32, 16 or 8), hold in v[0 .. count) the sums of registers base ..
// base + count - 1 (those < R) for column lane % kpad; count = ceil(R / 2), ceil(R / 4)
// or ceil(R / 8).  About R + 3 count shuffles instead of log2(64 / kpad) R, in a
// fixed order.  A lane's registers are those of its segment: each level halves the
// segment (the lower part ceil(size / 2)), so count = min(ceil, what remains) >= 0.
template <int R>
struct ScatterSizes {
    static constexpr int A = (R + 1) / 2, B = (A + 1) / 2, C = (B + 1) / 2;
};

struct ScatterOut {
    int base, count, low;
};

template <int R>
__device__ __forceinline__ ScatterOut wave_scatter_sum(double (&v)[R], int kpad) {
    constexpr int A = ScatterSizes<R>::A, B = ScatterSizes<R>::B, C = ScatterSizes<R>::C;
    const int lane = threadIdx.x % kWave;
    const bool b5 = (lane & 32) != 0, b4 = (lane & 16) != 0, b3 = (lane & 8) != 0;
    ScatterOut o{b5 ? A : 0, 0, 32};
    int end = b5 ? R : (A < R ? A : R);  // one past the lane's segment
#pragma unroll
    for (int i = 0; i < A; ++i) {
        const double lo = v[i], hi = i + A < R ? v[i + A] : 0.0;
        v[i] = (b5 ? hi : lo) + shfl_xor(b5 ? lo : hi, 32);
    }
    if (kpad <= 16) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const double lo = v[i], hi = i + B < A ? v[i + B] : 0.0;
            v[i] = (b4 ? hi : lo) + shfl_xor(b4 ? lo : hi, 16);
        }
        o.base += b4 ? B : 0;
        if (!b4) end = o.base + B < end ? o.base + B : end;
        o.low = 16;
    }
    if (kpad <= 8) {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const double lo = v[i], hi = i + C < B ? v[i + C] : 0.0;
            v[i] = (b3 ? hi : lo) + shfl_xor(b3 ? lo : hi, 8);
        }
        o.base += b3 ? C : 0;
        if (!b3) end = o.base + C < end ? o.base + C : end;
        o.low = 8;
    }
    for (int off = 4; off >= kpad; off >>= 1) {
#pragma unroll
        for (int i = 0; i < C; ++i) v[i] += shfl_xor(v[i], off);
    }
    o.count = end > o.base ? end - o.base : 0;
    return o;
}

// Sum over the `width` lanes of an aligned lane group (width a power of 2).
template <int WIDTH, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = WIDTH / 2; off > 0; off >>= 1) v += shfl_xor(v, off);
    return v;
}

}  // namespace tg
