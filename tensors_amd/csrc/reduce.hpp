// Deterministic column-wise reductions of n x k row-major blocks, shared by
// the solver kernels: thread (c, ro) handles column c of rows ro, ro + rows,
// ...; a fixed LDS tree combines the row lanes of a workgroup into
// partial[block * k + c]; one workgroup then adds the workgroup partials in a
// fixed order.  The result does not depend on scheduling.
#pragma once

#include "common.hpp"

namespace tg {

__device__ __forceinline__ void block_row_range(int n, int& i0, int& i1) {
    const int b = blockIdx.x, B = gridDim.x;
    i0 = (int)((int64_t)n * b / B);
    i1 = (int)((int64_t)n * (b + 1) / B);
}

// Thread layout of the column-blocked elementwise kernels: c = column lane,
// ro = row lane; kBlock / kpad rows are processed per sweep.
struct Lanes {
    int c, ro, rows;
};
__device__ __forceinline__ Lanes lanes_of(int kpad) {
    Lanes L;
    L.c = threadIdx.x % kpad;
    L.ro = threadIdx.x / kpad;
    L.rows = kBlock / kpad;
    return L;
}

// Reduce per-thread partials of the same column over the row lanes (fixed
// tree) and store the workgroup partial partial[block * k + c].
__device__ __forceinline__ void column_partial(double v, const Lanes& L, int k, int kpad, double* partial) {
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = L.rows / 2; s > 0; s >>= 1) {
        if (L.ro < s) sh[threadIdx.x] += sh[threadIdx.x + s * kpad];
        __syncthreads();
    }
    if (L.ro == 0 && L.c < k) partial[(int64_t)blockIdx.x * k + L.c] = sh[L.c];
}

// Threads of the one-workgroup kernels that finish a reduction.
constexpr int kFinalThreads = 1024;

// Sum of the B workgroup partials of each column in a fixed order, by one
// workgroup of kFinalThreads threads: thread (c, j), j < J = kFinalThreads / kpad,
// adds partials j, j+J, ... into eight running sums (partial j + qJ + 8mJ goes to
// sum q), combines them in a fixed tree, then a fixed tree over j.  Eight loads
// per thread are in flight at a time, so B = 1024-2048 partials of 32 columns
// take a few memory round trips.
__device__ __forceinline__ double column_total(const double* partial, int B, int k, int kpad, int c, int j,
                                               int J) {
    __shared__ double sh[kFinalThreads];
    double s[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = 0.0;
    if (c < k) {
        int b = j;
        for (; b + 7 * J < B; b += 8 * J) {
#pragma unroll
            for (int q = 0; q < 8; ++q) s[q] += partial[(int64_t)(b + q * J) * k + c];
        }
#pragma unroll
        for (int q = 0; q < 7; ++q)  // fewer than 8 left
            if (b + q * J < B) s[q] += partial[(int64_t)(b + q * J) * k + c];
    }
    sh[threadIdx.x] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    __syncthreads();
    for (int h = J / 2; h > 0; h >>= 1) {
        if (j < h) sh[threadIdx.x] += sh[threadIdx.x + h * kpad];
        __syncthreads();
    }
    const double t = sh[c];
    __syncthreads();
    return t;
}

}  // namespace tg
