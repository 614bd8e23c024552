// Deterministic synthetic inputs for the benchmark workloads (host code).
//
// Every random draw is a counter-based SplitMix64 hash of (seed, stream,
// index), so the output does not depend on the thread schedule:
//   h = sm64(sm64(sm64(seed) ^ stream) ^ index),  u = (h >> 11) * 2^-53.
// Streams: 1 row lengths, 2 column picks, 3 values, 4 dense blocks, 5 edge
// weights.
//
// Workloads (host buffers, int64 row pointers, int32 column indices):
//  * random CSR: row lengths Poisson(mean) clipped to [lo, hi] (inverse CDF),
//    columns uniform without repeats per row (Floyd's algorithm), sorted
//    ascending, values N(0,1) (Box-Muller);
//  * 3D 7-point graph Laplacian + shift on a g^3 grid: the GridLaplacian
//    convention (src/Sparse/GridLaplacian.hpp:11-77: diagonal = vertex degree
//    + mass, off-diagonal -1) extended to 3D, vertex (i*g + j)*g + k;
//  * weighted 2D/3D grid Laplacians: edge weights exp(N(0,1)), off-diagonal
//    -w_e, diagonal sum of incident weights + shift (SPD);
//  * 2D convection-diffusion: 5-point diffusion plus first-order upwind
//    convection with mesh Peclet number P along direction theta (nonsymmetric
//    M-matrix).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "tensors_amd.h"

namespace tg {
namespace gen {

static inline uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline uint64_t bits(uint64_t seed, uint64_t stream, uint64_t idx) {
    return sm64(sm64(sm64(seed) ^ stream) ^ idx);
}
static inline double u01(uint64_t seed, uint64_t stream, uint64_t idx) {
    return (double)(bits(seed, stream, idx) >> 11) * 0x1.0p-53;
}
static inline double normal(uint64_t seed, uint64_t stream, uint64_t idx) {
    const double u1 = u01(seed, stream, 2 * idx), u2 = u01(seed, stream, 2 * idx + 1);
    return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(6.283185307179586 * u2);
}

template <class F>
static void parallel_rows(int64_t n, F&& f) {
    unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) T = 1;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t)
        pool.emplace_back([&, t] {
            const int64_t b = n * t / T, e = n * (t + 1) / T;
            for (int64_t i = b; i < e; ++i) f(i);
        });
    for (auto& th : pool) th.join();
}

}  // namespace gen
}  // namespace tg

using namespace tg::gen;

extern "C" {

// Uniform[-1,1] fill: out[i] = 2 u(seed, 4, offset + i) - 1.
int tg_gen_uniform(uint64_t seed, uint64_t offset, int64_t count, tg_dtype dtype, void* host_out) {
    if (count < 0 || (!host_out && count > 0)) return TG_ERR_INVALID;
    if (dtype == TG_F64) {
        double* o = (double*)host_out;
        parallel_rows(count, [&](int64_t i) { o[i] = 2.0 * u01(seed, 4, offset + i) - 1.0; });
    } else if (dtype == TG_F32) {
        float* o = (float*)host_out;
        parallel_rows(count, [&](int64_t i) { o[i] = (float)(2.0 * u01(seed, 4, offset + i) - 1.0); });
    } else {
        return TG_ERR_DTYPE;
    }
    return TG_OK;
}

// Row pointers of the random CSR: rp[0..m] (host int64).  Returns nnz.
int64_t tg_gen_random_csr_rowptr(int64_t m, double mean, int32_t lo, int32_t hi, uint64_t seed, int64_t* rp) {
    if (m < 0 || lo < 0 || hi < lo || !rp) return TG_ERR_INVALID;
    // Poisson CDF table up to hi.
    std::vector<double> cdf(hi + 1);
    double p = std::exp(-mean), acc = 0.0;
    for (int k = 0; k <= hi; ++k) {
        acc += p;
        cdf[k] = acc;
        p *= mean / (k + 1);
    }
    std::vector<int32_t> len(m);
    parallel_rows(m, [&](int64_t i) {
        const double u = u01(seed, 1, (uint64_t)i);
        int k = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        len[i] = std::min(hi, std::max(lo, k));
    });
    rp[0] = 0;
    for (int64_t i = 0; i < m; ++i) rp[i + 1] = rp[i] + len[i];
    return rp[m];
}

// Columns and values of the random CSR for the given row pointers.  val64 /
// val32 may each be null; when both are given val32 is the float rounding of
// val64.
int tg_gen_random_csr_fill(int64_t m, int64_t n, const int64_t* rp, uint64_t seed, int32_t* ci, double* val64,
                           float* val32) {
    if (m < 0 || n <= 0 || !rp || !ci) return TG_ERR_INVALID;
    int bad = 0;
    parallel_rows(m, [&](int64_t i) {
        const int64_t b = rp[i], L = rp[i + 1] - rp[i];
        if (L > n || L > 1024) {
            bad = 1;
            return;
        }
        int32_t* row = ci + b;
        int64_t cnt = 0;
        // Floyd: for j = n-L .. n-1 draw t in [0, j]; take t unless taken, else j.
        for (int64_t d = 0; d < L; ++d) {
            const int64_t j = n - L + d;
            int64_t t = (int64_t)(u01(seed, 2, ((uint64_t)i << 10) | (uint64_t)d) * (double)(j + 1));
            if (t > j) t = j;
            bool taken = false;
            for (int64_t q = 0; q < cnt; ++q)
                if (row[q] == t) {
                    taken = true;
                    break;
                }
            row[cnt++] = (int32_t)(taken ? j : t);
        }
        std::sort(row, row + L);
        for (int64_t q = 0; q < L; ++q) {
            const double v = normal(seed, 3, (uint64_t)(b + q));
            if (val64) val64[b + q] = v;
            if (val32) val32[b + q] = (float)v;
        }
    });
    return bad ? TG_ERR_INVALID : TG_OK;
}

// 3D 7-point graph Laplacian + shift on a g x g x g grid.  Sizes: rp g^3+1,
// ci/val 7 g^3 - 6 g^2.  Returns nnz.
int64_t tg_gen_laplacian_3d(int32_t g, double shift, int64_t* rp, int32_t* ci, double* val) {
    if (g < 1 || !rp || !ci || !val) return TG_ERR_INVALID;
    const int64_t G = g, N = G * G * G;
    auto deg = [&](int64_t i, int64_t j, int64_t k) {
        return (int)(i > 0) + (i < G - 1) + (j > 0) + (j < G - 1) + (k > 0) + (k < G - 1);
    };
    rp[0] = 0;
    for (int64_t v = 0; v < N; ++v) {
        const int64_t i = v / (G * G), j = (v / G) % G, k = v % G;
        rp[v + 1] = rp[v] + 1 + deg(i, j, k);
    }
    parallel_rows(N, [&](int64_t v) {
        const int64_t i = v / (G * G), j = (v / G) % G, k = v % G;
        int64_t p = rp[v];
        auto put = [&](int64_t c, double a) {
            ci[p] = (int32_t)c;
            val[p] = a;
            ++p;
        };
        if (i > 0) put(v - G * G, -1.0);
        if (j > 0) put(v - G, -1.0);
        if (k > 0) put(v - 1, -1.0);
        put(v, deg(i, j, k) + shift);
        if (k < G - 1) put(v + 1, -1.0);
        if (j < G - 1) put(v + G, -1.0);
        if (i < G - 1) put(v + G * G, -1.0);
    });
    return rp[N];
}

// Weighted grid Laplacian (dims 2 or 3, sides s0 x s1 [x s2]): edge (v, v+e_d)
// gets weight exp(N(0,1)) drawn at index 3 v + d.  Sizes: rp N+1, ci/val at most
// (2 dims + 1) N.  Returns nnz.
int64_t tg_gen_weighted_grid(int32_t dims, int32_t s0, int32_t s1, int32_t s2, double shift, uint64_t seed,
                             int64_t* rp, int32_t* ci, double* val) {
    if ((dims != 2 && dims != 3) || s0 < 1 || s1 < 1 || !rp || !ci || !val) return TG_ERR_INVALID;
    if (dims == 2) s2 = 1;
    if (s2 < 1) return TG_ERR_INVALID;
    const int64_t S0 = s0, S1 = s1, S2 = s2, N = S0 * S1 * S2;
    const int64_t stride[3] = {S1 * S2, S2, 1};
    const int64_t side[3] = {S0, S1, S2};
    auto coord = [&](int64_t v, int d) { return (v / stride[d]) % side[d]; };
    auto w = [&](int64_t v, int d) { return std::exp(normal(seed, 5, (uint64_t)(3 * v + d))); };
    const int D = dims == 2 ? 2 : 3;
    // In 2D the two directions are the first two axes of (S0, S1, 1).
    rp[0] = 0;
    for (int64_t v = 0; v < N; ++v) {
        int c = 1;
        for (int d = 0; d < D; ++d) c += (coord(v, d) > 0) + (coord(v, d) < side[d] - 1);
        rp[v + 1] = rp[v] + c;
    }
    parallel_rows(N, [&](int64_t v) {
        int64_t p = rp[v];
        double diag = shift;
        // lower neighbours, farthest first (ascending column order)
        for (int d = 0; d < D; ++d)
            if (coord(v, d) > 0) {
                const double we = w(v - stride[d], d);
                ci[p] = (int32_t)(v - stride[d]);
                val[p] = -we;
                diag += we;
                ++p;
            }
        const int64_t pd = p++;
        for (int d = D - 1; d >= 0; --d)
            if (coord(v, d) < side[d] - 1) {
                const double we = w(v, d);
                ci[p] = (int32_t)(v + stride[d]);
                val[p] = -we;
                diag += we;
                ++p;
            }
        ci[pd] = (int32_t)v;
        val[pd] = diag;
    });
    return rp[N];
}

// 2D convection-diffusion on a g x g grid (Dirichlet boundary, rows in
// lexicographic order v = g i + j): 5-point diffusion (4, -1) plus upwind
// convection of strength 2 P along (cos theta, sin theta), P = mesh Peclet
// number.  Sizes: rp g^2+1, ci/val 5 g^2 - 4 g.  Returns nnz.
int64_t tg_gen_convection_diffusion_2d(int32_t g, double peclet, double theta, int64_t* rp, int32_t* ci,
                                       double* val) {
    if (g < 1 || !rp || !ci || !val) return TG_ERR_INVALID;
    const int64_t G = g, N = G * G;
    const double bx = 2.0 * peclet * std::cos(theta), by = 2.0 * peclet * std::sin(theta);
    rp[0] = 0;
    for (int64_t v = 0; v < N; ++v) {
        const int64_t i = v / G, j = v % G;
        rp[v + 1] = rp[v] + 1 + (i > 0) + (i < G - 1) + (j > 0) + (j < G - 1);
    }
    parallel_rows(N, [&](int64_t v) {
        const int64_t i = v / G, j = v % G;
        // upwind: flow along +x (j) takes from j-1, along +y (i) from i-1
        const double wN = 1.0 + (by > 0 ? by : 0.0), wS = 1.0 + (by < 0 ? -by : 0.0);
        const double wW = 1.0 + (bx > 0 ? bx : 0.0), wE = 1.0 + (bx < 0 ? -bx : 0.0);
        const double diag = 4.0 + std::fabs(bx) + std::fabs(by);
        int64_t p = rp[v];
        auto put = [&](int64_t c, double a) {
            ci[p] = (int32_t)c;
            val[p] = a;
            ++p;
        };
        if (i > 0) put(v - G, -wN);
        if (j > 0) put(v - 1, -wW);
        put(v, diag);
        if (j < G - 1) put(v + 1, -wE);
        if (i < G - 1) put(v + G, -wS);
    });
    return rp[N];
}

}  // extern "C"
