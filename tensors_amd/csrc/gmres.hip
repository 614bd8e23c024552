// Restarted GMRES(m) on the GPU:  X <- a * A^{-1} B + b * X  for nrhs columns.
//
// Follows GMRES::operator() / Solve / ArnoldiStep / ApplyGivensRotations
// (GMRES.hpp:127-260, 264-485, 526-601, 603-658): residual convention
// z = A x - b, left (default) or right diagonal preconditioning, Givens QR of
// the Hessenberg matrix per column, all columns stop together when every
// |beta_iter|_k <= tol * ||P b||_k, upper back substitution for y
// (GMRES.hpp:385-412, the CBLAS trsv of src/BLAS/trsv.hpp:7-48), solution
// -= Q y, restarts up to max_restarts (:244-250).
//
// Orthogonalisation: classical Gram-Schmidt applied twice (CGS2) instead of the
// reference's modified Gram-Schmidt applied twice (GMRES.hpp:555-585).  In exact
// arithmetic both give the same projection and the same H (each pass adds its
// coefficients into H); CGS2 needs O(1) passes over the basis per step instead
// of O(j) dependent dot/axpy pairs, which is what a GPU wants.  The Krylov
// vectors live in the vector type T (FP64 or FP32); the Hessenberg matrix,
// rotations, beta, y and all reductions are FP64.
//
// Kernels of Arnoldi step j (q = Q_{j+1}) in the plain CGS2 form (knob
// gmres_delayed = 0, and where the delayed form does not fit), eight launches:
//   spmm (left: z = A Q_j; right: z = P Q_j then q = A z)
//   dots1:  q = P z (left), h1_i = Q_i.q             i <= j
//   total:  h1 from the workgroup partials, H[.][j] += h1
//   upd1:   q -= sum h1_i Q_i,  h2_i = Q_i.q
//   total:  h2, H[.][j] += h2
//   upd2:   q -= sum h2_i Q_i,  |q|^2
//   givens: |q| from the partials, H[j+1][j] = |q|, rotations, beta, convergence
//           (one workgroup)
//   scale:  q /= |q|
// Every kernel returns at once when the device `done` flag is set, so the host
// enqueues steps in batches and polls between batches.

#include <string>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

constexpr int kGmresMaxCols = 32;  // nrhs limit (LDS of the multi-dot)
constexpr int kGmresMaxM = 63;     // restart length limit (m + 1 basis vectors <= 64)
constexpr int kGB = 512;           // workgroups of the vector kernels (at most; fewer for small n)
static int g_gmres_graph = 1;      // 1: cycles after the first run as one captured hipGraph
static int g_gmres_delayed = 1;    // 1: CGS2 with the second pass delayed to the next step (two basis passes per step)
static int g_gmres_split_dots = 1;  // 1: one-column dots pass of the 32-vector bucket with its two sides on alternate lanes

struct GmresHdr {
    int iter;
    int done;
    int succ;
    int m;
    int pad[12];
};

// Small FP64 arrays (offsets in doubles from the start of the small region).
struct GmresLayout {
    int m, k;
    size_t H, cs, sn, beta, y, TOL, bnorm, rnorm, h, inv, Hraw, dA, dE, dsc, total;
    GmresLayout(int m_, int k_) : m(m_), k(k_) {
        size_t o = 0;
        H = o;
        o += (size_t)(m + 1) * m * k;
        cs = o;
        o += (size_t)m * k;
        sn = o;
        o += (size_t)m * k;
        beta = o;
        o += (size_t)(m + 1) * k;
        y = o;
        o += (size_t)m * k;
        TOL = o;
        o += k;
        bnorm = o;
        o += k;
        rnorm = o;
        o += k;
        h = o;
        o += (size_t)(m + 1) * k;
        inv = o;
        o += k;
        Hraw = o;  // delayed form: H before the Givens rotations (the Arnoldi relation)
        o += (size_t)(m + 1) * m * k;
        dA = o;  // delayed form: coefficients of the update pass
        o += (size_t)(m + 1) * k;
        dE = o;
        o += (size_t)(m + 1) * k;
        dsc = o;  // per column: 1/rho, t_j/rho^2
        o += (size_t)2 * k;
        total = o;
    }
};

__device__ __forceinline__ void grow_range(int n, int& i0, int& i1) {
    const int b = blockIdx.x, B = gridDim.x;
    i0 = (int)((int64_t)n * b / B);
    i1 = (int)((int64_t)n * (b + 1) / B);
}

// Per-thread partial dots v[0..nq) of column c, reduced over the row lanes of the
// workgroup in a fixed order; writes partial[(block * nq + i) * k + c].  Uses
// (kBlock / kWave) * nq * k doubles of dynamic LDS.
template <int NQMAX>
__device__ __forceinline__ void multi_partial(double (&v)[NQMAX], int nq, int k, int kpad,
                                              double* __restrict__ partial) {
    extern __shared__ double gm_sh[];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    {
        // reduce-scatter over the row lanes (registers past nq hold zeros)
        const ScatterOut o = wave_scatter_sum<NQMAX>(v, kpad);
        const int c = lane % kpad;
        if ((lane % o.low) < kpad && c < k)
#pragma unroll
            for (int i = 0; i < ScatterSizes<NQMAX>::A; ++i)
                if (i < o.count && o.base + i < nq) gm_sh[(wave * nq + o.base + i) * k + c] = v[i];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nq * k; t += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += gm_sh[w * nq * k + t];
        partial[(int64_t)blockIdx.x * nq * k + t] = s;
    }
}

inline size_t multi_lds_bytes(int nq, int k) { return (size_t)(kBlock / kWave) * nq * k * sizeof(double); }

// One workgroup per (i, c): out[i*k + c] = sum_b partial[(b*nq + i)*k + c] (fixed tree).
__global__ __launch_bounds__(kBlock) void gm_total_kernel(int nq, int k, int B, const double* __restrict__ partial,
                                                          double* __restrict__ out, const GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    const int ic = blockIdx.x, i = ic / k, c = ic - i * k;
    __shared__ double sh[kBlock];
    double s = 0.0;
    for (int b = threadIdx.x; b < B; b += kBlock) s += partial[((int64_t)b * nq + i) * k + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[ic] = sh[0];
}

// gm_total_kernel followed by H[i][j] += out[i*k + c] (each Gram-Schmidt pass adds its
// coefficients, GMRES.hpp:564-568): one workgroup per (i, c), so each H entry has one writer.
__global__ __launch_bounds__(kBlock) void gm_total_addh_kernel(int nq, int k, int B, int j, int m,
                                                               const double* __restrict__ partial,
                                                               double* __restrict__ out, double* __restrict__ H,
                                                               const GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    const int ic = blockIdx.x, i = ic / k, c = ic - i * k;
    __shared__ double sh[kBlock];
    double s = 0.0;
    for (int b = threadIdx.x; b < B; b += kBlock) s += partial[((int64_t)b * nq + i) * k + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[ic] = sh[0];
        H[((int64_t)i * m + j) * k + c] += sh[0];
    }
}

// dots1: q = P z (left, diag) or q already holds A Q_j; h1_i = Q_i . q
template <typename T, int NQMAX>
__global__ __launch_bounds__(kBlock) void gm_dots1_kernel(int n, int k, int kpad, int nq, const T* __restrict__ Q,
                                                          int64_t qstride, const T* __restrict__ z,
                                                          const T* __restrict__ d, T* __restrict__ q,
                                                          double* __restrict__ partial, const GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    int i0, i1;
    grow_range(n, i0, i1);
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    double v[NQMAX];
#pragma unroll
    for (int i = 0; i < NQMAX; ++i) v[i] = 0.0;
    if (c < k)
        for (int r = i0 + ro; r < i1; r += rows) {
            const int64_t e = (int64_t)r * k + c;
            // all basis loads first, at clamped (valid) vectors: a condition around a
            // load would split the unrolled loop into one memory round trip per vector
            T qi[NQMAX];
#pragma unroll
            for (int i = 0; i < NQMAX; ++i) qi[i] = Q[(int64_t)min(i, nq - 1) * qstride + e];
            T qv;
            if (d) {
                qv = d[r] * z[e];
                q[e] = qv;
            } else {
                qv = q[e];
            }
#pragma unroll
            for (int i = 0; i < NQMAX; ++i)
                if (i < nq) v[i] += (double)qi[i] * (double)qv;
        }
    multi_partial<NQMAX>(v, nq, k, kpad, partial);
}

// upd: q -= sum_i h_i Q_i; then either h'_i = Q_i . q (norm_only == 0) or |q|^2.
template <typename T, int NQMAX>
__global__ __launch_bounds__(kBlock) void gm_update_kernel(int n, int k, int kpad, int nq, const T* __restrict__ Q,
                                                           int64_t qstride, T* __restrict__ q,
                                                           const double* __restrict__ h, int norm_only,
                                                           double* __restrict__ partial,
                                                           const GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    int i0, i1;
    grow_range(n, i0, i1);
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    double v[NQMAX];
#pragma unroll
    for (int i = 0; i < NQMAX; ++i) v[i] = 0.0;
    if (c < k) {
        T hc[NQMAX];
#pragma unroll
        for (int i = 0; i < NQMAX; ++i) hc[i] = i < nq ? (T)h[i * k + c] : T(0);
        for (int r = i0 + ro; r < i1; r += rows) {
            const int64_t e = (int64_t)r * k + c;
            T qi[NQMAX];
            T qv = q[e];
#pragma unroll
            for (int i = 0; i < NQMAX; ++i) qi[i] = Q[(int64_t)min(i, nq - 1) * qstride + e];  // loads first
#pragma unroll
            for (int i = 0; i < NQMAX; ++i)
                if (i < nq) qv -= hc[i] * qi[i];
            q[e] = qv;
            if (norm_only) {
                v[0] += (double)qv * (double)qv;
            } else {
#pragma unroll
                for (int i = 0; i < NQMAX; ++i)
                    if (i < nq) v[i] += (double)qi[i] * (double)qv;
            }
        }
    }
    multi_partial<NQMAX>(v, norm_only ? 1 : nq, k, kpad, partial);
}


// ---------------------------------------------------------------------------
// Delayed second pass (knob gmres_delayed, the default where it fits): classical
// Gram-Schmidt applied twice, the second pass of step j's new vector done in step
// j + 1 together with that step's projection (the delayed CGS2 of Swirydowicz et al.,
// "Low synchronization Gram-Schmidt and GMRES algorithms", 2020).  Step j starts with
// q_0..q_{j-1} final and u in slot j: the new vector of step j - 1 after its first
// pass (for j = 0, q_0 itself).
//   spmm:    v = A u into slot j + 1 (left Jacobi: z = A u, v = P z in the dots pass)
//   ddots:   per column a_i = q_i.u, b_i = q_i.v (i < j), alpha = u.u, beta = u.v
//   total:   the 2j + 2 totals
//   dscalar: rho = sqrt(alpha - |a|^2) = |u - Q a| (Q orthonormal); column j - 1 of H
//            gets += a (its second pass) and H[j][j-1] = rho, then its Givens
//            rotation, beta and the stop test (iter = j).  Unless the cycle stops:
//            with t = Q_{<=j}^T v (t_i = b_i, t_j = (beta - a.b) / rho) and
//            A q_j = (v - Q Hraw a) / rho (the Arnoldi relation of the finished
//            columns), column j of H gets its first pass g = (t - Hraw a) / rho and
//            the next vector is u' = A q_j - Q g = (v - Q t) / rho.
//   dupdate: q_j = (u - Q a) / rho and u' in one pass over the basis.
// A cycle of m steps ends with a finishing step (ddots, total, dscalar, dupdate
// without a new vector) that completes column m - 1 and q_m.  In exact arithmetic H
// and the basis are those of CGS2; the basis is read twice per step instead of three
// times and a step is five launches instead of eight.

// Workgroup partials of the 2j + 2 values [a_0..a_{j-1}, b_0..b_{j-1}, alpha, beta]
// held in v[i] (a_i), v[NQMAX + i] (b_i), v[2 NQMAX], v[2 NQMAX + 1].
template <int NQMAX>
__device__ __forceinline__ void dd_partial(double (&v)[2 * NQMAX + 2], int j, int k, int kpad,
                                           double* __restrict__ partial) {
    extern __shared__ double gm_sh[];
    const int nv = 2 * j + 2;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    // compact index of register r (-1: an unused register)
    auto compact = [&](int r) {
        return r < NQMAX ? (r < j ? r : -1) : r < 2 * NQMAX ? (r - NQMAX < j ? j + r - NQMAX : -1) : 2 * j + (r - 2 * NQMAX);
    };
    constexpr int R = 2 * NQMAX + 2;
    {
        // reduce-scatter over the row lanes (the unused registers hold zeros)
        const ScatterOut o = wave_scatter_sum<R>(v, kpad);
        const int c = lane % kpad;
        if ((lane % o.low) < kpad && c < k)
#pragma unroll
            for (int i = 0; i < ScatterSizes<R>::A; ++i) {
                const int t = i < o.count ? compact(o.base + i) : -1;
                if (t >= 0) gm_sh[(wave * nv + t) * k + c] = v[i];
            }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nv * k; t += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += gm_sh[w * nv * k + t];
        partial[(int64_t)blockIdx.x * nv * k + t] = s;
    }
}

template <typename T, int NQMAX>
__global__ __launch_bounds__(kBlock) void gm_ddots_kernel(int n, int k, int kpad, int j, int finish,
                                                          const T* __restrict__ Q, int64_t qstride,
                                                          const T* __restrict__ z, const T* __restrict__ d,
                                                          T* __restrict__ vslot, double* __restrict__ partial,
                                                          const GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    int i0, i1;
    grow_range(n, i0, i1);
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    double acc[2 * NQMAX + 2];
#pragma unroll
    for (int r = 0; r < 2 * NQMAX + 2; ++r) acc[r] = 0.0;
    const T* U = Q + (int64_t)j * qstride;
    // FP32: two rows per iteration, so the bytes in flight match FP64's (the FP64
    // accumulators bound the waves per SIMD); the second row is clamped to a valid one
    // and masked at the end of the range
    constexpr int RU = sizeof(T) == 4 ? 2 : 1;
    if (c < k)
        for (int r = i0 + ro; r < i1; r += rows * RU) {
            T qi[RU][NQMAX], uv[RU], vv[RU];
            int rr[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                rr[u] = r + u * rows < i1 ? r + u * rows : r;
                const int64_t e = (int64_t)rr[u] * k + c;
#pragma unroll
                for (int i = 0; i < NQMAX; ++i) qi[u][i] = Q[(int64_t)max(0, min(i, j - 1)) * qstride + e];  // loads first
                uv[u] = U[e];
            }
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                const int64_t e = (int64_t)rr[u] * k + c;
                const bool ok = r + u * rows < i1;
                if (finish) {
                    vv[u] = uv[u];
                } else if (d) {
                    vv[u] = d[rr[u]] * z[e];
                    if (ok) vslot[e] = vv[u];
                } else {
                    vv[u] = vslot[e];
                }
                if (!ok) uv[u] = vv[u] = T(0);
            }
#pragma unroll
            for (int u = 0; u < RU; ++u) {
#pragma unroll
                for (int i = 0; i < NQMAX; ++i)
                    if (i < j) {
                        acc[i] += (double)qi[u][i] * (double)uv[u];
                        acc[NQMAX + i] += (double)qi[u][i] * (double)vv[u];
                    }
                acc[2 * NQMAX] += (double)uv[u] * (double)uv[u];
                acc[2 * NQMAX + 1] += (double)uv[u] * (double)vv[u];
            }
        }
    dd_partial<NQMAX>(acc, j, k, kpad, partial);
}

// One column, 32-vector bucket: the dots pass with its two sides on alternate lanes.  Even lanes accumulate a_i = Q_i . u and alpha = u . u, odd lanes b_i = Q_i . v
// and beta = u . v, over the same rows (the pair's loads of a basis row are one
// address), so a lane holds NQMAX + 1 FP64 accumulators instead of 2 NQMAX + 2 and the
// reduce-scatter (over the lanes of one side: two "columns") half the values.  FP32: RU
// rows per iteration, 1 (111 VGPRs, 4 waves per SIMD) measured faster on systems of 1e4
// and 9e4 rows, 2 (179 VGPRs, 2 waves) on 1e6 rows.  FP64: 1 (144 VGPRs, 3 waves), below
// 2^19 rows.
template <int NQMAX>
__device__ __forceinline__ void dd_partial_split(double (&v)[NQMAX + 1], int j, double* __restrict__ partial) {
    extern __shared__ double gm_sh[];
    const int nv = 2 * j + 2;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    constexpr int R = NQMAX + 1;
    {
        const ScatterOut o = wave_scatter_sum<R>(v, 2);
        const int side = lane % 2;
        if ((lane % o.low) < 2)
#pragma unroll
            for (int i = 0; i < ScatterSizes<R>::A; ++i) {
                const int r = o.base + i;
                const int t = i >= o.count ? -1 : r < NQMAX ? (r < j ? side * j + r : -1) : 2 * j + side;
                if (t >= 0) gm_sh[wave * nv + t] = v[i];
            }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nv; t += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += gm_sh[w * nv + t];
        partial[(int64_t)blockIdx.x * nv + t] = s;
    }
}

template <typename T, int NQMAX, int RU>
__global__ __launch_bounds__(kBlock) void gm_ddots_split_kernel(int n, int j, int finish,
                                                                const T* __restrict__ Q, int64_t qstride,
                                                                const T* __restrict__ z,
                                                                const T* __restrict__ d, T* __restrict__ vslot,
                                                                double* __restrict__ partial,
                                                                const GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    int i0, i1;
    grow_range(n, i0, i1);
    const int side = threadIdx.x & 1, ro = threadIdx.x >> 1, rows = kBlock / 2;
    double acc[NQMAX + 1];
#pragma unroll
    for (int r = 0; r < NQMAX + 1; ++r) acc[r] = 0.0;
    const T* U = Q + (int64_t)j * qstride;
    for (int r = i0 + ro; r < i1; r += rows * RU) {
        T qi[RU][NQMAX], uv[RU], vv[RU];
        int rr[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            rr[u] = r + u * rows < i1 ? r + u * rows : r;
#pragma unroll
            for (int i = 0; i < NQMAX; ++i) qi[u][i] = Q[(int64_t)max(0, min(i, j - 1)) * qstride + rr[u]];
            uv[u] = U[rr[u]];
        }
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const bool ok = r + u * rows < i1;
            if (finish) {
                vv[u] = uv[u];
            } else if (d) {
                vv[u] = d[rr[u]] * z[rr[u]];
                if (ok && side) vslot[rr[u]] = vv[u];
            } else {
                vv[u] = vslot[rr[u]];
            }
            if (!ok) uv[u] = vv[u] = T(0);
        }
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const T w = side ? vv[u] : uv[u];
#pragma unroll
            for (int i = 0; i < NQMAX; ++i)
                if (i < j) acc[i] += (double)qi[u][i] * (double)w;
            acc[NQMAX] += (double)uv[u] * (double)w;
        }
    }
    dd_partial_split<NQMAX>(acc, j, partial);
}

// One workgroup: the scalar work of delayed step j (finish: the cycle's last column).
// The totals, column j - 1 of H and the rotations are staged in LDS; the per-column
// serial part (|a|^2, a.b, the Givens chain) reads only LDS, everything else is
// spread over (row, column) pairs.
constexpr int kDdMaxM = 32;  // restart length limit of the delayed form
__global__ __launch_bounds__(kBlock) void gm_dscalar_kernel(int j, int m, int k, int finish,
                                                            const double* __restrict__ tot, double* __restrict__ small,
                                                            GmresLayout L, GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    __shared__ double sv[(2 * kDdMaxM + 2) * kGmresMaxCols];    // totals [value][column]
    __shared__ double col[(kDdMaxM + 1) * kGmresMaxCols];       // column j - 1, then the new column j
    __shared__ double rot[2 * kDdMaxM * kGmresMaxCols];         // cs, sn of rotations 0..j-2
    __shared__ double sc[4 * kGmresMaxCols];                    // rho, 1/rho, a.b, t_j / rho^2
    __shared__ int conv;
    const int nv = 2 * j + 2, jj = j - 1;
    double* H = small + L.H;
    double* Hr = small + L.Hraw;
    auto at = [&](int r, int cc) -> int64_t { return (int64_t)r * m * k + cc; };  // + col * k
    for (int t = threadIdx.x; t < nv * k; t += kBlock) sv[t] = tot[t];
    if (j > 0) {
        for (int t = threadIdx.x; t < j * k; t += kBlock) {  // column jj (first pass) rows 0..jj
            const int i = t / k, c = t - i * k;
            col[t] = Hr[at(i, c) + (int64_t)jj * k];
        }
        for (int t = threadIdx.x; t < jj * k; t += kBlock) {
            rot[t] = small[L.cs + t];
            rot[jj * k + t] = small[L.sn + t];
        }
    }
    if (threadIdx.x == 0) conv = 1;
    __syncthreads();
    // per column: rho, a.b, column jj completed and rotated, beta, stop test
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        double rho = 1.0, ab = 0.0;
        if (j > 0) {
            const double alpha = sv[2 * j * k + c];
            double aa = 0.0;
            for (int i = 0; i < j; ++i) {
                const double ai = sv[i * k + c];
                aa += ai * ai;
                if (!finish) ab += ai * sv[(j + i) * k + c];
            }
            // |u - Q a|; at a breakdown (u in the span of Q to rounding) at least 2^-52 |u|
            rho = sqrt(fmax(alpha - aa, 0x1p-104 * alpha));
            for (int i = 0; i < j; ++i) col[i * k + c] += sv[i * k + c];  // second pass of column jj
            col[j * k + c] = rho;
        }
        sc[c] = rho;
        sc[k + c] = 1.0 / rho;
        sc[2 * k + c] = ab;
    }
    __syncthreads();
    if (j > 0) {
        // Hraw column jj (unrotated) to memory; then the Givens chain on a copy
        for (int t = threadIdx.x; t < (j + 1) * k; t += kBlock) {
            const int i = t / k, c = t - i * k;
            Hr[at(i, c) + (int64_t)jj * k] = col[t];
        }
        __syncthreads();
        for (int c = threadIdx.x; c < k; c += blockDim.x) {
            // Givens on column jj (GMRES.hpp:603-658)
            for (int i = 0; i < jj; ++i) {
                const double xi = col[i * k + c], eta = col[(i + 1) * k + c];
                const double co = rot[i * k + c], si = rot[jj * k + i * k + c];
                col[i * k + c] = co * xi + si * eta;
                col[(i + 1) * k + c] = -si * xi + co * eta;
            }
            const double xi = col[jj * k + c], eta = col[j * k + c];
            const double r = sqrt(xi * xi + eta * eta);
            double co, si;
            if (fabs(xi) <= 2.220446049250313e-16 * r) {
                co = 0.0;
                si = 1.0;
            } else {
                const double rinv = 1.0 / r;
                co = fabs(xi) * rinv;
                si = (xi / fabs(xi)) * eta * rinv;
            }
            small[L.cs + jj * k + c] = co;
            small[L.sn + jj * k + c] = si;
            col[jj * k + c] = co * xi + si * eta;
            col[j * k + c] = 0.0;
            const double bj = small[L.beta + jj * k + c];
            small[L.beta + j * k + c] = -si * bj;
            small[L.beta + jj * k + c] = co * bj;
            if (!(fabs(small[L.beta + j * k + c]) <= small[L.TOL + c])) atomicAnd(&conv, 0);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < (j + 1) * k; t += kBlock) {
            const int i = t / k, c = t - i * k;
            H[at(i, c) + (int64_t)jj * k] = col[t];
        }
    }
    // coefficients of q_j = u / rho - sum (a_i / rho) q_i
    for (int t = threadIdx.x; t < j * k; t += kBlock) small[L.dA + t] = sv[t] * sc[k + t % k];
    for (int c = threadIdx.x; c < k; c += blockDim.x) small[L.dsc + c] = sc[k + c];
    __syncthreads();
    const bool stop = j > 0 && (conv || j >= hd->m);
    if (!finish && !stop) {
        // column j, first pass: g_l = (t_l - (Hraw a)_l) / rho, t_l = b_l (l < j),
        // t_j = (beta - a.b) / rho; Hraw[l][i] = 0 for l > i + 1
        for (int t = threadIdx.x; t < (j + 1) * k; t += kBlock) {
            const int l = t / k, c = t - l * k;
            const double irho = sc[k + c];
            double ha = 0.0;
#pragma unroll 4
            for (int i = l > 0 ? l - 1 : 0; i < j; ++i) ha += Hr[at(l, c) + (int64_t)i * k] * sv[i * k + c];
            const double tl = l < j ? sv[(j + l) * k + c] : (j > 0 ? (sv[(2 * j + 1) * k + c] - sc[2 * k + c]) * irho
                                                                    : sv[(2 * j + 1) * k + c]);
            const double g = (tl - ha) * irho;
            Hr[at(l, c) + (int64_t)j * k] = g;
            H[at(l, c) + (int64_t)j * k] = g;
            if (l == j) sc[3 * k + c] = tl * irho * irho;
        }
        __syncthreads();
        // u' = v / rho - (t_j / rho^2) u - sum (t_i / rho - t_j a_i / rho^2) q_i
        for (int t = threadIdx.x; t < j * k; t += kBlock) {
            const int c = t % k;
            small[L.dE + t] = sv[j * k + t] * sc[k + c] - sc[3 * k + c] * sv[t];
        }
        for (int c = threadIdx.x; c < k; c += blockDim.x) small[L.dsc + k + c] = sc[3 * k + c];
    }
    if (threadIdx.x == 0 && j > 0) {
        hd->iter = j;
        hd->succ = conv;
        hd->done = stop ? 1 : 0;
    }
}

// q_j = u / rho - sum a_i / rho q_i (slot j, in place) and, while the cycle goes on,
// u' = v / rho - (t_j / rho^2) u - sum (t_i / rho - t_j a_i / rho^2) q_i (slot j + 1,
// in place of v).  Runs for the step whose scalar kernel ended the cycle (q_j only).
template <typename T, int NQMAX>
__global__ __launch_bounds__(kBlock) void gm_dupdate_kernel(int n, int k, int kpad, int j, T* __restrict__ Q,
                                                            int64_t qstride, const double* __restrict__ small,
                                                            GmresLayout L, const GmresHdr* __restrict__ hd) {
    if (hd->done && hd->iter != j) return;
    const bool next = !hd->done;
    int i0, i1;
    grow_range(n, i0, i1);
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    if (c >= k) return;
    T ca[NQMAX], ce[NQMAX];
#pragma unroll
    for (int i = 0; i < NQMAX; ++i) {
        ca[i] = i < j ? (T)small[L.dA + i * k + c] : T(0);
        ce[i] = (next && i < j) ? (T)small[L.dE + i * k + c] : T(0);
    }
    const T ir = (T)small[L.dsc + c];
    const T cj = next ? (T)small[L.dsc + k + c] : T(0);
    T* U = Q + (int64_t)j * qstride;
    T* V = Q + (int64_t)(j + 1) * qstride;
    for (int r = i0 + ro; r < i1; r += rows) {
        const int64_t e = (int64_t)r * k + c;
        T qi[NQMAX];
#pragma unroll
        for (int i = 0; i < NQMAX; ++i) qi[i] = Q[(int64_t)max(0, min(i, j - 1)) * qstride + e];  // loads first
        const T uv = U[e];
        const T vv = next ? V[e] : T(0);
        T qv = uv * ir;
        T nv = vv * ir - cj * uv;
#pragma unroll
        for (int i = 0; i < NQMAX; ++i)
            if (i < j) {
                qv -= ca[i] * qi[i];
                nv -= ce[i] * qi[i];
            }
        if (j > 0) U[e] = qv;
        if (next) V[e] = nv;
    }
}

// One column, 32-vector bucket, FP32: gm_dupdate_kernel with its two outputs on
// alternate lanes (even: q_j, odd: the next pending vector) over the same rows, the
// pair's loads of a basis row being one address; each output's operations and order are
// those of gm_dupdate_kernel (the compiler's contraction of the products may differ, so
// the vectors agree to rounding).  Mixed GMRES 1e4 / 9e4 / 1e6 rows 1-3 % faster; in FP64
// the unsplit kernel measured faster on 1e6 rows.
template <typename T, int NQMAX>
__global__ __launch_bounds__(kBlock) void gm_dupdate_split_kernel(int n, int j, T* __restrict__ Q, int64_t qstride,
                                                                  const double* __restrict__ small, GmresLayout L,
                                                                  const GmresHdr* __restrict__ hd) {
    if (hd->done && hd->iter != j) return;
    const bool next = !hd->done;
    const int side = threadIdx.x & 1;
    if (side && !next) return;  // no pending vector to form
    int i0, i1;
    grow_range(n, i0, i1);
    const int ro = threadIdx.x >> 1, rows = kBlock / 2;
    T co[NQMAX];
#pragma unroll
    for (int i = 0; i < NQMAX; ++i) co[i] = i < j ? (T)small[(side ? L.dE : L.dA) + i] : T(0);
    const T ir = (T)small[L.dsc];
    const T cj = next ? (T)small[L.dsc + 1] : T(0);
    T* U = Q + (int64_t)j * qstride;
    T* V = Q + (int64_t)(j + 1) * qstride;
    for (int r = i0 + ro; r < i1; r += rows) {
        T qi[NQMAX];
#pragma unroll
        for (int i = 0; i < NQMAX; ++i) qi[i] = Q[(int64_t)max(0, min(i, j - 1)) * qstride + r];  // loads first
        const T uv = U[r];
        T out;
        if (side) {
            const T vv = V[r];
            out = vv * ir - cj * uv;
        } else {
            out = uv * ir;
        }
#pragma unroll
        for (int i = 0; i < NQMAX; ++i)
            if (i < j) out -= co[i] * qi[i];
        if (side)
            V[r] = out;
        else if (j > 0)
            U[r] = out;
    }
}

// H[j+1][j] = |q|; Givens (GMRES.hpp:603-658); beta; convergence; ++iter.  |q|^2 per
// column is first summed from the B workgroup partials in gm_total_kernel's order (one
// workgroup of kBlock threads, one column at a time).
__global__ __launch_bounds__(kBlock) void gm_givens_kernel(int j, int m, int k, int B, const double* __restrict__ partial,
                                                           double* __restrict__ qn2, double* __restrict__ H,
                                                           double* __restrict__ cs, double* __restrict__ sn,
                                                           double* __restrict__ beta, const double* __restrict__ TOL,
                                                           double* __restrict__ inv, GmresHdr* __restrict__ hd) {
    if (hd->done) return;
    __shared__ double sh[kBlock];
    for (int c = 0; c < k; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < B; b += kBlock) s += partial[(int64_t)b * k + c];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int h = kBlock / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) qn2[c] = sh[0];
        __syncthreads();
    }
    __shared__ int conv;
    if (threadIdx.x == 0) conv = 1;
    __syncthreads();
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        const double qn = sqrt(qn2[c]);
        auto Hx = [&](int r) -> double& { return H[((int64_t)r * m + j) * k + c]; };
        Hx(j + 1) = qn;
        inv[c] = 1.0 / qn;
        for (int i = 0; i < j; ++i) {
            const double xi = Hx(i), eta = Hx(i + 1);
            const double co = cs[i * k + c], si = sn[i * k + c];
            Hx(i) = co * xi + si * eta;
            Hx(i + 1) = -si * xi + co * eta;
        }
        const double xi = Hx(j), eta = Hx(j + 1);
        const double r = sqrt(xi * xi + eta * eta);
        double co, si;
        if (fabs(xi) <= 2.220446049250313e-16 * r) {
            co = 0.0;
            si = 1.0;
        } else {
            const double rinv = 1.0 / r;
            co = fabs(xi) * rinv;
            si = (xi / fabs(xi)) * eta * rinv;
        }
        cs[j * k + c] = co;
        sn[j * k + c] = si;
        Hx(j) = co * xi + si * eta;
        Hx(j + 1) = 0.0;
        const double bj = beta[j * k + c];
        beta[(j + 1) * k + c] = -si * bj;
        beta[j * k + c] = co * bj;
        if (!(fabs(beta[(j + 1) * k + c]) <= TOL[c])) atomicAnd(&conv, 0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        hd->iter = j + 1;
        hd->succ = conv;
        hd->done = (conv || j + 1 >= hd->m) ? 1 : 0;
    }
}

// q *= inv (per column); runs even when the step just set `done` is not needed,
// so it keeps the guard.
template <typename T>
__global__ __launch_bounds__(kBlock) void gm_scale_kernel(int n, int k, T* __restrict__ q,
                                                          const double* __restrict__ inv, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock)
        q[e] = q[e] * (T)inv[e % k];
}

// q *= inv after Arnoldi step j: runs when step j ran, including the step that ended
// the cycle (the reference normalises every new basis vector, GMRES.hpp:589-598, so
// GetBasis sees it normalised); returns at once for steps after the cycle ended.
template <typename T>
__global__ __launch_bounds__(kBlock) void gm_scale_step_kernel(int n, int k, T* __restrict__ q,
                                                               const double* __restrict__ inv,
                                                               const GmresHdr* __restrict__ hd, int j) {
    if (hd->done && hd->iter != j + 1) return;
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock)
        q[e] = q[e] * (T)inv[e % k];
}

// v = P u (diag) or copy
template <typename T>
__global__ __launch_bounds__(kBlock) void gm_prec_kernel(int n, int k, const T* __restrict__ u, const T* __restrict__ d,
                                                         T* __restrict__ v, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock)
        v[e] = d ? d[e / k] * u[e] : u[e];
}

// z = s_x * x + s_b * B (row-major ld for B), used for z = A x - b and z = -b.
template <typename T>
__global__ __launch_bounds__(kBlock) void gm_residual_kernel(int n, int k, const T* __restrict__ Bm, int64_t ldB,
                                                             T* __restrict__ z, int has_ax) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        const T bv = Bm[i * ldB + (e - i * k)];
        z[e] = has_ax ? z[e] + (-bv) : -bv;
    }
}

// |v|^2 per column (for b norms and r norms) -> partial
template <typename T>
__global__ __launch_bounds__(kBlock) void gm_norm_kernel(int n, int k, int kpad, const T* __restrict__ v,
                                                         double* __restrict__ partial) {
    int i0, i1;
    grow_range(n, i0, i1);
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    double acc[1] = {0.0};
    if (c < k)
        for (int r = i0 + ro; r < i1; r += rows) {
            const double x = (double)v[(int64_t)r * k + c];
            acc[0] += x * x;
        }
    multi_partial<1>(acc, 1, k, kpad, partial);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void gm_zero_kernel(int64_t count, T* __restrict__ d) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        d[e] = T(0);
}

// After the cycle's first residual: r_norm, beta[0] = r_norm, inv = 1/r_norm,
// H = 0, TOL (first call), convergence at iteration 0 (GMRES.hpp:352-375).
__global__ void gm_start_kernel(int m, int k, const double* __restrict__ nn2, double* __restrict__ small,
                                GmresLayout L, double tol, int set_tol, GmresHdr* __restrict__ hd) {
    __shared__ int conv;
    if (threadIdx.x == 0) conv = 1;
    __syncthreads();
    for (int t = threadIdx.x; t < (m + 1) * m * k; t += blockDim.x) small[L.H + t] = 0.0;
    for (int t = threadIdx.x; t < (m + 1) * k; t += blockDim.x) small[L.beta + t] = 0.0;
    __syncthreads();
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        const double nv = sqrt(nn2[c]);
        if (set_tol) {
            small[L.bnorm + c] = nv;
            small[L.TOL + c] = nv * tol;
        } else {
            small[L.rnorm + c] = nv;
            small[L.beta + c] = nv;
            small[L.inv + c] = 1.0 / nv;
            if (!(fabs(nv) <= small[L.TOL + c])) atomicAnd(&conv, 0);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && !set_tol) {
        hd->iter = 0;
        hd->succ = conv;
        hd->m = m;
        hd->done = (conv || m <= 0) ? 1 : 0;
    }
}

// y = H^{-1} beta per column (upper back substitution, iter x iter).  The triangle of
// H and beta of a chunk of columns are first staged in LDS by the whole workgroup, so
// the serial substitution of each column reads LDS only; the order of the
// subtractions is the plain row-oriented one.
constexpr int kLsqLds = 6144;  // doubles of staged triangle per chunk of columns
__global__ __launch_bounds__(kBlock) void gm_lsq_kernel(int m, int k, double* __restrict__ small, GmresLayout L,
                                                        const GmresHdr* __restrict__ hd) {
    const int it = hd->iter;
    if (it <= 0) return;
    __shared__ double tri[kLsqLds];
    __shared__ double sb[(kGmresMaxM + 1) * kGmresMaxCols];  // beta, then y, [row][chunk column]
    const int ntri = it * (it + 1) / 2;
    const int kc = max(1, min(k, kLsqLds / ntri));
    auto tix = [&](int i, int j) { return i * it - i * (i - 1) / 2 + (j - i); };  // i <= j
    for (int c0 = 0; c0 < k; c0 += kc) {
        const int nc = min(kc, k - c0);
        for (int t = threadIdx.x; t < it * it * nc; t += blockDim.x) {
            const int cc = t % nc, ij = t / nc, i = ij / it, j = ij % it;
            if (j >= i) tri[tix(i, j) * nc + cc] = small[L.H + ((int64_t)i * m + j) * k + c0 + cc];
        }
        for (int t = threadIdx.x; t < it * nc; t += blockDim.x)
            sb[(t / nc) * kGmresMaxCols + t % nc] = small[L.beta + (t / nc) * k + c0 + t % nc];
        __syncthreads();
        for (int cc = threadIdx.x; cc < nc; cc += blockDim.x)
            for (int i = it - 1; i >= 0; --i) {
                double s = sb[i * kGmresMaxCols + cc];
                for (int j = i + 1; j < it; ++j) s -= tri[tix(i, j) * nc + cc] * sb[j * kGmresMaxCols + cc];
                sb[i * kGmresMaxCols + cc] = s / tri[tix(i, i) * nc + cc];
            }
        __syncthreads();
        for (int t = threadIdx.x; t < it * nc; t += blockDim.x)
            small[L.y + (t / nc) * k + c0 + t % nc] = sb[(t / nc) * kGmresMaxCols + t % nc];
        __syncthreads();
    }
}

// z = sum_{j < iter} y_j Q_j;  (left) x = z;  solution = (implicit ? 0 : solution) - x.
// Right preconditioning applies P to z first (x = P z, GMRES.hpp:444-448).
template <typename T>
__global__ __launch_bounds__(kBlock) void gm_synth_kernel(int n, int k, const T* __restrict__ Q, int64_t qstride,
                                                          const double* __restrict__ small, GmresLayout L,
                                                          const T* __restrict__ d_right, T* __restrict__ sol,
                                                          int implicit_zero, const GmresHdr* __restrict__ hd) {
    const int it = hd->iter;
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int c = (int)(e % k);
        T zv = T(0);
        for (int j = 0; j < it; ++j) zv += (T)small[L.y + j * k + c] * Q[j * qstride + e];
        if (d_right) zv = d_right[e / k] * zv;
        sol[e] = implicit_zero ? -zv : sol[e] - zv;
    }
}

// solution = (implicit ? 0 : solution) + v, where v = P(-Q y)
template <typename T>
__global__ __launch_bounds__(kBlock) void gm_sub_kernel(int64_t count, const T* __restrict__ v, T* __restrict__ sol,
                                                        int implicit_zero) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        sol[e] = implicit_zero ? v[e] : sol[e] + v[e];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void gm_copy_kernel(int64_t count, const T* __restrict__ s, T* __restrict__ d) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        d[e] = s[e];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void gm_load_kernel(int n, int k, const T* __restrict__ X, int64_t ldX,
                                                         T* __restrict__ x) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        x[e] = X[i * ldX + (e - i * k)];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void gm_write_kernel(int n, int k, T a, const T* __restrict__ x, T b,
                                                          T* __restrict__ X, int64_t ldX, int a_one, int b_mode) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        T* Xp = X + i * ldX + (e - i * k);
        T v = a_one ? x[e] : a * x[e];
        if (b_mode == 1)
            v = v + *Xp;
        else if (b_mode == 2)
            v = v + b * *Xp;
        *Xp = v;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void gm_zero_or_scale_kernel(int n, int k, T b, T* __restrict__ X, int64_t ldX) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        T* Xp = X + i * ldX + (e - i * k);
        *Xp = (b == T(0)) ? T(0) : *Xp * b;
    }
}

// ---------------------------------------------------------------------------
// Host driver.
// ---------------------------------------------------------------------------
static int pow2ceil_g(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

template <typename T, int NQMAX>
static void launch_dots1(hipStream_t s, int gb, int n, int k, int kpad, int nq, const T* Q, int64_t qs, const T* z,
                         const T* d, T* q, double* partial, const GmresHdr* hd) {
    hipLaunchKernelGGL((gm_dots1_kernel<T, NQMAX>), dim3(gb), dim3(kBlock), multi_lds_bytes(nq, k), s, n, k, kpad,
                       nq, Q, qs, z, d, q, partial, hd);
}
template <typename T, int NQMAX>
static void launch_update(hipStream_t s, int gb, int n, int k, int kpad, int nq, const T* Q, int64_t qs, T* q,
                          const double* h, int norm_only, double* partial, const GmresHdr* hd) {
    hipLaunchKernelGGL((gm_update_kernel<T, NQMAX>), dim3(gb), dim3(kBlock), multi_lds_bytes(nq, k), s, n, k, kpad,
                       nq, Q, qs, q, h, norm_only, partial, hd);
}
template <typename T>
static void dots1(hipStream_t s, int gb, int n, int k, int kpad, int nq, const T* Q, int64_t qs, const T* z, const T* d,
                  T* q, double* partial, const GmresHdr* hd) {
    if (nq <= 4) launch_dots1<T, 4>(s, gb, n, k, kpad, nq, Q, qs, z, d, q, partial, hd);
    else if (nq <= 8) launch_dots1<T, 8>(s, gb, n, k, kpad, nq, Q, qs, z, d, q, partial, hd);
    else if (nq <= 16) launch_dots1<T, 16>(s, gb, n, k, kpad, nq, Q, qs, z, d, q, partial, hd);
    else if (nq <= 32) launch_dots1<T, 32>(s, gb, n, k, kpad, nq, Q, qs, z, d, q, partial, hd);
    else launch_dots1<T, 64>(s, gb, n, k, kpad, nq, Q, qs, z, d, q, partial, hd);
}
template <typename T>
static void update(hipStream_t s, int gb, int n, int k, int kpad, int nq, const T* Q, int64_t qs, T* q, const double* h,
                   int norm_only, double* partial, const GmresHdr* hd) {
    if (nq <= 4) launch_update<T, 4>(s, gb, n, k, kpad, nq, Q, qs, q, h, norm_only, partial, hd);
    else if (nq <= 8) launch_update<T, 8>(s, gb, n, k, kpad, nq, Q, qs, q, h, norm_only, partial, hd);
    else if (nq <= 16) launch_update<T, 16>(s, gb, n, k, kpad, nq, Q, qs, q, h, norm_only, partial, hd);
    else if (nq <= 32) launch_update<T, 32>(s, gb, n, k, kpad, nq, Q, qs, q, h, norm_only, partial, hd);
    else launch_update<T, 64>(s, gb, n, k, kpad, nq, Q, qs, q, h, norm_only, partial, hd);
}

template <typename T, int NQMAX>
static void launch_ddots(hipStream_t s, int gb, int n, int k, int kpad, int j, int finish, const T* Q, int64_t qs,
                         const T* z, const T* d, T* v, double* partial, const GmresHdr* hd) {
    const size_t lds = (size_t)(kBlock / kWave) * (2 * j + 2) * k * sizeof(double);
    if constexpr (NQMAX == 32) {
        // FP64 on 2^19 rows and more: the unsplit pass (1e6 rows: 210 against 214 ms)
        if (k == 1 && g_gmres_split_dots && (sizeof(T) == 4 || n < (1 << 19))) {
            bool two = false;
            if constexpr (sizeof(T) == 4) two = n >= (1 << 19);
            if (two) {
                if constexpr (sizeof(T) == 4)
                    hipLaunchKernelGGL((gm_ddots_split_kernel<T, NQMAX, 2>), dim3(gb), dim3(kBlock), lds, s, n, j,
                                       finish, Q, qs, z, d, v, partial, hd);
            } else {
                hipLaunchKernelGGL((gm_ddots_split_kernel<T, NQMAX, 1>), dim3(gb), dim3(kBlock), lds, s, n, j, finish,
                                   Q, qs, z, d, v, partial, hd);
            }
            return;
        }
    }
    hipLaunchKernelGGL((gm_ddots_kernel<T, NQMAX>), dim3(gb), dim3(kBlock), lds, s, n, k, kpad, j, finish, Q, qs, z,
                       d, v, partial, hd);
}
template <typename T>
static void ddots(hipStream_t s, int gb, int n, int k, int kpad, int j, int finish, const T* Q, int64_t qs,
                  const T* z, const T* d, T* v, double* partial, const GmresHdr* hd) {
    if (j <= 4) launch_ddots<T, 4>(s, gb, n, k, kpad, j, finish, Q, qs, z, d, v, partial, hd);
    else if (j <= 8) launch_ddots<T, 8>(s, gb, n, k, kpad, j, finish, Q, qs, z, d, v, partial, hd);
    else if (j <= 16) launch_ddots<T, 16>(s, gb, n, k, kpad, j, finish, Q, qs, z, d, v, partial, hd);
    else launch_ddots<T, 32>(s, gb, n, k, kpad, j, finish, Q, qs, z, d, v, partial, hd);
}
template <typename T>
static void dupdate(hipStream_t s, int gb, int n, int k, int kpad, int j, T* Q, int64_t qs, const double* small,
                    const GmresLayout& L, const GmresHdr* hd) {
    if (j <= 4)
        hipLaunchKernelGGL((gm_dupdate_kernel<T, 4>), dim3(gb), dim3(kBlock), 0, s, n, k, kpad, j, Q, qs, small, L, hd);
    else if (j <= 8)
        hipLaunchKernelGGL((gm_dupdate_kernel<T, 8>), dim3(gb), dim3(kBlock), 0, s, n, k, kpad, j, Q, qs, small, L, hd);
    else if (j <= 16)
        hipLaunchKernelGGL((gm_dupdate_kernel<T, 16>), dim3(gb), dim3(kBlock), 0, s, n, k, kpad, j, Q, qs, small, L,
                           hd);
    else if (sizeof(T) == 4 && k == 1 && g_gmres_split_dots)  // FP64: the unsplit pass measured faster
        hipLaunchKernelGGL((gm_dupdate_split_kernel<T, 32>), dim3(gb), dim3(kBlock), 0, s, n, j, Q, qs, small, L, hd);
    else
        hipLaunchKernelGGL((gm_dupdate_kernel<T, 32>), dim3(gb), dim3(kBlock), 0, s, n, k, kpad, j, Q, qs, small, L,
                           hd);
}

static int grid_elems(int64_t count) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((count + kBlock - 1) / kBlock, 8192));
}

struct GmresWs {
    GmresHdr* hd;
    double* small;
    double* partial;
    double* tot;  // (2m+2) * k totals
    void* Q;
    void* x;
    void* z;
    void* sol;
    size_t bytes;
};

static size_t gmres_layout(int32_t n, int32_t m, int32_t k, size_t vs, GmresWs* w, char* base) {
    GmresLayout L(m, k);
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t o = off;
        off += align256(b);
        return base ? (void*)(base + o) : nullptr;
    };
    GmresWs t;
    t.hd = (GmresHdr*)take(sizeof(GmresHdr));
    t.small = (double*)take(L.total * sizeof(double));
    t.partial = (double*)take((size_t)kGB * (2 * m + 2) * k * sizeof(double));  // delayed form: 2m + 2 values
    t.tot = (double*)take((size_t)(2 * m + 2) * k * sizeof(double));
    t.Q = take((size_t)(m + 1) * n * k * vs);
    t.x = take((size_t)n * k * vs);
    t.z = take((size_t)n * k * vs);
    t.sol = take((size_t)n * k * vs);
    t.bytes = off;
    if (w) *w = t;
    return off;
}

size_t gmres_workspace_bytes(int32_t n, int32_t max_iter, int32_t k, size_t vs) {
    const int m = std::max(0, std::min(max_iter, n + 1));
    return gmres_layout(n, std::max(m, 1), k, vs, nullptr, nullptr);
}

template <typename T>
int gmres_solve_t(const CsrView& A, const OpRef& Aop, int prec_kind, const T* d, const OpRef& Pop, int side,
                  int32_t k, int32_t max_iter, double a, const T* Bm, int64_t ldB, double b, T* X, int64_t ldX,
                  double tol, int32_t max_restarts, int use_guess, void* ws, size_t ws_bytes, GmresResult* res,
                  hipStream_t s, GmresReuse* reuse) {
    const int n = A.n;
    res->succeeded = 0;
    res->iterations = 0;
    res->restarts = 0;
    if (k < 1 || k > kGmresMaxCols) return fail(TG_ERR_INVALID, "gmres: nrhs must be in [1, %d]", kGmresMaxCols);
    if (A.m != n || n < 1) return fail(TG_ERR_SHAPE, "gmres: the operator must be square and non-empty");
    const int m = std::min(max_iter, n + 1);  // GMRES.hpp:89
    if (m < 0 || m > kGmresMaxM) return fail(TG_ERR_INVALID, "gmres: restart length must be in [0, %d]", kGmresMaxM);
    const int mm = std::max(m, 1);
    if (!ws || ws_bytes < gmres_layout(n, mm, k, sizeof(T), nullptr, nullptr))
        return fail(TG_ERR_WORKSPACE, "gmres: workspace missing or smaller than tg_gmres_workspace_bytes()");
    if (prec_kind == 1 && !d) return fail(TG_ERR_INVALID, "gmres: diagonal preconditioner without a diagonal");
    if (prec_kind == 2 && !Pop.fn) return fail(TG_ERR_INVALID, "gmres: callback preconditioner without a function");
    const bool host_ops = Aop.is_callback() || prec_kind == 2;
    GmresWs w;
    gmres_layout(n, mm, k, sizeof(T), &w, (char*)ws);
    GmresLayout L(mm, k);
    const T* dd = prec_kind == 1 ? d : nullptr;
    // v = P u: device scaling (identity / diagonal) or the host callback
    auto applyP_on = [&](hipStream_t st, const T* u_, T* v_, const int* skip) -> int {
        if (prec_kind == 2) return op_apply(Pop, u_, v_, k, skip, st);
        hipLaunchKernelGGL((gm_prec_kernel<T>), dim3(grid_elems((int64_t)n * k)), dim3(kBlock), 0, st, n, k, u_, dd,
                           v_, skip);
        return TG_OK;
    };
    auto applyP = [&](const T* u_, T* v_, const int* skip) -> int { return applyP_on(s, u_, v_, skip); };
    const int64_t N = (int64_t)n * k;
    const int64_t qstride = N;
    T* Q = (T*)w.Q;
    T* x = (T*)w.x;
    T* z = (T*)w.z;
    T* sol = (T*)w.sol;
    const int kpad = pow2ceil_g(k);
    const int ge = grid_elems(N);
    // workgroups of the reduction kernels: enough for one row per thread, at most kGB
    // (a 1e4-row system gets 40: fewer idle lanes and 40 partials per total, not 512)
    const int gb = (int)std::max<int64_t>(1, std::min<int64_t>(kGB, ((int64_t)n + kBlock / kpad - 1) / (kBlock / kpad)));

    // b norms: x = B; z = P x (left) or z = B (right)  (GMRES.hpp:168-185)
    GmresHdr h0{};
    h0.m = m;
    h0.done = 0;
    TG_HIP(hipMemcpyAsync(w.hd, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((gm_load_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, Bm, ldB, x);
    if (side == 0) {
        int rcp = applyP((const T*)x, z, nullptr);
        if (rcp) return rcp;
    } else {
        hipLaunchKernelGGL((gm_copy_kernel<T>), dim3(ge), dim3(kBlock), 0, s, N, (const T*)x, z);
    }
    hipLaunchKernelGGL((gm_norm_kernel<T>), dim3(gb), dim3(kBlock), multi_lds_bytes(1, k), s, n, k, kpad, (const T*)z,
                       w.partial);
    hipLaunchKernelGGL(gm_total_kernel, dim3(k), dim3(kBlock), 0, s, 1, k, gb, (const double*)w.partial, w.tot,
                       (const GmresHdr*)w.hd);
    hipLaunchKernelGGL(gm_start_kernel, dim3(1), dim3(kBlock), 0, s, mm, k, (const double*)w.tot, w.small, L, tol, 1,
                       w.hd);
    TG_LAUNCH_CHECK();
    std::vector<double> bn(k);
    if (!(reuse && reuse->checked)) {
        TG_HIP(hipMemcpyAsync(bn.data(), w.small + L.bnorm, k * sizeof(double), hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        double bmax = 0.0;
        for (int c = 0; c < k; ++c) {
            if (std::isnan(bn[c])) {  // GMRES.hpp:193-204
                fail(TG_OK, "gmres: right-hand side contains NaNs. Doing nothing.");
                return TG_OK;
            }
            bmax = std::max(bmax, bn[c]);
        }
        if (!(bmax > 0.0)) {  // :206-229
            hipLaunchKernelGGL((gm_zero_or_scale_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, (T)b, X, ldX);
            TG_LAUNCH_CHECK();
            res->succeeded = 1;
            return TG_OK;
        }
    }
    if (use_guess) hipLaunchKernelGGL((gm_load_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, (const T*)X, ldX, sol);
    else hipLaunchKernelGGL((gm_zero_kernel<T>), dim3(ge), dim3(kBlock), 0, s, N, sol);

    // Arnoldi step j on stream st (GMRES.hpp:526-601): eight launches, each returning at
    // once when the device flag says the cycle is done.
    auto arnoldi_step = [&](hipStream_t st, int j) -> int {
        const int nq = j + 1;
        T* Qj = Q + j * qstride;
        T* q = Q + (int64_t)(j + 1) * qstride;
        int rc;
        if (side == 0) {
            // z = A Q_j; q = P z (diagonal P is folded into dots1)
            rc = op_apply(Aop, (const T*)Qj, (dd || prec_kind == 2) ? z : q, k, &w.hd->done, st);
            if (!rc && prec_kind == 2) rc = applyP_on(st, (const T*)z, q, &w.hd->done);
        } else {
            rc = applyP_on(st, (const T*)Qj, z, &w.hd->done);
            if (!rc) rc = op_apply(Aop, (const T*)z, q, k, &w.hd->done, st);
        }
        if (rc) return rc;
        dots1<T>(st, gb, n, k, kpad, nq, (const T*)Q, qstride, (const T*)z, side == 0 ? dd : nullptr, q, w.partial,
                 (const GmresHdr*)w.hd);
        hipLaunchKernelGGL(gm_total_addh_kernel, dim3(nq * k), dim3(kBlock), 0, st, nq, k, gb, j, mm,
                           (const double*)w.partial, w.tot, w.small + L.H, (const GmresHdr*)w.hd);
        update<T>(st, gb, n, k, kpad, nq, (const T*)Q, qstride, q, (const double*)w.tot, 0, w.partial,
                  (const GmresHdr*)w.hd);
        hipLaunchKernelGGL(gm_total_addh_kernel, dim3(nq * k), dim3(kBlock), 0, st, nq, k, gb, j, mm,
                           (const double*)w.partial, w.tot, w.small + L.H, (const GmresHdr*)w.hd);
        update<T>(st, gb, n, k, kpad, nq, (const T*)Q, qstride, q, (const double*)w.tot, 1, w.partial,
                  (const GmresHdr*)w.hd);
        hipLaunchKernelGGL(gm_givens_kernel, dim3(1), dim3(kBlock), 0, st, j, mm, k, gb, (const double*)w.partial,
                           w.tot, w.small + L.H, w.small + L.cs, w.small + L.sn, w.small + L.beta,
                           (const double*)(w.small + L.TOL), w.small + L.inv, w.hd);
        hipLaunchKernelGGL((gm_scale_step_kernel<T>), dim3(ge), dim3(kBlock), 0, st, n, k, q,
                           (const double*)(w.small + L.inv), (const GmresHdr*)w.hd, j);
        return TG_OK;
    };
    // Delayed second pass: two basis passes and five launches per step, and a finishing
    // step per cycle.  Used where its partial sums fit the LDS (2m + 2 values of k columns)
    // and its registers (m <= 32).
    const bool delayed = g_gmres_delayed && mm <= kDdMaxM && (2 * mm + 2) * k <= 2048;
    auto delayed_step = [&](hipStream_t st, int j, bool finish) -> int {
        T* U = Q + (int64_t)j * qstride;
        T* V = finish ? U : Q + (int64_t)(j + 1) * qstride;
        if (!finish) {
            int rc;
            if (side == 0) {
                // v = A u (diagonal P folded into the dots pass)
                rc = op_apply(Aop, (const T*)U, (dd || prec_kind == 2) ? z : V, k, &w.hd->done, st);
                if (!rc && prec_kind == 2) rc = applyP_on(st, (const T*)z, V, &w.hd->done);
            } else {
                rc = applyP_on(st, (const T*)U, z, &w.hd->done);
                if (!rc) rc = op_apply(Aop, (const T*)z, V, k, &w.hd->done, st);
            }
            if (rc) return rc;
        }
        const int nv = 2 * j + 2;
        ddots<T>(st, gb, n, k, kpad, j, finish ? 1 : 0, (const T*)Q, qstride, (const T*)z,
                 (side == 0 && !finish) ? dd : nullptr, V, w.partial, (const GmresHdr*)w.hd);
        hipLaunchKernelGGL(gm_total_kernel, dim3(nv * k), dim3(kBlock), 0, st, nv, k, gb, (const double*)w.partial,
                           w.tot, (const GmresHdr*)w.hd);
        hipLaunchKernelGGL(gm_dscalar_kernel, dim3(1), dim3(kBlock), 0, st, j, mm, k, finish ? 1 : 0,
                           (const double*)w.tot, w.small, L, w.hd);
        dupdate<T>(st, gb, n, k, kpad, j, Q, qstride, (const double*)w.small, L, (const GmresHdr*)w.hd);
        return TG_OK;
    };
    auto step = [&](hipStream_t st, int j) -> int { return delayed ? delayed_step(st, j, false) : arnoldi_step(st, j); };
    auto finish_cycle = [&](hipStream_t st) -> int { return delayed ? delayed_step(st, m, true) : TG_OK; };
    // the captured cycle: this call's own, or the caller's (kept across calls)
    GmresReuse own;
    GmresReuse& cyc = reuse ? *reuse : own;
    int cycles = 0;

    auto solve = [&](bool implicit_zero, int& succ_out) -> int {
        // residual: z = A x - b (or -b), x = solution   (GMRES.hpp:282-350)
        T* target = (side == 0) ? z : Q;
        if (implicit_zero) {
            hipLaunchKernelGGL((gm_residual_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, Bm, ldB, target, 0);
        } else {
            int rc = op_apply(Aop, (const T*)sol, target, k, nullptr, s);
            if (rc) return rc;
            hipLaunchKernelGGL((gm_residual_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, Bm, ldB, target, 1);
        }
        if (side == 0) {
            int rcp = applyP((const T*)z, Q, nullptr);
            if (rcp) return rcp;
        }
        hipLaunchKernelGGL((gm_norm_kernel<T>), dim3(gb), dim3(kBlock), multi_lds_bytes(1, k), s, n, k, kpad,
                           (const T*)Q, w.partial);
        TG_LAUNCH_CHECK();
        // the totals kernel returns early while `done` is set: clear it first
        GmresHdr clr{};
        clr.m = m;
        TG_HIP(hipMemcpyAsync(w.hd, &clr, sizeof(clr), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(gm_total_kernel, dim3(k), dim3(kBlock), 0, s, 1, k, gb, (const double*)w.partial, w.tot,
                           (const GmresHdr*)w.hd);
        hipLaunchKernelGGL(gm_start_kernel, dim3(1), dim3(kBlock), 0, s, mm, k, (const double*)w.tot, w.small, L, tol,
                           0, w.hd);
        hipLaunchKernelGGL((gm_scale_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, Q,
                           (const double*)(w.small + L.inv), (const int*)nullptr);
        TG_LAUNCH_CHECK();
        GmresHdr hh{};
        if (!host_ops && g_gmres_graph && (cycles > 0 || reuse)) {
            // every cycle after the first (with a caller's reuse state, every cycle): its
            // m steps as one captured graph (captured once, replayed after that; steps
            // after convergence return at once on the device flag)
            if (!cyc.ex) {
                if (!cyc.cs) TG_HIP(hipStreamCreateWithFlags(&cyc.cs, hipStreamNonBlocking));
                TG_HIP(hipStreamBeginCapture(cyc.cs, hipStreamCaptureModeThreadLocal));
                int rc = TG_OK;
                for (int j = 0; j < m && rc == TG_OK; ++j) rc = step(cyc.cs, j);
                if (rc == TG_OK) rc = finish_cycle(cyc.cs);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(cyc.cs, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (ec != hipSuccess) return fail(TG_ERR_HIP, "gmres: graph capture failed: %s", hipGetErrorString(ec));
                const hipError_t ei = hipGraphInstantiate(&cyc.ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    cyc.ex = nullptr;
                    return fail(TG_ERR_HIP, "gmres: graph instantiation failed: %s", hipGetErrorString(ei));
                }
            }
            TG_HIP(hipGraphLaunch(cyc.ex, s));
        } else {
            int j = 0, batch = host_ops ? 1 : 4;
            while (j < m) {
                for (int t = 0; t < batch && j < m; ++t, ++j) {
                    int rc = step(s, j);
                    if (rc) return rc;
                }
                TG_LAUNCH_CHECK();
                TG_HIP(hipMemcpyAsync(&hh, w.hd, sizeof(hh), hipMemcpyDeviceToHost, s));
                TG_HIP(hipStreamSynchronize(s));
                if (hh.done) break;
                if (!host_ops) batch = std::min(batch * 2, 16);
            }
            int rcf = finish_cycle(s);  // returns at once on the device when the cycle stopped early
            if (rcf) return rcf;
            TG_LAUNCH_CHECK();
        }
        ++cycles;
        TG_HIP(hipMemcpyAsync(&hh, w.hd, sizeof(hh), hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        succ_out = hh.succ;
        res->iterations = hh.iter;
        if (hh.succ && hh.iter == 0) return TG_OK;  // converged before any step: solution untouched (:371-375)
        hipLaunchKernelGGL(gm_lsq_kernel, dim3(1), dim3(kBlock), 0, s, mm, k, w.small, L, (const GmresHdr*)w.hd);
        if (side == 1 && prec_kind == 2) {
            // z = Q y;  x = P z;  solution -= x   (GMRES.hpp:416-480)
            hipLaunchKernelGGL((gm_synth_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, (const T*)Q, qstride,
                               (const double*)w.small, L, (const T*)nullptr, z, 1, (const GmresHdr*)w.hd);
            int rcp = applyP((const T*)z, x, nullptr);
            if (rcp) return rcp;
            hipLaunchKernelGGL((gm_sub_kernel<T>), dim3(ge), dim3(kBlock), 0, s, N, (const T*)x, sol,
                               implicit_zero ? 1 : 0);
        } else {
            hipLaunchKernelGGL((gm_synth_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, (const T*)Q, qstride,
                               (const double*)w.small, L, side == 1 ? dd : nullptr, sol, implicit_zero ? 1 : 0,
                               (const GmresHdr*)w.hd);
        }
        TG_LAUNCH_CHECK();
        return TG_OK;
    };

    int restarts = 0, succ = 0;
    int rc = solve(!use_guess, succ);
    if (rc) return rc;
    ++restarts;
    while (!succ && restarts <= max_restarts) {
        rc = solve(false, succ);
        if (rc) return rc;
        ++restarts;
    }
    --restarts;
    const int a_one = (a == 1.0);
    const int b_mode = (b == 0.0) ? 0 : (b == 1.0) ? 1 : 2;
    hipLaunchKernelGGL((gm_write_kernel<T>), dim3(ge), dim3(kBlock), 0, s, n, k, (T)a, (const T*)sol, (T)b, X, ldX,
                       a_one, b_mode);
    TG_LAUNCH_CHECK();
    std::vector<double> beta((size_t)(mm + 1) * k);
    TG_HIP(hipMemcpyAsync(beta.data(), w.small + L.beta, beta.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    if (reuse && reuse->checked)  // the norms were not read before the cycles
        TG_HIP(hipMemcpyAsync(bn.data(), w.small + L.bnorm, k * sizeof(double), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    res->succeeded = succ;
    res->restarts = restarts;
    for (int c = 0; c < k; ++c) res->rel_res[c] = std::fabs(beta[(size_t)res->iterations * k + c]) / bn[c];
    return TG_OK;
}

template int gmres_solve_t<double>(const CsrView&, const OpRef&, int, const double*, const OpRef&, int, int32_t,
                                   int32_t, double, const double*, int64_t, double, double*, int64_t, double, int32_t,
                                   int, void*, size_t, GmresResult*, hipStream_t, GmresReuse*);
template int gmres_solve_t<float>(const CsrView&, const OpRef&, int, const float*, const OpRef&, int, int32_t,
                                  int32_t, double, const float*, int64_t, double, float*, int64_t, double, int32_t, int,
                                  void*, size_t, GmresResult*, hipStream_t, GmresReuse*);

int gmres_tune_set(const char* key, int32_t value, bool* known) {
    const std::string k = key;
    *known = k == "gmres_graph" || k == "gmres_delayed" || k == "gmres_split_dots";
    if (!*known) return TG_OK;
    if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: %s must be 0 or 1", key);
    (k == "gmres_graph" ? g_gmres_graph : k == "gmres_delayed" ? g_gmres_delayed : g_gmres_split_dots) = value;
    return TG_OK;
}

int32_t gmres_tune_get(const char* key) {
    const std::string k = key;
    return k == "gmres_graph"        ? g_gmres_graph
           : k == "gmres_delayed"    ? g_gmres_delayed
           : k == "gmres_split_dots" ? g_gmres_split_dots
                                     : -1;
}

}  // namespace tg

using namespace tg;

extern "C" size_t tg_gmres_workspace_bytes(int32_t n, int32_t max_iter, int32_t nrhs, tg_dtype val_type) {
    if (n < 1 || nrhs < 1) return 0;
    return gmres_workspace_bytes(n, max_iter, nrhs, val_type == TG_F32 ? 4 : 8);
}

template <typename T>
static int gmres_entry(const CsrView& A, const OpRef& Aop, int32_t prec_kind, const void* prec_diag,
                       const OpRef& Pop, int32_t side, int32_t nrhs, int32_t max_iter, double a_coef, const void* B,
                       int64_t ldB, double b_coef, void* X, int64_t ldX, double tol, int32_t max_restarts,
                       int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                       int32_t* iterations, int32_t* restarts, double* rel_residuals, hipStream_t s) {
    GmresResult res{};
    const int rc = gmres_solve_t<T>(A, Aop, prec_kind, (const T*)prec_diag, Pop, side, nrhs, max_iter, a_coef,
                                    (const T*)B, ldB, b_coef, (T*)X, ldX, tol, max_restarts, use_guess, workspace,
                                    workspace_bytes, &res, s);
    if (rc) return rc;
    if (succeeded) *succeeded = res.succeeded;
    if (iterations) *iterations = res.iterations;
    if (restarts) *restarts = res.restarts;
    if (rel_residuals)
        for (int c = 0; c < nrhs; ++c) rel_residuals[c] = res.rel_res[c];
    return TG_OK;
}

extern "C" int tg_gmres_solve(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                              const void* a, tg_dtype val_type, int32_t prec_kind, const void* prec_diag,
                              int32_t side, int32_t nrhs, int32_t max_iter, double a_coef, const void* B,
                              int64_t ldB, double b_coef, void* X, int64_t ldX, double tol, int32_t max_restarts,
                              int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                              int32_t* iterations, int32_t* restarts, double* rel_residuals, void* stream) {
    if (succeeded) *succeeded = 0;
    if (iterations) *iterations = 0;
    if (restarts) *restarts = 0;
    if (n < 1 || nrhs < 1 || !B || !X || ldB < nrhs || ldX < nrhs || (prec_kind != 0 && prec_kind != 1) ||
        (side != 0 && side != 1))
        return fail(TG_ERR_INVALID, "tg_gmres_solve: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_gmres_solve: rp_type must be I32/I64");
    CsrView A;
    A.m = n;
    A.n = n;
    A.nnz = nnz;
    A.rp = rp;
    A.rp_type = rp_type;
    A.ci = ci;
    A.val = a;
    A.val_type = val_type;
    OpRef Aop;
    Aop.csr = &A;
    OpRef Pop;
    hipStream_t s = as_stream(stream);
    if (spmm_may_slab(A, nrhs)) A.band = csr_band(A, s);
    if (val_type == TG_F64)
        return gmres_entry<double>(A, Aop, prec_kind, prec_diag, Pop, side, nrhs, max_iter, a_coef, B, ldB, b_coef, X,
                                   ldX, tol, max_restarts, use_guess, workspace, workspace_bytes, succeeded,
                                   iterations, restarts, rel_residuals, s);
    if (val_type == TG_F32)
        return gmres_entry<float>(A, Aop, prec_kind, prec_diag, Pop, side, nrhs, max_iter, a_coef, B, ldB, b_coef, X,
                                  ldX, tol, max_restarts, use_guess, workspace, workspace_bytes, succeeded, iterations,
                                  restarts, rel_residuals, s);
    return fail(TG_ERR_DTYPE, "tg_gmres_solve: val_type must be F64/F32");
}

extern "C" int tg_gmres_solve_op(int32_t n, tg_apply_fn apply_A, void* ctx_A, tg_apply_fn apply_P, void* ctx_P,
                                 tg_dtype val_type, int32_t side, int32_t nrhs, int32_t max_iter, double a_coef,
                                 const void* B, int64_t ldB, double b_coef, void* X, int64_t ldX, double tol,
                                 int32_t max_restarts, int32_t use_guess, void* workspace, size_t workspace_bytes,
                                 int32_t* succeeded, int32_t* iterations, int32_t* restarts, double* rel_residuals,
                                 void* stream) {
    if (succeeded) *succeeded = 0;
    if (iterations) *iterations = 0;
    if (restarts) *restarts = 0;
    if (n < 1 || nrhs < 1 || !B || !X || ldB < nrhs || ldX < nrhs || !apply_A || (side != 0 && side != 1))
        return fail(TG_ERR_INVALID, "tg_gmres_solve_op: bad argument");
    CsrView A;
    A.m = n;
    A.n = n;
    A.val_type = val_type;
    OpRef Aop;
    Aop.fn = apply_A;
    Aop.ctx = ctx_A;
    OpRef Pop;
    Pop.fn = apply_P;
    Pop.ctx = ctx_P;
    const int prec_kind = apply_P ? 2 : 0;
    hipStream_t s = as_stream(stream);
    if (val_type == TG_F64)
        return gmres_entry<double>(A, Aop, prec_kind, nullptr, Pop, side, nrhs, max_iter, a_coef, B, ldB, b_coef, X,
                                   ldX, tol, max_restarts, use_guess, workspace, workspace_bytes, succeeded,
                                   iterations, restarts, rel_residuals, s);
    if (val_type == TG_F32)
        return gmres_entry<float>(A, Aop, prec_kind, nullptr, Pop, side, nrhs, max_iter, a_coef, B, ldB, b_coef, X,
                                  ldX, tol, max_restarts, use_guess, workspace, workspace_bytes, succeeded, iterations,
                                  restarts, rel_residuals, s);
    return fail(TG_ERR_DTYPE, "tg_gmres_solve_op: val_type must be F64/F32");
}

// Device views into a GMRES workspace (the members GMRES::GetBasis,
// GetHessenbergMatrix, Solution and beta of GMRES.hpp:51-69, 820-833), valid
// after a solve with the same n, max_iter, nrhs and value type.  m = min(max_iter,
// n + 1) (at least 1): Q is (m+1) x n x nrhs in val_type, H (m+1) x m x nrhs and
// beta (m+1) x nrhs in FP64 (H holds the Givens-rotated Hessenberg matrix, as the
// reference's H after ApplyGivensRotations), solution n x nrhs in val_type.
extern "C" int tg_gmres_workspace_parts(void* workspace, size_t workspace_bytes, int32_t n, int32_t max_iter,
                                        int32_t nrhs, tg_dtype val_type, void** Q, double** H, double** beta,
                                        void** solution) {
    if (!workspace || n < 1 || nrhs < 1 || (val_type != TG_F64 && val_type != TG_F32))
        return fail(TG_ERR_INVALID, "tg_gmres_workspace_parts: bad argument");
    const int m = std::max(1, std::min(max_iter, n + 1));
    GmresWs w;
    const size_t need = gmres_layout(n, m, nrhs, val_type == TG_F64 ? 8 : 4, &w, (char*)workspace);
    if (workspace_bytes < need) return fail(TG_ERR_WORKSPACE, "tg_gmres_workspace_parts: workspace too small");
    const GmresLayout L(m, nrhs);
    if (Q) *Q = w.Q;
    if (H) *H = w.small + L.H;
    if (beta) *beta = w.small + L.beta;
    if (solution) *solution = w.sol;
    return TG_OK;
}
