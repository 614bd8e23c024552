// Batched, continuously fed GMRES(m) engine: many nonsymmetric systems of uneven
// size, each with 1..8 right-hand sides, advanced together.
//
// Every system runs the recurrences of GMRES::operator() / Solve / ArnoldiStep /
// ApplyGivensRotations (GMRES.hpp:127-260, 264-485, 526-601, 603-658) with
// left diagonal or identity preconditioning, its own restart length m,
// tolerance and restart limit: residual z = A x - b (:282-350), Arnoldi with
// two orthogonalisation passes (classical Gram-Schmidt twice, as the
// single-system solver; the reference applies modified Gram-Schmidt twice),
// Givens QR of the Hessenberg matrix per column, stop when every
// |beta_iter|_k <= tol ||P b||_k (:183-185), upper back substitution for y
// (:385-412) and solution -= Q y (:420-480).
//
// All systems run restart cycles in lock step.  One cycle is, for ALL active
// systems at once, launches over the row tiles (kGbTileRows rows of one system)
// of every active system:
//   start1   z = A x - b (or -b on a first cycle without guess), Q_0 = P z,
//            tile partials of |Q_0|^2
//   start2   Q_0 /= |Q_0|; tile 0: beta_0, H = 0, convergence before any step
//   for j = 0 .. m_max - 1 (a system past its own m, or converged, skips):
//     step1  q = P A Q_j on the tile's rows; tile partials of Q_i.q, i <= j
//     step2  h1 = sum of those partials; q -= sum_i h1_i Q_i; partials of Q_i.q;
//            tile 0: H[.][j] = h1
//     step3  h2 likewise; q -= sum_i h2_i Q_i; partials of |q|^2; tile 0: H += h2
//     step4  q /= |q|; tile 0: H[j+1][j] = |q|, the Givens rotations, beta,
//            the stop test
//   lsq      tile 0: y = H^{-1} beta (upper back substitution)
//   update   x -= Q y on the tile's rows
//   end      tile 0: cycle count, finished / succeeded
// With the delayed second pass (knob gbatch_delayed = 1 when the engine is
// created, the default) a step is two launches and two passes over the basis,
// as the single-system solver's delayed form: step j (j = 0 .. m; j = m is a
// system's finishing step) runs
//   dstep1   v = P A u on the tile's rows (u = Q_j, the step j - 1 vector after one
//            pass; no product on the finishing step), tile partials of
//            a_i = Q_i.u, b_i = Q_i.v (i < j), u.u, u.v
//   dstep2   every workgroup adds the partials and computes the same scalars:
//            rho = sqrt(u.u - |a|^2), column j - 1 of H completed (first pass G
//            plus a, H[j][j-1] = rho), its Givens rotation and the stop test, and
//            unless the cycle stops, column j's first pass G = (t - Hc a) / rho
//            (t = Q^T v from the totals, Hc the unrotated columns); tile 0 writes
//            them (Hc, H, cs, sn, the carried right-hand side beta, G, the state)
//            into arrays no workgroup reads in the same launch; then q_j =
//            (u - Q a) / rho and the next vector (v - Q t) / rho on the tile's rows.
// Every workgroup of a system sums that system's tile partials itself, in one
// fixed order, so all of them see bitwise the same scalars, and no workgroup
// reads what another writes in the same launch.  A system's arithmetic depends
// only on the system, so its result is bitwise the same alone, in any mix and
// when it joins while others run (new systems start at the next cycle).  A
// cycle of all systems is one captured hipGraph, captured again when the tile
// list or the longest restart length changes.

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

constexpr int kGbMaxK = 8;
constexpr int kGbMaxM = 63;
constexpr int kGbTileRows = 1024;

struct GbSys {
    const int32_t* rp;
    const int32_t* ci;
    const double* val;
    const double* d;  // left Jacobi diagonal (z = d * r) or null
    const double* b;
    double* x;     // the solution, n x k (ld k)
    double* Q;     // (m+1) x n x k basis
    double* H;     // (m+1) x m x k, entry (i, j, c) at (i*m + j)*k + c
    double* cs;    // m x k
    double* sn;    // m x k
    double* beta;  // (m+1) x k
    double* y;     // m x k
    double* p1;    // tiles x (m+1) x k partials of the first Gram-Schmidt pass
    double* p2;    // ... of the second pass
    double* pn;    // tiles x k partials of a squared norm
    double* tot;   // (2 kGbMaxM + 2) x kGbMaxK totals of the last reduction (big systems)
    double* G;     // delayed form: (m+1) x m x k first-pass coefficients, column j from step j
    double* Hc;    // delayed form: (m+1) x m x k completed, unrotated columns (the Arnoldi relation)
    int n, k, kpad, tiles, tile_rows, m;
    int big;       // more than g_gb_big_tiles tiles: the totals come from gb_totals_kernel
    int delayed;   // delayed second pass (p1 holds 2m + 2 values per tile, beta the carried rhs)
};

struct GbState {
    double TOL[kGbMaxK], bnorm[kGbMaxK], rel[kGbMaxK];
    int done;        // finished: converged, restarts used up, or a special case
    int succ;        // converged
    int nan;         // NaN right-hand side: nothing was written
    int zero_rhs;    // every right-hand side is zero: x = 0, succeeded
    int active;      // the current cycle is still stepping
    int conv;        // the current cycle met the stop test
    int iters;       // Arnoldi steps of the current (last) cycle
    int cycles;      // completed cycles
    int max_cycles;  // max_restarts + 1
    int zero_cycle;  // the cycle that starts from x = 0 without reading x (0), or -1
    int pad[6];
};

struct GbTile {
    int slot, r0, r1, t;
};

// ---- reductions ---------------------------------------------------------------
// Thread layout: c = column lane, ro = row lane (kBlock / kpad rows per sweep).

// Sums NQ per-thread values per column over the workgroup (fixed order: wave
// butterflies over the row lanes, then the waves in order) and writes
// part[(t * stride + i) * k + c] for i < nq.  ws: 4 * NQ * kGbMaxK doubles of LDS.
template <int NQ>
__device__ __forceinline__ void multi_partial(double (&acc)[NQ], int nq, int k, int kpad, double* part, int t,
                                              int stride, double* ws) {
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave, c = threadIdx.x % kpad;
    const ScatterOut o = wave_scatter_sum<NQ>(acc, kpad);  // kpad <= kGbMaxK = 8: C values per lane
    if ((lane & 7) < kpad)
#pragma unroll
        for (int i = 0; i < ScatterSizes<NQ>::C; ++i)
            if (i < o.count) ws[(w * NQ + o.base + i) * kGbMaxK + c] = acc[i];
    __syncthreads();
    for (int p = threadIdx.x; p < nq * kpad; p += kBlock) {
        const int i = p / kpad, cc = p % kpad;
        if (cc < k) {
            double s = ws[(0 * NQ + i) * kGbMaxK + cc];
            for (int ww = 1; ww < kBlock / kWave; ++ww) s += ws[(ww * NQ + i) * kGbMaxK + cc];
            part[((int64_t)t * stride + i) * k + cc] = s;
        }
    }
    __syncthreads();
}

// Totals over a system's tiles of part[(t * stride + i) * k + c] for i < nq, in a
// fixed order that depends only on (tiles, nq, kpad): pair (i, c) gets J threads
// that add tiles j, j + J, ..., then a fixed tree over j.  Result in
// out[i * kGbMaxK + c] (LDS) for every thread after the call.  sh: kBlock doubles.
__device__ __forceinline__ void system_totals(const double* part, int tiles, int stride, int nq, int k, int kpad,
                                              double* out, double* sh) {
    const int P = nq * kpad;
    int P2 = 1;
    while (P2 < P) P2 <<= 1;
    if (P2 <= kBlock) {
        const int J = kBlock / P2, p = threadIdx.x % P2, j = threadIdx.x / P2;
        const int i = p / kpad, c = p % kpad;
        double s = 0.0;
        if (p < P && c < k)
            for (int t = j; t < tiles; t += J) s += part[((int64_t)t * stride + i) * k + c];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int h = J / 2; h > 0; h >>= 1) {
            if (j < h) sh[threadIdx.x] += sh[threadIdx.x + h * P2];
            __syncthreads();
        }
        if (j == 0 && p < P) out[i * kGbMaxK + c] = sh[p];
    } else {
        for (int p = threadIdx.x; p < P; p += kBlock) {
            const int i = p / kpad, c = p % kpad;
            double s = 0.0;
            if (c < k)
                for (int t = 0; t < tiles; ++t) s += part[((int64_t)t * stride + i) * k + c];
            out[i * kGbMaxK + c] = s;
        }
    }
    __syncthreads();
}

// The totals a consumer kernel needs: for a big system the ones gb_totals_kernel
// published (the same system_totals, same block size: bitwise the same numbers),
// otherwise summed here from the tile partials.
__device__ __forceinline__ void gb_totals(const GbSys& S, const double* part, int stride, int nq, double* out,
                                          double* sh) {
    if (S.big) {
        for (int p = threadIdx.x; p < nq * kGbMaxK; p += kBlock)
            if (p % kGbMaxK < S.k) out[p] = S.tot[p];
        __syncthreads();
    } else {
        system_totals(part, S.tiles, stride, nq, S.k, S.kpad, out, sh);
    }
}

// ---- setup ----------------------------------------------------------------------
__global__ void gb_install_kernel(GbSys S, GbState z, int slot, GbSys* __restrict__ sys, GbState* __restrict__ st) {
    if (threadIdx.x == 0) {
        sys[slot] = S;
        st[slot] = z;
    }
}

// partials of |P b|^2 (grid: the system's tiles)
__global__ __launch_bounds__(kBlock) void gb_bnorm_kernel(int slot, const GbSys* __restrict__ sys) {
    const GbSys S = sys[slot];
    const int t = blockIdx.x, r0 = t * S.tile_rows, r1 = min(S.n, r0 + S.tile_rows);
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    double acc[1] = {0.0};
    if (c < S.k)
        for (int row = r0 + ro; row < r1; row += rows) {
            double v = S.b[(int64_t)row * S.k + c];
            if (S.d) v = S.d[row] * v;
            acc[0] += v * v;
        }
    __shared__ double ws[4 * 1 * kGbMaxK];
    multi_partial<1>(acc, 1, S.k, S.kpad, S.pn, t, 1, ws);
}

// |P b|, TOL, the special cases (GMRES.hpp:168-229)
__global__ __launch_bounds__(kBlock) void gb_setup_kernel(int slot, const GbSys* __restrict__ sys,
                                                          GbState* __restrict__ st, double tol) {
    const GbSys S = sys[slot];
    __shared__ double sh[kBlock];
    __shared__ double tot[kGbMaxK * kGbMaxK];
    system_totals(S.pn, S.tiles, 1, 1, S.k, S.kpad, tot, sh);
    if (threadIdx.x == 0) {
        GbState& s = st[slot];
        double bmax = 0.0;
        int nan = 0;
        for (int c = 0; c < S.k; ++c) {
            const double bn = sqrt(tot[c]);
            s.bnorm[c] = bn;
            s.TOL[c] = bn * tol;
            nan |= isnan(bn) ? 1 : 0;
            bmax = fmax(bmax, bn);
        }
        if (nan) {
            s.nan = 1;
            s.done = 1;
        } else if (!(bmax > 0.0)) {
            s.zero_rhs = 1;
            s.done = 1;
            s.succ = 1;
        }
    }
}

// x = 0 for a zero right-hand side (grid: the system's tiles)
__global__ __launch_bounds__(kBlock) void gb_zero_x_kernel(int slot, const GbSys* __restrict__ sys,
                                                           const GbState* __restrict__ st) {
    if (!st[slot].zero_rhs) return;
    const GbSys S = sys[slot];
    const int64_t e0 = (int64_t)blockIdx.x * S.tile_rows * S.k;
    const int64_t e1 = min((int64_t)S.n * S.k, e0 + (int64_t)S.tile_rows * S.k);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kBlock) S.x[e] = 0.0;
}

// ---- one restart cycle ------------------------------------------------------------
#define GB_TILE_PROLOGUE                         \
    const GbTile T = tiles[blockIdx.x];          \
    const GbSys S = sys[T.slot];                 \
    const GbState* __restrict__ s = st + T.slot; \
    const int k = S.k, kpad = S.kpad;            \
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;

// z = A x - b (or -b), Q_0 = P z, partials of |Q_0|^2
__global__ __launch_bounds__(kBlock) void gb_start1_kernel(const GbTile* __restrict__ tiles,
                                                           const GbSys* __restrict__ sys,
                                                           const GbState* __restrict__ st) {
    GB_TILE_PROLOGUE
    if (s->done) return;
    const bool zero = s->zero_cycle == s->cycles;
    double acc[1] = {0.0};
    if (c < k)
        for (int row = T.r0 + ro; row < T.r1; row += rows) {
            const int64_t e = (int64_t)row * k + c;
            double z;
            if (zero) {
                z = -S.b[e];
            } else {
                double a = 0.0;
                for (int32_t l = S.rp[row]; l < S.rp[row + 1]; ++l) a += S.val[l] * S.x[(int64_t)S.ci[l] * k + c];
                z = a + (-S.b[e]);
            }
            if (S.d) z = S.d[row] * z;
            S.Q[e] = z;
            acc[0] += z * z;
        }
    __shared__ double ws[4 * 1 * kGbMaxK];
    multi_partial<1>(acc, 1, k, kpad, S.pn, T.t, 1, ws);
}

// Q_0 *= 1 / |Q_0|; tile 0: beta, H = 0, stop test before any step (GMRES.hpp:353-375)
__global__ __launch_bounds__(kBlock) void gb_start2_kernel(const GbTile* __restrict__ tiles,
                                                           const GbSys* __restrict__ sys, GbState* __restrict__ st) {
    GB_TILE_PROLOGUE
    if (s->done) return;
    __shared__ double sh[kBlock];
    __shared__ double tot[kGbMaxK * kGbMaxK];
    gb_totals(S, S.pn, 1, 1, tot, sh);
    if (c < k) {
        const double inv = 1.0 / sqrt(tot[c]);
        for (int row = T.r0 + ro; row < T.r1; row += rows) S.Q[(int64_t)row * k + c] *= inv;
    }
    if (T.t == 0) {
        const int m = S.m;
        for (int e = threadIdx.x; e < (m + 1) * m * k; e += kBlock) S.H[e] = 0.0;
        for (int e = threadIdx.x; e < (m + 1) * k; e += kBlock) S.beta[e] = e < k ? sqrt(tot[e]) : 0.0;
        if (threadIdx.x == 0) {
            GbState& w = st[T.slot];
            int conv = 1;
            for (int cc = 0; cc < k; ++cc) conv &= (fabs(sqrt(tot[cc])) <= w.TOL[cc]) ? 1 : 0;
            w.conv = conv;
            w.iters = 0;
            w.active = conv ? 0 : 1;
        }
    }
}

// q = P A Q_j on the tile's rows, partials of Q_i.q (i <= j)
template <int NQ>
__global__ __launch_bounds__(kBlock) void gb_step1_kernel(const GbTile* __restrict__ tiles,
                                                          const GbSys* __restrict__ sys,
                                                          const GbState* __restrict__ st, int j) {
    GB_TILE_PROLOGUE
    if (s->done || !s->active || j >= S.m) return;
    const int nq = j + 1;
    const int64_t N = (int64_t)S.n * k;
    const double* __restrict__ Qj = S.Q + j * N;
    double* __restrict__ q = S.Q + (int64_t)(j + 1) * N;
    double acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = 0.0;
    if (c < k)
        for (int row = T.r0 + ro; row < T.r1; row += rows) {
            const int64_t e = (int64_t)row * k + c;
            double w = 0.0;
            for (int32_t l = S.rp[row]; l < S.rp[row + 1]; ++l) w += S.val[l] * Qj[(int64_t)S.ci[l] * k + c];
            if (S.d) w = S.d[row] * w;
            q[e] = w;
            double qv[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i) qv[i] = S.Q[(int64_t)min(i, nq - 1) * N + e];
#pragma unroll
            for (int i = 0; i < NQ; ++i) acc[i] += (i < nq) ? qv[i] * w : 0.0;
        }
    __shared__ double ws[4 * NQ * kGbMaxK];
    multi_partial<NQ>(acc, nq, k, kpad, S.p1, T.t, S.m + 1, ws);
}

// PASS 0: h1 from the first partials, q -= sum h1_i Q_i, partials of Q_i.q, tile 0
// writes H[.][j] = h1.  PASS 1: h2 from the second partials, q -= sum h2_i Q_i,
// partials of |q|^2, tile 0 adds h2 into H.  (H starts the cycle at zero, so the
// two passes add into H as GMRES.hpp:555-585 does.)
template <int NQ, int PASS>
__global__ __launch_bounds__(kBlock) void gb_step23_kernel(const GbTile* __restrict__ tiles,
                                                           const GbSys* __restrict__ sys,
                                                           const GbState* __restrict__ st, int j) {
    GB_TILE_PROLOGUE
    if (s->done || !s->active || j >= S.m) return;
    const int nq = j + 1;
    const int64_t N = (int64_t)S.n * k;
    double* __restrict__ q = S.Q + (int64_t)(j + 1) * N;
    __shared__ double sh[kBlock];
    __shared__ double h[kGbMaxM + 1][kGbMaxK];
    gb_totals(S, PASS == 0 ? S.p1 : S.p2, S.m + 1, nq, &h[0][0], sh);
    if (T.t == 0)
        for (int p = threadIdx.x; p < nq * k; p += kBlock) {
            const int i = p / k, cc = p % k;
            double& Hij = S.H[((int64_t)i * S.m + j) * k + cc];
            Hij = PASS == 0 ? 0.0 + h[i][cc] : Hij + h[i][cc];
        }
    double acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = 0.0;
    double accn = 0.0;
    if (c < k)
        for (int row = T.r0 + ro; row < T.r1; row += rows) {
            const int64_t e = (int64_t)row * k + c;
            double qv[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i) qv[i] = S.Q[(int64_t)min(i, nq - 1) * N + e];
            double w = q[e];
#pragma unroll
            for (int i = 0; i < NQ; ++i) w -= (i < nq) ? qv[i] * h[i][c] : 0.0;
            q[e] = w;
            if (PASS == 0) {
#pragma unroll
                for (int i = 0; i < NQ; ++i) acc[i] += (i < nq) ? qv[i] * w : 0.0;
            } else {
                accn += w * w;
            }
        }
    if (PASS == 0) {
        __shared__ double ws[4 * NQ * kGbMaxK];
        multi_partial<NQ>(acc, nq, k, kpad, S.p2, T.t, S.m + 1, ws);
    } else {
        double a1[1] = {accn};
        __shared__ double ws1[4 * 1 * kGbMaxK];
        multi_partial<1>(a1, 1, k, kpad, S.pn, T.t, 1, ws1);
    }
}

// q *= 1 / |q|; tile 0: H[j+1][j] = |q|, the Givens rotations (GMRES.hpp:603-658),
// beta, the stop test, the end of the cycle at convergence or at j + 1 == m
__global__ __launch_bounds__(kBlock) void gb_step4_kernel(const GbTile* __restrict__ tiles,
                                                          const GbSys* __restrict__ sys, GbState* __restrict__ st,
                                                          int j) {
    GB_TILE_PROLOGUE
    if (s->done || !s->active || j >= S.m) return;
    const int64_t N = (int64_t)S.n * k;
    double* __restrict__ q = S.Q + (int64_t)(j + 1) * N;
    __shared__ double sh[kBlock];
    __shared__ double tot[kGbMaxK * kGbMaxK];
    gb_totals(S, S.pn, 1, 1, tot, sh);
    if (c < k) {
        const double inv = 1.0 / sqrt(tot[c]);
        for (int row = T.r0 + ro; row < T.r1; row += rows) q[(int64_t)row * k + c] *= inv;
    }
    if (T.t != 0) return;  // block-uniform
    const int m = S.m;
    if ((int)threadIdx.x < k) {
        const int cc = threadIdx.x;
        auto Hx = [&](int r) -> double& { return S.H[((int64_t)r * m + j) * k + cc]; };
        Hx(j + 1) = sqrt(tot[cc]);
        for (int i = 0; i < j; ++i) {
            const double xi = Hx(i), eta = Hx(i + 1);
            const double co = S.cs[i * k + cc], si = S.sn[i * k + cc];
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
        S.cs[j * k + cc] = co;
        S.sn[j * k + cc] = si;
        Hx(j) = co * xi + si * eta;
        Hx(j + 1) = 0.0;
        const double bj = S.beta[j * k + cc];
        S.beta[(j + 1) * k + cc] = -si * bj;
        S.beta[j * k + cc] = co * bj;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        GbState& w = st[T.slot];
        int conv = 1;
        for (int cc = 0; cc < k; ++cc) conv &= (fabs(S.beta[(j + 1) * k + cc]) <= w.TOL[cc]) ? 1 : 0;
        w.iters = j + 1;
        w.conv = conv;
        if (conv || j + 1 >= m) w.active = 0;
    }
}

// ---- delayed second pass ----------------------------------------------------------
// Tile partials of the 2j + 2 values [a_0..a_{j-1}, b_0..b_{j-1}, u.u, u.v] held in
// acc[i] (a_i), acc[NQ + i] (b_i), acc[2 NQ], acc[2 NQ + 1]; the same fixed order as
// multi_partial.  ws: 4 * (2 NQ + 2) * kGbMaxK doubles of LDS.
template <int NQ>
__device__ __forceinline__ void dd_partial(double (&acc)[2 * NQ + 2], int j, int k, int kpad, double* part, int t,
                                           int stride, double* ws) {
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave, c = threadIdx.x % kpad;
    constexpr int R = 2 * NQ + 2;
    const ScatterOut o = wave_scatter_sum<R>(acc, kpad);  // kpad <= kGbMaxK = 8: C values per lane
    if ((lane & 7) < kpad)
#pragma unroll
        for (int i = 0; i < ScatterSizes<R>::C; ++i)
            if (i < o.count) ws[(w * R + o.base + i) * kGbMaxK + c] = acc[i];
    __syncthreads();
    const int nv = 2 * j + 2;
    for (int p = threadIdx.x; p < nv * kpad; p += kBlock) {
        const int i = p / kpad, cc = p % kpad;
        const int r = i < j ? i : i < 2 * j ? NQ + (i - j) : 2 * NQ + (i - 2 * j);
        if (cc < k) {
            double v = ws[(0 * (2 * NQ + 2) + r) * kGbMaxK + cc];
            for (int ww = 1; ww < kBlock / kWave; ++ww) v += ws[(ww * (2 * NQ + 2) + r) * kGbMaxK + cc];
            part[((int64_t)t * stride + i) * k + cc] = v;
        }
    }
    __syncthreads();
}

// v = P A u on the tile's rows (u = Q_j) into Q_{j+1}, partials of the 2j + 2 dots.
// On a system's finishing step (j == m) there is no product: v = u, nothing written.
template <int NQ>
__global__ __launch_bounds__(kBlock, 2) void gb_dstep1_kernel(const GbTile* __restrict__ tiles,
                                                           const GbSys* __restrict__ sys,
                                                           const GbState* __restrict__ st, int j) {
    GB_TILE_PROLOGUE
    if (s->done || !s->active || j > S.m) return;
    const bool fin = j == S.m;
    const int64_t N = (int64_t)S.n * k;
    const double* __restrict__ U = S.Q + j * N;
    double* __restrict__ V = fin ? nullptr : S.Q + (int64_t)(j + 1) * N;
    double acc[2 * NQ + 2];
#pragma unroll
    for (int r = 0; r < 2 * NQ + 2; ++r) acc[r] = 0.0;
    if (c < k)
        for (int row = T.r0 + ro; row < T.r1; row += rows) {
            const int64_t e = (int64_t)row * k + c;
            double qv[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i) qv[i] = S.Q[(int64_t)max(0, min(i, j - 1)) * N + e];
            const double uv = U[e];
            double vv = uv;
            if (!fin) {
                double w = 0.0;
                for (int32_t l = S.rp[row]; l < S.rp[row + 1]; ++l) w += S.val[l] * U[(int64_t)S.ci[l] * k + c];
                if (S.d) w = S.d[row] * w;
                V[e] = w;
                vv = w;
            }
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                if (i < j) {
                    acc[i] += qv[i] * uv;
                    acc[NQ + i] += qv[i] * vv;
                }
            acc[2 * NQ] += uv * uv;
            acc[2 * NQ + 1] += uv * vv;
        }
    __shared__ double ws[4 * (2 * NQ + 2) * kGbMaxK];
    dd_partial<NQ>(acc, j, k, kpad, S.p1, T.t, 2 * S.m + 2, ws);
}

// The scalars of step j from the totals (the same in every workgroup of the system),
// tile 0's writes, then q_j and the next vector on the tile's rows.
template <int NQ>
__global__ __launch_bounds__(kBlock) void gb_dstep2_kernel(const GbTile* __restrict__ tiles,
                                                           const GbSys* __restrict__ sys, GbState* __restrict__ st,
                                                           int j) {
    GB_TILE_PROLOGUE
    if (s->done || !s->active || j > S.m) return;
    const int m = S.m, nv = 2 * j + 2, jj = j - 1;
    const bool fin = j == m;
    const int64_t N = (int64_t)S.n * k;
    __shared__ double sh[kBlock];
    __shared__ double tv[(2 * kGbMaxM + 2) * kGbMaxK];  // totals [value][column]
    __shared__ double colc[(kGbMaxM + 1) * kGbMaxK];    // column j - 1 completed, unrotated
    __shared__ double colr[(kGbMaxM + 1) * kGbMaxK];    // ... rotated
    __shared__ double dA[kGbMaxM * kGbMaxK], dE[kGbMaxM * kGbMaxK];
    __shared__ double rho_s[kGbMaxK], irho_s[kGbMaxK], ab_s[kGbMaxK], cj_s[kGbMaxK];
    __shared__ double rcs[kGbMaxM * kGbMaxK], rsn[kGbMaxM * kGbMaxK];  // rotations of earlier columns
    __shared__ int conv_s;
    for (int p = threadIdx.x; p < jj * k; p += kBlock) {
        const int i = p / k, cc = p % k;
        rcs[i * kGbMaxK + cc] = S.cs[p];
        rsn[i * kGbMaxK + cc] = S.sn[p];
    }
    gb_totals(S, S.p1, 2 * m + 2, nv, tv, sh);
    auto a = [&](int i, int cc) { return tv[i * kGbMaxK + cc]; };
    auto b = [&](int i, int cc) { return tv[(j + i) * kGbMaxK + cc]; };
    if (threadIdx.x == 0) conv_s = 1;
    if ((int)threadIdx.x < k) {
        const int cc = threadIdx.x;
        double rho = 1.0, ab = 0.0;
        if (j > 0) {
            const double alpha = tv[2 * j * kGbMaxK + cc];
            double aa = 0.0;
            for (int i = 0; i < j; ++i) {
                aa += a(i, cc) * a(i, cc);
                ab += a(i, cc) * b(i, cc);
            }
            // |u - Q a|; at a breakdown (u in the span of Q to rounding) at least 2^-52 |u|
            rho = sqrt(fmax(alpha - aa, 0x1p-104 * alpha));
            for (int l = 0; l < j; ++l) colc[l * kGbMaxK + cc] = S.G[((int64_t)l * m + jj) * k + cc] + a(l, cc);
            colc[j * kGbMaxK + cc] = rho;
        }
        rho_s[cc] = rho;
        irho_s[cc] = 1.0 / rho;
        ab_s[cc] = ab;
    }
    __syncthreads();
    if (j > 0 && (int)threadIdx.x < k) {
        // Givens on column j - 1 (GMRES.hpp:603-658) with the rotations of earlier launches
        const int cc = threadIdx.x;
        for (int l = 0; l <= j; ++l) colr[l * kGbMaxK + cc] = colc[l * kGbMaxK + cc];
        for (int i = 0; i < jj; ++i) {
            const double xi = colr[i * kGbMaxK + cc], eta = colr[(i + 1) * kGbMaxK + cc];
            const double co = rcs[i * kGbMaxK + cc], si = rsn[i * kGbMaxK + cc];
            colr[i * kGbMaxK + cc] = co * xi + si * eta;
            colr[(i + 1) * kGbMaxK + cc] = -si * xi + co * eta;
        }
        const double xi = colr[jj * kGbMaxK + cc], eta = colr[j * kGbMaxK + cc];
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
        colr[jj * kGbMaxK + cc] = co * xi + si * eta;
        colr[j * kGbMaxK + cc] = 0.0;
        const double carry = -si * S.beta[jj * k + cc];  // beta holds the carried right-hand side
        if (!(fabs(carry) <= s->TOL[cc])) atomicAnd(&conv_s, 0);
        if (T.t == 0) {
            S.cs[jj * k + cc] = co;
            S.sn[jj * k + cc] = si;
            S.beta[j * k + cc] = carry;
        }
    }
    __syncthreads();
    const bool stop = fin || (j > 0 && conv_s);
    if (T.t == 0 && j > 0) {
        for (int p = threadIdx.x; p < (j + 1) * k; p += kBlock) {
            const int l = p / k, cc = p % k;
            S.Hc[((int64_t)l * m + jj) * k + cc] = colc[l * kGbMaxK + cc];
            S.H[((int64_t)l * m + jj) * k + cc] = colr[l * kGbMaxK + cc];
        }
        if (threadIdx.x == 0) {
            GbState& w = st[T.slot];
            w.iters = j;
            w.conv = conv_s;
            if (stop) w.active = 0;
        }
    }
    if (stop) return;  // block-uniform; the solution update uses q_0 .. q_{j-1}
    // column j, first pass: g_l = (t_l - (Hc a)_l) / rho, t_l = b_l (l < j),
    // t_j = (u.v - a.b) / rho; Hc[l][i] = 0 for l > i + 1, column j - 1 from colc
    for (int p = threadIdx.x; p < (j + 1) * k; p += kBlock) {
        const int l = p / k, cc = p % k;
        const double irho = irho_s[cc];
        double ha = 0.0;
#pragma unroll 8
        for (int i = l > 0 ? l - 1 : 0; i < jj; ++i) ha += S.Hc[((int64_t)l * m + i) * k + cc] * a(i, cc);
        if (j > 0) ha += colc[l * kGbMaxK + cc] * a(jj, cc);
        const double uv = tv[(2 * j + 1) * kGbMaxK + cc];
        const double tl = l < j ? b(l, cc) : (j > 0 ? (uv - ab_s[cc]) * irho : uv);
        if (T.t == 0) S.G[((int64_t)l * m + j) * k + cc] = (tl - ha) * irho;
        if (l == j) cj_s[cc] = tl * irho * irho;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < j * k; p += kBlock) {
        const int i = p / k, cc = p % k;
        dA[i * kGbMaxK + cc] = a(i, cc) * irho_s[cc];
        dE[i * kGbMaxK + cc] = b(i, cc) * irho_s[cc] - cj_s[cc] * a(i, cc);
    }
    __syncthreads();
    if (c >= k) return;
    double ca[NQ], ce[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        ca[i] = i < j ? dA[i * kGbMaxK + c] : 0.0;
        ce[i] = i < j ? dE[i * kGbMaxK + c] : 0.0;
    }
    const double ir = irho_s[c], cj = cj_s[c];
    double* __restrict__ U = S.Q + j * N;
    double* __restrict__ V = S.Q + (int64_t)(j + 1) * N;
    for (int row = T.r0 + ro; row < T.r1; row += rows) {
        const int64_t e = (int64_t)row * k + c;
        double qv[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) qv[i] = S.Q[(int64_t)max(0, min(i, j - 1)) * N + e];
        const double uv = U[e], vv = V[e];
        double qj = uv * ir, nx = vv * ir - cj * uv;
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            if (i < j) {
                qj -= ca[i] * qv[i];
                nx -= ce[i] * qv[i];
            }
        if (j > 0) U[e] = qj;
        V[e] = nx;
    }
}

// Totals of one reduction of each big system, one workgroup per system (grid: the
// big active systems): phase 0 after start1 (|Q_0|^2), 1 after step1 (first-pass
// dots), 2 after the first step23 (second-pass dots), 3 after the second (|q|^2),
// 4 after dstep1 (the delayed form's 2j + 2 dots).
// Each big system's tile workgroups then read 1 + nq values instead of adding all
// tiles' partials each (tiles^2 reads per reduction).
__global__ __launch_bounds__(kBlock) void gb_totals_kernel(const int* __restrict__ bigs,
                                                           const GbSys* __restrict__ sys,
                                                           const GbState* __restrict__ st, int phase, int j) {
    const int slot = bigs[blockIdx.x];
    const GbSys S = sys[slot];
    const GbState* __restrict__ s = st + slot;
    if (s->done) return;
    if (phase > 0 && (!s->active || (phase == 4 ? j > S.m : j >= S.m))) return;
    const int nq = phase == 4 ? 2 * j + 2 : (phase == 1 || phase == 2) ? j + 1 : 1;
    const double* part = phase == 0 || phase == 3 ? S.pn : phase == 1 || phase == 4 ? S.p1 : S.p2;
    const int stride = phase == 4 ? 2 * S.m + 2 : (phase == 1 || phase == 2) ? S.m + 1 : 1;
    __shared__ double sh[kBlock];
    __shared__ double out[(2 * kGbMaxM + 2) * kGbMaxK];
    system_totals(part, S.tiles, stride, nq, S.k, S.kpad, out, sh);
    for (int p = threadIdx.x; p < nq * kGbMaxK; p += kBlock)
        if (p % kGbMaxK < S.k) S.tot[p] = out[p];
}

// tile 0: y = H^{-1} beta over the cycle's steps (BLAS::trsv, src/BLAS/trsv.hpp:7-48)
__global__ __launch_bounds__(kBlock) void gb_lsq_kernel(const GbTile* __restrict__ tiles,
                                                        const GbSys* __restrict__ sys,
                                                        const GbState* __restrict__ st) {
    const GbTile T = tiles[blockIdx.x];
    if (T.t != 0) return;
    const GbSys S = sys[T.slot];
    const GbState* s = st + T.slot;
    if (s->done || s->iters == 0) return;
    const int it = s->iters, m = S.m, k = S.k;
    // the triangle of H and the right-hand side staged in LDS (chunks of columns), then
    // the serial substitution per column from LDS; plain row-oriented order
    constexpr int kTri = 6144;
    __shared__ double tri[kTri];
    __shared__ double sb[(kGbMaxM + 1) * kGbMaxK];
    const int ntri = it * (it + 1) / 2;
    const int kc = max(1, min(k, kTri / ntri));
    auto tix = [&](int i, int j) { return i * it - i * (i - 1) / 2 + (j - i); };  // i <= j
    for (int c0 = 0; c0 < k; c0 += kc) {
        const int nc = min(kc, k - c0);
        for (int t = threadIdx.x; t < it * it * nc; t += blockDim.x) {
            const int cc = t % nc, ij = t / nc, i = ij / it, j = ij % it;
            if (j >= i) tri[tix(i, j) * nc + cc] = S.H[((int64_t)i * m + j) * k + c0 + cc];
        }
        for (int t = threadIdx.x; t < it * nc; t += blockDim.x) {
            const int i = t / nc, c = c0 + t % nc;
            // the delayed form keeps beta as the carried right-hand side: rotate it here
            sb[i * kGbMaxK + t % nc] = S.delayed ? S.cs[i * k + c] * S.beta[i * k + c] : S.beta[i * k + c];
        }
        __syncthreads();
        for (int cc = threadIdx.x; cc < nc; cc += blockDim.x)
            for (int i = it - 1; i >= 0; --i) {
                double v = sb[i * kGbMaxK + cc];
                for (int jj = i + 1; jj < it; ++jj) v -= tri[tix(i, jj) * nc + cc] * sb[jj * kGbMaxK + cc];
                sb[i * kGbMaxK + cc] = v / tri[tix(i, i) * nc + cc];
            }
        __syncthreads();
        for (int t = threadIdx.x; t < it * nc; t += blockDim.x) S.y[(t / nc) * k + c0 + t % nc] = sb[(t / nc) * kGbMaxK + t % nc];
        __syncthreads();
    }
}

// x -= Q y on the tile's rows (x = -Q y on a cycle that started from x = 0)
__global__ __launch_bounds__(kBlock) void gb_update_kernel(const GbTile* __restrict__ tiles,
                                                           const GbSys* __restrict__ sys,
                                                           const GbState* __restrict__ st) {
    GB_TILE_PROLOGUE
    if (s->done || s->iters == 0) return;  // converged before any step: x untouched (:371-375)
    const int it = s->iters;
    const bool zero = s->zero_cycle == s->cycles;
    const int64_t N = (int64_t)S.n * k;
    if (c < k)
        for (int row = T.r0 + ro; row < T.r1; row += rows) {
            const int64_t e = (int64_t)row * k + c;
            double z = 0.0;
            for (int jj = 0; jj < it; ++jj) z += S.y[jj * k + c] * S.Q[jj * N + e];
            S.x[e] = zero ? -z : S.x[e] - z;
        }
}

// tile 0: the cycle's bookkeeping (GMRES.hpp:240-250)
__global__ void gb_end_kernel(const GbTile* __restrict__ tiles, const GbSys* __restrict__ sys,
                              GbState* __restrict__ st) {
    const GbTile T = tiles[blockIdx.x];
    if (T.t != 0 || threadIdx.x != 0) return;
    GbState& s = st[T.slot];
    if (s.done) return;
    const GbSys S = sys[T.slot];
    s.cycles += 1;
    for (int c = 0; c < S.k; ++c) s.rel[c] = fabs(S.beta[s.iters * S.k + c]) / s.bnorm[c];
    if (s.conv) {
        s.done = 1;
        s.succ = 1;
    } else if (s.cycles >= s.max_cycles) {
        s.done = 1;
    }
}

#undef GB_TILE_PROLOGUE

static int gb_ceil_div(int a, int b) { return (a + b - 1) / b; }
static int gb_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

size_t gb_system_bytes(int n, int k, int m, int tile_rows) {
    const int tiles = gb_ceil_div(n, tile_rows);
    const size_t N = (size_t)n * k;
    return align256((size_t)(m + 1) * N * 8) + align256((size_t)(m + 1) * m * k * 8) + 2 * align256((size_t)m * k * 8) +
           align256((size_t)(m + 1) * k * 8) + align256((size_t)m * k * 8) +
           align256((size_t)tiles * (2 * m + 2) * k * 8) + align256((size_t)tiles * (m + 1) * k * 8) +
           align256((size_t)tiles * k * 8) + align256((size_t)(2 * kGbMaxM + 2) * kGbMaxK * 8) +
           2 * align256((size_t)(m + 1) * m * k * 8);
}

// Systems of more than this many tiles get their totals from gb_totals_kernel
// (tg_tune_set "gbatch_big_tiles", read when a system is added); bitwise the same.
static int g_gb_big_tiles = 64;
// 1: the delayed second Gram-Schmidt pass (two launches per step); read when an
// engine is created (tg_tune_set "gbatch_delayed")
static int g_gb_delayed = 1;

int gbatch_tune_set(const char* key, int32_t value, bool* known) {
    const std::string kk = key;
    *known = kk == "gbatch_big_tiles" || kk == "gbatch_delayed";
    if (!*known) return TG_OK;
    if (kk == "gbatch_delayed") {
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: gbatch_delayed must be 0 or 1");
        g_gb_delayed = value;
        return TG_OK;
    }
    if (value < 0) return fail(TG_ERR_INVALID, "tg_tune_set: gbatch_big_tiles must be >= 0");
    g_gb_big_tiles = value;
    return TG_OK;
}

int32_t gbatch_tune_get(const char* key) {
    const std::string kk = key;
    return kk == "gbatch_big_tiles" ? g_gb_big_tiles : kk == "gbatch_delayed" ? g_gb_delayed : -1;
}

class GmresBatch {
  public:
    explicit GmresBatch(hipStream_t s) : delayed_(g_gb_delayed), s_(s) {}
    ~GmresBatch() { release(); }

    int add(int n, int64_t nnz, const int32_t* rp, const int32_t* ci, const double* val, const double* d, int k,
            const double* b, double* x, double tol, int restart, int max_restarts, int* slot_out) {
        if (n < 1 || k < 1 || k > kGbMaxK) return fail(TG_ERR_INVALID, "gmres batch: n >= 1 and 1 <= nrhs <= 8");
        if (!rp || (nnz > 0 && (!ci || !val)) || !b || !x) return fail(TG_ERR_INVALID, "gmres batch: null array");
        const int m = std::min(restart, n + 1);  // GMRES.hpp:89
        if (m < 1 || m > kGbMaxM) return fail(TG_ERR_INVALID, "gmres batch: restart length must be in [1, %d]", kGbMaxM);
        if (nnz >= (1ll << 31)) return fail(TG_ERR_INVALID, "gmres batch: nnz must fit int32 row pointers");
        int slot = -1;
        for (int i = 0; i < (int)used_.size(); ++i)
            if (!used_[i]) {
                slot = i;
                break;
            }
        if (slot < 0) {
            slot = (int)used_.size();
            int rc = grow(slot + 1);
            if (rc) return rc;
        }
        GbSys S{};
        S.rp = rp;
        S.ci = ci;
        S.val = val;
        S.d = d;
        S.b = b;
        S.x = x;
        S.n = n;
        S.k = k;
        S.kpad = gb_pow2(k);
        S.m = m;
        S.tile_rows = kGbTileRows;
        S.tiles = gb_ceil_div(n, S.tile_rows);
        const size_t N = (size_t)n * k;
        const size_t bytes = gb_system_bytes(n, k, m, S.tile_rows);
        char* mem = nullptr;
        TG_HIP(hipMallocAsync((void**)&mem, bytes, s_));
        size_t off = 0;
        auto take = [&](size_t b_) {
            char* p = mem + off;
            off += align256(b_);
            return (double*)p;
        };
        S.Q = take((size_t)(m + 1) * N * 8);
        S.H = take((size_t)(m + 1) * m * k * 8);
        S.cs = take((size_t)m * k * 8);
        S.sn = take((size_t)m * k * 8);
        S.beta = take((size_t)(m + 1) * k * 8);
        S.y = take((size_t)m * k * 8);
        S.p1 = take((size_t)S.tiles * (2 * m + 2) * k * 8);
        S.p2 = take((size_t)S.tiles * (m + 1) * k * 8);
        S.pn = take((size_t)S.tiles * k * 8);
        S.tot = take((size_t)(2 * kGbMaxM + 2) * kGbMaxK * 8);
        S.G = take((size_t)(m + 1) * m * k * 8);
        S.Hc = take((size_t)(m + 1) * m * k * 8);
        S.big = S.tiles > g_gb_big_tiles ? 1 : 0;
        S.delayed = delayed_;
        mem_[slot] = mem;
        bytes_[slot] = bytes;
        used_[slot] = 1;
        host_[slot] = S;
        GbState z{};
        z.max_cycles = std::max(0, max_restarts) + 1;
        z.zero_cycle = 0;  // no initial guess: the first cycle starts from x = 0
        // setup without host waits: descriptor and state by kernel arguments, |P b|,
        // the special cases; the system joins at the next cycle
        hipLaunchKernelGGL(gb_install_kernel, dim3(1), dim3(kWave), 0, s_, S, z, slot, dsys_, dst_);
        hipLaunchKernelGGL(gb_bnorm_kernel, dim3((unsigned)S.tiles), dim3(kBlock), 0, s_, slot, (const GbSys*)dsys_);
        hipLaunchKernelGGL(gb_setup_kernel, dim3(1), dim3(kBlock), 0, s_, slot, (const GbSys*)dsys_, dst_, tol);
        hipLaunchKernelGGL(gb_zero_x_kernel, dim3((unsigned)S.tiles), dim3(kBlock), 0, s_, slot, (const GbSys*)dsys_,
                           (const GbState*)dst_);
        TG_LAUNCH_CHECK();
        dirty_ = true;
        *slot_out = slot;
        return TG_OK;
    }

    // Runs `cycles` restart cycles of every active system.
    int step(int cycles) {
        if (dirty_) {
            int rc = rebuild();
            if (rc) return rc;
        }
        if (ntiles_ == 0) return TG_OK;
        const GraphKey key{dtiles_, dsys_, dst_, dbig_, ntiles_, nbig_, mmax_, cycles};
        if (!gexec_ || !(key == gkey_)) {
            if (gexec_) (void)hipGraphExecDestroy(gexec_);
            gexec_ = nullptr;
            if (!cap_s_) TG_HIP(hipStreamCreateWithFlags(&cap_s_, hipStreamNonBlocking));
            TG_HIP(hipStreamBeginCapture(cap_s_, hipStreamCaptureModeThreadLocal));
            for (int cy = 0; cy < cycles; ++cy) launch_cycle(cap_s_);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(cap_s_, &g);
            if (ec != hipSuccess) return fail(TG_ERR_HIP, "gmres batch: graph capture failed: %s", hipGetErrorString(ec));
            const hipError_t ei = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                gexec_ = nullptr;
                return fail(TG_ERR_HIP, "gmres batch: graph instantiation failed: %s", hipGetErrorString(ei));
            }
            gkey_ = key;
        }
        TG_HIP(hipGraphLaunch(gexec_, s_));
        return TG_OK;
    }

    // Reads the states of all used slots (synchronises).
    int poll(std::vector<GbState>& out) {
        out.resize(used_.size());
        if (!used_.empty())
            TG_HIP(hipMemcpyAsync(out.data(), dst_, used_.size() * sizeof(GbState), hipMemcpyDeviceToHost, s_));
        TG_HIP(hipStreamSynchronize(s_));
        return TG_OK;
    }

    int retire(int slot) {
        if (slot < 0 || slot >= (int)used_.size() || !used_[slot]) return fail(TG_ERR_INVALID, "gmres batch: bad slot");
        TG_HIP(hipFreeAsync(mem_[slot], s_));
        mem_[slot] = nullptr;
        bytes_[slot] = 0;
        used_[slot] = 0;
        dirty_ = true;
        return TG_OK;
    }

    // A finished system's tiles leave the list at the next step.
    void finished(int) { dirty_ = true; }

    size_t bytes_in_use() const {
        size_t s = 0;
        for (size_t b : bytes_) s += b;
        return s;
    }
    hipStream_t stream() const { return s_; }

  private:
    void launch_cycle(hipStream_t s) {
        const dim3 g((unsigned)ntiles_), b(kBlock);
        // totals of the big systems after each reduction (no launch without big systems)
        auto totals = [&](int phase, int j) {
            if (nbig_ > 0)
                hipLaunchKernelGGL(gb_totals_kernel, dim3((unsigned)nbig_), b, 0, s, (const int*)dbig_, dsys_,
                                   (const GbState*)dst_, phase, j);
        };
        hipLaunchKernelGGL(gb_start1_kernel, g, b, 0, s, dtiles_, dsys_, dst_);
        totals(0, 0);
        hipLaunchKernelGGL(gb_start2_kernel, g, b, 0, s, dtiles_, dsys_, dst_);
        // delayed form: steps 0 .. m (m: the finishing step), two launches each
        for (int j = 0; delayed_ && j <= mmax_; ++j) {
#define TG_GB_DSTEPS(NQ)                                                                     \
    hipLaunchKernelGGL((gb_dstep1_kernel<NQ>), g, b, 0, s, dtiles_, dsys_, dst_, j);       \
    totals(4, j);                                                                          \
    hipLaunchKernelGGL((gb_dstep2_kernel<NQ>), g, b, 0, s, dtiles_, dsys_, dst_, j);
            if (j <= 4) {
                TG_GB_DSTEPS(4)
            } else if (j <= 8) {
                TG_GB_DSTEPS(8)
            } else if (j <= 16) {
                TG_GB_DSTEPS(16)
            } else if (j <= 32) {
                TG_GB_DSTEPS(32)
            } else {
                TG_GB_DSTEPS(64)
            }
#undef TG_GB_DSTEPS
        }
        for (int j = 0; !delayed_ && j < mmax_; ++j) {
            const int nq = j + 1;
#define TG_GB_STEPS(NQ)                                                                                 \
    hipLaunchKernelGGL((gb_step1_kernel<NQ>), g, b, 0, s, dtiles_, dsys_, dst_, j);                     \
    totals(1, j);                                                                                       \
    hipLaunchKernelGGL((gb_step23_kernel<NQ, 0>), g, b, 0, s, dtiles_, dsys_, dst_, j);                 \
    totals(2, j);                                                                                       \
    hipLaunchKernelGGL((gb_step23_kernel<NQ, 1>), g, b, 0, s, dtiles_, dsys_, dst_, j);                 \
    totals(3, j);
            if (nq <= 4) {
                TG_GB_STEPS(4)
            } else if (nq <= 8) {
                TG_GB_STEPS(8)
            } else if (nq <= 16) {
                TG_GB_STEPS(16)
            } else if (nq <= 32) {
                TG_GB_STEPS(32)
            } else {
                TG_GB_STEPS(64)
            }
#undef TG_GB_STEPS
            hipLaunchKernelGGL(gb_step4_kernel, g, b, 0, s, dtiles_, dsys_, dst_, j);
        }
        hipLaunchKernelGGL(gb_lsq_kernel, g, b, 0, s, dtiles_, dsys_, dst_);
        hipLaunchKernelGGL(gb_update_kernel, g, b, 0, s, dtiles_, dsys_, dst_);
        hipLaunchKernelGGL(gb_end_kernel, g, dim3(kWave), 0, s, dtiles_, dsys_, dst_);
    }

    int grow(int need) {
        int cap = std::max(need, 2 * (int)cap_);
        if (cap <= (int)cap_) return TG_OK;
        GbSys* nsys = nullptr;
        GbState* nst = nullptr;
        TG_HIP(hipMallocAsync((void**)&nsys, cap * sizeof(GbSys), s_));
        TG_HIP(hipMallocAsync((void**)&nst, cap * sizeof(GbState), s_));
        if (cap_) {
            TG_HIP(hipMemcpyAsync(nsys, dsys_, cap_ * sizeof(GbSys), hipMemcpyDeviceToDevice, s_));
            TG_HIP(hipMemcpyAsync(nst, dst_, cap_ * sizeof(GbState), hipMemcpyDeviceToDevice, s_));
            TG_HIP(hipFreeAsync(dsys_, s_));
            TG_HIP(hipFreeAsync(dst_, s_));
        }
        dsys_ = nsys;
        dst_ = nst;
        cap_ = cap;
        used_.resize(cap, 0);
        mem_.resize(cap, nullptr);
        bytes_.resize(cap, 0);
        host_.resize(cap);
        return TG_OK;
    }

    // Tile list of every used, not finished system (the host reads the flags).
    int rebuild() {
        std::vector<GbState> states;
        int rc = poll(states);
        if (rc) return rc;
        std::vector<GbTile> tl;
        std::vector<int> bl;  // big active systems
        int mmax = 0;
        for (int i = 0; i < (int)used_.size(); ++i)
            if (used_[i] && !states[i].done) {
                const GbSys& S = host_[i];
                for (int t = 0; t < S.tiles; ++t)
                    tl.push_back(GbTile{i, t * S.tile_rows, std::min(S.n, (t + 1) * S.tile_rows), t});
                if (S.big) bl.push_back(i);
                mmax = std::max(mmax, S.m);
            }
        if (tl.size() > tiles_cap_) {
            if (dtiles_) TG_HIP(hipFreeAsync(dtiles_, s_));
            tiles_cap_ = std::max(tl.size(), 2 * tiles_cap_);
            TG_HIP(hipMallocAsync((void**)&dtiles_, tiles_cap_ * sizeof(GbTile), s_));
        }
        if (bl.size() > big_cap_) {
            if (dbig_) TG_HIP(hipFreeAsync(dbig_, s_));
            big_cap_ = std::max(bl.size(), 2 * big_cap_);
            TG_HIP(hipMallocAsync((void**)&dbig_, big_cap_ * sizeof(int), s_));
        }
        if (!tl.empty())
            TG_HIP(hipMemcpyAsync(dtiles_, tl.data(), tl.size() * sizeof(GbTile), hipMemcpyHostToDevice, s_));
        if (!bl.empty()) TG_HIP(hipMemcpyAsync(dbig_, bl.data(), bl.size() * sizeof(int), hipMemcpyHostToDevice, s_));
        TG_HIP(hipStreamSynchronize(s_));  // the host lists go out of scope
        ntiles_ = (int)tl.size();
        nbig_ = (int)bl.size();
        mmax_ = mmax;
        dirty_ = false;
        return TG_OK;
    }

    void release() {
        for (auto* m : mem_)
            if (m) (void)hipFreeAsync(m, s_);
        if (dsys_) (void)hipFreeAsync(dsys_, s_);
        if (dst_) (void)hipFreeAsync(dst_, s_);
        if (dtiles_) (void)hipFreeAsync(dtiles_, s_);
        if (dbig_) (void)hipFreeAsync(dbig_, s_);
        (void)hipStreamSynchronize(s_);
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
        if (cap_s_) (void)hipStreamDestroy(cap_s_);
        cap_s_ = nullptr;
    }

    const int delayed_;  // the form of every system of this engine (knob gbatch_delayed at creation)

    struct GraphKey {
        const void* tiles;
        const void* sys;
        const void* st;
        const void* big;
        int ntiles, nbig, mmax, cycles;
        bool operator==(const GraphKey& o) const {
            return tiles == o.tiles && sys == o.sys && st == o.st && big == o.big && ntiles == o.ntiles &&
                   nbig == o.nbig && mmax == o.mmax && cycles == o.cycles;
        }
    };
    hipStream_t cap_s_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    GraphKey gkey_{};

    hipStream_t s_;
    size_t cap_ = 0;
    std::vector<char> used_;
    std::vector<char*> mem_;
    std::vector<size_t> bytes_;
    std::vector<GbSys> host_;
    GbSys* dsys_ = nullptr;
    GbState* dst_ = nullptr;
    GbTile* dtiles_ = nullptr;
    size_t tiles_cap_ = 0;
    int ntiles_ = 0;
    int* dbig_ = nullptr;  // slots of the big active systems
    size_t big_cap_ = 0;
    int nbig_ = 0;
    int mmax_ = 0;
    bool dirty_ = true;
};

struct GbRecord {
    int slot = -1;
    int k = 0;
    int finished = 0, succeeded = 0, iterations = 0, restarts = 0;
    double rel[kGbMaxK] = {};
};

}  // namespace tg

using namespace tg;

struct tg_gbatch_s {
    std::unique_ptr<GmresBatch> eng;
    std::map<int64_t, GbRecord> sys;  // by id
    std::vector<int64_t> finished;
    int64_t next_id = 1;
};

extern "C" int tg_gbatch_create(tg_gbatch_t* out, void* stream) {
    if (!out) return fail(TG_ERR_INVALID, "tg_gbatch_create: null output");
    auto* h = new tg_gbatch_s();
    h->eng = std::make_unique<GmresBatch>(as_stream(stream));
    *out = h;
    return TG_OK;
}

extern "C" int tg_gbatch_destroy(tg_gbatch_t h) {
    delete h;
    return TG_OK;
}

extern "C" int tg_gbatch_add(tg_gbatch_t h, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                             const double* val, const double* inv_diag, int32_t nrhs, const double* B, double* X,
                             double tol, int32_t restart, int32_t max_restarts, int64_t* id_out) {
    if (!h || !id_out) return fail(TG_ERR_INVALID, "tg_gbatch_add: null handle or id");
    int slot = -1;
    const int rc = h->eng->add(n, nnz, rp, ci, val, inv_diag, nrhs, B, X, tol, restart, max_restarts, &slot);
    if (rc) return rc;
    GbRecord r;
    r.slot = slot;
    r.k = nrhs;
    const int64_t id = h->next_id++;
    h->sys[id] = r;
    *id_out = id;
    return TG_OK;
}

extern "C" int tg_gbatch_step(tg_gbatch_t h, int32_t cycles) {
    if (!h || cycles < 1) return fail(TG_ERR_INVALID, "tg_gbatch_step: bad argument");
    int rc = h->eng->step(cycles);
    if (rc) return rc;
    std::vector<GbState> st;
    rc = h->eng->poll(st);
    if (rc) return rc;
    for (auto& kv : h->sys) {
        GbRecord& r = kv.second;
        if (r.finished || r.slot < 0) continue;
        const GbState& s = st[r.slot];
        if (!s.done) continue;
        r.finished = 1;
        r.succeeded = s.succ;
        r.iterations = s.iters;
        r.restarts = std::max(0, s.cycles - 1);
        for (int c = 0; c < r.k; ++c) r.rel[c] = (s.nan || s.zero_rhs) ? (s.zero_rhs ? 0.0 : NAN) : s.rel[c];
        h->finished.push_back(kv.first);
        h->eng->finished(r.slot);
    }
    return TG_OK;
}

extern "C" int tg_gbatch_finished(tg_gbatch_t h, int64_t* ids, int32_t cap, int32_t* count) {
    if (!h || !count || (cap > 0 && !ids)) return fail(TG_ERR_INVALID, "tg_gbatch_finished: bad argument");
    const int c = (int)std::min<size_t>(h->finished.size(), (size_t)std::max(0, cap));
    for (int i = 0; i < c; ++i) ids[i] = h->finished[i];
    h->finished.erase(h->finished.begin(), h->finished.begin() + c);
    *count = c;
    return TG_OK;
}

extern "C" int tg_gbatch_result(tg_gbatch_t h, int64_t id, int32_t* finished, int32_t* succeeded, int32_t* iterations,
                                int32_t* restarts, double* rel_residuals) {
    if (!h) return fail(TG_ERR_INVALID, "tg_gbatch_result: null handle");
    auto it = h->sys.find(id);
    if (it == h->sys.end()) return fail(TG_ERR_INVALID, "tg_gbatch_result: unknown id");
    const GbRecord& r = it->second;
    if (finished) *finished = r.finished;
    if (succeeded) *succeeded = r.succeeded;
    if (iterations) *iterations = r.iterations;
    if (restarts) *restarts = r.restarts;
    if (rel_residuals)
        for (int c = 0; c < r.k; ++c) rel_residuals[c] = r.rel[c];
    return TG_OK;
}

extern "C" int tg_gbatch_release(tg_gbatch_t h, int64_t id) {
    if (!h) return fail(TG_ERR_INVALID, "tg_gbatch_release: null handle");
    auto it = h->sys.find(id);
    if (it == h->sys.end()) return fail(TG_ERR_INVALID, "tg_gbatch_release: unknown id");
    if (!it->second.finished) return fail(TG_ERR_INVALID, "tg_gbatch_release: system still running");
    const int rc = h->eng->retire(it->second.slot);
    h->sys.erase(it);
    return rc;
}

extern "C" int tg_gbatch_active(tg_gbatch_t h, int32_t* systems, int64_t* bytes_in_use) {
    if (!h) return fail(TG_ERR_INVALID, "tg_gbatch_active: null handle");
    int c = 0;
    for (auto& kv : h->sys) c += kv.second.finished ? 0 : 1;
    if (systems) *systems = c;
    if (bytes_in_use) *bytes_in_use = (int64_t)h->eng->bytes_in_use();
    return TG_OK;
}

extern "C" size_t tg_gbatch_request_bytes(int32_t n, int32_t nrhs, int32_t restart) {
    if (n < 1 || nrhs < 1) return 0;
    const int m = std::max(1, std::min(restart, n + 1));
    return gb_system_bytes(n, nrhs, m, kGbTileRows);
}
