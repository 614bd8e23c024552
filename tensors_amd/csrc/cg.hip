// Block preconditioned conjugate gradient on the GPU:
//   X <- a * A^{-1} B + b * X    for nrhs right-hand sides at once.
//
// The recurrences, special cases and stopping rule are those of
// ConjugateGradient::operator() (ConjugateGradient.hpp:105-353): per-column
// alpha_k, beta_k, rho_k; stop when every column has |r.z|_k <= tol^2 |b.Pb|_k
// (:176-181, 470-480) or after max_iter iterations.  One departure: x starts
// at zero (the reference leaves it uninitialised when the loop never runs
// without an initial guess, :214-250, 342-346).
//
// Layout: every n x nrhs block is row-major with ld = nrhs (the reference's
// Tensor2 workspaces r, u, p, x, z, :38-42).  Kernels per iteration:
//   spmm (u = A p)  |  dot(p,u) -> alpha  |  x,r update + z = P r + dot(r,z)
//   -> rho, beta, convergence  |  p = z + beta p
// Dots are deterministic: each thread sums its rows in order, the workgroup
// combines threads in a fixed tree, and a one-workgroup kernel adds the
// workgroup partials in index order (replacing ComputeScalarProducts +
// ThreadTensor2::AddReduce, ConjugateGradient.hpp:395-438).
// Convergence lives on the device: once `done` is set every kernel returns at
// once, so the host enqueues iterations in batches and polls the flag between
// batches instead of synchronising every iteration.

#include <string>

#include "common.hpp"
#include "krylov.hpp"
#include "reduce.hpp"

namespace tg {

// --- setup ----------------------------------------------------------------
// r = B;  z = P r;  x = 0 (unless guessing);  partial(r.z)
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_init_kernel(int n, int k, int kpad, const T* __restrict__ Bm,
                                                         int64_t ldB, const T* __restrict__ d, T* __restrict__ r,
                                                         T* __restrict__ z, T* __restrict__ x, int zero_x,
                                                         int zmode, double* __restrict__ partial) {
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k)
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e = (int64_t)i * k + L.c;
            const T rv = Bm[(int64_t)i * ldB + L.c];
            r[e] = rv;
            T zv = rv;
            if (d) {
                zv = d[i] * rv;
                z[e] = zv;
            }
            if (zero_x) x[e] = T(0);
            if (zmode != 2) acc += (double)rv * (double)zv;
        }
    column_partial(acc, L, k, kpad, partial);
}

// With an initial guess: r -= u (u = A x);  z = P r;  partial(r.z)   (:214-246)
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_guess_kernel(int n, int k, int kpad, const T* __restrict__ u,
                                                          const T* __restrict__ d, T* __restrict__ r,
                                                          T* __restrict__ z, int zmode, double* __restrict__ partial) {
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k)
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e = (int64_t)i * k + L.c;
            const T rv = r[e] - u[e];
            r[e] = rv;
            T zv = rv;
            if (d) {
                zv = d[i] * rv;
                z[e] = zv;
            }
            if (zmode != 2) acc += (double)rv * (double)zv;
        }
    column_partial(acc, L, k, kpad, partial);
}

// mode 0: rho = sum, bsq = |rho|, TOL = bsq * tol^2, nan/zero flags (:171-211)
// mode 1: rho = sum (after the initial-guess residual)
__global__ __launch_bounds__(kFinalThreads) void cg_setup_finalize_kernel(int k, int kpad, int B, int mode, double tol2,
                                                                   const double* __restrict__ partial,
                                                                   CGState* __restrict__ st) {
    const int c = threadIdx.x % kpad, j = threadIdx.x / kpad, J = kFinalThreads / kpad;
    const double tot = column_total(partial, B, k, kpad, c, j, J);
    if (j == 0 && c < k) {
        st->rho[c] = tot;
        if (mode == 0) {
            st->bsq[c] = fabs(tot);
            st->TOL[c] = fabs(tot) * tol2;
        }
    }
}

// After p = z: succeeded = all |rho| <= TOL; done if succeeded or max_iter == 0.
__global__ void cg_check_kernel(int k, CGState* __restrict__ st) {
    if (threadIdx.x != 0) return;
    bool ok = true;
    for (int c = 0; c < k; ++c) ok = ok && (fabs(st->rho[c]) <= st->TOL[c]);
    st->succeeded = ok ? 1 : 0;
    st->iter = 0;
    st->done = (ok || st->max_iter <= 0) ? 1 : 0;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void cg_copy_kernel(int64_t count, const T* __restrict__ s, T* __restrict__ d) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        d[e] = s[e];
}

// x = X (row-major ld) for the initial guess
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_load_guess_kernel(int n, int k, const T* __restrict__ X, int64_t ldX,
                                                               T* __restrict__ x) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        x[e] = X[i * ldX + (e - i * k)];
    }
}

// --- iteration --------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_dot_pu_kernel(int n, int k, int kpad, const T* __restrict__ p,
                                                           const T* __restrict__ u, double* __restrict__ partial,
                                                           const CGState* __restrict__ st) {
    if (st && st->done) return;
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k)
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e = (int64_t)i * k + L.c;
            acc += (double)p[e] * (double)u[e];
        }
    column_partial(acc, L, k, kpad, partial);
}

// alpha = rho / (p.u)   (:255-260); also kept in the alpha history (slot iter % xdefer)
// for the deferred x update
__global__ __launch_bounds__(kFinalThreads) void cg_alpha_kernel(int k, int kpad, int B, const double* __restrict__ partial,
                                                          double* __restrict__ ahist, CGState* __restrict__ st) {
    if (st->done) return;
    const int c = threadIdx.x % kpad, j = threadIdx.x / kpad, J = kFinalThreads / kpad;
    const double pu = column_total(partial, B, k, kpad, c, j, J);
    // guard: a column whose direction has vanished (p.u <= 0, an FP32 iteration past
    // the point its Krylov space is used up) stops moving instead of producing NaN
    if (j == 0 && c < k) {
        const double al = (st->guard && !(pu > 0.0)) ? 0.0 : st->rho[c] / pu;
        st->alpha[c] = al;
        ahist[(int64_t)(st->iter % st->xdefer) * k + c] = al;
    }
}

template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
    if constexpr (sizeof(T) == 8)
        return __builtin_fma(a, b, c);
    else
        return __builtin_fmaf(a, b, c);
}

// x += alpha p with the first iteration's x = alpha p (x starts at zero, :262-266)
template <typename T>
__device__ __forceinline__ T x_step(bool first, T x, T al, T pv) {
    return first ? al * pv : fma_t(al, pv, x);
}

constexpr int kRowsInFlight = 4;  // rows per thread whose loads are issued together

// The thread's rows of the r update, kRowsInFlight at a time: all loads first (at
// clamped, always valid rows), then the arithmetic and stores in row order, so the
// loads of four rows are in flight together and the partial adds its terms in the
// same order as a plain row loop.
template <typename T, bool HAS_D, bool ZDOT>
__device__ __forceinline__ double update_rows(int i0, int i1, const Lanes& L, int k, T al, const T* __restrict__ u,
                                              const T* __restrict__ d, T* __restrict__ r, T* __restrict__ z) {
    double acc = 0.0;
    for (int ib = i0 + L.ro; ib < i1; ib += kRowsInFlight * L.rows) {
        T rv[kRowsInFlight], uv[kRowsInFlight], dv[kRowsInFlight];
#pragma unroll
        for (int q = 0; q < kRowsInFlight; ++q) {
            const int i = min(ib + q * L.rows, i1 - 1);
            const int64_t e = (int64_t)i * k + L.c;
            rv[q] = r[e];
            uv[q] = u[e];
            if constexpr (HAS_D) dv[q] = d[i];
        }
#pragma unroll
        for (int q = 0; q < kRowsInFlight; ++q) {
            const int i = ib + q * L.rows;
            if (i < i1) {
                const int64_t e = (int64_t)i * k + L.c;
                const T rn = rv[q] - al * uv[q];
                r[e] = rn;
                T zv = rn;
                if constexpr (HAS_D) {
                    zv = dv[q] * rn;
                    z[e] = zv;
                }
                if constexpr (ZDOT) acc += (double)rn * (double)zv;
            }
        }
    }
    return acc;
}

// r -= alpha u;  z = P r;  partial(r.z)   (:262-313).  The reference updates x here as
// well; x += alpha p is left to the p step (cg_p_kernel), which reads p anyway: three
// passes over n x nrhs here instead of six.
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(int n, int k, int kpad, const T* __restrict__ u,
                                                           const T* __restrict__ d, T* __restrict__ r,
                                                           T* __restrict__ z, int zmode, double* __restrict__ partial,
                                                           const CGState* __restrict__ st) {
    if (st->done) return;
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k) {
        const T al = (T)st->alpha[L.c];
        if (d)
            acc = zmode != 2 ? update_rows<T, true, true>(i0, i1, L, k, al, u, d, r, z)
                             : update_rows<T, true, false>(i0, i1, L, k, al, u, d, r, z);
        else
            acc = zmode != 2 ? update_rows<T, false, true>(i0, i1, L, k, al, u, d, r, z)
                             : update_rows<T, false, false>(i0, i1, L, k, al, u, d, r, z);
    }
    column_partial(acc, L, k, kpad, partial);
}

// rho_old = rho; rho = r.z; beta = rho / rho_old; convergence; ++iter  (:309-338)
__global__ __launch_bounds__(kFinalThreads) void cg_rho_kernel(int k, int kpad, int B, const double* __restrict__ partial,
                                                        CGState* __restrict__ st) {
    if (st->done) return;
    const int c = threadIdx.x % kpad, j = threadIdx.x / kpad, J = kFinalThreads / kpad;
    const double rz = column_total(partial, B, k, kpad, c, j, J);
    __shared__ int conv;
    if (threadIdx.x == 0) conv = 1;
    __syncthreads();
    if (j == 0 && c < k) {
        const double old = st->rho[c];
        st->rho_old[c] = old;
        st->rho[c] = rz;
        st->beta[c] = (st->guard && !(old != 0.0)) ? 0.0 : rz / old;  // guard: a column whose r vanished
        if (!(fabs(rz) <= st->TOL[c])) atomicAnd(&conv, 0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        st->succeeded = conv;
        const int it = st->iter + 1;
        st->iter = it;
        const int dn = (conv || it >= st->max_iter) ? 1 : 0;
        st->done = dn;
        // x += alpha p of the iterations [xfrom, it): in this iteration's p step once
        // xdefer of them are pending, else later (after a stop, cg_xfinish_kernel: the p
        // step does not run)
        int cnt = 0;
        if (!dn && it - st->xfrom >= st->xdefer) {
            cnt = it - st->xfrom;
            st->xfirst = st->xfrom;
            st->xfrom = it;
        }
        st->xcnt = cnt;
    }
}

constexpr int kMaxDefer = kCgRingSlots;  // ring slots of p for the deferred x update

// The x update of the iterations [xf, xf + cnt) for the thread's R rows: x += alpha_t p_t
// in iteration order, each the same fused multiply-add the per-iteration update makes (x
// starts at zero without a guess: the first update is alpha p), so x is bitwise the same
// for any deferral depth.  The p of the newest iteration, xf + cnt - 1, may be passed in
// `own` (already loaded).
template <typename T, int R>
__device__ __forceinline__ void x_updates(int cnt, int xf, int D, bool first0, const T (&al)[kMaxDefer],
                                          const T* ring, int64_t pstride, const int64_t (&e)[R], const T* own,
                                          T (&xv)[R]) {
    T q[kMaxDefer][R];
#pragma unroll
    for (int j = 0; j < kMaxDefer; ++j)
        if (j < cnt && !(own && j == cnt - 1)) {
            const T* pj = ring + (int64_t)((xf + j) % D) * pstride;
#pragma unroll
            for (int r = 0; r < R; ++r) q[j][r] = pj[e[r]];
        }
#pragma unroll
    for (int j = 0; j < kMaxDefer; ++j)
        if (j < cnt) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const T pv = (own && j == cnt - 1) ? own[r] : q[j][r];
                xv[r] = x_step(first0 && j == 0, xv[r], al[j], pv);
            }
        }
}

// p = z + beta p   (:321-335), and the deferred x += alpha p (the update the r step
// leaves out) of the iterations the rho step released.  rho has advanced iter, so
// this is iteration t = iter - 1: p_t is read from ring slot t % D and p_{t+1} written
// to slot (t + 1) % D, whose p_{t+1-D} is in x by then or added here.
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_p_kernel(int n, int k, int kpad, const T* __restrict__ z, T* ring,
                                                      int64_t pstride, T* __restrict__ x,
                                                      const double* __restrict__ ahist,
                                                      const CGState* __restrict__ st) {
    if (st->done) return;
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    const int D = st->xdefer, t = st->iter - 1;
    const int cnt = st->xcnt, xf = st->xfirst;
    const bool first0 = xf == 0 && !st->use_guess;
    const T* pc = ring + (int64_t)(t % D) * pstride;
    T* pn = ring + (int64_t)((t + 1) % D) * pstride;
    if (L.c < k) {
        const T be = (T)st->beta[L.c];
        T al[kMaxDefer];
#pragma unroll
        for (int j = 0; j < kMaxDefer; ++j)
            al[j] = j < cnt ? (T)ahist[(int64_t)((xf + j) % D) * k + L.c] : T(0);
        for (int ib = i0 + L.ro; ib < i1; ib += kRowsInFlight * L.rows) {
            T pv[kRowsInFlight], zv[kRowsInFlight], xv[kRowsInFlight];
            int64_t e[kRowsInFlight];
#pragma unroll
            for (int q = 0; q < kRowsInFlight; ++q) {  // loads of four rows first (clamped rows)
                e[q] = (int64_t)min(ib + q * L.rows, i1 - 1) * k + L.c;
                pv[q] = pc[e[q]];
                zv[q] = z[e[q]];
                if (cnt > 0) xv[q] = x[e[q]];
            }
            if (cnt > 0) x_updates<T, kRowsInFlight>(cnt, xf, D, first0, al, ring, pstride, e, pv, xv);
#pragma unroll
            for (int q = 0; q < kRowsInFlight; ++q) {
                if (ib + q * L.rows < i1) {
                    if (cnt > 0) x[e[q]] = xv[q];
                    pn[e[q]] = zv[q] + be * pv[q];
                }
            }
        }
    }
}

// The deferred x updates still pending, [xfrom, iter) (after a stop the p step did
// not run; at the end of a batch of iterations the last group may be short), so x is
// complete before the host looks at it.  cg_xclear_kernel then advances xfrom.
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_xfinish_kernel(int n, int k, int kpad, const T* __restrict__ ring,
                                                            int64_t pstride, T* __restrict__ x,
                                                            const double* __restrict__ ahist,
                                                            const CGState* __restrict__ st) {
    const int xf = st->xfrom, cnt = st->iter - xf, D = st->xdefer;
    if (cnt <= 0) return;
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    const bool first0 = xf == 0 && !st->use_guess;
    if (L.c < k) {
        T al[kMaxDefer];
#pragma unroll
        for (int j = 0; j < kMaxDefer; ++j)
            al[j] = j < cnt ? (T)ahist[(int64_t)((xf + j) % D) * k + L.c] : T(0);
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e[1] = {(int64_t)i * k + L.c};
            T xv[1] = {x[e[0]]};
            x_updates<T, 1>(cnt, xf, D, first0, al, ring, pstride, e, nullptr, xv);
            x[e[0]] = xv[0];
        }
    }
}

__global__ void cg_xclear_kernel(CGState* __restrict__ st) { st->xfrom = st->iter; }

// --- fused iteration forms (CSR operator, P = I or Jacobi) ------------------
// Form 1 keeps the four-kernel order but computes partial(p.u) in the SpMM
// epilogue (cg_dir_spmm_kernel with FORM = false), saving the separate p.u pass.
// Form 2 has two vector kernels per iteration instead of four:
//   K1 (cg_dir_spmm_kernel):  p = z + beta p_prev;  u = A p;  partial(p.u)
//   K2 (cg_step_kernel):      x += alpha p;  r -= alpha u;  partial(r.z)
// with z = r (P = I) or z = d r (Jacobi) formed where it is needed and never
// stored.  p is double-buffered by the parity of the iteration count: K1 of
// iteration t writes p[t & 1] and reads p[(t - 1) & 1]; p[1] is zero before the
// first iteration, where beta = 0, so p = z exactly as after the reference's
// copy (ConjugateGradient.hpp:247-253).  A row needs p of the rows its columns
// name, which other workgroups write in the same launch; it recomputes them from
// z and p_prev with the same fused multiply-add, so the values are bitwise those
// the owners store and no workgroup reads what another writes.

// One row per lane group of G lanes; each lane holds V consecutive columns of a
// G*V-column tile (k > G*V only for G = 64: further tiles).  The workgroup's p.u
// partial is summed per (wave, lane group, column) in row order in LDS, then by a
// fixed tree over the lane groups: deterministic.
// Rows of a workgroup: panel = 0, a contiguous 1/B of [0, n); panel = P > 0, the
// panel order: [0, n) cut into panels of P rows and the workgroup taking the same
// 1/B of the offsets [0, P) in every panel, panel after panel.  With P the matrix
// band (the plane of a 3D grid), a row's neighbours lie at nearby offsets of the
// panels before and after, which the workgroups of the same XCD (contiguous
// offsets after xcd_swizzle) work on at about the same time, so p rows are reused
// from the L2 instead of being fetched again for each neighbouring plane.
// FORM = false is the plain "u = A p, partial(p.u)" step of the three-pass
// iteration (p formed by cg_p_kernel in ring slot iter % xdefer from p0 on, pstride
// elements apart; r, d, p1 unused).
template <typename T, typename RP, int G, int V, bool FORM, bool HAS_D>
__global__ __launch_bounds__(kBlock) void cg_dir_spmm_kernel(int n, int k, const RP* __restrict__ rp,
                                                             const int32_t* __restrict__ ci, const T* __restrict__ a,
                                                             const T* __restrict__ r, const T* __restrict__ d,
                                                             T* __restrict__ p0, T* __restrict__ p1,
                                                             int64_t pstride, T* __restrict__ u,
                                                             double* __restrict__ partial,
                                                             const CGState* __restrict__ st, int panel) {
    if (st->done) return;
    constexpr int S = kWave / G;  // rows in flight per wave
    constexpr int KT = G * V;     // columns per tile
    constexpr int WAVES = kBlock / kWave;
    // gathers in flight per lane and round (8 measured slower on the 128^3
    // Laplacian: 86 VGPRs and 5 waves per SIMD instead of 56 and 8)
    constexpr int U = 4;
    __shared__ double sh[WAVES * kWave * 4];  // WAVES * S * k <= 1024 (k <= G V, or G = 64 and k <= 256)
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int s = lane / G, g = lane % G;
    const int slot = wave * S + s;
    const int slots = WAVES * S;
    for (int i = threadIdx.x; i < slots * k; i += kBlock) sh[i] = 0.0;
    __syncthreads();
    const int B = gridDim.x;
    const int lb = xcd_swizzle(blockIdx.x, B);
    // rows: q in [0, cnt) -> first + (q / sw) * P + q % sw (panel order) or first + q
    int first, sw, cnt, P;
    if (panel > 0) {
        P = panel;
        first = (int)((int64_t)P * lb / B);
        sw = (int)((int64_t)P * (lb + 1) / B) - first;
        cnt = sw * ((n + P - 1) / P);
    } else {
        P = 0;
        first = (int)((int64_t)n * lb / B);
        sw = 1;
        cnt = (int)((int64_t)n * (lb + 1) / B) - first;
    }
    const bool odd = FORM && (st->iter & 1) != 0;
    // FORM: written by the previous iteration; else ring slot iter % xdefer (p0 = slot 0)
    const T* __restrict__ pp = FORM ? (odd ? p0 : p1) : p0 + (int64_t)(st->iter % st->xdefer) * pstride;
    T* __restrict__ pn = odd ? p1 : p0;
    for (int qb = wave * S; qb < cnt; qb += WAVES * S) {
        const int q = qb + s;
        const int row = P > 0 ? first + (q / sw) * P + q % sw : first + q;
        const bool rok = q < cnt && row < n;
        const RP lo = rok ? rp[row] : RP(0), hi = rok ? rp[row + 1] : RP(0);
        const int len = (int)(hi - lo);
        int wmax = len;  // longest row of the wave (uniform loop bound)
#pragma unroll
        for (int off = G; off < kWave; off <<= 1) wmax = max(wmax, __shfl_xor(wmax, off, kWave));
        const T di = (HAS_D && rok) ? d[row] : T(1);
        for (int c0 = 0; c0 < k; c0 += KT) {
            const int col = c0 + g * V;
            const bool colok = col < k;
            const int colc = colok ? col : 0;
            T be[V];
#pragma unroll
            for (int i = 0; i < V; ++i) be[i] = FORM ? (T)st->beta[colc + i] : T(0);
            T acc[V];
#pragma unroll
            for (int i = 0; i < V; ++i) acc[i] = T(0);
            // the row's own r and p_prev (p for FORM = false), issued before the
            // gathers so they are in flight with them (clamped to a valid row)
            const int64_t own = (int64_t)(rok ? row : first) * k + colc;
            T rv_own[V], pv_own[V];
            if constexpr (FORM) load_vec<T, V>(r + own, rv_own);
            load_vec<T, V>(pp + own, pv_own);
            for (int base = 0; base < wmax; base += G) {
                const int e = base + g;
                int32_t myc = 0;
                T mya = T(0);
                if (e < len) {
                    myc = __builtin_nontemporal_load(ci + lo + e);
                    mya = a ? __builtin_nontemporal_load(a + lo + e) : T(1);
                }
                const int cnt = min(G, len - base);  // this group's pairs in the chunk
                const int wcnt = min(G, wmax - base);
                for (int u0 = 0; u0 < wcnt; u0 += U) {
                    // U (column, value) pairs from the group's lanes, then all their
                    // gathers in flight together, then the arithmetic
                    int cj[U];
                    T av[U];
                    bool live[U];
#pragma unroll
                    for (int uu = 0; uu < U; ++uu) {
                        const int j = u0 + uu;
                        const int src = s * G + (j < G ? j : 0);
                        const int32_t c = __shfl(myc, src, kWave);
                        av[uu] = __shfl(mya, src, kWave);
                        live[uu] = (j < cnt) && colok;
                        cj[uu] = live[uu] ? c : 0;
                    }
                    T rv[U][V], pv[U][V], dj[U];
#pragma unroll
                    for (int uu = 0; uu < U; ++uu) {
                        const int64_t o = (int64_t)cj[uu] * k + colc;
                        if constexpr (FORM) load_vec<T, V>(r + o, rv[uu]);
                        load_vec<T, V>(pp + o, pv[uu]);
                        dj[uu] = HAS_D ? d[cj[uu]] : T(1);
                    }
#pragma unroll
                    for (int uu = 0; uu < U; ++uu) {
#pragma unroll
                        for (int i = 0; i < V; ++i) {
                            T pj = pv[uu][i];
                            if constexpr (FORM) {
                                const T z = HAS_D ? dj[uu] * rv[uu][i] : rv[uu][i];
                                pj = fma_t(be[i], pv[uu][i], z);
                            }
                            acc[i] += live[uu] ? av[uu] * pj : T(0);
                        }
                    }
                }
            }
            if (rok && colok) {
                const int64_t o = (int64_t)row * k + col;
                T pi[V];
                double* shs = sh + (int64_t)slot * k + col;
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    pi[i] = pv_own[i];
                    if constexpr (FORM) {
                        const T z = HAS_D ? di * rv_own[i] : rv_own[i];
                        pi[i] = fma_t(be[i], pv_own[i], z);
                    }
                    shs[i] += (double)pi[i] * (double)acc[i];
                }
                if constexpr (FORM) store_vec<T, V>(pn + o, pi);
                store_vec<T, V>(u + o, acc);
            }
        }
    }
    __syncthreads();
    // fixed tree over the lane-group slots (slots is a power of two)
    for (int h = slots / 2; h > 0; h >>= 1) {
        for (int t = threadIdx.x; t < h * k; t += kBlock) sh[t] += sh[t + h * k];
        __syncthreads();
    }
    for (int c = threadIdx.x; c < k; c += kBlock) partial[(int64_t)lb * k + c] = sh[c];
}

// x = (first ? 0 : x) + alpha p;  r -= alpha u;  partial(r.z), z = r or d r
template <typename T, bool HAS_D>
__global__ __launch_bounds__(kBlock) void cg_step_kernel(int n, int k, int kpad, const T* __restrict__ p0,
                                                         const T* __restrict__ p1, const T* __restrict__ u,
                                                         const T* __restrict__ d, T* __restrict__ x,
                                                         T* __restrict__ r, double* __restrict__ partial,
                                                         const CGState* __restrict__ st) {
    if (st->done) return;
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    const T* __restrict__ p = (st->iter & 1) ? p1 : p0;
    const bool first = (st->iter == 0) && !st->use_guess;
    double acc = 0.0;
    if (L.c < k) {
        const T al = (T)st->alpha[L.c];
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e = (int64_t)i * k + L.c;
            x[e] = x_step(first, x[e], al, p[e]);
            const T rv = r[e] - al * u[e];
            r[e] = rv;
            const T zv = HAS_D ? d[i] * rv : rv;
            acc += (double)rv * (double)zv;
        }
    }
    column_partial(acc, L, k, kpad, partial);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void cg_zero_kernel(int64_t count, T* __restrict__ v) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        v[e] = T(0);
}

// X <- a x + b X  with combine_matrices_auto's shortcuts (:342-346)
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_write_kernel(int n, int k, T a, const T* __restrict__ x, T b,
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

// X <- b X literally (scale_matrix at :151-153, :201-203)
template <typename T>
__global__ __launch_bounds__(kBlock) void scale_kernel(int n, int k, T b, T* __restrict__ X, int64_t ldX) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        X[i * ldX + (e - i * k)] *= b;
    }
}

// ---------------------------------------------------------------------------
// Host driver.
// ---------------------------------------------------------------------------
static int pow2ceil_i(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Workgroup partials of the dots: kSolverBlocks per column for the vector
// kernels, kFusedBlocks for the fused direction + SpMM kernel.
constexpr int kPartialBlocks = kFusedBlocks > kSolverBlocks ? kFusedBlocks : kSolverBlocks;

// Layout: state, partials, alpha history, then r, u, x (+ z) and the ring of p
// (slot 0 = p; slot 1 = the fused iteration's second p; `slots` = 2 or kMaxDefer).
size_t cg_workspace_bytes(int32_t n, int32_t k, size_t vs, int prec_kind, int slots) {
    const size_t vec = align256((size_t)n * k * vs);
    const size_t nvec = (prec_kind ? 4 : 3) + (size_t)std::max(2, slots);
    return align256(sizeof(CGState)) + align256((size_t)kPartialBlocks * k * sizeof(double)) +
           align256((size_t)kMaxDefer * k * sizeof(double)) + nvec * vec;
}

// Iteration form for CSR operators with P = I or Jacobi, short rows (mean length
// <= 16) and no column slabs ("cg_fused"): 0 = four vector kernels after the
// SpMM; 1 = p.u in the SpMM epilogue (kSpmmDot); 2 = direction, SpMM and p.u in
// one kernel (kFormDir); 3 = default (form 2 up to 2^20 entries n * nrhs, form 1
// above).  On the 128^3 Laplacian with 32 right-hand sides form 1 measured fastest in
// FP64 and FP32: gathering both z and p_prev per nonzero (form 2) cost more than
// streaming p once; on small blocks form 2's one launch less per iteration wins.  Other
// operators always take form 0.
enum CgForm : int { kFourKernels = 0, kSpmmDot = 1, kFormDir = 2, kFormAuto = 3 };
static int g_cg_fused = kFormAuto;
static int g_cg_graph = 1;              // 1: blocks of kCgGraphIters iterations as captured hipGraphs
constexpr int kCgGraphIters = 16;
// Deferral depth of form 1's x update ("cg_xdefer": 1, 2 or 4): x += alpha p of
// `depth` iterations in one pass, from a ring of p, where the workspace holds the ring
// (2 slots always, 4 in tg_cg_workspace_bytes).  Depth 4 skips the x read and write in
// three iterations of four: 17 instead of 20 n x nrhs passes of the p step.
static int g_cg_xdefer = kMaxDefer;
// rows of the forms' SpMM kernel: 1 = panel order where the band allows, 0 = contiguous chunks
static int g_cg_panel = 1;

int cg_tune_set(const char* key, int32_t value, bool* known) {
    const std::string k(key);
    *known = k == "cg_fused" || k == "cg_panel" || k == "cg_graph" || k == "cg_xdefer";
    if (!*known) return TG_OK;
    if (k == "cg_xdefer") {
        if (value != 1 && value != 2 && value != kMaxDefer)
            return fail(TG_ERR_INVALID, "tg_tune_set: cg_xdefer must be 1, 2 or %d", kMaxDefer);
        g_cg_xdefer = value;
        return TG_OK;
    }
    if (k == "cg_panel" || k == "cg_graph") {
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: %s must be 0 or 1", key);
        (k == "cg_panel" ? g_cg_panel : g_cg_graph) = value;
        return TG_OK;
    }
    if (value < 0 || value > 3) return fail(TG_ERR_INVALID, "tg_tune_set: cg_fused must be 0, 1, 2 or 3 (auto)");
    g_cg_fused = value;
    return TG_OK;
}

// Whether the panel order could apply to an n-row system (so the caller measures the band).
bool cg_panel_wants_band(int64_t n) { return g_cg_panel != 0 && n >= 16 * (int64_t)kFusedBlocks; }

int32_t cg_tune_get(const char* key) {
    const std::string k(key);
    if (k == "cg_fused") return g_cg_fused;
    if (k == "cg_panel") return g_cg_panel;
    if (k == "cg_graph") return g_cg_graph;
    if (k == "cg_xdefer") return g_cg_xdefer;
    return -1;
}

template <typename T, typename RP, int V>
static int launch_dir_v(int G, int form, bool has_d, hipStream_t s, int n, int k, const RP* rp, const int32_t* ci,
                        const T* a, const T* r, const T* d, T* p0, T* p1, int64_t pstride, T* u, double* partial,
                        const CGState* st, int panel) {
#define TG_DIRK(GG, FORM, HD)                                                                                  \
    hipLaunchKernelGGL((cg_dir_spmm_kernel<T, RP, GG, V, FORM, HD>), dim3(kFusedBlocks), dim3(kBlock), 0, s, n, k, \
                       rp, ci, a, r, d, p0, p1, pstride, u, partial, st, panel)
#define TG_DIR(GG)                     \
    case GG:                           \
        if (form != kFormDir)          \
            TG_DIRK(GG, false, false); \
        else if (has_d)                \
            TG_DIRK(GG, true, true);   \
        else                           \
            TG_DIRK(GG, true, false);  \
        break;
    switch (G) {
        TG_DIR(1)
        TG_DIR(2)
        TG_DIR(4)
        TG_DIR(8)
        TG_DIR(16)
        TG_DIR(32)
        TG_DIR(64)
        default:
            return fail(TG_ERR_INVALID, "internal: bad lane group %d", G);
    }
#undef TG_DIR
#undef TG_DIRK
    return TG_OK;
}

// The SpMM kernel of forms 1 and 2: V columns per lane (16-byte loads where k
// allows), G = enough lanes for all k columns (at most 64; then k / (64 V) tiles).
template <typename T, typename RP>
static int launch_dir(int form, hipStream_t s, int n, int k, const RP* rp, const int32_t* ci, const T* a,
                      const T* r, const T* d, T* p0, T* p1, int64_t pstride, T* u, double* partial,
                      const CGState* st, int panel) {
    int V = 1;
    if (sizeof(T) == 8) {
        V = (k % 2 == 0) ? 2 : 1;  // 4 doubles per lane measured slower on the 128^3 Laplacian
    } else {
        V = (k % 4 == 0) ? 4 : (k % 2 == 0) ? 2 : 1;
    }
    int G = 1;
    while (G < 64 && G * V < k) G *= 2;
    const bool has_d = d != nullptr;
    switch (V) {
        case 1:
            return launch_dir_v<T, RP, 1>(G, form, has_d, s, n, k, rp, ci, a, r, d, p0, p1, pstride, u, partial, st,
                                                  panel);
        case 2:
            return launch_dir_v<T, RP, 2>(G, form, has_d, s, n, k, rp, ci, a, r, d, p0, p1, pstride, u, partial, st,
                                                  panel);
        default:
            if constexpr (sizeof(T) == 4)
                return launch_dir_v<T, RP, 4>(G, form, has_d, s, n, k, rp, ci, a, r, d, p0, p1, pstride, u, partial, st,
                                                  panel);
            return fail(TG_ERR_INVALID, "internal: bad vector width");
    }
}

template <typename T>
static int grid_for(int64_t count) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((count + kBlock - 1) / kBlock, 8192));
}

template <typename T>
int cg_solve_t(const CsrView& A, const OpRef& Aop, int prec_kind, const T* d, const OpRef& Pop, int32_t k,
               int32_t max_iter, double a, const T* Bm, int64_t ldB, double b, T* X, int64_t ldX, double tol,
               int use_guess, void* ws, size_t ws_bytes, CGResult* res, hipStream_t s, const CGHook* hook,
               bool allow_graphs) {
    const int n = A.n;
    res->succeeded = 0;
    res->iterations = 0;
    if (k < 1 || k > kBlock) return fail(TG_ERR_INVALID, "cg: nrhs must be in [1, %d]", kBlock);
    if (n < 0 || A.m != n) return fail(TG_ERR_SHAPE, "cg: the operator must be square");
    if (!ws || ws_bytes < cg_workspace_bytes(n, k, sizeof(T), prec_kind, 2))
        return fail(TG_ERR_WORKSPACE, "cg: workspace missing or smaller than tg_cg_workspace_bytes()");
    const int slots = ws_bytes >= cg_workspace_bytes(n, k, sizeof(T), prec_kind, kMaxDefer) ? kMaxDefer : 2;
    if (prec_kind == 1 && !d) return fail(TG_ERR_INVALID, "cg: diagonal preconditioner without a diagonal");
    if (prec_kind == 2 && !Pop.fn) return fail(TG_ERR_INVALID, "cg: callback preconditioner without a function");
    max_iter = std::max(0, max_iter);
    const int64_t N = (int64_t)n * k;
    if (a == 0.0) {  // :147-163
        if (N > 0) hipLaunchKernelGGL((scale_kernel<T>), dim3(grid_for<T>(N)), dim3(kBlock), 0, s, n, k, (T)b, X, ldX);
        TG_LAUNCH_CHECK();
        res->succeeded = 1;
        return TG_OK;
    }
    char* w = (char*)ws;
    CGState* st = (CGState*)w;
    w += align256(sizeof(CGState));
    double* partial = (double*)w;
    w += align256((size_t)kPartialBlocks * k * sizeof(double));
    double* ahist = (double*)w;
    w += align256((size_t)kMaxDefer * k * sizeof(double));
    const size_t vec = align256((size_t)N * sizeof(T));
    const int64_t pstride = (int64_t)(vec / sizeof(T));  // elements between ring slots
    T* r = (T*)w;
    T* u = (T*)(w + vec);
    T* x = (T*)(w + 2 * vec);
    T* z = prec_kind ? (T*)(w + 3 * vec) : r;
    T* p = (T*)(w + (prec_kind ? 4 : 3) * vec);  // ring slot 0
    T* p1 = p + pstride;                          // ring slot 1
    const T* dd = prec_kind == 1 ? d : nullptr;
    const int zmode = prec_kind;  // 0: z is r, 1: z = d r in the kernels, 2: z = P(r) by callback
    const int kpad = pow2ceil_i(k);
    const int B = kSolverBlocks;
    const bool host_ops = Aop.is_callback() || prec_kind == 2;
    // forms 1 / 2: CSR operator, P = I or Jacobi, short rows, no column slabs
    const CsrView& Ac = Aop.csr ? *Aop.csr : A;
    const bool csr_ok = !host_ops && Aop.csr && n > 0 && Ac.m == n && Ac.n == n && Ac.nnz <= 16 * (int64_t)n &&
                        !spmm_uses_slabs(Ac, k) && Ac.rp && (Ac.nnz == 0 || Ac.ci) &&
                        Ac.val_type == (sizeof(T) == 8 ? TG_F64 : TG_F32);
    // auto: form 2 (four launches per iteration) for blocks of up to 2^20 entries, where
    // an iteration is bound by the kernels' fixed costs (22^3 and 45^3 Laplacians: 12 %
    // faster), form 1 above (its SpMM gathers one vector instead of two)
    int form = g_cg_fused == kFormAuto ? ((int64_t)n * k <= (int64_t(1) << 20) ? kFormDir : kSpmmDot) : g_cg_fused;
    if (!csr_ok) form = kFourKernels;
    const bool fused = form == kFormDir;
    // deferred x update: form 1 only (form 0's operator reads p at a fixed address, form 2
    // updates x in its step kernel)
    const int xdefer = form == kSpmmDot ? std::min(g_cg_xdefer, slots) : 1;
    // panel order for the SpMM kernel of forms 1 / 2 (cg_dir_spmm_kernel): panels of
    // one band, when that gives every workgroup at least 4 rows per panel and the
    // matrix at least 4 panels (the band as the caller measured it; none measured: off)
    int panel = 0;
    if (form != kFourKernels && g_cg_panel && Ac.band >= 4 * (int64_t)kFusedBlocks && (int64_t)n >= 4 * Ac.band)
        panel = (int)Ac.band;
    auto dir_spmm_on = [&](hipStream_t ss) -> int {
        if (Ac.rp_type == TG_I64)
            return launch_dir<T, int64_t>(form, ss, n, k, (const int64_t*)Ac.rp, Ac.ci, (const T*)Ac.val, r, dd, p,
                                          p1, pstride, u, partial, st, panel);
        return launch_dir<T, int32_t>(form, ss, n, k, (const int32_t*)Ac.rp, Ac.ci, (const T*)Ac.val, r, dd, p, p1,
                                      pstride, u, partial, st, panel);
    };

    // z = P r by callback, then partial(r.z) -- the external-preconditioner step
    auto apply_P = [&]() -> int {
        int rc = op_apply(Pop, r, z, k, nullptr, s);
        if (rc) return rc;
        hipLaunchKernelGGL((cg_dot_pu_kernel<T>), dim3(B), dim3(kBlock), 0, s, n, k, kpad, (const T*)r, (const T*)z,
                           partial, (const CGState*)nullptr);
        return TG_OK;
    };

    CGState h{};
    h.max_iter = max_iter;
    h.use_guess = use_guess ? 1 : 0;
    h.guard = hook ? hook->guard : 0;
    h.xdefer = xdefer;
    TG_HIP(hipMemcpyAsync(st, &h, sizeof(CGState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((cg_init_kernel<T>), dim3(B), dim3(kBlock), 0, s, n, k, kpad, Bm, ldB, dd, r, z, x,
                       use_guess ? 0 : 1, zmode, partial);
    if (zmode == 2) {
        int rc = apply_P();
        if (rc) return rc;
    }
    hipLaunchKernelGGL(cg_setup_finalize_kernel, dim3(1), dim3(kFinalThreads), 0, s, k, kpad, B, 0, tol * tol, partial, st);
    TG_LAUNCH_CHECK();
    CGState hs;
    TG_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    double bmax = 0.0;
    for (int c = 0; c < k; ++c) {
        if (std::isnan(hs.bsq[c])) {  // :185-195
            fail(TG_OK, "cg: right-hand side contains NaNs. Doing nothing.");
            return TG_OK;
        }
        bmax = std::max(bmax, hs.bsq[c]);
    }
    if (!(bmax > 0.0)) {  // :197-211
        if (N > 0) hipLaunchKernelGGL((scale_kernel<T>), dim3(grid_for<T>(N)), dim3(kBlock), 0, s, n, k, (T)b, X, ldX);
        TG_LAUNCH_CHECK();
        res->succeeded = 1;
        return TG_OK;
    }
    if (use_guess) {
        hipLaunchKernelGGL((cg_load_guess_kernel<T>), dim3(grid_for<T>(N)), dim3(kBlock), 0, s, n, k, X, ldX, x);
        int rc = op_apply(Aop, x, u, k, nullptr, s);
        if (rc) return rc;
        hipLaunchKernelGGL((cg_guess_kernel<T>), dim3(B), dim3(kBlock), 0, s, n, k, kpad, u, dd, r, z, zmode, partial);
        if (zmode == 2) {
            rc = apply_P();
            if (rc) return rc;
        }
        hipLaunchKernelGGL(cg_setup_finalize_kernel, dim3(1), dim3(kFinalThreads), 0, s, k, kpad, B, 1, 0.0, partial, st);
    }
    if (fused)  // p_prev of the first iteration (beta = 0 there)
        hipLaunchKernelGGL((cg_zero_kernel<T>), dim3(grid_for<T>(N)), dim3(kBlock), 0, s, N, p1);
    else
        hipLaunchKernelGGL((cg_copy_kernel<T>), dim3(grid_for<T>(N)), dim3(kBlock), 0, s, N, (const T*)z, p);
    hipLaunchKernelGGL(cg_check_kernel, dim3(1), dim3(64), 0, s, k, st);
    TG_LAUNCH_CHECK();

    // Iterations.  With device operators every kernel checks `done` itself and
    // the host polls in growing batches; host callbacks cannot check it, so
    // then the host polls after every iteration.
    // The p the next iteration reads (ring slot iter % xdefer), for the hook.
    auto p_next = [&](const CGState& hs) { return p + (int64_t)(hs.iter % xdefer) * pstride; };
    int launched = 0, batch = host_ops ? 1 : hook ? std::max(1, hook->every) : 4;
    int hflag[2] = {0, 0};
    TG_HIP(hipMemcpyAsync(hflag, &st->iter, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    int rc = TG_OK;
    if (hook && hflag[1]) {  // converged before the first iteration: let the hook see it
        TG_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        bool stop = false;
        rc = hook->fn(hs, r, z, x, fused ? nullptr : p_next(hs), partial, st, &stop);
        if (rc) return rc;
        if (!stop) {
            TG_HIP(hipMemcpyAsync(hflag, &st->iter, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
            TG_HIP(hipStreamSynchronize(s));
        }
    }
    // One iteration's launches on stream ss (every kernel returns at once when `done` is set).
    auto one_iteration = [&](hipStream_t ss) -> int {
        if (fused) {
            int rc1 = dir_spmm_on(ss);
            if (rc1) return rc1;
            hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kFinalThreads), 0, ss, k, kpad, kFusedBlocks,
                               (const double*)partial, ahist, st);
            if (dd)
                hipLaunchKernelGGL((cg_step_kernel<T, true>), dim3(B), dim3(kBlock), 0, ss, n, k, kpad, (const T*)p,
                                   (const T*)p1, (const T*)u, dd, x, r, partial, (const CGState*)st);
            else
                hipLaunchKernelGGL((cg_step_kernel<T, false>), dim3(B), dim3(kBlock), 0, ss, n, k, kpad, (const T*)p,
                                   (const T*)p1, (const T*)u, dd, x, r, partial, (const CGState*)st);
            hipLaunchKernelGGL(cg_rho_kernel, dim3(1), dim3(kFinalThreads), 0, ss, k, kpad, B, (const double*)partial,
                               st);
            return TG_OK;
        }
        if (form == kSpmmDot) {  // u = A p and partial(p.u) in one kernel
            int rc1 = dir_spmm_on(ss);
            if (rc1) return rc1;
            hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kFinalThreads), 0, ss, k, kpad, kFusedBlocks,
                               (const double*)partial, ahist, st);
        } else {
            int rc1 = op_apply(Aop, p, u, k, &st->done, ss);
            if (rc1) return rc1;
            hipLaunchKernelGGL((cg_dot_pu_kernel<T>), dim3(B), dim3(kBlock), 0, ss, n, k, kpad, (const T*)p,
                               (const T*)u, partial, (const CGState*)st);
            hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kFinalThreads), 0, ss, k, kpad, B,
                               (const double*)partial, ahist, st);
        }
        hipLaunchKernelGGL((cg_update_kernel<T>), dim3(B), dim3(kBlock), 0, ss, n, k, kpad, (const T*)u, dd, r, z,
                           zmode, partial, (const CGState*)st);
        if (zmode == 2) {
            int rc1 = apply_P();
            if (rc1) return rc1;
        }
        hipLaunchKernelGGL(cg_rho_kernel, dim3(1), dim3(kFinalThreads), 0, ss, k, kpad, B, (const double*)partial, st);
        hipLaunchKernelGGL((cg_p_kernel<T>), dim3(B), dim3(kBlock), 0, ss, n, k, kpad, (const T*)z, p, pstride, x,
                           (const double*)ahist, (const CGState*)st);
        return TG_OK;
    };
    // Blocks of kCgGraphIters iterations as one captured hipGraph (the CSR forms, whose
    // launches never allocate), captured after the first batch and replayed; batches
    // run as whole blocks where they can, the rest eagerly.  The same kernels in the
    // same order as the eager launches.
    const bool graphs = allow_graphs && g_cg_graph && form != kFourKernels && !host_ops &&
                        (!hook || std::max(1, hook->every) % kCgGraphIters == 0);
    struct IterGraph {
        hipStream_t cs = nullptr;
        hipGraphExec_t ex = nullptr;
        ~IterGraph() {
            if (ex) (void)hipGraphExecDestroy(ex);
            if (cs) (void)hipStreamDestroy(cs);
        }
    } ig;
    while (!hflag[1] && launched < max_iter) {
        int todo = std::min(batch, max_iter - launched);
        if (graphs && launched >= 4 && todo >= kCgGraphIters) {
            if (!ig.ex) {
                if (!ig.cs) TG_HIP(hipStreamCreateWithFlags(&ig.cs, hipStreamNonBlocking));
                TG_HIP(hipStreamBeginCapture(ig.cs, hipStreamCaptureModeThreadLocal));
                for (int it = 0; it < kCgGraphIters && rc == TG_OK; ++it) rc = one_iteration(ig.cs);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(ig.cs, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    break;
                }
                if (ec != hipSuccess) return fail(TG_ERR_HIP, "cg: graph capture failed: %s", hipGetErrorString(ec));
                const hipError_t ei = hipGraphInstantiate(&ig.ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    ig.ex = nullptr;
                    return fail(TG_ERR_HIP, "cg: graph instantiation failed: %s", hipGetErrorString(ei));
                }
            }
            for (; todo >= kCgGraphIters; todo -= kCgGraphIters, launched += kCgGraphIters)
                TG_HIP(hipGraphLaunch(ig.ex, s));
        }
        for (; todo > 0 && rc == TG_OK; --todo, ++launched) rc = one_iteration(s);
        if (rc) break;
        if (!fused) {  // x complete before the host looks at the state
            hipLaunchKernelGGL((cg_xfinish_kernel<T>), dim3(B), dim3(kBlock), 0, s, n, k, kpad, (const T*)p, pstride,
                               x, (const double*)ahist, (const CGState*)st);
            hipLaunchKernelGGL(cg_xclear_kernel, dim3(1), dim3(1), 0, s, st);
        }
        if (hipGetLastError() != hipSuccess) {
            rc = fail(TG_ERR_HIP, "cg: kernel launch failed");
            break;
        }
        if (hipMemcpyAsync(hflag, &st->iter, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(TG_ERR_HIP, "cg: polling the convergence flag failed");
            break;
        }
        if (hook) {
            if (hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = fail(TG_ERR_HIP, "cg: reading the solver state failed");
                break;
            }
            bool stop = false;
            rc = hook->fn(hs, r, z, x, fused ? nullptr : p_next(hs), partial, st, &stop);
            if (rc || stop) break;
            if (hipMemcpyAsync(hflag, &st->iter, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = fail(TG_ERR_HIP, "cg: polling the convergence flag failed");
                break;
            }
            continue;
        }
        if (!host_ops) batch = std::min(batch * 2, 64);
    }
    if (rc) return rc;
    const int a_one = (a == 1.0);
    const int b_mode = (b == 0.0) ? 0 : (b == 1.0) ? 1 : 2;
    hipLaunchKernelGGL((cg_write_kernel<T>), dim3(grid_for<T>(N)), dim3(kBlock), 0, s, n, k, (T)a, (const T*)x, (T)b,
                       X, ldX, a_one, b_mode);
    TG_LAUNCH_CHECK();
    TG_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    res->succeeded = hs.succeeded;
    res->iterations = hs.iter;
    for (int c = 0; c < k && c < kMaxCols; ++c) res->rel_res[c] = std::sqrt(std::fabs(hs.rho[c]) / hs.bsq[c]);
    return TG_OK;
}

template int cg_solve_t<double>(const CsrView&, const OpRef&, int, const double*, const OpRef&, int32_t, int32_t,
                                double, const double*, int64_t, double, double*, int64_t, double, int, void*, size_t,
                                CGResult*, hipStream_t, const CGHook*, bool);
template int cg_solve_t<float>(const CsrView&, const OpRef&, int, const float*, const OpRef&, int32_t, int32_t,
                               double, const float*, int64_t, double, float*, int64_t, double, int, void*, size_t,
                               CGResult*, hipStream_t, const CGHook*, bool);

}  // namespace tg

using namespace tg;

extern "C" size_t tg_cg_workspace_bytes(int32_t n, int32_t nrhs, tg_dtype val_type, int32_t prec_kind) {
    if (n < 0 || nrhs < 1) return 0;
    return cg_workspace_bytes(n, nrhs, val_type == TG_F32 ? 4 : 8, prec_kind ? 1 : 0, kMaxDefer);
}

template <typename T>
static int cg_entry(const CsrView& A, const OpRef& Aop, int32_t prec_kind, const void* prec_diag, const OpRef& Pop,
                    int32_t nrhs, int32_t max_iter, double a_coef, const void* B, int64_t ldB, double b_coef, void* X,
                    int64_t ldX, double tol, int32_t use_guess, void* workspace, size_t workspace_bytes,
                    int32_t* succeeded, int32_t* iterations, double* rel_residuals, hipStream_t s) {
    CGResult res{};
    const int rc = cg_solve_t<T>(A, Aop, prec_kind, (const T*)prec_diag, Pop, nrhs, max_iter, a_coef, (const T*)B,
                                 ldB, b_coef, (T*)X, ldX, tol, use_guess, workspace, workspace_bytes, &res, s);
    if (rc) return rc;
    if (succeeded) *succeeded = res.succeeded;
    if (iterations) *iterations = res.iterations;
    if (rel_residuals)
        for (int c = 0; c < nrhs; ++c) rel_residuals[c] = res.rel_res[c];
    return TG_OK;
}

extern "C" int tg_cg_solve(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci, const void* a,
                           tg_dtype val_type, int32_t prec_kind, const void* prec_diag, int32_t nrhs, int32_t max_iter,
                           double a_coef, const void* B, int64_t ldB, double b_coef, void* X, int64_t ldX, double tol,
                           int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                           int32_t* iterations, double* rel_residuals, void* stream) {
    if (succeeded) *succeeded = 0;
    if (iterations) *iterations = 0;
    if (n < 0 || nrhs < 1 || !B || !X || ldB < nrhs || ldX < nrhs || (prec_kind != 0 && prec_kind != 1))
        return fail(TG_ERR_INVALID, "tg_cg_solve: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_cg_solve: rp_type must be I32/I64");
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
    if (spmm_may_slab(A, nrhs) || cg_panel_wants_band(n)) A.band = csr_band(A, s);
    if (val_type == TG_F64)
        return cg_entry<double>(A, Aop, prec_kind, prec_diag, Pop, nrhs, max_iter, a_coef, B, ldB, b_coef, X, ldX, tol,
                                use_guess, workspace, workspace_bytes, succeeded, iterations, rel_residuals, s);
    if (val_type == TG_F32)
        return cg_entry<float>(A, Aop, prec_kind, prec_diag, Pop, nrhs, max_iter, a_coef, B, ldB, b_coef, X, ldX, tol,
                               use_guess, workspace, workspace_bytes, succeeded, iterations, rel_residuals, s);
    return fail(TG_ERR_DTYPE, "tg_cg_solve: val_type must be F64/F32");
}

extern "C" int tg_cg_solve_op(int32_t n, tg_apply_fn apply_A, void* ctx_A, tg_apply_fn apply_P, void* ctx_P,
                              tg_dtype val_type, int32_t nrhs, int32_t max_iter, double a_coef, const void* B,
                              int64_t ldB, double b_coef, void* X, int64_t ldX, double tol, int32_t use_guess,
                              void* workspace, size_t workspace_bytes, int32_t* succeeded, int32_t* iterations,
                              double* rel_residuals, void* stream) {
    if (succeeded) *succeeded = 0;
    if (iterations) *iterations = 0;
    if (n < 0 || nrhs < 1 || !B || !X || ldB < nrhs || ldX < nrhs || !apply_A)
        return fail(TG_ERR_INVALID, "tg_cg_solve_op: bad argument");
    CsrView A;  // shape only
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
        return cg_entry<double>(A, Aop, prec_kind, nullptr, Pop, nrhs, max_iter, a_coef, B, ldB, b_coef, X, ldX, tol,
                                use_guess, workspace, workspace_bytes, succeeded, iterations, rel_residuals, s);
    if (val_type == TG_F32)
        return cg_entry<float>(A, Aop, prec_kind, nullptr, Pop, nrhs, max_iter, a_coef, B, ldB, b_coef, X, ldX, tol,
                               use_guess, workspace, workspace_bytes, succeeded, iterations, rel_residuals, s);
    return fail(TG_ERR_DTYPE, "tg_cg_solve_op: val_type must be F64/F32");
}

// Residuals of the last solve on `workspace`: sqrt|r.z| per column
// (ConjugateGradient::Residuals, ConjugateGradient.hpp:447-456) and |b.Pb|
// (b_squared_norms), read from the device state at the head of the workspace.
extern "C" int tg_cg_residuals(const void* workspace, size_t workspace_bytes, int32_t nrhs, double* residuals,
                               double* b_squared_norms, void* stream) {
    if (!workspace || nrhs < 1 || nrhs > kMaxCols || workspace_bytes < sizeof(CGState))
        return fail(TG_ERR_INVALID, "tg_cg_residuals: bad argument");
    hipStream_t s = as_stream(stream);
    CGState hs;
    TG_HIP(hipMemcpyAsync(&hs, workspace, sizeof(CGState), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nrhs; ++c) {
        if (residuals) residuals[c] = std::sqrt(std::fabs(hs.rho[c]));
        if (b_squared_norms) b_squared_norms[c] = hs.bsq[c];
    }
    return TG_OK;
}
