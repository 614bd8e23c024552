// Batched, continuously fed CG engine: many SPD systems of uneven size with
// 1..8 right-hand sides each, advanced together.
//
// Every system runs the recurrences of ConjugateGradient::operator()
// (ConjugateGradient.hpp:105-353) with its own per-column scalars, Jacobi or
// identity preconditioner, tolerance and iteration limit.  One iteration for ALL
// active systems is two launches over the row tiles (kTileRows rows of one
// system) of every active system:
//   K1  beta = rz_i / rz_{i-1}, convergence test, p_i = z + beta p_{i-1}
//       (own rows written, neighbour rows recomputed on the fly), u = A p_i,
//       tile partials of p.u
//   K2  alpha = rz_i / (p.u), x += alpha p, r -= alpha u, z = P r, tile
//       partials of r.z
// The per-system scalars are not kept in device state between launches: every
// block of a system sums that system's tile partials itself, in one fixed order,
// so all blocks see bitwise the same alpha / beta / convergence decision and no
// block reads a value another block of the same launch writes.  p and the r.z
// partials are double-buffered by iteration parity.  A system leaves the tile
// list at the next host poll after it converged; until then its blocks return at
// once.  Systems can be added at any poll (continuous batching).
//
// Determinism: a system's reductions depend only on the system itself (its n, k
// and tile partials), so its result is bitwise the same alone and in any mix.
//
// Mixed systems run the same recurrences in FP32 in a second engine, with FP64
// residual replacement at checkpoints every 16 of the system's own iterations
// (see the replacement section below); the stopping test is on the FP64 true
// residual.

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

constexpr int kBatchMaxK = 8;
constexpr int kTileRows = 1024;  // default rows per tile (tg_tune_set "batch_tile_rows")
constexpr int kMinTileRows = 256;
static int g_batch_tile_rows = kTileRows;
static int g_batch_stream = 0;  // SpMM kernel of the engine: 0 = a row per lane (measured faster), 1 = streaming chunks
static int g_batch_split = 1;   // 1: direction pass + SpMM of p (three launches per iteration); 0: fused K1
static int g_batch_idx16 = 1;   // 1: 16-bit column offsets in the split SpMM where they fit (read at add)
static int g_batch_graph = 1;   // 1: the iterations of a step run as one captured hipGraph
// Tile order over the XCDs.  The kernels take tile xcd_swizzle(b) in workgroup b, which
// runs on XCD b % 8, so with the system-major list as it is (0, the default) each XCD
// works through a contiguous run of it and a tile's stencil neighbours (the next tiles of
// its system) are in flight on the same XCD.  1: the list permuted once more, which
// spreads each XCD's tiles over the whole list; measured 8-9% slower on 2 000 uneven
// 2D/3D systems (the neighbours then come from other XCDs' L2s).
static int g_batch_xcd_order = 0;
// SpMM kernel of the split iteration: 0 = row loop, 1 = software-pipelined row loop,
// 2 = indices and values of a 256-row segment staged through LDS with coalesced loads
// (two rows per lane in flight measured slower: 102 VGPRs, 4 waves per SIMD)
static int g_batch_spmm = 2;
constexpr int kGather = 8;     // nonzeros per gather chunk in the SpMM kernel
constexpr int kRowUnroll = 4;  // rows per step in the update kernel

template <typename T> struct SysDev {
    const int32_t* rp;
    const int32_t* ci;
    const int16_t* cid;  // ci - row as 16 bits when every |ci - row| <= 32767 (else null)
    const T* val;
    const T* d;  // Jacobi diagonal (inverse) or null
    const T* b;
    T* x;
    T* r;
    T* u;
    T* p[2];         // direction, by iteration parity (z = d * r is never stored)
    double* ppu;     // tiles * k partials of p.u
    double* prz[2];  // tiles * k partials of r.z, by parity
    int* xit;        // per tile: the iteration whose deferred x update the tile has applied
    int n, k, kpad, tiles, tile_rows;
    int check_every;  // > 0: pause at every multiple of this many iterations for the host
};

struct SysState {
    double rho[kBatchMaxK], TOL[kBatchMaxK], bsq[kBatchMaxK];
    int iter, done, succ, max_iter, nan;
    int it_a;  // iteration index, written by K1 (tile 0), read by K2
    int it_b;     // completed iterations, written by K2 (tile 0), read by K1
    int paused;   // done because of a checkpoint (written by K1 tile 0, read by the host)
    int resumed;  // iteration at which the host last resumed the system (no pause there)
    int restart;  // iteration at which the direction restarts (p = z, as at iteration 0), or -1
};

struct Tile {
    int slot, r0, r1, t;
};

// Pointer into the global address space.
template <typename T> using gptr = __attribute__((address_space(1))) T*;

// Parity selection without indexing a local array (which would live in scratch).
template <typename P> __host__ __device__ __forceinline__ P pick(P const (&a)[2], int it) {
    const P a0 = a[0], a1 = a[1];
    return (it & 1) ? a1 : a0;
}

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// ---- reductions ----------------------------------------------------------------
// Block tree over the kBlock/kpad row groups; row group 0 column c writes partial[t*k + c].
// sh: kBlock doubles of LDS (a kernel's own buffer, or tile_partial's).
__device__ __forceinline__ void tile_partial_in(double v, int k, int kpad, double* partial, int t, double* sh) {
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = rows / 2; s > 0; s >>= 1) {
        if (ro < s) sh[threadIdx.x] += sh[threadIdx.x + s * kpad];
        __syncthreads();
    }
    if (ro == 0 && c < k) partial[(int64_t)t * k + c] = sh[c];
}

__device__ __forceinline__ void tile_partial(double v, int k, int kpad, double* partial, int t) {
    __shared__ double sh[kBlock];
    tile_partial_in(v, k, kpad, partial, t, sh);
}

// Fixed-order sum over a system's tile partials (column c); every thread gets its column's total.
__device__ __forceinline__ double sys_total(const double* partial, int tiles, int k, int kpad, int c, int j, int J) {
    __shared__ double sh[kBlock];
    double s = 0.0;
    if (c < k)
        for (int t = j; t < tiles; t += J) s += partial[(int64_t)t * k + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = J / 2; h > 0; h >>= 1) {
        if (j < h) sh[threadIdx.x] += sh[threadIdx.x + h * kpad];
        __syncthreads();
    }
    const double v = sh[c];
    __syncthreads();
    return v;
}

// ---- setup: r = b, z = P r, x = 0, partial r.z (parity 0) -------------------------
// A new system's descriptor, state and tol^2 into the slot tables (arguments by
// value, so no host buffer has to outlive the launch); the 16-bit-offset flag starts
// set (bk_delta16_kernel clears it).
template <typename T>
__global__ void bk_install_kernel(SysDev<T> S, SysState z, double t2, int slot, SysDev<T>* __restrict__ sys,
                                  SysState* __restrict__ st, double* __restrict__ tol2, int* __restrict__ ok16) {
    if (threadIdx.x == 0) {
        sys[slot] = S;
        st[slot] = z;
        tol2[slot] = t2;
        *ok16 = 1;
    }
}

// The 16-bit column offsets go live when every offset fitted.
template <typename T>
__global__ void bk_cid_commit_kernel(int slot, SysDev<T>* __restrict__ sys, const int* __restrict__ ok16,
                                     const int16_t* cid) {
    if (threadIdx.x == 0 && *ok16) sys[slot].cid = cid;
}

// One workgroup per row tile of the system in `slot` (tile blockIdx.x).
template <typename T>
__global__ __launch_bounds__(kBlock) void bk_init_kernel(int slot, const SysDev<T>* __restrict__ sys) {
    const SysDev<T>& S = sys[slot];
    const int t = blockIdx.x;
    const Tile tl{slot, t * S.tile_rows, min(S.n, (t + 1) * S.tile_rows), t};
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    double acc = 0.0;
    if (c < S.k)
        for (int i = tl.r0 + ro; i < tl.r1; i += rows) {
            const int64_t e = (int64_t)i * S.k + c;
            const T rv = S.b[e];
            S.r[e] = rv;
            const T zv = S.d ? S.d[i] * rv : rv;
            S.x[e] = T(0);
            acc += (double)rv * (double)zv;
        }
    if (threadIdx.x == 0) S.xit[tl.t] = 0;
    tile_partial(acc, S.k, S.kpad, S.prz[0], tl.t);
}

// Setup totals of one system (one workgroup): rho_0, TOL = tol^2 |r.z|, initial test.
template <typename T>
__global__ __launch_bounds__(kBlock) void bk_setup_kernel(int slot, const SysDev<T>* __restrict__ sys,
                                                          SysState* __restrict__ st, const double* __restrict__ tol2) {
    SysState& ss = st[slot];
    const SysDev<T>& S = sys[slot];
    const int c = threadIdx.x % S.kpad, j = threadIdx.x / S.kpad, J = kBlock / S.kpad;
    const double tot = sys_total(S.prz[0], S.tiles, S.k, S.kpad, c, j, J);
    __shared__ int flag, nanf;
    if (threadIdx.x == 0) {
        flag = 1;
        nanf = 0;
    }
    __syncthreads();
    if (j == 0 && c < S.k) {
        ss.rho[c] = tot;
        ss.bsq[c] = fabs(tot);
        ss.TOL[c] = fabs(tot) * tol2[slot];
        if (isnan(tot)) nanf = 1;
        if (!(fabs(tot) <= ss.TOL[c])) flag = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ss.iter = 0;
        ss.it_a = 0;
        ss.it_b = 0;
        ss.paused = 0;
        ss.resumed = 0;
        ss.restart = -1;
        ss.nan = nanf;
        ss.succ = nanf ? 0 : flag;
        ss.done = (nanf || flag || ss.max_iter <= 0) ? 1 : 0;
    }
}

// ---- K1 body, streaming form ------------------------------------------------------
// The tile's rows are taken in chunks whose nonzeros fit kSlots * kBlock / kpad; a
// lane fills kSlots (nonzero, column) slots of the chunk -- first all column
// indices and values (consecutive nonzeros on consecutive lane groups, so the
// loads coalesce), then all neighbour directions d_j r_j + beta p_j -- and leaves
// the products in LDS; then the (row, column) lanes add their row's products in
// nonzero order, form their own row's direction and write p and u.  Compared with
// a row per lane, every lane has kSlots independent gathers in flight and the
// index/value loads are coalesced.  A row longer than a chunk is done by its own
// lanes.  All threads of the workgroup must call it (it synchronises).
constexpr int kSlots = 4;
constexpr int kRowSeg = 1024;  // rows of the tile whose row pointers sit in LDS at a time

template <typename T>
__device__ double k1_stream(const SysDev<T>& S, const Tile& tl, int c, int ro, int rows, bool first, T be,
                            const T* __restrict__ pold, T* __restrict__ pnew) {
    constexpr int CHUNK = kBlock * kSlots;
    __shared__ T prod[CHUNK];
    __shared__ int srp[kRowSeg + 1];
    __shared__ int sb;
    const int k = S.k, kpad = S.kpad, cap = CHUNK / kpad;
    const int32_t* __restrict__ rp = S.rp;
    const int32_t* __restrict__ ci = S.ci;
    const T* __restrict__ val = S.val;
    const T* __restrict__ r = S.r;
    const T* __restrict__ d = S.d;
    const bool col = c < k;
    double acc = 0.0;
    for (int g0 = tl.r0; g0 < tl.r1; g0 += kRowSeg) {
        const int nr = min(tl.r1 - g0, kRowSeg);
        for (int i = threadIdx.x; i <= nr; i += kBlock) srp[i] = rp[g0 + i];
        __syncthreads();
        for (int a = 0; a < nr;) {
            if (threadIdx.x == 0) {  // last row b with srp[b] - srp[a] <= cap (b > a)
                const int base = srp[a];
                int lo = a + 1, hi = nr;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (srp[mid] - base <= cap)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                sb = lo;
            }
            __syncthreads();
            const int b = sb;
            const int L0 = srp[a], cnt = srp[b] - L0;
            if (cnt <= cap) {
                int32_t jn[kSlots];
                T v[kSlots], pj[kSlots];
#pragma unroll
                for (int q = 0; q < kSlots; ++q) {
                    const int e = ro + q * rows;
                    const bool ok = col && e < cnt;
                    jn[q] = ok ? ci[L0 + e] : 0;
                    v[q] = ok ? val[L0 + e] : T(0);
                }
#pragma unroll
                for (int q = 0; q < kSlots; ++q) {
                    const int e = ro + q * rows;
                    if (col && e < cnt) {
                        const int64_t f = (int64_t)jn[q] * k + c;
                        const T zj = d ? d[jn[q]] * r[f] : r[f];
                        pj[q] = first ? zj : fma_t(be, pold[f], zj);
                    } else {
                        pj[q] = T(0);
                    }
                }
#pragma unroll
                for (int q = 0; q < kSlots; ++q) prod[threadIdx.x + q * kBlock] = v[q] * pj[q];
                __syncthreads();
                if (col)
                    for (int i = a + ro; i < b; i += rows) {
                        T sum = T(0);
                        for (int l = srp[i]; l < srp[i + 1]; ++l) sum += prod[(l - L0) * kpad + c];
                        const int gi = g0 + i;
                        const int64_t e = (int64_t)gi * k + c;
                        const T zi = d ? d[gi] * r[e] : r[e];
                        const T pi = first ? zi : fma_t(be, pold[e], zi);
                        pnew[e] = pi;
                        S.u[e] = sum;
                        acc += (double)pi * (double)sum;
                    }
            } else if (col && ro == 0) {  // one row longer than a chunk (b == a + 1)
                const int gi = g0 + a;
                T sum = T(0);
                for (int l = srp[a]; l < srp[a + 1]; ++l) {
                    const int32_t j = ci[l];
                    const int64_t f = (int64_t)j * k + c;
                    const T zj = d ? d[j] * r[f] : r[f];
                    sum += val[l] * (first ? zj : fma_t(be, pold[f], zj));
                }
                const int64_t e = (int64_t)gi * k + c;
                const T zi = d ? d[gi] * r[e] : r[e];
                const T pi = first ? zi : fma_t(be, pold[e], zi);
                pnew[e] = pi;
                S.u[e] = sum;
                acc += (double)pi * (double)sum;
            }
            __syncthreads();  // prod and sb are reused by the next chunk
            a = b;
        }
    }
    return acc;
}

// ---- 16-bit column offsets ---------------------------------------------------------
// cid[l] = ci[l] - row for every nonzero; *ok is cleared if an offset does not fit
// (the system then keeps its int32 indices).  Grid operators, the engine's
// workload, have offsets up to a plane of the grid.
__global__ __launch_bounds__(kBlock) void bk_delta16_kernel(int n, int64_t nnz, const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ ci, int16_t* __restrict__ cid,
                                                            int* __restrict__ ok) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        for (int l = rp[i]; l < rp[i + 1]; ++l) {
            if (l >= nnz) {  // rp does not match the nnz given: keep the int32 indices
                atomicAnd(ok, 0);
                break;
            }
            int dlt = ci[l] - i;
            if (dlt < -32768 || dlt > 32767) {
                atomicAnd(ok, 0);
                dlt = 0;
            }
            cid[l] = (int16_t)dlt;
        }
}

// ---- K1 body, a row per (row, column) lane -------------------------------------------
// A row's nonzeros in chunks of kGather.  The loop body has no branches: the
// index/value loads are made at clamped (always valid) positions and masked
// afterwards, and the uniform choices (Jacobi or not, first iteration or not) are
// template arguments, so the compiler issues all loads of a chunk before the first
// wait (branches inside the unrolled body split it into blocks that each wait for
// their loads).  The sum runs in nonzero order.
template <typename T, bool HAS_D, bool FIRST>
__device__ __forceinline__ double k1_rows(const SysDev<T>& S, const Tile& tl, int c, int ro, int rows, T be,
                                          const T* __restrict__ pold_, T* __restrict__ pnew_) {
    // global-address-space views: the pointers come from a table in memory, so the
    // compiler cannot tell them from LDS pointers and would use flat accesses
    const gptr<const T> pold = (gptr<const T>)pold_;
    const gptr<T> pnew = (gptr<T>)pnew_;
    const gptr<const T> r = (gptr<const T>)S.r;
    const gptr<const T> d = (gptr<const T>)S.d;
    const gptr<const int32_t> rp = (gptr<const int32_t>)S.rp;
    const gptr<const int32_t> ci = (gptr<const int32_t>)S.ci;
    const gptr<const T> val = (gptr<const T>)S.val;
    const gptr<T> u = (gptr<T>)S.u;
    const int k = S.k;
    double acc = 0.0;
    for (int i = tl.r0 + ro; i < tl.r1; i += rows) {
        const int lb = rp[i], le = rp[i + 1];
        T sum = T(0);
        for (int l0 = lb; l0 < le; l0 += kGather) {
            int32_t jj[kGather];
            T vv[kGather], pj[kGather];
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const int lq = min(l0 + q, le - 1);
                const int32_t cq = ci[lq];
                const T vq = val[lq];
                const bool ok = l0 + q < le;
                jj[q] = ok ? cq : i;
                vv[q] = ok ? vq : T(0);
            }
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const int64_t f = (int64_t)jj[q] * k + c;
                T zj = r[f];
                if constexpr (HAS_D) zj = d[jj[q]] * zj;
                if constexpr (FIRST)
                    pj[q] = zj;
                else
                    pj[q] = fma_t(be, pold[f], zj);
            }
#pragma unroll
            for (int q = 0; q < kGather; ++q) sum += (l0 + q < le) ? vv[q] * pj[q] : T(0);
        }
        const int64_t e = (int64_t)i * k + c;
        T zi = r[e];
        if constexpr (HAS_D) zi = d[i] * zi;
        T pi;
        if constexpr (FIRST)
            pi = zi;
        else
            pi = fma_t(be, pold[e], zi);
        pnew[e] = pi;
        u[e] = sum;
        acc += (double)pi * (double)sum;
    }
    return acc;
}

// ---- K1: direction + SpMM + partial p.u -----------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void bk_dir_spmm_kernel(const Tile* __restrict__ tiles, const SysDev<T>* __restrict__ sys,
                                                             SysState* __restrict__ st, int streaming) {
    const Tile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    SysState& ss = st[tl.slot];
    const SysDev<T>& S = sys[tl.slot];
    const int it = ss.it_b;
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    const double rz = sys_total(pick(S.prz, it), S.tiles, S.k, S.kpad, c, ro, rows);
    const double rz_prev = it > 0 ? sys_total(pick(S.prz, it + 1), S.tiles, S.k, S.kpad, c, ro, rows) : 1.0;
    __shared__ int conv;
    if (threadIdx.x == 0) conv = 1;
    __syncthreads();
    if (ro == 0 && c < S.k && !(fabs(rz) <= ss.TOL[c])) conv = 0;
    __syncthreads();
    if (tl.t == 0 && ro == 0 && c < S.k) ss.rho[c] = rz;  // current r.z, read by the host between polls
    // Checkpoints depend only on the system's own iteration count, so the host's
    // decisions there do not depend on when it polls or what else runs.
    const bool pause = S.check_every > 0 && it > 0 && it % S.check_every == 0 && it != ss.resumed;
    if (conv || it >= ss.max_iter || pause) {
        if (tl.t == 0 && threadIdx.x == 0) {
            ss.iter = it;
            ss.succ = conv;
            ss.paused = (!conv && it < ss.max_iter) ? 1 : 0;
            ss.done = 1;
        }
        return;
    }
    if (tl.t == 0 && threadIdx.x == 0) ss.it_a = it;
    double acc = 0.0;
    if (streaming) {
        acc = k1_stream<T>(S, tl, c, ro, rows, it == 0 || it == ss.restart, (T)(rz / rz_prev), pick(S.p, it + 1),
                           pick(S.p, it));
    } else if (c < S.k) {
        const T be = (T)(rz / rz_prev);
        const T* pold = pick(S.p, it + 1);
        T* pnew = pick(S.p, it);
        if (S.d)
            acc = (it == 0 || it == ss.restart) ? k1_rows<T, true, true>(S, tl, c, ro, rows, be, pold, pnew)
                          : k1_rows<T, true, false>(S, tl, c, ro, rows, be, pold, pnew);
        else
            acc = (it == 0 || it == ss.restart) ? k1_rows<T, false, true>(S, tl, c, ro, rows, be, pold, pnew)
                          : k1_rows<T, false, false>(S, tl, c, ro, rows, be, pold, pnew);
    }
    tile_partial(acc, S.k, S.kpad, S.ppu, tl.t);
}

// ---- split form: Kdir (direction) + Kspmm (u = A p, partial p.u) ----------------
// The fused K1 gathers r, p_prev (and d) for every nonzero to rebuild p of the
// neighbour rows; the split form writes p for the system's own rows in a
// streaming pass and gathers only p in the SpMM, which halves the gathered bytes
// for one more launch per iteration.  The arithmetic of p, u and the partials is
// the same as the fused form's (p = fma(beta, p_prev, z) per entry; the row sums
// in nonzero order), so both forms give bitwise the same iterates.
// XUPD: also x += alpha p_old, the x update of the previous iteration that the
// update kernel leaves to this pass (which reads p_old anyway).
template <typename T, bool HAS_D, bool FIRST, bool XUPD>
__device__ __forceinline__ void dir_rows(const SysDev<T>& S, const Tile& tl, int c, int ro, int rows, T be, T al,
                                         const T* __restrict__ pold_, T* __restrict__ pnew_) {
    const gptr<const T> pold = (gptr<const T>)pold_;
    const gptr<T> pnew = (gptr<T>)pnew_;
    const gptr<const T> r = (gptr<const T>)S.r;
    const gptr<const T> d = (gptr<const T>)S.d;
    const gptr<T> x = (gptr<T>)S.x;
    const int k = S.k;
    for (int i0 = tl.r0 + ro; i0 < tl.r1; i0 += kRowUnroll * rows) {
        T rv[kRowUnroll], pv[kRowUnroll], dv[kRowUnroll], xv[kRowUnroll];
#pragma unroll
        for (int q = 0; q < kRowUnroll; ++q) {
            const int i = min(i0 + q * rows, tl.r1 - 1);  // clamped: always a valid row
            const int64_t e = (int64_t)i * k + c;
            rv[q] = r[e];
            if constexpr (!FIRST) pv[q] = pold[e];
            if constexpr (HAS_D) dv[q] = d[i];
            if constexpr (XUPD) xv[q] = x[e];
        }
#pragma unroll
        for (int q = 0; q < kRowUnroll; ++q) {
            const int i = i0 + q * rows;
            if (i < tl.r1) {
                const int64_t e = (int64_t)i * k + c;
                T z = rv[q];
                if constexpr (HAS_D) z = dv[q] * z;
                T pi;
                if constexpr (FIRST)
                    pi = z;
                else
                    pi = fma_t(be, pv[q], z);
                pnew[e] = pi;
                if constexpr (XUPD) x[e] = fma_t(al, pv[q], xv[q]);
            }
        }
    }
}

// x += alpha p_old alone: the deferred x update of a system that stops or pauses here.
template <typename T>
__device__ __forceinline__ void x_rows(const SysDev<T>& S, const Tile& tl, int c, int ro, int rows, T al,
                                       const T* __restrict__ pold_) {
    const gptr<const T> pold = (gptr<const T>)pold_;
    const gptr<T> x = (gptr<T>)S.x;
    const int k = S.k;
    for (int i = tl.r0 + ro; i < tl.r1; i += rows) {
        const int64_t e = (int64_t)i * k + c;
        x[e] = fma_t(al, pold[e], x[e]);
    }
}

// IDX16: column indices from the 16-bit offsets (2 bytes per nonzero instead of 4).
template <typename T, bool IDX16>
__device__ __forceinline__ double spmm_dot_rows(const SysDev<T>& S, const Tile& tl, int c, int ro, int rows,
                                                const T* __restrict__ p_, double acc = 0.0) {
    const gptr<const T> p = (gptr<const T>)p_;
    const gptr<const int32_t> rp = (gptr<const int32_t>)S.rp;
    const gptr<const int32_t> ci = (gptr<const int32_t>)S.ci;
    const gptr<const int16_t> cid = (gptr<const int16_t>)S.cid;
    const gptr<const T> val = (gptr<const T>)S.val;
    const gptr<T> u = (gptr<T>)S.u;
    const int k = S.k;
    for (int i = tl.r0 + ro; i < tl.r1; i += rows) {
        const int lb = rp[i], le = rp[i + 1];
        const int64_t e = (int64_t)i * k + c;
        const T pi = p[e];  // issued with the index loads
        T sum = T(0);
        for (int l0 = lb; l0 < le; l0 += kGather) {
            int32_t jj[kGather];
            T vv[kGather], pj[kGather];
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const int lq = min(l0 + q, le - 1);
                int32_t cq;
                if constexpr (IDX16)
                    cq = i + (int32_t)cid[lq];
                else
                    cq = ci[lq];
                const T vq = val[lq];
                const bool ok = l0 + q < le;
                jj[q] = ok ? cq : i;
                vv[q] = ok ? vq : T(0);
            }
#pragma unroll
            for (int q = 0; q < kGather; ++q) pj[q] = p[(int64_t)jj[q] * k + c];
#pragma unroll
            for (int q = 0; q < kGather; ++q) sum += (l0 + q < le) ? vv[q] * pj[q] : T(0);
        }
        u[e] = sum;
        acc += (double)pi * (double)sum;
    }
    return acc;
}

// The software-pipelined form of spmm_dot_rows.  The row pointers of a segment of
// the tile sit in LDS (rps, staged by the caller), and a lane issues the column
// indices and values of its next row before the gathers of the current one, so a
// row costs about one memory latency instead of three (row pointers, then
// indices, then gathers).  Same operations in the same order as spmm_dot_rows:
// bitwise the same u and partials.
template <typename T, bool IDX16>
__device__ __forceinline__ void row_chunk_indices(const SysDev<T>& S, int i, int lb, int le, int32_t (&jj)[kGather],
                                                  T (&vv)[kGather]) {
    const gptr<const int32_t> ci = (gptr<const int32_t>)S.ci;
    const gptr<const int16_t> cid = (gptr<const int16_t>)S.cid;
    const gptr<const T> val = (gptr<const T>)S.val;
    if (le > lb) {
#pragma unroll
        for (int q = 0; q < kGather; ++q) {
            const int lq = min(lb + q, le - 1);
            int32_t cq;
            if constexpr (IDX16)
                cq = i + (int32_t)cid[lq];
            else
                cq = ci[lq];
            const T vq = val[lq];
            const bool ok = lb + q < le;
            jj[q] = ok ? cq : i;
            vv[q] = ok ? vq : T(0);
        }
    } else {
#pragma unroll
        for (int q = 0; q < kGather; ++q) {
            jj[q] = i;
            vv[q] = T(0);
        }
    }
}

template <typename T, bool IDX16>
__device__ __forceinline__ double spmm_dot_rows_pipe(const SysDev<T>& S, int s0, int s1, const int* rps, int c,
                                                     int ro, int rows, const T* __restrict__ p_, double acc) {
    const gptr<const T> p = (gptr<const T>)p_;
    const gptr<const int32_t> ci = (gptr<const int32_t>)S.ci;
    const gptr<const int16_t> cid = (gptr<const int16_t>)S.cid;
    const gptr<const T> val = (gptr<const T>)S.val;
    const gptr<T> u = (gptr<T>)S.u;
    const int k = S.k;
    int i = s0 + ro;
    if (i >= s1) return acc;
    int32_t jA[kGather];
    T vA[kGather];
    row_chunk_indices<T, IDX16>(S, i, rps[i - s0], rps[i - s0 + 1], jA, vA);
    for (; i < s1; i += rows) {
        const int in = i + rows;
        int32_t jB[kGather];
        T vB[kGather];
        if (in < s1) row_chunk_indices<T, IDX16>(S, in, rps[in - s0], rps[in - s0 + 1], jB, vB);
        const int lb = rps[i - s0], le = rps[i - s0 + 1];
        const int64_t e = (int64_t)i * k + c;
        const T pi = p[e];
        T pj[kGather];
#pragma unroll
        for (int q = 0; q < kGather; ++q) pj[q] = p[(int64_t)jA[q] * k + c];
        T sum = T(0);
#pragma unroll
        for (int q = 0; q < kGather; ++q) sum += (lb + q < le) ? vA[q] * pj[q] : T(0);
        for (int l0 = lb + kGather; l0 < le; l0 += kGather) {  // rows longer than one chunk
            int32_t jj[kGather];
            T vv[kGather];
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const int lq = min(l0 + q, le - 1);
                int32_t cq;
                if constexpr (IDX16)
                    cq = i + (int32_t)cid[lq];
                else
                    cq = ci[lq];
                const T vq = val[lq];
                const bool ok = l0 + q < le;
                jj[q] = ok ? cq : i;
                vv[q] = ok ? vq : T(0);
            }
#pragma unroll
            for (int q = 0; q < kGather; ++q) pj[q] = p[(int64_t)jj[q] * k + c];
#pragma unroll
            for (int q = 0; q < kGather; ++q) sum += (l0 + q < le) ? vv[q] * pj[q] : T(0);
        }
        u[e] = sum;
        acc += (double)pi * (double)sum;
#pragma unroll
        for (int q = 0; q < kGather; ++q) {
            jA[q] = jB[q];
            vA[q] = vB[q];
        }
    }
    return acc;
}

constexpr int kPipeSeg = 2048;  // rows per LDS segment of row pointers in the pipelined SpMM
constexpr int kLdsRows = 256;   // rows per segment of the LDS-staged SpMM (a multiple of every kBlock / kpad)
constexpr int kLdsNnz = 1792;   // nonzeros of a segment that fit the staging buffers (7 per row)

// The rows [s0, s1) of a segment whose 16-bit column offsets and values sit in LDS
// from nonzero rps[0] on: the gathers and the arithmetic of spmm_dot_rows, the index
// loads from LDS.  acc continues the lane's partial.
template <typename T>
__device__ __forceinline__ double spmm_dot_rows_lds(const SysDev<T>& S, int s0, int s1, const int* rps,
                                                    const int16_t* cls, const T* vls, int c, int ro, int rows,
                                                    const T* __restrict__ p_, double acc) {
    const gptr<const T> p = (gptr<const T>)p_;
    const gptr<T> u = (gptr<T>)S.u;
    const int k = S.k;
    const int base = rps[0];
    for (int i = s0 + ro; i < s1; i += rows) {
        const int lb = rps[i - s0] - base, le = rps[i - s0 + 1] - base;
        const int64_t e = (int64_t)i * k + c;
        const T pi = p[e];
        T sum = T(0);
        for (int l0 = lb; l0 < le; l0 += kGather) {
            T pj[kGather];
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const int lq = min(l0 + q, le - 1);
                const int32_t jj = l0 + q < le ? i + (int32_t)cls[lq] : i;
                // 32-bit byte offset from the block-uniform base (n k sizeof(T) < 4 GiB is
                // checked at add): one address register per gather instead of two
                const uint32_t off = ((uint32_t)jj * (uint32_t)k + (uint32_t)c) * (uint32_t)sizeof(T);
                pj[q] = *(gptr<const T>)((gptr<const char>)p + off);
            }
            // the values only now, from LDS: fewer registers live across the gathers
#pragma unroll
            for (int q = 0; q < kGather; ++q) {
                const bool ok = l0 + q < le;
                const T vv = ok ? vls[min(l0 + q, le - 1)] : T(0);
                sum += ok ? vv * pj[q] : T(0);
            }
        }
        u[e] = sum;
        acc += (double)pi * (double)sum;
    }
    return acc;
}

// The rows of spmm_dot_rows one nonzero at a time, with few registers: the rare paths
// of the LDS-staged kernel (systems with int32 indices, segments with more nonzeros
// than the staging buffers) would otherwise set that kernel's register count.  The
// same products and additions in the same order (unfused: the chunked loop adds a
// masked product through a select, which the compiler does not contract); it only
// skips the +0 that loop adds for masked slots, which changes no value but the sign
// of a zero sum.
template <typename T, bool IDX16>
__device__ __attribute__((noinline)) double spmm_dot_rows_lean(const SysDev<T>& S, int r0, int r1, int c, int ro,
                                                               int rows, const T* __restrict__ p_, double acc) {
    const gptr<const T> p = (gptr<const T>)p_;
    const gptr<const int32_t> rp = (gptr<const int32_t>)S.rp;
    const gptr<const int32_t> ci = (gptr<const int32_t>)S.ci;
    const gptr<const int16_t> cid = (gptr<const int16_t>)S.cid;
    const gptr<const T> val = (gptr<const T>)S.val;
    const gptr<T> u = (gptr<T>)S.u;
    const int k = S.k;
    for (int i = r0 + ro; i < r1; i += rows) {
        const int64_t e = (int64_t)i * k + c;
        const T pi = p[e];
        T sum = T(0);
        for (int l = rp[i], le = rp[i + 1]; l < le; ++l) {
#pragma clang fp contract(off)
            const int32_t j = IDX16 ? i + (int32_t)cid[l] : ci[l];
            sum += val[l] * p[(int64_t)j * k + c];
        }
        u[e] = sum;
        acc += (double)pi * (double)sum;
    }
    return acc;
}

// Kdir: beta, the stop test and the checkpoint pause as in K1; then p for the own
// rows, together with the previous iteration's x += alpha p (deferred from the
// update kernel).  A tile applies that x update once per iteration (xit), also
// when the system stops or pauses here; before a checkpoint the update kernel has
// done it already.  So x is whole whenever the host reads it.
template <typename T>
__global__ __launch_bounds__(kBlock) void bk_dir_kernel(const Tile* __restrict__ tiles, const SysDev<T>* __restrict__ sys,
                                                        SysState* __restrict__ st) {
    const Tile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    SysState& ss = st[tl.slot];
    const SysDev<T>& S = sys[tl.slot];
    const int it = ss.it_b;
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    const double rz = sys_total(pick(S.prz, it), S.tiles, S.k, S.kpad, c, ro, rows);
    const double rz_prev = it > 0 ? sys_total(pick(S.prz, it + 1), S.tiles, S.k, S.kpad, c, ro, rows) : 1.0;
    __shared__ int conv, xdo;
    if (threadIdx.x == 0) {
        conv = 1;
        xdo = it > 0 && S.xit[tl.t] != it;
    }
    __syncthreads();
    if (threadIdx.x == 0 && xdo) S.xit[tl.t] = it;
    // alpha of iteration it - 1 from the same partials and sums the update kernel used
    const double pu = xdo ? sys_total(S.ppu, S.tiles, S.k, S.kpad, c, ro, rows) : 1.0;
    const T al = (T)(rz_prev / pu);
    if (ro == 0 && c < S.k && !(fabs(rz) <= ss.TOL[c])) conv = 0;
    __syncthreads();
    if (tl.t == 0 && ro == 0 && c < S.k) ss.rho[c] = rz;
    const bool pause = S.check_every > 0 && it > 0 && it % S.check_every == 0 && it != ss.resumed;
    const T* pold = pick(S.p, it + 1);
    if (conv || it >= ss.max_iter || pause) {
        if (tl.t == 0 && threadIdx.x == 0) {
            ss.iter = it;
            ss.succ = conv;
            ss.paused = (!conv && it < ss.max_iter) ? 1 : 0;
            ss.done = 1;
        }
        if (xdo && c < S.k) x_rows<T>(S, tl, c, ro, rows, al, pold);
        return;
    }
    if (tl.t == 0 && threadIdx.x == 0) ss.it_a = it;
    // p = z at iteration 0 and where the host restarted the direction (after a pause,
    // so the deferred x update is done: xdo is false there)
    const bool first = it == 0 || (it == ss.restart && !xdo);
    if (c < S.k) {
        const T be = (T)(rz / rz_prev);
        T* pnew = pick(S.p, it);
        if (S.d) {
            if (first)
                dir_rows<T, true, true, false>(S, tl, c, ro, rows, be, al, pold, pnew);
            else if (xdo)
                dir_rows<T, true, false, true>(S, tl, c, ro, rows, be, al, pold, pnew);
            else
                dir_rows<T, true, false, false>(S, tl, c, ro, rows, be, al, pold, pnew);
        } else {
            if (first)
                dir_rows<T, false, true, false>(S, tl, c, ro, rows, be, al, pold, pnew);
            else if (xdo)
                dir_rows<T, false, false, true>(S, tl, c, ro, rows, be, al, pold, pnew);
            else
                dir_rows<T, false, false, false>(S, tl, c, ro, rows, be, al, pold, pnew);
        }
    }
}

// Kspmm: u = A p and the tile partial of p.u (systems stopped by Kdir skip).
template <typename T>
__global__ __launch_bounds__(kBlock) void bk_spmm_dot_kernel(const Tile* __restrict__ tiles,
                                                             const SysDev<T>* __restrict__ sys,
                                                             SysState* __restrict__ st) {
    const Tile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    SysState& ss = st[tl.slot];
    if (ss.done) return;
    const SysDev<T>& S = sys[tl.slot];
    const int it = ss.it_a;
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    double acc = 0.0;
    if (c < S.k)
        acc = S.cid ? spmm_dot_rows<T, true>(S, tl, c, ro, rows, pick(S.p, it))
                    : spmm_dot_rows<T, false>(S, tl, c, ro, rows, pick(S.p, it));
    tile_partial(acc, S.k, S.kpad, S.ppu, tl.t);
}

// Kspmm with the column indices and values staged through LDS.  A row per lane
// group keeps the gathers of neighbouring rows in the same cache lines, but its
// index and value loads are strided (each lane reads its own row's nonzeros), so
// every load instruction touches many lines and each line is touched once per
// slot; here the workgroup first copies a 256-row segment's nonzeros with
// consecutive lanes on consecutive nonzeros, then the row loop reads them from LDS.
// Systems with 16-bit column offsets only (the others, and a segment with more than
// kLdsNnz nonzeros, take the global loads of spmm_dot_rows_lean).  Rows, lanes and
// arithmetic are those of spmm_dot_rows: bitwise the same u and partials.
template <typename T>
__global__ __launch_bounds__(kBlock) void bk_spmm_dot_lds_kernel(const Tile* __restrict__ tiles,
                                                                 const SysDev<T>* __restrict__ sys,
                                                                 SysState* __restrict__ st) {
    __shared__ int rps[kLdsRows + 1];
    __shared__ int16_t cls[kLdsNnz];
    // the staged values; after the last segment the tile partial's reduction buffer
    __shared__ __align__(16) char vbuf[kLdsNnz * sizeof(T) > kBlock * sizeof(double) ? kLdsNnz * sizeof(T)
                                                                                      : kBlock * sizeof(double)];
    T* vls = (T*)vbuf;
    const Tile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    SysState& ss = st[tl.slot];
    if (ss.done) return;
    const SysDev<T>& S = sys[tl.slot];
    const int it = ss.it_a;
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    const gptr<const int32_t> rp = (gptr<const int32_t>)S.rp;
    const gptr<const int16_t> cid = (gptr<const int16_t>)S.cid;
    const gptr<const T> val = (gptr<const T>)S.val;
    const T* p = pick(S.p, it);
    double acc = 0.0;
    if (!S.cid) {
        if (c < S.k) acc = spmm_dot_rows_lean<T, false>(S, tl.r0, tl.r1, c, ro, rows, p, acc);
        tile_partial_in(acc, S.k, S.kpad, S.ppu, tl.t, (double*)vbuf);
        return;
    }
    for (int s0 = tl.r0; s0 < tl.r1; s0 += kLdsRows) {
        const int s1 = min(s0 + kLdsRows, tl.r1);
        for (int j = threadIdx.x; j <= s1 - s0; j += kBlock) rps[j] = rp[s0 + j];
        __syncthreads();
        const int base = rps[0], cnt = rps[s1 - s0] - base;
        if (cnt <= kLdsNnz) {
            for (int t = threadIdx.x; t < cnt; t += kBlock) {
                cls[t] = cid[base + t];
                vls[t] = val[base + t];
            }
            __syncthreads();
            if (c < S.k) acc = spmm_dot_rows_lds<T>(S, s0, s1, rps, cls, vls, c, ro, rows, p, acc);
        } else if (c < S.k) {
            acc = spmm_dot_rows_lean<T, true>(S, s0, s1, c, ro, rows, p, acc);
        }
        __syncthreads();
    }
    tile_partial_in(acc, S.k, S.kpad, S.ppu, tl.t, (double*)vbuf);
}

// Kspmm, software-pipelined: the tile's row pointers go through LDS in segments.
// A segment's rows are those of spmm_dot_rows in the same per-lane order (a lane
// takes rows ro, ro + rows, ... of the tile; segments are multiples of rows long),
// so each lane adds the same terms to its partial in the same order.
template <typename T>
__global__ __launch_bounds__(kBlock) void bk_spmm_dot_pipe_kernel(const Tile* __restrict__ tiles,
                                                                  const SysDev<T>* __restrict__ sys,
                                                                  SysState* __restrict__ st) {
    __shared__ int rps[kPipeSeg + 1];
    const Tile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    SysState& ss = st[tl.slot];
    if (ss.done) return;
    const SysDev<T>& S = sys[tl.slot];
    const int it = ss.it_a;
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    const gptr<const int32_t> rp = (gptr<const int32_t>)S.rp;
    const T* p = pick(S.p, it);
    double acc = 0.0;
    for (int s0 = tl.r0; s0 < tl.r1; s0 += kPipeSeg) {
        const int s1 = min(s0 + kPipeSeg, tl.r1);
        for (int j = threadIdx.x; j <= s1 - s0; j += kBlock) rps[j] = rp[s0 + j];
        __syncthreads();
        if (c < S.k)  // the lane's partial continues across segments, as in one pass over the tile
            acc = S.cid ? spmm_dot_rows_pipe<T, true>(S, s0, s1, rps, c, ro, rows, p, acc)
                        : spmm_dot_rows_pipe<T, false>(S, s0, s1, rps, c, ro, rows, p, acc);
        __syncthreads();
    }
    tile_partial(acc, S.k, S.kpad, S.ppu, tl.t);
}

// ---- K2 body: kRowUnroll rows per step, all loads first, then the stores ----------
// DEFER_X: x is left to the next direction pass (bk_dir_kernel): r and u only.
template <typename T, bool HAS_D, bool DEFER_X>
__device__ __forceinline__ double k2_rows(const SysDev<T>& S, const Tile& tl, int c, int ro, int rows, T al,
                                          const T* __restrict__ p_) {
    const gptr<const T> p = (gptr<const T>)p_;
    const gptr<const T> u = (gptr<const T>)S.u;
    const gptr<const T> d = (gptr<const T>)S.d;
    const gptr<T> x = (gptr<T>)S.x;
    const gptr<T> r = (gptr<T>)S.r;
    const int k = S.k;
    double acc = 0.0;
    for (int i0 = tl.r0 + ro; i0 < tl.r1; i0 += kRowUnroll * rows) {
        T pv[kRowUnroll], xv[kRowUnroll], rv[kRowUnroll], uv[kRowUnroll], dv[kRowUnroll];
#pragma unroll
        for (int q = 0; q < kRowUnroll; ++q) {
            const int i = min(i0 + q * rows, tl.r1 - 1);  // clamped: always a valid row
            const int64_t e = (int64_t)i * k + c;
            if constexpr (!DEFER_X) {
                pv[q] = p[e];
                xv[q] = x[e];
            }
            rv[q] = r[e];
            uv[q] = u[e];
            if constexpr (HAS_D) dv[q] = d[i];
        }
#pragma unroll
        for (int q = 0; q < kRowUnroll; ++q) {
            const int i = i0 + q * rows;
            if (i < tl.r1) {
                const int64_t e = (int64_t)i * k + c;
                if constexpr (!DEFER_X) x[e] = fma_t(al, pv[q], xv[q]);
                const T rn = rv[q] - al * uv[q];
                r[e] = rn;
                T zv = rn;
                if constexpr (HAS_D) zv = dv[q] * rn;
                acc += (double)rn * (double)zv;
            }
        }
    }
    return acc;
}

// ---- K2: x, r, z update + partial r.z (next parity) -------------------------------
template <typename T, bool DEFER_X>
__global__ __launch_bounds__(kBlock) void bk_update_kernel(const Tile* __restrict__ tiles, const SysDev<T>* __restrict__ sys,
                                                           SysState* __restrict__ st) {
    const Tile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    SysState& ss = st[tl.slot];
    if (ss.done) return;
    const SysDev<T>& S = sys[tl.slot];
    const int it = ss.it_a;
    const int c = threadIdx.x % S.kpad, ro = threadIdx.x / S.kpad, rows = kBlock / S.kpad;
    const double rz = sys_total(pick(S.prz, it), S.tiles, S.k, S.kpad, c, ro, rows);
    const double pu = sys_total(S.ppu, S.tiles, S.k, S.kpad, c, ro, rows);
    // Before a checkpoint the x update is done here: the host may read x there
    // without the direction pass having run (a poll that lands on the checkpoint).
    const bool xnow = !DEFER_X || (S.check_every > 0 && (it + 1) % S.check_every == 0);
    double acc = 0.0;
    if (c < S.k) {
        const T al = (T)(rz / pu);
        if (xnow)
            acc = S.d ? k2_rows<T, true, false>(S, tl, c, ro, rows, al, pick(S.p, it))
                      : k2_rows<T, false, false>(S, tl, c, ro, rows, al, pick(S.p, it));
        else
            acc = S.d ? k2_rows<T, true, true>(S, tl, c, ro, rows, al, pick(S.p, it))
                      : k2_rows<T, false, true>(S, tl, c, ro, rows, al, pick(S.p, it));
    }
    if (DEFER_X && xnow && threadIdx.x == 0) S.xit[tl.t] = it + 1;  // the direction pass of it + 1 skips it
    tile_partial(acc, S.k, S.kpad, pick(S.prz, it + 1), tl.t);
    if (tl.t == 0 && threadIdx.x == 0) ss.it_b = it + 1;
}

struct Resume {
    int slot;
    int it;  // the system's completed iterations: no checkpoint pause there
    double scale;
    int restart;  // 1: the direction restarts there (p = z)
};

// One workgroup per resumed system: tolerance scaled, pause cleared, iterating again.
__global__ void bk_resume_kernel(const Resume* __restrict__ list, SysState* __restrict__ st) {
    const Resume rs = list[blockIdx.x];
    SysState& ss = st[rs.slot];
    if (threadIdx.x < kBatchMaxK) ss.TOL[threadIdx.x] *= rs.scale;
    if (threadIdx.x == 0) {
        ss.resumed = rs.it;
        if (rs.restart) ss.restart = rs.it;
        ss.paused = 0;
        ss.done = 0;
        ss.succ = 0;
    }
}

// ---- host engine ------------------------------------------------------------------
static int ceil_div(int a, int b) { return (a + b - 1) / b; }
static int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// One engine per vector type.  Slots are fixed per system while it is active.
template <typename T> class CGBatch {
  public:
    explicit CGBatch(hipStream_t s) : s_(s) {}
    ~CGBatch() { release(); }

    // Adds a system; its setup runs at once.  b, x: n x k device blocks (ld = k).
    int add(int n, int64_t nnz, const int32_t* rp, const int32_t* ci, const T* val, const T* d, int k, const T* b,
            T* x, double tol, int max_iter, int* slot_out, int check_every = 0) {
        if (n < 1 || k < 1 || k > kBatchMaxK) return fail(TG_ERR_INVALID, "batch: n >= 1 and 1 <= nrhs <= 8");
        if ((uint64_t)n * (uint64_t)k * sizeof(T) >= (1ull << 32))  // the SpMM addresses vectors by 32-bit byte offsets
            return fail(TG_ERR_INVALID, "batch: n * nrhs * sizeof(scalar) must be below 4 GiB (use the single-system solver)");
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
        SysDev<T> S{};
        S.rp = rp;
        S.ci = ci;
        S.val = val;
        S.d = d;
        S.b = b;
        S.x = x;
        S.n = n;
        S.k = k;
        S.check_every = check_every;
        S.kpad = pow2_at_least(k);
        S.tile_rows = g_batch_tile_rows;
        S.tiles = ceil_div(n, S.tile_rows);
        const size_t N = (size_t)n * k;
        const size_t nvec = 4;  // r, u, p0, p1
        const size_t vb = align256(N * sizeof(T)), pb = align256((size_t)S.tiles * k * sizeof(double));
        const size_t ib = align256((size_t)std::max<int64_t>(nnz, 1) * sizeof(int16_t));
        const size_t xb = align256((size_t)S.tiles * sizeof(int));
        const size_t bytes = nvec * vb + 3 * pb + ib + 256 + xb;
        char* mem = nullptr;
        TG_HIP(hipMallocAsync((void**)&mem, bytes, s_));
        int16_t* cid = (int16_t*)(mem + nvec * vb + 3 * pb);
        int* dok = (int*)(mem + nvec * vb + 3 * pb + ib);
        S.xit = (int*)(mem + nvec * vb + 3 * pb + ib + 256);
        S.r = (T*)mem;
        S.u = (T*)(mem + vb);
        S.p[0] = (T*)(mem + 2 * vb);
        S.p[1] = (T*)(mem + 3 * vb);
        S.ppu = (double*)(mem + nvec * vb);
        S.prz[0] = (double*)(mem + nvec * vb + pb);
        S.prz[1] = (double*)(mem + nvec * vb + 2 * pb);
        mem_[slot] = mem;
        bytes_[slot] = bytes;
        used_[slot] = 1;
        parked_[slot] = 0;  // a reused slot starts iterating again
        host_[slot] = S;
        SysState z{};
        z.max_iter = std::max(0, max_iter);
        // Setup for this system alone, all stream-ordered with no host wait: the
        // descriptor and state by kernel arguments, r = b, x = 0, the r.z partials
        // and totals, then the 16-bit column offsets (used by the split iteration's
        // SpMM when every offset fits; the device decides).
        hipLaunchKernelGGL((bk_install_kernel<T>), dim3(1), dim3(kWave), 0, s_, S, z, tol * tol, slot, dsys_, dst_,
                           dtol2_, dok);
        hipLaunchKernelGGL((bk_init_kernel<T>), dim3((unsigned)S.tiles), dim3(kBlock), 0, s_, slot,
                           (const SysDev<T>*)dsys_);
        hipLaunchKernelGGL((bk_setup_kernel<T>), dim3(1), dim3(kBlock), 0, s_, slot, (const SysDev<T>*)dsys_, dst_,
                           (const double*)dtol2_);
        hipLaunchKernelGGL(bk_delta16_kernel, dim3((unsigned)std::min(ceil_div(n, kBlock), 4096)), dim3(kBlock), 0,
                           s_, n, nnz, rp, ci, cid, dok);
        if (g_batch_idx16)
            hipLaunchKernelGGL((bk_cid_commit_kernel<T>), dim3(1), dim3(kWave), 0, s_, slot, dsys_, (const int*)dok,
                               (const int16_t*)cid);
        TG_LAUNCH_CHECK();
        dirty_ = true;
        *slot_out = slot;
        return TG_OK;
    }

    // Enqueues `iters` iterations for every active (not done) system, then polls.
    int flush_resumes() {
        if (resumes_.empty()) return TG_OK;
        if (resumes_.size() > resume_cap_) {
            if (dresume_) TG_HIP(hipFreeAsync(dresume_, s_));
            resume_cap_ = std::max(resumes_.size(), 2 * resume_cap_);
            TG_HIP(hipMallocAsync((void**)&dresume_, resume_cap_ * sizeof(Resume), s_));
        }
        TG_HIP(hipMemcpyAsync(dresume_, resumes_.data(), resumes_.size() * sizeof(Resume), hipMemcpyHostToDevice, s_));
        hipLaunchKernelGGL(bk_resume_kernel, dim3((unsigned)resumes_.size()), dim3(kWave), 0, s_, dresume_, dst_);
        TG_LAUNCH_CHECK();
        // the host list is read by the copy: keep it until the stream passes it
        TG_HIP(hipStreamSynchronize(s_));
        resumes_.clear();
        return TG_OK;
    }

    int step(int iters) {
        int rc0 = flush_resumes();
        if (rc0) return rc0;
        if (dirty_) {
            int rc = rebuild();
            if (rc) return rc;
        }
        if (ntiles_ == 0) return TG_OK;
        if (!g_batch_graph) {
            launch_iters(s_, iters);
            TG_LAUNCH_CHECK();
            return TG_OK;
        }
        // The iterations as one hipGraph: captured on a private stream (capture is
        // not allowed on the legacy default stream a caller may pass), launched on
        // s_, captured again when the tile list, the buffers or the shape change.
        const GraphKey key{dtiles_, dsys_, dst_, ntiles_, iters, g_batch_split, g_batch_stream, g_batch_spmm};
        if (!gexec_ || !(key == gkey_)) {
            if (gexec_) (void)hipGraphExecDestroy(gexec_);
            gexec_ = nullptr;
            if (!cap_s_) TG_HIP(hipStreamCreateWithFlags(&cap_s_, hipStreamNonBlocking));
            TG_HIP(hipStreamBeginCapture(cap_s_, hipStreamCaptureModeThreadLocal));
            launch_iters(cap_s_, iters);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(cap_s_, &g);
            if (ec != hipSuccess) return fail(TG_ERR_HIP, "batch: graph capture failed: %s", hipGetErrorString(ec));
            const hipError_t ei = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                gexec_ = nullptr;
                return fail(TG_ERR_HIP, "batch: graph instantiation failed: %s", hipGetErrorString(ei));
            }
            gkey_ = key;
        }
        TG_HIP(hipGraphLaunch(gexec_, s_));
        return TG_OK;
    }

    void launch_iters(hipStream_t s, int iters) {
        for (int it = 0; it < iters; ++it) {
            if (g_batch_split) {
                hipLaunchKernelGGL((bk_dir_kernel<T>), dim3(ntiles_), dim3(kBlock), 0, s, dtiles_, dsys_, dst_);
                if (g_batch_spmm == 2)
                    hipLaunchKernelGGL((bk_spmm_dot_lds_kernel<T>), dim3(ntiles_), dim3(kBlock), 0, s, dtiles_,
                                       dsys_, dst_);
                else if (g_batch_spmm == 1)
                    hipLaunchKernelGGL((bk_spmm_dot_pipe_kernel<T>), dim3(ntiles_), dim3(kBlock), 0, s, dtiles_,
                                       dsys_, dst_);
                else
                    hipLaunchKernelGGL((bk_spmm_dot_kernel<T>), dim3(ntiles_), dim3(kBlock), 0, s, dtiles_, dsys_,
                                       dst_);
            } else {
                hipLaunchKernelGGL((bk_dir_spmm_kernel<T>), dim3(ntiles_), dim3(kBlock), 0, s, dtiles_, dsys_, dst_,
                                   g_batch_stream);
            }
            if (g_batch_split)
                hipLaunchKernelGGL((bk_update_kernel<T, true>), dim3(ntiles_), dim3(kBlock), 0, s, dtiles_, dsys_, dst_);
            else
                hipLaunchKernelGGL((bk_update_kernel<T, false>), dim3(ntiles_), dim3(kBlock), 0, s, dtiles_, dsys_, dst_);
        }
    }

    // Reads the states of all used slots (synchronises).
    int poll(std::vector<SysState>& out) {
        int rc0 = flush_resumes();
        if (rc0) return rc0;
        out.resize(used_.size());
        if (!used_.empty())
            TG_HIP(hipMemcpyAsync(out.data(), dst_, used_.size() * sizeof(SysState), hipMemcpyDeviceToHost, s_));
        TG_HIP(hipStreamSynchronize(s_));
        return TG_OK;
    }

    // Frees a slot (after its result has been consumed).
    int retire(int slot) {
        if (slot < 0 || slot >= (int)used_.size() || !used_[slot]) return fail(TG_ERR_INVALID, "batch: bad slot");
        TG_HIP(hipFreeAsync(mem_[slot], s_));
        mem_[slot] = nullptr;
        bytes_[slot] = 0;
        used_[slot] = 0;
        dirty_ = true;
        return TG_OK;
    }

    // Marks a slot as no longer iterating (its tiles leave the list).
    void park(int slot) {
        if (slot >= 0 && slot < (int)parked_.size()) parked_[slot] = 1;
        dirty_ = true;
    }
    void unpark(int slot) {
        if (slot >= 0 && slot < (int)parked_.size()) parked_[slot] = 0;
        dirty_ = true;
    }

    size_t bytes_in_use() const {
        size_t s = 0;
        for (size_t b : bytes_) s += b;
        return s;
    }
    int active_tiles() const { return ntiles_; }
    int capacity() const { return (int)cap_; }
    const SysDev<T>& desc(int slot) const { return host_[slot]; }

    // Continues a system the device marked done, with its tolerance scaled by `scale`.
    // Queued; applied by one launch before the next iterations (or poll).
    void resume(int slot, int it, double scale, bool restart = false) {
        resumes_.push_back(Resume{slot, it, scale, restart ? 1 : 0});
    }

    void mark_done_host(int slot) {
        if (slot >= 0 && slot < (int)parked_.size()) parked_[slot] = 1;
        dirty_ = true;
    }

  private:
    void tiles_of(int slot, std::vector<Tile>& tl) const {
        const SysDev<T>& S = host_[slot];
        for (int t = 0; t < S.tiles; ++t)
            tl.push_back(Tile{slot, t * S.tile_rows, std::min(S.n, (t + 1) * S.tile_rows), t});
    }

    int grow(int need) {
        int cap = std::max(need, 2 * (int)cap_);
        if (cap <= (int)cap_) return TG_OK;
        SysDev<T>* nsys = nullptr;
        SysState* nst = nullptr;
        double* ntol = nullptr;
        TG_HIP(hipMallocAsync((void**)&nsys, cap * sizeof(SysDev<T>), s_));
        TG_HIP(hipMallocAsync((void**)&nst, cap * sizeof(SysState), s_));
        TG_HIP(hipMallocAsync((void**)&ntol, cap * sizeof(double), s_));
        if (cap_) {
            TG_HIP(hipMemcpyAsync(nsys, dsys_, cap_ * sizeof(SysDev<T>), hipMemcpyDeviceToDevice, s_));
            TG_HIP(hipMemcpyAsync(nst, dst_, cap_ * sizeof(SysState), hipMemcpyDeviceToDevice, s_));
            TG_HIP(hipMemcpyAsync(ntol, dtol2_, cap_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
            TG_HIP(hipFreeAsync(dsys_, s_));
            TG_HIP(hipFreeAsync(dst_, s_));
            TG_HIP(hipFreeAsync(dtol2_, s_));
        }
        dsys_ = nsys;
        dst_ = nst;
        dtol2_ = ntol;
        cap_ = cap;
        used_.resize(cap, 0);
        parked_.resize(cap, 0);
        mem_.resize(cap, nullptr);
        bytes_.resize(cap, 0);
        host_.resize(cap);
        return TG_OK;
    }

    // Tile list of every used, not parked, not done system (host reads the flags).
    int rebuild() {
        std::vector<SysState> states;
        int rc = poll(states);
        if (rc) return rc;
        std::vector<Tile> tl;
        for (int i = 0; i < (int)used_.size(); ++i)
            if (used_[i] && !parked_[i] && !states[i].done) tiles_of(i, tl);
        if (g_batch_xcd_order && tl.size() > (size_t)kXcds) {
            // The spread order: position b of the list gets the tile at position
            // xcd_swizzle(b) of the system-major order, the same mapping the kernels apply
            // once more, so each XCD's tiles come from all eight eighths of the list.  Every
            // reduction adds tile partials by tile index, so the results do not change.
            // (the positions b < nt with b % 8 == x are exactly q + [x < r] many, so this
            // is a permutation)
            const int nt = (int)tl.size(), q = nt / kXcds, r = nt % kXcds;
            std::vector<Tile> perm(tl.size());
            for (int b = 0; b < nt; ++b) {
                const int x = b % kXcds, idx = b / kXcds;
                perm[b] = tl[x * q + std::min(x, r) + idx];
            }
            tl.swap(perm);
        }
        if (tl.size() > tiles_cap_) {
            if (dtiles_) TG_HIP(hipFreeAsync(dtiles_, s_));
            tiles_cap_ = std::max(tl.size(), 2 * tiles_cap_);
            TG_HIP(hipMallocAsync((void**)&dtiles_, tiles_cap_ * sizeof(Tile), s_));
        }
        if (!tl.empty()) TG_HIP(hipMemcpyAsync(dtiles_, tl.data(), tl.size() * sizeof(Tile), hipMemcpyHostToDevice, s_));
        TG_HIP(hipStreamSynchronize(s_));  // host vectors go out of scope
        ntiles_ = (int)tl.size();
        dirty_ = false;
        return TG_OK;
    }

    void release() {
        for (auto* m : mem_)
            if (m) (void)hipFreeAsync(m, s_);
        if (dsys_) (void)hipFreeAsync(dsys_, s_);
        if (dst_) (void)hipFreeAsync(dst_, s_);
        if (dtol2_) (void)hipFreeAsync(dtol2_, s_);
        if (dtiles_) (void)hipFreeAsync(dtiles_, s_);
        if (dresume_) (void)hipFreeAsync(dresume_, s_);
        (void)hipStreamSynchronize(s_);
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
        if (cap_s_) (void)hipStreamDestroy(cap_s_);
        cap_s_ = nullptr;
    }

    struct GraphKey {
        const void* tiles;
        const void* sys;
        const void* st;
        int ntiles, iters, split, stream, pipe;
        bool operator==(const GraphKey& o) const {
            return tiles == o.tiles && sys == o.sys && st == o.st && ntiles == o.ntiles && iters == o.iters &&
                   split == o.split && stream == o.stream && pipe == o.pipe;
        }
    };
    hipStream_t cap_s_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    GraphKey gkey_{};

    hipStream_t s_;
    size_t cap_ = 0;
    std::vector<Resume> resumes_;
    Resume* dresume_ = nullptr;
    size_t resume_cap_ = 0;
    std::vector<char> used_, parked_;
    std::vector<char*> mem_;
    std::vector<size_t> bytes_;
    std::vector<SysDev<T>> host_;
    SysDev<T>* dsys_ = nullptr;
    SysState* dst_ = nullptr;
    double* dtol2_ = nullptr;
    Tile* dtiles_ = nullptr;
    size_t tiles_cap_ = 0;
    int ntiles_ = 0;
    bool dirty_ = true;
};

// ---- FP64 residual replacement for mixed systems ----------------------------------
// A mixed system runs FP32 CG in the engine (A32, r, z, p, u and the increment x32
// in FP32; reductions in FP64).  Between polls the host replaces the recursive
// residual by the true one (reliable updates: van der Vorst & Ye, "Residual
// replacement strategies for Krylov subspace iterative methods for the
// convergence of true residuals", SIAM J. Sci. Comput. 22 (2000); the
// mixed-precision form of Clark et al., Comput. Phys. Commun. 181 (2010)):
//   r = b - A64 (x64 + x32) in FP64, r32 = float(r), new r.z
//   partials; then x64 += x32, x32 = 0.
// The search direction p is kept, so CG continues instead of restarting.

// r = b - A (x64 + x32) for every system being replaced, one workgroup per (system,
// row tile of the engine's tiling); writes r32, the r.z partials of the engine's
// current parity and the ||r||^2 partials.  x64 and x32 are only read here.
struct ReplaceDesc {
    const int32_t* rp;
    const int32_t* ci;
    const double* val;
    const double* b;
    double* x64;
    float* x32;
    const float* d32;
    float* r32;      // the FP32 engine's r (mixed systems)
    double* r64;     // the FP64 engine's r (FP64 systems with the true-residual stop)
    const double* d64;
    double* prz;     // the engine's r.z partials of the current parity
    double* pnorm;   // ||r||^2 partials
    double* norms;   // k true-residual norms out
    int n, k, kpad, tiles;
};

struct ReplaceTile {
    int desc, r0, r1, t;
};

// F32: mixed systems (x = x64 + x32, the replaced residual goes to the FP32
// engine's r); otherwise FP64 systems (x = x64 is the engine's x, r in FP64).
template <bool F32>
__global__ __launch_bounds__(kBlock) void bm_replace_many_kernel(const ReplaceTile* __restrict__ tiles,
                                                                 const ReplaceDesc* __restrict__ descs) {
    const ReplaceTile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    const ReplaceDesc& D = descs[tl.desc];
    const int k = D.k, kpad = D.kpad;
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    double acc_rz = 0.0, acc_n = 0.0;
    if (c < k) {
        const int32_t* __restrict__ rp = D.rp;
        const int32_t* __restrict__ ci = D.ci;
        const double* __restrict__ val = D.val;
        const double* __restrict__ x64 = D.x64;
        const float* __restrict__ x32 = D.x32;
        for (int i = tl.r0 + ro; i < tl.r1; i += rows) {
            const int lb = rp[i], le = rp[i + 1];
            double ax = 0.0;
            for (int l0 = lb; l0 < le; l0 += kGather) {
                int32_t jj[kGather];
                double vv[kGather], xv[kGather];
#pragma unroll
                for (int q = 0; q < kGather; ++q) {  // clamped loads, masked values: no branches
                    const int lq = min(l0 + q, le - 1);
                    const int32_t cq = ci[lq];
                    const double vq = val[lq];
                    const bool ok = l0 + q < le;
                    jj[q] = ok ? cq : i;
                    vv[q] = ok ? vq : 0.0;
                }
#pragma unroll
                for (int q = 0; q < kGather; ++q) {
                    const int64_t f = (int64_t)jj[q] * k + c;
                    xv[q] = F32 ? x64[f] + (double)x32[f] : x64[f];
                }
#pragma unroll
                for (int q = 0; q < kGather; ++q) ax += (l0 + q < le) ? vv[q] * xv[q] : 0.0;
            }
            const int64_t e = (int64_t)i * k + c;
            const double r = D.b[e] - ax;
            if constexpr (F32) {
                const float rf = (float)r;
                D.r32[e] = rf;
                const float zf = D.d32 ? D.d32[i] * rf : rf;
                acc_rz += (double)rf * (double)zf;
            } else {
                D.r64[e] = r;
                const double z = D.d64 ? D.d64[i] * r : r;
                acc_rz += r * z;
            }
            acc_n += r * r;
        }
    }
    tile_partial(acc_rz, k, kpad, D.prz, tl.t);
    __syncthreads();
    tile_partial(acc_n, k, kpad, D.pnorm, tl.t);
}

// x64 += x32, x32 = 0 on the replaced systems' tiles
__global__ __launch_bounds__(kBlock) void bm_flush_many_kernel(const ReplaceTile* __restrict__ tiles,
                                                               const ReplaceDesc* __restrict__ descs) {
    const ReplaceTile tl = tiles[xcd_swizzle(blockIdx.x, gridDim.x)];
    const ReplaceDesc& D = descs[tl.desc];
    const int64_t e0 = (int64_t)tl.r0 * D.k, e1 = (int64_t)tl.r1 * D.k;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
        D.x64[e] += (double)D.x32[e];
        D.x32[e] = 0.0f;
    }
}

// one workgroup per replaced system: sqrt of the tile-partial totals
__global__ __launch_bounds__(kBlock) void bm_norm_many_kernel(const ReplaceDesc* __restrict__ descs) {
    const ReplaceDesc& D = descs[blockIdx.x];
    const int c = threadIdx.x % D.kpad, j = threadIdx.x / D.kpad, J = kBlock / D.kpad;
    const double t = sys_total(D.pnorm, D.tiles, D.k, D.kpad, c, j, J);
    if (j == 0 && c < D.k) D.norms[c] = sqrt(t);
}

__global__ __launch_bounds__(kBlock) void bm_norm_total_kernel(int tiles, int k, int kpad, const double* __restrict__ partial,
                                                               double* __restrict__ out) {
    const int c = threadIdx.x % kpad, j = threadIdx.x / kpad, J = kBlock / kpad;
    const double t = sys_total(partial, tiles, k, kpad, c, j, J);
    if (j == 0 && c < k) out[c] = sqrt(t);
}

// ||b||^2 partials per tile
__global__ __launch_bounds__(kBlock) void bm_bnorm_kernel(int n, int k, int kpad, const double* __restrict__ b,
                                                          double* __restrict__ partial) {
    const int r0 = blockIdx.x * kTileRows, r1 = min(n, r0 + kTileRows);
    const int c = threadIdx.x % kpad, ro = threadIdx.x / kpad, rows = kBlock / kpad;
    double acc = 0.0;
    if (c < k)
        for (int i = r0 + ro; i < r1; i += rows) {
            const double v = b[(int64_t)i * k + c];
            acc += v * v;
        }
    tile_partial(acc, k, kpad, partial, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void bm_zero_kernel(int64_t count, double* __restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        x[e] = 0.0;
}

__global__ __launch_bounds__(kBlock) void bm_tofloat_kernel(int64_t count, const double* __restrict__ a,
                                                            float* __restrict__ b) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        b[e] = (float)a[e];
}

constexpr double kReplaceDrop = 0.1;  // replace when the residual estimate fell by this factor
constexpr int kFlatChecks = 8;        // ... or did not fall at this many checkpoints in a row
constexpr int kCheckEvery = 16;       // mixed systems pause every 16 iterations for that test
// knobs "batch_replace_drop_pct" / "batch_check_every": both are read when a system is added
static double g_replace_drop = kReplaceDrop;
static int g_check_every = kCheckEvery;

// ---- the batch object behind the C ABI ------------------------------------------
struct BatchSystem {
    int64_t id;
    int mixed;  // 0: FP64, the reference's stop test; 1: mixed; 2: FP64, true-residual stop
    int n, k;
    int64_t nnz;
    const int32_t* rp;
    const int32_t* ci;
    const double* val64;
    const double* d64;
    const double* b64;
    double* x64;  // caller's X
    double tol;
    int max_iter;
    int slot = -1;  // engine slot (FP64 engine for plain systems, FP32 engine for mixed ones)
    // mixed
    int replacements = 0;
    double est_last = 1.0;   // preconditioned residual estimate at the last replacement
    double est_check = 0.0;  // the estimate at the previous checkpoint
    int flat = 0;            // consecutive checkpoints at which the estimate did not fall
    bool restart_next = false;  // the pending replacement was forced by stagnation: restart p
    double replace_drop = kReplaceDrop;
    int check_every = kCheckEvery;
    char* mem = nullptr;     // a32, d32, b32, x32, norm partials, norms
    size_t mem_bytes = 0;
    float* a32 = nullptr;
    float* d32 = nullptr;
    float* b32 = nullptr;
    float* x32 = nullptr;
    double* pnorm = nullptr;
    double* norms = nullptr;  // k
    std::vector<double> bnorm;
    bool bnorm_pending = false;  // ||b|| computed on the device, not read yet
    // results
    int finished = 0, succeeded = 0, iterations = 0;
    std::vector<double> rel;
};

// Collects the residual replacements of one poll; run() launches them all at once
// (replacement, fold of x32 into x64, norms) and reads every norm back in one copy.
struct ReplaceBatch {
    std::vector<ReplaceDesc> descs;
    std::vector<ReplaceTile> tiles;  // tiles of the mixed systems, then of the FP64 ones
    size_t tiles32 = 0;              // how many of them are mixed
    std::vector<ReplaceTile> tiles64;
    std::vector<BatchSystem*> sys;
    ReplaceDesc* ddesc = nullptr;
    ReplaceTile* dtile = nullptr;
    size_t desc_cap = 0, tile_cap = 0;

    // a mixed system (FP32 engine)
    void add(const SysDev<float>& E, double* norms_out, BatchSystem& S, int it) {
        ReplaceDesc D{};
        D.x32 = S.x32;
        D.d32 = S.d32;
        D.r32 = E.r;
        add_common(D, E.prz, E.tiles, E.tile_rows, norms_out, S, it, tiles);
    }
    // an FP64 system with the true-residual stop (FP64 engine)
    void add(const SysDev<double>& E, double* norms_out, BatchSystem& S, int it) {
        ReplaceDesc D{};
        D.r64 = E.r;
        D.d64 = S.d64;
        add_common(D, E.prz, E.tiles, E.tile_rows, norms_out, S, it, tiles64);
    }
    void add_common(ReplaceDesc& D, double* const (&prz)[2], int etiles, int tile_rows, double* norms_out,
                    BatchSystem& S, int it, std::vector<ReplaceTile>& out) {
        D.rp = S.rp;
        D.ci = S.ci;
        D.val = S.val64;
        D.b = S.b64;
        D.x64 = S.x64;
        D.prz = pick(prz, it);
        D.pnorm = S.pnorm;
        D.norms = norms_out;
        D.n = S.n;
        D.k = S.k;
        D.kpad = pow2_at_least(S.k);
        D.tiles = etiles;
        const int di = (int)descs.size();
        for (int t = 0; t < etiles; ++t)
            out.push_back(ReplaceTile{di, t * tile_rows, std::min(S.n, (t + 1) * tile_rows), t});
        descs.push_back(D);
        sys.push_back(&S);
        ++S.replacements;
    }

    int run(hipStream_t s) {
        if (descs.empty()) return TG_OK;
        tiles32 = tiles.size();
        tiles.insert(tiles.end(), tiles64.begin(), tiles64.end());
        if (descs.size() > desc_cap) {
            if (ddesc) TG_HIP(hipFreeAsync(ddesc, s));
            desc_cap = std::max(descs.size(), 2 * desc_cap);
            TG_HIP(hipMallocAsync((void**)&ddesc, desc_cap * sizeof(ReplaceDesc), s));
        }
        if (tiles.size() > tile_cap) {
            if (dtile) TG_HIP(hipFreeAsync(dtile, s));
            tile_cap = std::max(tiles.size(), 2 * tile_cap);
            TG_HIP(hipMallocAsync((void**)&dtile, tile_cap * sizeof(ReplaceTile), s));
        }
        TG_HIP(hipMemcpyAsync(ddesc, descs.data(), descs.size() * sizeof(ReplaceDesc), hipMemcpyHostToDevice, s));
        TG_HIP(hipMemcpyAsync(dtile, tiles.data(), tiles.size() * sizeof(ReplaceTile), hipMemcpyHostToDevice, s));
        if (tiles32) {
            hipLaunchKernelGGL(bm_replace_many_kernel<true>, dim3((unsigned)tiles32), dim3(kBlock), 0, s,
                               (const ReplaceTile*)dtile, (const ReplaceDesc*)ddesc);
            hipLaunchKernelGGL(bm_flush_many_kernel, dim3((unsigned)tiles32), dim3(kBlock), 0, s,
                               (const ReplaceTile*)dtile, (const ReplaceDesc*)ddesc);
        }
        if (tiles.size() > tiles32)
            hipLaunchKernelGGL(bm_replace_many_kernel<false>, dim3((unsigned)(tiles.size() - tiles32)), dim3(kBlock),
                               0, s, (const ReplaceTile*)(dtile + tiles32), (const ReplaceDesc*)ddesc);
        hipLaunchKernelGGL(bm_norm_many_kernel, dim3((unsigned)descs.size()), dim3(kBlock), 0, s,
                           (const ReplaceDesc*)ddesc);
        TG_LAUNCH_CHECK();
        return TG_OK;
    }

    void clear() {
        descs.clear();
        tiles.clear();
        tiles64.clear();
        tiles32 = 0;
        sys.clear();
    }

    void release(hipStream_t s) {
        if (ddesc) (void)hipFreeAsync(ddesc, s);
        if (dtile) (void)hipFreeAsync(dtile, s);
        ddesc = nullptr;
        dtile = nullptr;
    }
};

struct Batch {
    hipStream_t s;
    CGBatch<double> b64;
    CGBatch<float> b32;
    std::map<int64_t, std::unique_ptr<BatchSystem>> sys;
    std::vector<int64_t> finished_queue;
    int64_t next_id = 1;
    double* dnorms = nullptr;  // kBatchMaxK true-residual norms per FP32-engine slot
    int dnorms_cap = 0;
    std::vector<double> hnorms;
    double* dnorms64 = nullptr;  // the same per FP64-engine slot (true-residual FP64 systems)
    int dnorms64_cap = 0;
    std::vector<double> hnorms64;
    ReplaceBatch rb;
    explicit Batch(hipStream_t st) : s(st), b64(st), b32(st) {}
    ~Batch() {
        if (dnorms) (void)hipFreeAsync(dnorms, s);
        if (dnorms64) (void)hipFreeAsync(dnorms64, s);
        rb.release(s);
    }
};


constexpr int kMaxReplacements = 200;

static int finish(Batch& B, BatchSystem& S, int succeeded, int iterations) {
    S.finished = 1;
    S.succeeded = succeeded;
    S.iterations = iterations;
    B.finished_queue.push_back(S.id);
    return TG_OK;
}

static int grid_of(int64_t c) { return (int)std::max<int64_t>(1, std::min<int64_t>((c + kBlock - 1) / kBlock, 8192)); }

static int mixed_start(Batch& B, BatchSystem& S) {
    const size_t N = (size_t)S.n * S.k;
    const int tiles = ceil_div(S.n, kTileRows);        // the ||b|| pass
    const int ptiles = ceil_div(S.n, kMinTileRows);    // any engine tiling
    const size_t bytes = align256((size_t)S.nnz * 4) + (S.d64 ? align256((size_t)S.n * 4) : 0) + 2 * align256(N * 4) +
                         align256((size_t)ptiles * S.k * 8) + align256((size_t)S.k * 8);
    TG_HIP(hipMallocAsync((void**)&S.mem, bytes, B.s));
    S.mem_bytes = bytes;
    char* p = S.mem;
    S.a32 = (float*)p;
    p += align256((size_t)S.nnz * 4);
    if (S.d64) {
        S.d32 = (float*)p;
        p += align256((size_t)S.n * 4);
    }
    S.b32 = (float*)p;
    p += align256(N * 4);
    S.x32 = (float*)p;
    p += align256(N * 4);
    S.pnorm = (double*)p;
    p += align256((size_t)ptiles * S.k * 8);
    S.norms = (double*)p;
    hipLaunchKernelGGL(bm_tofloat_kernel, dim3(grid_of(S.nnz)), dim3(kBlock), 0, B.s, S.nnz, S.val64, S.a32);
    if (S.d64) hipLaunchKernelGGL(bm_tofloat_kernel, dim3(grid_of(S.n)), dim3(kBlock), 0, B.s, (int64_t)S.n, S.d64, S.d32);
    hipLaunchKernelGGL(bm_tofloat_kernel, dim3(grid_of((int64_t)N)), dim3(kBlock), 0, B.s, (int64_t)N, S.b64, S.b32);
    hipLaunchKernelGGL(bm_zero_kernel, dim3(grid_of((int64_t)N)), dim3(kBlock), 0, B.s, (int64_t)N, S.x64);
    const int kpad = pow2_at_least(S.k);
    hipLaunchKernelGGL(bm_bnorm_kernel, dim3(tiles), dim3(kBlock), 0, B.s, S.n, S.k, kpad, S.b64, S.pnorm);
    hipLaunchKernelGGL(bm_norm_total_kernel, dim3(1), dim3(kBlock), 0, B.s, tiles, S.k, kpad, (const double*)S.pnorm,
                       S.norms);
    TG_LAUNCH_CHECK();
    S.bnorm.assign(S.k, 0.0);
    S.bnorm_pending = true;  // read at the next poll (a NaN in b stops the engine's setup itself)
    // The engine's own test (preconditioned FP32 estimate <= tol) only pauses the
    // system; the host then decides on the FP64 true residual.
    return B.b32.add(S.n, S.nnz, S.rp, S.ci, S.a32, S.d32, S.k, S.b32, S.x32, S.tol, S.max_iter, &S.slot,
                     S.check_every);
}

// FP64 system with the true-residual stop: the FP64 engine iterates on the caller's
// arrays with checkpoints; the host replaces the residual there (FP64) and stops
// on ||b - A x|| / ||b|| <= tol, exactly as for mixed systems.
static int true_start(Batch& B, BatchSystem& S) {
    const int tiles = ceil_div(S.n, kTileRows);
    const int ptiles = ceil_div(S.n, kMinTileRows);
    const size_t bytes = align256((size_t)ptiles * S.k * 8) + align256((size_t)S.k * 8);
    TG_HIP(hipMallocAsync((void**)&S.mem, bytes, B.s));
    S.mem_bytes = bytes;
    S.pnorm = (double*)S.mem;
    S.norms = (double*)(S.mem + align256((size_t)ptiles * S.k * 8));
    const int kpad = pow2_at_least(S.k);
    hipLaunchKernelGGL(bm_bnorm_kernel, dim3(tiles), dim3(kBlock), 0, B.s, S.n, S.k, kpad, S.b64, S.pnorm);
    hipLaunchKernelGGL(bm_norm_total_kernel, dim3(1), dim3(kBlock), 0, B.s, tiles, S.k, kpad, (const double*)S.pnorm,
                       S.norms);
    TG_LAUNCH_CHECK();
    S.bnorm.assign(S.k, 0.0);
    S.bnorm_pending = true;
    return B.b64.add(S.n, S.nnz, S.rp, S.ci, S.val64, S.d64, S.k, S.b64, S.x64, S.tol, S.max_iter, &S.slot,
                     S.check_every);
}

static int batch_advance(Batch& B, int iters) {
    int rc = B.b64.step(iters);
    if (rc) return rc;
    rc = B.b32.step(iters);
    if (rc) return rc;
    std::vector<SysState> s64, s32;
    rc = B.b64.poll(s64);
    if (rc) return rc;
    rc = B.b32.poll(s32);
    if (rc) return rc;
    // ||b|| of the systems added since the last poll (before any replacement needs it)
    bool pend = false;
    for (auto& kv : B.sys) {
        BatchSystem& S = *kv.second;
        if (!S.bnorm_pending || S.finished) continue;
        TG_HIP(hipMemcpyAsync(S.bnorm.data(), S.norms, S.k * 8, hipMemcpyDeviceToHost, B.s));
        pend = true;
    }
    if (pend) {
        TG_HIP(hipStreamSynchronize(B.s));
        for (auto& kv : B.sys) kv.second->bnorm_pending = false;
    }
    if (B.dnorms_cap < B.b32.capacity()) {  // sized before any replacement of this poll writes it
        if (B.dnorms) TG_HIP(hipFreeAsync(B.dnorms, B.s));
        B.dnorms_cap = std::max(B.b32.capacity(), 2 * B.dnorms_cap);
        TG_HIP(hipMallocAsync((void**)&B.dnorms, (size_t)B.dnorms_cap * kBatchMaxK * sizeof(double), B.s));
    }
    if (B.dnorms64_cap < B.b64.capacity()) {
        if (B.dnorms64) TG_HIP(hipFreeAsync(B.dnorms64, B.s));
        B.dnorms64_cap = std::max(B.b64.capacity(), 2 * B.dnorms64_cap);
        TG_HIP(hipMallocAsync((void**)&B.dnorms64, (size_t)B.dnorms64_cap * kBatchMaxK * sizeof(double), B.s));
    }
    // park / resume in the engine the system runs in
    const auto park = [&](BatchSystem& S) {
        if (S.mixed == 1)
            B.b32.park(S.slot);
        else
            B.b64.park(S.slot);
    };
    const auto resume = [&](BatchSystem& S, int it, double scale, bool restart) {
        if (S.mixed == 1)
            B.b32.resume(S.slot, it, scale, restart);
        else
            B.b64.resume(S.slot, it, scale, restart);
    };
    for (auto& kv : B.sys) {
        BatchSystem& S = *kv.second;
        if (S.finished || S.slot < 0) continue;
        if (S.mixed == 0) {
            const SysState& st = s64[S.slot];
            if (st.done) {
                for (int c = 0; c < S.k; ++c) S.rel[c] = st.bsq[c] > 0 ? std::sqrt(std::fabs(st.rho[c]) / st.bsq[c]) : 0.0;
                B.b64.park(S.slot);
                rc = finish(B, S, st.nan ? 0 : st.succ, st.iter);
                if (rc) return rc;
            }
            continue;
        }
        const SysState& st = S.mixed == 1 ? s32[S.slot] : s64[S.slot];
        if (st.nan) {
            park(S);
            rc = finish(B, S, 0, st.iter);
            if (rc) return rc;
            continue;
        }
        // Decisions only at the system's own checkpoints (every check_every iterations:
        // paused there by the device, or standing exactly there at this poll) and stops.
        const bool stopped = st.done && !st.paused;
        const bool at_check = st.done ? (bool)st.paused
                                      : (st.it_b > 0 && st.it_b % S.check_every == 0 && st.it_b != st.resumed);
        if (!at_check && !stopped) continue;
        double est = 0.0;
        for (int c = 0; c < S.k; ++c) est = std::max(est, st.bsq[c] > 0 ? std::sqrt(std::fabs(st.rho[c]) / st.bsq[c]) : 0.0);
        // a replacement when the estimate fell replace_drop-fold since the last one, or
        // when it stagnated (no decrease at kFlatChecks checkpoints in a row: the FP32
        // recursion at its rounding floor, which only a fresh true residual moves on)
        const bool flat_now = !stopped && S.est_check > 0.0 && !(est < S.est_check);
        S.flat = flat_now ? S.flat + 1 : 0;
        S.est_check = est;
        S.restart_next = !stopped && !(est < S.replace_drop * S.est_last);  // forced by stagnation
        if (S.restart_next && S.flat < kFlatChecks) {
            resume(S, st.it_b, 1.0, false);
            continue;
        }
        if (S.mixed == 1)
            B.rb.add(B.b32.desc(S.slot), B.dnorms + (size_t)S.slot * kBatchMaxK, S, st.it_b);
        else
            B.rb.add(B.b64.desc(S.slot), B.dnorms64 + (size_t)S.slot * kBatchMaxK, S, st.it_b);
        S.est_last = est;
        S.flat = 0;
        S.est_check = 0.0;  // the replaced residual's estimate is not known here: start over
    }
    if (B.rb.sys.empty()) return TG_OK;
    rc = B.rb.run(B.s);
    if (rc) return rc;
    B.hnorms.resize((size_t)B.dnorms_cap * kBatchMaxK);
    TG_HIP(hipMemcpyAsync(B.hnorms.data(), B.dnorms, B.hnorms.size() * sizeof(double), hipMemcpyDeviceToHost, B.s));
    B.hnorms64.resize((size_t)B.dnorms64_cap * kBatchMaxK);
    TG_HIP(hipMemcpyAsync(B.hnorms64.data(), B.dnorms64, B.hnorms64.size() * sizeof(double), hipMemcpyDeviceToHost,
                          B.s));
    TG_HIP(hipStreamSynchronize(B.s));
    std::vector<BatchSystem*> replaced = B.rb.sys;
    B.rb.clear();
    for (BatchSystem* Sp : replaced) {
        BatchSystem& S = *Sp;
        const SysState& st = S.mixed == 1 ? s32[S.slot] : s64[S.slot];
        const std::vector<double>& hn = S.mixed == 1 ? B.hnorms : B.hnorms64;
        double worst = 0.0;
        for (int c = 0; c < S.k; ++c) {
            const double rn = hn[(size_t)S.slot * kBatchMaxK + c];
            S.rel[c] = S.bnorm[c] > 0 ? rn / S.bnorm[c] : 0.0;
            worst = std::max(worst, S.rel[c]);
        }
        if (worst <= S.tol) {
            park(S);
            rc = finish(B, S, 1, st.it_b);
        } else if (st.it_b >= S.max_iter || S.replacements >= kMaxReplacements || std::isnan(worst)) {
            park(S);
            rc = finish(B, S, 0, st.it_b);
        } else if (st.done && !st.paused) {
            // the engine's estimate met the tolerance but the true residual did not: tighten
            // it and restart the direction (the old p has vanished; the replaced residual's
            // larger r.z over its tiny p.u would throw x far off)
            const double f = S.tol / worst;
            resume(S, st.it_b, std::min(0.25, std::max(1e-4, 0.25 * f * f)), true);
        } else {
            // a replacement forced by stagnation also restarts the direction (p = z): the
            // old p no longer spans anything new (a small system's Krylov space is used up)
            resume(S, st.it_b, 1.0, S.restart_next);
        }
        if (rc) return rc;
    }
    return TG_OK;
}

int batch_tune_set(const char* key, int32_t value, bool* known) {
    *known = false;
    if (std::string(key) == "batch_replace_drop_pct") {
        *known = true;
        if (value < 1 || value > 99) return fail(TG_ERR_INVALID, "tg_tune_set: batch_replace_drop_pct in [1, 99]");
        g_replace_drop = value / 100.0;
        return TG_OK;
    }
    if (std::string(key) == "batch_check_every") {
        *known = true;
        if (value < 1 || value > 1024) return fail(TG_ERR_INVALID, "tg_tune_set: batch_check_every in [1, 1024]");
        g_check_every = value;
        return TG_OK;
    }
    if (std::string(key) == "batch_xcd_order") {
        *known = true;
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: batch_xcd_order must be 0 or 1");
        g_batch_xcd_order = value;
        return TG_OK;
    }
    if (std::string(key) == "batch_graph") {
        *known = true;
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: batch_graph must be 0 or 1");
        g_batch_graph = value;
        return TG_OK;
    }
    if (std::string(key) == "batch_idx16") {
        *known = true;
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: batch_idx16 must be 0 or 1");
        g_batch_idx16 = value;
        return TG_OK;
    }
    if (std::string(key) == "batch_split") {
        *known = true;
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: batch_split must be 0 or 1");
        g_batch_split = value;
        return TG_OK;
    }
    if (std::string(key) == "batch_spmm") {
        *known = true;
        if (value < 0 || value > 2) return fail(TG_ERR_INVALID, "tg_tune_set: batch_spmm must be 0, 1 or 2");
        g_batch_spmm = value;
        return TG_OK;
    }
    if (std::string(key) == "batch_stream") {
        *known = true;
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: batch_stream must be 0 or 1");
        g_batch_stream = value;
        return TG_OK;
    }
    if (std::string(key) != "batch_tile_rows") return TG_OK;
    *known = true;
    if (value < kMinTileRows || value > 8192 || (value & (value - 1)))
        return fail(TG_ERR_INVALID, "tg_tune_set: batch_tile_rows must be a power of two in [256, 8192]");
    g_batch_tile_rows = value;
    return TG_OK;
}

int32_t batch_tune_get(const char* key) {
    if (std::string(key) == "batch_replace_drop_pct") return (int32_t)std::lround(g_replace_drop * 100.0);
    if (std::string(key) == "batch_check_every") return g_check_every;
    if (std::string(key) == "batch_split") return g_batch_split;
    if (std::string(key) == "batch_idx16") return g_batch_idx16;
    if (std::string(key) == "batch_graph") return g_batch_graph;
    if (std::string(key) == "batch_xcd_order") return g_batch_xcd_order;
    if (std::string(key) == "batch_stream") return g_batch_stream;
    if (std::string(key) == "batch_spmm") return g_batch_spmm;
    return std::string(key) == "batch_tile_rows" ? g_batch_tile_rows : -1;
}

}  // namespace tg

using namespace tg;

struct tg_batch_s {
    Batch* b;
};

extern "C" int tg_batch_create(tg_batch_t* out, void* stream) {
    if (!out) return fail(TG_ERR_INVALID, "tg_batch_create: null handle");
    *out = new tg_batch_s{new Batch(as_stream(stream))};
    return TG_OK;
}

extern "C" int tg_batch_destroy(tg_batch_t h) {
    if (!h) return TG_OK;
    for (auto& kv : h->b->sys)
        if (kv.second->mem) (void)hipFreeAsync(kv.second->mem, h->b->s);
    delete h->b;
    delete h;
    return TG_OK;
}

extern "C" int tg_batch_add_cg(tg_batch_t h, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                               const double* val, const double* diag, int32_t nrhs, const double* B, double* X,
                               double tol, int32_t max_iter, int32_t mixed, int64_t* id_out) {
    if (!h || !id_out || n < 1 || nrhs < 1 || nrhs > kBatchMaxK || !rp || !ci || !val || !B || !X || mixed < 0 ||
        mixed > 2)
        return fail(TG_ERR_INVALID, "tg_batch_add_cg: bad argument (1 <= nrhs <= 8, mode 0 / 1 / 2, device pointers)");
    Batch& Bt = *h->b;
    auto S = std::make_unique<BatchSystem>();
    S->id = Bt.next_id++;
    S->mixed = mixed;
    S->replace_drop = g_replace_drop;
    S->check_every = g_check_every;
    S->n = n;
    S->k = nrhs;
    S->nnz = nnz;
    S->rp = rp;
    S->ci = ci;
    S->val64 = val;
    S->d64 = diag;
    S->b64 = B;
    S->x64 = X;
    S->tol = tol;
    S->max_iter = max_iter;
    S->rel.assign(nrhs, 0.0);
    BatchSystem& ref = *S;
    Bt.sys[S->id] = std::move(S);
    *id_out = ref.id;
    if (ref.mixed == 0) return Bt.b64.add(n, nnz, rp, ci, val, diag, nrhs, B, X, tol, max_iter, &ref.slot);
    if (ref.mixed == 2) return true_start(Bt, ref);
    return mixed_start(Bt, ref);
}

extern "C" int tg_batch_step(tg_batch_t h, int32_t iterations) {
    if (!h || iterations < 1) return fail(TG_ERR_INVALID, "tg_batch_step: bad argument");
    return batch_advance(*h->b, iterations);
}

extern "C" int tg_batch_finished(tg_batch_t h, int64_t* ids, int32_t cap, int32_t* count) {
    if (!h || !count || (cap > 0 && !ids)) return fail(TG_ERR_INVALID, "tg_batch_finished: bad argument");
    auto& q = h->b->finished_queue;
    const int32_t c = (int32_t)std::min<size_t>(q.size(), (size_t)std::max(cap, 0));
    for (int32_t i = 0; i < c; ++i) ids[i] = q[i];
    q.erase(q.begin(), q.begin() + c);
    *count = c;
    return TG_OK;
}

extern "C" int tg_batch_result(tg_batch_t h, int64_t id, int32_t* finished, int32_t* succeeded, int32_t* iterations,
                               int32_t* outer, double* rel_residuals) {
    if (!h) return fail(TG_ERR_INVALID, "tg_batch_result: null handle");
    auto it = h->b->sys.find(id);
    if (it == h->b->sys.end()) return fail(TG_ERR_INVALID, "tg_batch_result: unknown id");
    const BatchSystem& S = *it->second;
    if (finished) *finished = S.finished;
    if (succeeded) *succeeded = S.succeeded;
    if (iterations) *iterations = S.iterations;
    if (outer) *outer = S.replacements;
    if (rel_residuals)
        for (int c = 0; c < S.k; ++c) rel_residuals[c] = S.rel[c];
    return TG_OK;
}

// Frees the device workspace of a finished system (its X stays with the caller).
extern "C" int tg_batch_release(tg_batch_t h, int64_t id) {
    if (!h) return fail(TG_ERR_INVALID, "tg_batch_release: null handle");
    Batch& Bt = *h->b;
    auto it = Bt.sys.find(id);
    if (it == Bt.sys.end()) return fail(TG_ERR_INVALID, "tg_batch_release: unknown id");
    BatchSystem& S = *it->second;
    if (!S.finished) return fail(TG_ERR_INVALID, "tg_batch_release: system %lld still running", (long long)id);
    if (S.slot >= 0) {
        int rc = S.mixed == 1 ? Bt.b32.retire(S.slot) : Bt.b64.retire(S.slot);
        if (rc) return rc;
    }
    if (S.mem) TG_HIP(hipFreeAsync(S.mem, Bt.s));
    Bt.sys.erase(it);
    return TG_OK;
}

// Device bytes the engine allocates for one system (its own workspace, not the
// caller's matrix and vectors).
extern "C" size_t tg_batch_request_bytes(int32_t n, int64_t nnz, int32_t nrhs, int32_t has_diag, int32_t mixed) {
    if (n < 1 || nrhs < 1 || nrhs > kBatchMaxK || nnz < 0) return 0;
    const size_t N = (size_t)n * nrhs, tiles = (size_t)ceil_div(n, g_batch_tile_rows);
    const size_t pb = align256(tiles * nrhs * sizeof(double));
    const size_t nvec = 4;
    const size_t ib = align256((size_t)std::max<int64_t>(nnz, 1) * 2) + 256 +  // 16-bit column offsets, flag,
                      align256(tiles * sizeof(int));                              // per-tile x-update marks
    if (mixed == 0) return nvec * align256(N * sizeof(double)) + 3 * pb + ib;
    if (mixed == 2)  // + the ||r|| partials and norms of the true-residual stop
        return nvec * align256(N * sizeof(double)) + 3 * pb + ib +
               align256((size_t)ceil_div(n, kMinTileRows) * nrhs * 8) + align256((size_t)nrhs * 8);
    return nvec * align256(N * sizeof(float)) + 3 * pb + ib + align256((size_t)nnz * 4) + (has_diag ? align256((size_t)n * 4) : 0) +
           2 * align256(N * 4) + align256((size_t)ceil_div(n, kMinTileRows) * nrhs * 8) + align256((size_t)2 * nrhs * 8);
}

extern "C" int tg_batch_active(tg_batch_t h, int32_t* systems, int64_t* bytes_in_use) {
    if (!h) return fail(TG_ERR_INVALID, "tg_batch_active: null handle");
    int c = 0;
    size_t extra = 0;
    for (auto& kv : h->b->sys) {
        c += kv.second->finished ? 0 : 1;
        extra += kv.second->mem_bytes;
    }
    if (systems) *systems = c;
    if (bytes_in_use) *bytes_in_use = (int64_t)(h->b->b64.bytes_in_use() + h->b->b32.bytes_in_use() + extra);
    return TG_OK;
}
