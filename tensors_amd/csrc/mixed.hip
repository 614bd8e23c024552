// Mixed-precision CG: the ConjugateGradient.hpp recurrences run in FP32 (FP32
// copy of A, FP32 Krylov vectors, FP64 reductions) on the correction equation,
// with FP64 residual replacement (reliable updates: van der Vorst & Ye, "Residual
// replacement strategies for Krylov subspace iterative methods for the
// convergence of true residuals", SIAM J. Sci. Comput. 22 (2000); the mixed
// precision form of Clark et al., Comput. Phys. Commun. 181 (2010)):
//
//   x64 = guess or 0;  r = float(b - A64 x64);  FP32 CG on A32 d = r, with d = 0
//   every 16 iterations, when the FP32 residual estimate fell 10x since the last
//   replacement (or the FP32 CG met its own test):
//          x64 += d, d = 0                (FP64 solution, the FP32 part restarts at 0)
//          r = float(b - A64 x64)         (FP64 SpMM and residual)
//          stop when every column has ||r||_2 / ||b||_2 <= tol   (FP64 norms)
//          rho = r.z for the replaced r   (the search direction p is kept)
//   X <- a x + b X
//
// Keeping p makes the FP32 iteration continue the same Krylov process instead
// of restarting it for every refinement step, so the iteration count stays
// close to FP64 CG's.  A32 is the FP32 rounding of the FP64 values on the same
// pattern (the reference's external-values form of Dot, MatrixCSR.hpp:1026-1086).

#include "common.hpp"
#include "krylov.hpp"
#include "reduce.hpp"

namespace tg {

// partial ||B_c||^2  (B row-major with ldB)
__global__ __launch_bounds__(kBlock) void mx_bnorm_kernel(int n, int k, int kpad, const double* __restrict__ Bm,
                                                          int64_t ldB, double* __restrict__ partial) {
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k)
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const double v = Bm[(int64_t)i * ldB + L.c];
            acc += v * v;
        }
    column_partial(acc, L, k, kpad, partial);
}

// r32 = float(B - u) (u = A x, or absent on the first sweep); partial ||B - u||^2
__global__ __launch_bounds__(kBlock) void mx_residual_kernel(int n, int k, int kpad, const double* __restrict__ Bm,
                                                             int64_t ldB, const double* __restrict__ u,
                                                             float* __restrict__ r32, double* __restrict__ partial) {
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k)
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e = (int64_t)i * k + L.c;
            const double r = Bm[(int64_t)i * ldB + L.c] - (u ? u[e] : 0.0);
            r32[e] = (float)r;
            acc += r * r;
        }
    column_partial(acc, L, k, kpad, partial);
}

__global__ __launch_bounds__(kFinalThreads) void mx_total_kernel(int k, int kpad, int B, const double* __restrict__ partial,
                                                          double* __restrict__ out) {
    const int c = threadIdx.x % kpad, j = threadIdx.x / kpad, J = kFinalThreads / kpad;
    const double t = column_total(partial, B, k, kpad, c, j, J);
    if (j == 0 && c < k) out[c] = sqrt(t);
}

// x += double(d)
__global__ __launch_bounds__(kBlock) void mx_update_kernel(int64_t count, const float* __restrict__ d,
                                                           double* __restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        x[e] += (double)d[e];
}

__global__ __launch_bounds__(kBlock) void mx_to_float_kernel(int64_t count, const double* __restrict__ a,
                                                             float* __restrict__ b) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        b[e] = (float)a[e];
}

__global__ __launch_bounds__(kBlock) void mx_load_kernel(int n, int k, const double* __restrict__ X, int64_t ldX,
                                                         double* __restrict__ x, int zero) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        x[e] = zero ? 0.0 : X[i * ldX + (e - i * k)];
    }
}

__global__ __launch_bounds__(kBlock) void mx_write_kernel(int n, int k, double a, const double* __restrict__ x,
                                                          double b, double* __restrict__ X, int64_t ldX, int b_mode) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        double* Xp = X + i * ldX + (e - i * k);
        double v = (a == 1.0) ? x[e] : a * x[e];
        if (b_mode == 1)
            v = v + *Xp;
        else if (b_mode == 2)
            v = v + b * *Xp;
        *Xp = v;
    }
}

__global__ __launch_bounds__(kBlock) void mx_scale_kernel(int n, int k, double b, double* __restrict__ X, int64_t ldX) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        X[i * ldX + (e - i * k)] *= b;
    }
}

// x64 += d, d = 0
__global__ __launch_bounds__(kBlock) void mx_flush_kernel(int64_t count, float* __restrict__ d, double* __restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock) {
        x[e] += (double)d[e];
        d[e] = 0.0f;
    }
}

// z = P r (if a diagonal is given); partial(r.z)
__global__ __launch_bounds__(kBlock) void mx_rz_kernel(int n, int k, int kpad, const float* __restrict__ r,
                                                       const float* __restrict__ diag, float* __restrict__ z,
                                                       double* __restrict__ partial) {
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k)
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e = (int64_t)i * k + L.c;
            const float rv = r[e];
            float zv = rv;
            if (diag) {
                zv = diag[i] * rv;
                z[e] = zv;
            }
            acc += (double)rv * (double)zv;
        }
    column_partial(acc, L, k, kpad, partial);
}

// rho = r.z of the replaced residual, beta = rho / rho_old for the next direction
// p = z + beta p (iterations after the first); with `scale` != 1 the device had
// stopped on the FP32 estimate: tighten its tolerance and let it continue.
__global__ __launch_bounds__(kFinalThreads) void mx_replace_state_kernel(int k, int kpad, int B,
                                                                  const double* __restrict__ partial, double scale,
                                                                  int restart, CGState* __restrict__ st) {
    const int c = threadIdx.x % kpad, j = threadIdx.x / kpad, J = kFinalThreads / kpad;
    const double tot = column_total(partial, B, k, kpad, c, j, J);
    if (j == 0 && c < k) {
        st->rho[c] = tot;
        // restart: beta = 0, so the fused-direction form builds p = z in its next SpMM (the
        // other forms already hold the next p, which the host reset to z)
        if (st->iter > 0) st->beta[c] = (!restart && st->rho_old[c] != 0.0) ? tot / st->rho_old[c] : 0.0;
        st->TOL[c] *= scale;
    }
    if (threadIdx.x == 0) {
        st->done = 0;
        st->succeeded = 0;
    }
}

static int pow2ceil_m(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
static int grid_m(int64_t count) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((count + kBlock - 1) / kBlock, 8192));
}

constexpr int kReplaceCheck = 16;     // iterations between residual-replacement checks
constexpr double kReplaceDrop = 0.1;  // default: replace when the FP32 residual estimate fell 10x
constexpr int kFlatChecks = 8;        // ... or when it did not fall at this many checks in a row

struct MixedWs {
    double* partial;
    double* norms;  // 2k: ||b||, ||r||
    double* x64;
    double* u64;
    float* r32;
    float* d32;
    float* a32;
    float* diag32;
    void* inner;
    size_t inner_bytes;
    size_t total;
};

static size_t mixed_layout(int32_t n, int64_t nnz, int32_t k, int prec_kind, bool need_a32, char* base, MixedWs* w) {
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t o = off;
        off += align256(std::max<size_t>(b, 8));
        return base ? (void*)(base + o) : nullptr;
    };
    MixedWs t;
    const size_t N = (size_t)n * k;
    t.partial = (double*)take((size_t)kSolverBlocks * k * 8);
    t.norms = (double*)take((size_t)2 * k * 8);
    t.x64 = (double*)take(N * 8);
    t.u64 = (double*)take(N * 8);
    t.r32 = (float*)take(N * 4);
    t.d32 = (float*)take(N * 4);
    t.a32 = need_a32 ? (float*)take((size_t)nnz * 4) : nullptr;
    t.diag32 = prec_kind ? (float*)take((size_t)n * 4) : nullptr;
    t.inner_bytes = cg_workspace_bytes(n, k, 4, prec_kind);
    t.inner = take(t.inner_bytes);
    t.total = off;
    if (w) *w = t;
    return off;
}

struct MixedResult {
    int succeeded;
    int outer;
    int inner_total;
    double rel_res[kMaxCols];
};

static int cg_mixed(const CsrView& A64, const float* a32_in, int prec_kind, const double* d64, int32_t k,
                    int32_t max_outer, int32_t max_inner, double replace_drop, double a, const double* Bm, int64_t ldB,
                    double b, double* X, int64_t ldX, double tol, int use_guess, void* ws, size_t ws_bytes,
                    MixedResult* res, hipStream_t s) {
    const int n = A64.n;
    res->succeeded = 0;
    res->outer = 0;
    res->inner_total = 0;
    if (k < 1 || k > kMaxCols) return fail(TG_ERR_INVALID, "cg_mixed: nrhs must be in [1, %d]", kMaxCols);
    if (A64.m != n || A64.val_type != TG_F64 || !A64.val)
        return fail(TG_ERR_INVALID, "cg_mixed: A must be square with FP64 values");
    if (prec_kind == 1 && !d64) return fail(TG_ERR_INVALID, "cg_mixed: diagonal preconditioner without a diagonal");
    MixedWs w;
    if (!ws || ws_bytes < mixed_layout(n, A64.nnz, k, prec_kind, a32_in == nullptr, nullptr, nullptr))
        return fail(TG_ERR_WORKSPACE, "cg_mixed: workspace missing or smaller than tg_cg_mixed_workspace_bytes()");
    mixed_layout(n, A64.nnz, k, prec_kind, a32_in == nullptr, (char*)ws, &w);
    const int64_t N = (int64_t)n * k;
    const int kpad = pow2ceil_m(k);
    const int Bk = kSolverBlocks;
    if (a == 0.0) {
        hipLaunchKernelGGL(mx_scale_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, b, X, ldX);
        TG_LAUNCH_CHECK();
        res->succeeded = 1;
        return TG_OK;
    }
    // FP32 copies of the operator values and the preconditioner diagonal
    const float* a32 = a32_in;
    if (!a32) {
        hipLaunchKernelGGL(mx_to_float_kernel, dim3(grid_m(A64.nnz)), dim3(kBlock), 0, s, A64.nnz,
                           (const double*)A64.val, w.a32);
        a32 = w.a32;
    }
    if (prec_kind) hipLaunchKernelGGL(mx_to_float_kernel, dim3(grid_m(n)), dim3(kBlock), 0, s, (int64_t)n, d64, w.diag32);
    CsrView A32 = A64;
    A32.val = a32;
    A32.val_type = TG_F32;
    OpRef A32op;
    A32op.csr = &A32;
    OpRef none;

    // ||b||
    hipLaunchKernelGGL(mx_bnorm_kernel, dim3(Bk), dim3(kBlock), 0, s, n, k, kpad, Bm, ldB, w.partial);
    hipLaunchKernelGGL(mx_total_kernel, dim3(1), dim3(kFinalThreads), 0, s, k, kpad, Bk, (const double*)w.partial, w.norms);
    TG_LAUNCH_CHECK();
    std::vector<double> bn(k), rn(k);
    TG_HIP(hipMemcpyAsync(bn.data(), w.norms, k * 8, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    double bmax = 0.0;
    for (int c = 0; c < k; ++c) {
        if (std::isnan(bn[c])) {
            fail(TG_OK, "cg_mixed: right-hand side contains NaNs. Doing nothing.");
            return TG_OK;
        }
        bmax = std::max(bmax, bn[c]);
    }
    if (!(bmax > 0.0)) {
        hipLaunchKernelGGL(mx_scale_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, b, X, ldX);
        TG_LAUNCH_CHECK();
        res->succeeded = 1;
        return TG_OK;
    }
    hipLaunchKernelGGL(mx_load_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, (const double*)X, ldX, w.x64,
                       use_guess ? 0 : 1);
    // r = float(b - A64 x64)   (true residual norms in w.norms + k)
    auto residual = [&](bool x_is_zero) -> int {
        if (!x_is_zero) {
            int rc = csr_apply<double>(A64, w.x64, w.u64, k, s);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(mx_residual_kernel, dim3(Bk), dim3(kBlock), 0, s, n, k, kpad, Bm, ldB,
                           x_is_zero ? (const double*)nullptr : (const double*)w.u64, w.r32, w.partial);
        hipLaunchKernelGGL(mx_total_kernel, dim3(1), dim3(kFinalThreads), 0, s, k, kpad, Bk, (const double*)w.partial,
                           w.norms + k);
        TG_LAUNCH_CHECK();
        TG_HIP(hipMemcpyAsync(rn.data(), w.norms + k, k * 8, hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        return TG_OK;
    };
    auto worst_rel = [&]() {
        double worst = 0.0;
        for (int c = 0; c < k; ++c) {
            res->rel_res[c] = bn[c] > 0 ? rn[c] / bn[c] : 0.0;
            worst = std::max(worst, std::isnan(res->rel_res[c]) ? INFINITY : res->rel_res[c]);
        }
        return worst;
    };
    int rc = residual(!use_guess);
    if (rc) return rc;
    bool ok = worst_rel() <= tol;
    int replacements = 0;
    if (!ok) {
        double est_last = 1.0, est_check = 0.0;
        int flat = 0;  // consecutive checks at which the estimate did not fall
        const double drop = (replace_drop > 0.0 && replace_drop < 1.0) ? replace_drop : kReplaceDrop;
        CGHook hook;
        hook.every = kReplaceCheck;
        hook.guard = 1;  // a column whose FP32 direction vanished stops instead of turning NaN
        hook.fn = [&](const CGState& hs, void* r, void* z, void* d, void* p, double* partial, CGState* st,
                      bool* stop) -> int {
            double est = 0.0;
            for (int c = 0; c < k; ++c) est = std::max(est, hs.bsq[c] > 0 ? std::sqrt(std::fabs(hs.rho[c]) / hs.bsq[c]) : 0.0);
            const bool stopped = hs.done != 0;
            // a replacement when the estimate fell drop-fold since the last one, when the
            // FP32 iteration stopped, or when the estimate did not fall at kFlatChecks checks
            // in a row (the FP32 recursion at its rounding floor, or a small system's
            // Krylov space used up); the last one also restarts the direction
            flat = (!stopped && est_check > 0.0 && !(est < est_check)) ? flat + 1 : 0;
            est_check = est;
            const bool forced = !stopped && !(est < drop * est_last);
            if (forced && flat < kFlatChecks) return TG_OK;
            flat = 0;
            est_check = 0.0;
            // fold the FP32 correction into x64, then the true residual
            hipLaunchKernelGGL(mx_flush_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, N, (float*)d, w.x64);
            int rc2 = residual(false);
            if (rc2) return rc2;
            ++replacements;
            est_last = est;
            const double worst = worst_rel();
            if (worst <= tol) {
                ok = true;
                *stop = true;
                return TG_OK;
            }
            if (std::isinf(worst) || replacements >= std::max(1, max_outer) || hs.iter >= max_inner) {
                *stop = true;
                return TG_OK;
            }
            // continue the FP32 process from the replaced residual (copy it into the
            // iteration's r); a stop on the FP32 estimate tightens its tolerance
            TG_HIP(hipMemcpyAsync(r, w.r32, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(mx_rz_kernel, dim3(Bk), dim3(kBlock), 0, s, n, k, kpad, (const float*)r,
                               (const float*)w.diag32, (float*)z, partial);
            // restart (p = z of the replaced residual) after a stop or a stagnation: there
            // the old direction has vanished, and the replaced residual's larger r.z over
            // its tiny p.u would throw x far off.  Forms 0 / 1 hold the next p already (reset
            // here); form 2 (p null) forms it from beta, which the state kernel zeroes.
            const int restart = (forced || stopped) ? 1 : 0;
            if (restart && p) TG_HIP(hipMemcpyAsync(p, z, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
            double scale = 1.0;
            if (stopped) {
                const double f = tol / worst;
                scale = std::min(0.25, std::max(1e-4, 0.25 * f * f));
            }
            hipLaunchKernelGGL(mx_replace_state_kernel, dim3(1), dim3(kFinalThreads), 0, s, k, kpad, Bk,
                               (const double*)partial, scale, restart, st);
            TG_LAUNCH_CHECK();
            return TG_OK;
        };
        CGResult cr{};
        rc = cg_solve_t<float>(A32, A32op, prec_kind, w.diag32, none, k, max_inner, 1.0, w.r32, k, 0.0, w.d32, k, tol,
                               0, w.inner, w.inner_bytes, &cr, s, &hook);
        if (rc) return rc;
        res->inner_total = cr.iterations;
        // the FP32 part left after the last replacement (zero if the solve stopped there)
        hipLaunchKernelGGL(mx_update_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, N, (const float*)w.d32, w.x64);
        TG_LAUNCH_CHECK();
        if (!ok) {
            rc = residual(false);
            if (rc) return rc;
            ok = worst_rel() <= tol;
        }
    }
    const int outer = replacements;
    const int b_mode = (b == 0.0) ? 0 : (b == 1.0) ? 1 : 2;
    hipLaunchKernelGGL(mx_write_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, a, (const double*)w.x64, b, X, ldX,
                       b_mode);
    TG_LAUNCH_CHECK();
    TG_HIP(hipStreamSynchronize(s));
    res->succeeded = ok ? 1 : 0;
    res->outer = outer;
    return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" size_t tg_cg_mixed_workspace_bytes(int32_t n, int64_t nnz, int32_t nrhs, int32_t prec_kind,
                                              int32_t with_a32) {
    if (n < 0 || nnz < 0 || nrhs < 1) return 0;
    return mixed_layout(n, nnz, nrhs, prec_kind ? 1 : 0, with_a32 == 0, nullptr, nullptr);
}

extern "C" int tg_cg_solve_mixed(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                                 const double* a64, const float* a32, int32_t prec_kind, const double* prec_diag,
                                 int32_t nrhs, int32_t max_outer, int32_t max_inner, double replace_drop, double a_coef,
                                 const double* B, int64_t ldB, double b_coef, double* X, int64_t ldX, double tol,
                                 int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                                 int32_t* outer_iterations, int32_t* inner_iterations, double* rel_residuals,
                                 void* stream) {
    if (succeeded) *succeeded = 0;
    if (outer_iterations) *outer_iterations = 0;
    if (inner_iterations) *inner_iterations = 0;
    if (n < 0 || nrhs < 1 || !B || !X || !a64 || ldB < nrhs || ldX < nrhs || (prec_kind != 0 && prec_kind != 1) ||
        max_outer < 0 || max_inner < 0)
        return fail(TG_ERR_INVALID, "tg_cg_solve_mixed: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_cg_solve_mixed: rp_type must be I32/I64");
    CsrView A;
    A.m = n;
    A.n = n;
    A.nnz = nnz;
    A.rp = rp;
    A.rp_type = rp_type;
    A.ci = ci;
    A.val = a64;
    A.val_type = TG_F64;
    if (spmm_may_slab(A, nrhs) || cg_panel_wants_band(n)) A.band = csr_band(A, as_stream(stream));
    MixedResult res{};
    const int rc = cg_mixed(A, a32, prec_kind, prec_diag, nrhs, max_outer, max_inner, replace_drop, a_coef, B, ldB,
                            b_coef, X, ldX, tol, use_guess, workspace, workspace_bytes, &res, as_stream(stream));
    if (rc) return rc;
    if (succeeded) *succeeded = res.succeeded;
    if (outer_iterations) *outer_iterations = res.outer;
    if (inner_iterations) *inner_iterations = res.inner_total;
    if (rel_residuals)
        for (int c = 0; c < nrhs; ++c) rel_residuals[c] = res.rel_res[c];
    return TG_OK;
}

// ---------------------------------------------------------------------------
// Mixed-precision GMRES(m): restarted GMRES whose Arnoldi cycles run in FP32
// (FP32 values a32 on the same pattern, FP32 basis Q and Krylov vectors), while
// everything that decides accuracy stays FP64: the Hessenberg matrix, Givens
// rotations, beta and the least squares of each cycle (the FP64 arrays of
// gmres_solve_t), and at every restart the solution update x64 += d and the
// true residual b - A64 x64 with its norm, which is the stop test.  Compared
// with GMRES::operator() (GMRES.hpp:127-260) the restart residual z = A x - b
// (:282-350) is the FP64 one, so a cycle solves the correction equation
// A d = r in FP32 and the restarts are iterative refinement steps:
//
//   x64 = guess or 0
//   for cycle = 0 .. max_restarts:
//       r = b - A64 x64                                  (FP64)
//       stop if every column has ||P r|| <= tol ||P b||  (P only for Side::Left,
//                                                         as GMRES.hpp:183-185)
//       d = one GMRES(m) cycle of A32 d = float(r)       (FP32 vectors, FP64 H)
//       x64 += d
//   X <- a x64 + b X
//
// A cycle stops early once its own residual fell below what the outer test
// still needs (at most to 1e-6, near the FP32 rounding floor of the cycle).
// ---------------------------------------------------------------------------

namespace tg {

// r = B - u (u null: B); r32 = float(r) when given; partial ||D r||^2 with D the
// preconditioner diagonal when given (else the identity)
__global__ __launch_bounds__(kBlock) void mg_residual_kernel(int n, int k, int kpad, const double* __restrict__ Bm,
                                                             int64_t ldB, const double* __restrict__ u,
                                                             const double* __restrict__ dg, float* __restrict__ r32,
                                                             double* __restrict__ partial) {
    int i0, i1;
    block_row_range(n, i0, i1);
    const Lanes L = lanes_of(kpad);
    double acc = 0.0;
    if (L.c < k)
        for (int i = i0 + L.ro; i < i1; i += L.rows) {
            const int64_t e = (int64_t)i * k + L.c;
            const double r = Bm[(int64_t)i * ldB + L.c] - (u ? u[e] : 0.0);
            if (r32) r32[e] = (float)r;
            const double v = dg ? dg[i] * r : r;
            acc += v * v;
        }
    column_partial(acc, L, k, kpad, partial);
}

// X <- b X, or zeros when b == 0 (GMRES.hpp:206-229)
__global__ __launch_bounds__(kBlock) void mg_zero_or_scale_kernel(int n, int k, double b, double* __restrict__ X,
                                                                  int64_t ldX) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / k;
        double* Xp = X + i * ldX + (e - i * k);
        *Xp = (b == 0.0) ? 0.0 : *Xp * b;
    }
}

struct MixedGmresWs {
    double* partial;
    double* norms;  // 2k: ||P b||, ||P r||
    double* x64;
    double* u64;
    float* r32;
    float* d32;
    float* a32;
    float* diag32;
    void* inner;
    size_t inner_bytes;
    size_t total;
};

static size_t mixed_gmres_layout(int32_t n, int64_t nnz, int32_t m, int32_t k, int prec_kind, bool need_a32,
                                 char* base, MixedGmresWs* w) {
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t o = off;
        off += align256(std::max<size_t>(b, 8));
        return base ? (void*)(base + o) : nullptr;
    };
    MixedGmresWs t;
    const size_t N = (size_t)n * k;
    t.partial = (double*)take((size_t)kSolverBlocks * k * 8);
    t.norms = (double*)take((size_t)2 * k * 8);
    t.x64 = (double*)take(N * 8);
    t.u64 = (double*)take(N * 8);
    t.r32 = (float*)take(N * 4);
    t.d32 = (float*)take(N * 4);
    t.a32 = need_a32 ? (float*)take((size_t)nnz * 4) : nullptr;
    t.diag32 = prec_kind ? (float*)take((size_t)n * 4) : nullptr;
    t.inner_bytes = gmres_workspace_bytes(n, m, k, 4);
    t.inner = take(t.inner_bytes);
    t.total = off;
    if (w) *w = t;
    return off;
}

struct MixedGmresResult {
    int succeeded;
    int restarts;
    int inner_total;
    double rel_res[kMaxCols];
};

static int gmres_mixed(const CsrView& A64, const float* a32_in, int prec_kind, const double* d64, int side,
                       int32_t k, int32_t max_iter, int32_t max_restarts, double a, const double* Bm, int64_t ldB,
                       double b, double* X, int64_t ldX, double tol, int use_guess, void* ws, size_t ws_bytes,
                       MixedGmresResult* res, hipStream_t s) {
    const int n = A64.n;
    res->succeeded = 0;
    res->restarts = 0;
    res->inner_total = 0;
    if (k < 1 || k > 32) return fail(TG_ERR_INVALID, "gmres_mixed: nrhs must be in [1, 32]");
    if (A64.m != n || n < 1 || A64.val_type != TG_F64 || !A64.val)
        return fail(TG_ERR_INVALID, "gmres_mixed: A must be square, non-empty, with FP64 values");
    const int m = std::min(max_iter, n + 1);  // GMRES.hpp:89
    if (m < 1 || m > 63) return fail(TG_ERR_INVALID, "gmres_mixed: restart length must be in [1, 63]");
    if (prec_kind == 1 && !d64) return fail(TG_ERR_INVALID, "gmres_mixed: diagonal preconditioner without a diagonal");
    MixedGmresWs w;
    if (!ws || ws_bytes < mixed_gmres_layout(n, A64.nnz, m, k, prec_kind, a32_in == nullptr, nullptr, nullptr))
        return fail(TG_ERR_WORKSPACE, "gmres_mixed: workspace missing or smaller than tg_gmres_mixed_workspace_bytes()");
    mixed_gmres_layout(n, A64.nnz, m, k, prec_kind, a32_in == nullptr, (char*)ws, &w);
    const int64_t N = (int64_t)n * k;
    const int kpad = pow2ceil_m(k);
    const int Bk = kSolverBlocks;
    if (a == 0.0) {
        hipLaunchKernelGGL(mx_scale_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, b, X, ldX);
        TG_LAUNCH_CHECK();
        res->succeeded = 1;
        return TG_OK;
    }
    const float* a32 = a32_in;
    if (!a32) {
        hipLaunchKernelGGL(mx_to_float_kernel, dim3(grid_m(A64.nnz)), dim3(kBlock), 0, s, A64.nnz,
                           (const double*)A64.val, w.a32);
        a32 = w.a32;
    }
    if (prec_kind) hipLaunchKernelGGL(mx_to_float_kernel, dim3(grid_m(n)), dim3(kBlock), 0, s, (int64_t)n, d64, w.diag32);
    CsrView A32 = A64;
    A32.val = a32;
    A32.val_type = TG_F32;
    OpRef A32op;
    A32op.csr = &A32;
    OpRef none;
    // the stop test's norms carry P on the left side only (GMRES.hpp:168-185)
    const double* dnorm = (prec_kind == 1 && side == 0) ? d64 : nullptr;

    hipLaunchKernelGGL(mg_residual_kernel, dim3(Bk), dim3(kBlock), 0, s, n, k, kpad, Bm, ldB, (const double*)nullptr,
                       dnorm, (float*)nullptr, w.partial);
    hipLaunchKernelGGL(mx_total_kernel, dim3(1), dim3(kFinalThreads), 0, s, k, kpad, Bk, (const double*)w.partial, w.norms);
    TG_LAUNCH_CHECK();
    std::vector<double> bn(k), rn(k);
    TG_HIP(hipMemcpyAsync(bn.data(), w.norms, k * 8, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    double bmax = 0.0;
    for (int c = 0; c < k; ++c) {
        if (std::isnan(bn[c])) {  // GMRES.hpp:193-204
            fail(TG_OK, "gmres_mixed: right-hand side contains NaNs. Doing nothing.");
            return TG_OK;
        }
        bmax = std::max(bmax, bn[c]);
    }
    if (!(bmax > 0.0)) {  // :206-229
        hipLaunchKernelGGL(mg_zero_or_scale_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, b, X, ldX);
        TG_LAUNCH_CHECK();
        res->succeeded = 1;
        return TG_OK;
    }
    hipLaunchKernelGGL(mx_load_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, (const double*)X, ldX, w.x64,
                       use_guess ? 0 : 1);
    // r = b - A64 x64 in FP64, r32 = float(r), norms of P r into w.norms + k
    auto residual = [&](bool x_is_zero) -> int {
        if (!x_is_zero) {
            int rc = csr_apply<double>(A64, w.x64, w.u64, k, s);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(mg_residual_kernel, dim3(Bk), dim3(kBlock), 0, s, n, k, kpad, Bm, ldB,
                           x_is_zero ? (const double*)nullptr : (const double*)w.u64, dnorm, w.r32, w.partial);
        hipLaunchKernelGGL(mx_total_kernel, dim3(1), dim3(kFinalThreads), 0, s, k, kpad, Bk, (const double*)w.partial,
                           w.norms + k);
        TG_LAUNCH_CHECK();
        TG_HIP(hipMemcpyAsync(rn.data(), w.norms + k, k * 8, hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        return TG_OK;
    };
    bool ok = false;
    int cycles = 0;
    bool x_zero = !use_guess;
    // the FP32 cycles: one captured cycle for all of them, and no read-back of the
    // right-hand side's norms before each (r is checked finite and non-zero here)
    GmresReuse reuse;
    reuse.checked = true;
    for (;;) {
        int rc = residual(x_zero);
        if (rc) return rc;
        x_zero = false;
        double worst = 0.0, need = 1.0;
        for (int c = 0; c < k; ++c) {
            res->rel_res[c] = bn[c] > 0 ? rn[c] / bn[c] : 0.0;
            worst = std::max(worst, std::isnan(res->rel_res[c]) ? INFINITY : res->rel_res[c]);
            if (rn[c] > 0.0) need = std::min(need, tol * bn[c] / rn[c]);
        }
        if (worst <= tol) {
            ok = true;
            break;
        }
        if (std::isinf(worst) || cycles > max_restarts) break;
        // one FP32 cycle on A32 d = r; it stops once its residual fell by what the outer
        // test still needs (half of it, for margin), but not below the FP32 floor
        const double tol_in = std::max(0.5 * need, 1e-6);
        GmresResult gr{};
        rc = gmres_solve_t<float>(A32, A32op, prec_kind, (const float*)w.diag32, none, side, k, m, 1.0,
                                  (const float*)w.r32, k, 0.0, w.d32, k, tol_in, 0, 0, w.inner, w.inner_bytes, &gr, s,
                                  &reuse);
        if (rc) return rc;
        res->inner_total += gr.iterations;
        hipLaunchKernelGGL(mx_update_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, N, (const float*)w.d32, w.x64);
        TG_LAUNCH_CHECK();
        ++cycles;
    }
    const int b_mode = (b == 0.0) ? 0 : (b == 1.0) ? 1 : 2;
    hipLaunchKernelGGL(mx_write_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, n, k, a, (const double*)w.x64, b, X, ldX,
                       b_mode);
    TG_LAUNCH_CHECK();
    TG_HIP(hipStreamSynchronize(s));
    res->succeeded = ok ? 1 : 0;
    res->restarts = std::max(0, cycles - 1);
    return TG_OK;
}

}  // namespace tg

extern "C" size_t tg_gmres_mixed_workspace_bytes(int32_t n, int64_t nnz, int32_t max_iter, int32_t nrhs,
                                                 int32_t prec_kind, int32_t with_a32) {
    if (n < 1 || nnz < 0 || nrhs < 1) return 0;
    const int m = std::max(1, std::min(max_iter, n + 1));
    return tg::mixed_gmres_layout(n, nnz, m, nrhs, prec_kind ? 1 : 0, with_a32 == 0, nullptr, nullptr);
}

extern "C" int tg_gmres_solve_mixed(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                                    const double* a64, const float* a32, int32_t prec_kind, const double* prec_diag,
                                    int32_t side, int32_t nrhs, int32_t max_iter, int32_t max_restarts, double a_coef,
                                    const double* B, int64_t ldB, double b_coef, double* X, int64_t ldX, double tol,
                                    int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                                    int32_t* restarts, int32_t* inner_iterations, double* rel_residuals,
                                    void* stream) {
    using namespace tg;
    if (succeeded) *succeeded = 0;
    if (restarts) *restarts = 0;
    if (inner_iterations) *inner_iterations = 0;
    if (n < 1 || nrhs < 1 || !B || !X || !a64 || ldB < nrhs || ldX < nrhs || (prec_kind != 0 && prec_kind != 1) ||
        (side != 0 && side != 1) || max_restarts < 0)
        return fail(TG_ERR_INVALID, "tg_gmres_solve_mixed: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_gmres_solve_mixed: rp_type must be I32/I64");
    CsrView A;
    A.m = n;
    A.n = n;
    A.nnz = nnz;
    A.rp = rp;
    A.rp_type = rp_type;
    A.ci = ci;
    A.val = a64;
    A.val_type = TG_F64;
    if (spmm_may_slab(A, nrhs)) A.band = csr_band(A, as_stream(stream));
    MixedGmresResult res{};
    const int rc = gmres_mixed(A, a32, prec_kind, prec_diag, side, nrhs, max_iter, max_restarts, a_coef, B, ldB,
                               b_coef, X, ldX, tol, use_guess, workspace, workspace_bytes, &res, as_stream(stream));
    if (rc) return rc;
    if (succeeded) *succeeded = res.succeeded;
    if (restarts) *restarts = res.restarts;
    if (inner_iterations) *inner_iterations = res.inner_total;
    if (rel_residuals)
        for (int c = 0; c < nrhs; ++c) rel_residuals[c] = res.rel_res[c];
    return TG_OK;
}
