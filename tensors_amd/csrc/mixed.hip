// Mixed-precision CG: FP32 inner CG sweeps inside an FP64 iterative-refinement
// loop.
//
//   x = guess or 0
//   loop:  r = b - A64 x                  (FP64 SpMM, FP64 residual)
//          stop when every column has ||r||_2 / ||b||_2 <= tol   (FP64 norms)
//          r32 = float(r)                 (A32 and the Krylov vectors are FP32)
//          A32 d = r32 by FP32 CG         (the ConjugateGradient.hpp recurrences,
//                                          stopping at |r.z| <= tol_in^2 |b.Pb|)
//          x += double(d)                 (FP64 solution update)
//   X <- a x + b X
//
// A32 is the FP32 rounding of the FP64 values on the same pattern (the
// reference's external-values form of Dot, MatrixCSR.hpp:1026-1086).  The
// stopping test of the refinement is the true FP64 residual, kept distinct
// from the inner CG's recursive preconditioned residual.  Each inner solve is
// asked for just the reduction still needed (never below inner_tol, the FP32
// floor), so the last sweep does not over-solve.

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

__global__ __launch_bounds__(kBlock) void mx_total_kernel(int k, int kpad, int B, const double* __restrict__ partial,
                                                          double* __restrict__ out) {
    const int c = threadIdx.x % kpad, j = threadIdx.x / kpad, J = kBlock / kpad;
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

static int pow2ceil_m(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
static int grid_m(int64_t count) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((count + kBlock - 1) / kBlock, 8192));
}

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
                    int32_t max_outer, int32_t max_inner, double inner_tol, double a, const double* Bm, int64_t ldB,
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
    hipLaunchKernelGGL(mx_total_kernel, dim3(1), dim3(kBlock), 0, s, k, kpad, Bk, (const double*)w.partial, w.norms);
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
    bool first = !use_guess;
    int outer = 0;
    bool ok = false;
    while (true) {
        // r = b - A x  (FP64)
        if (!first) {
            int rc = csr_apply<double>(A64, w.x64, w.u64, k, s);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(mx_residual_kernel, dim3(Bk), dim3(kBlock), 0, s, n, k, kpad, Bm, ldB,
                           first ? (const double*)nullptr : (const double*)w.u64, w.r32, w.partial);
        hipLaunchKernelGGL(mx_total_kernel, dim3(1), dim3(kBlock), 0, s, k, kpad, Bk, (const double*)w.partial,
                           w.norms + k);
        TG_LAUNCH_CHECK();
        TG_HIP(hipMemcpyAsync(rn.data(), w.norms + k, k * 8, hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        first = false;
        double worst = 0.0;
        for (int c = 0; c < k; ++c) {
            res->rel_res[c] = bn[c] > 0 ? rn[c] / bn[c] : 0.0;
            worst = std::max(worst, res->rel_res[c]);
        }
        if (worst <= tol) {
            ok = true;
            break;
        }
        if (outer >= max_outer) break;
        // inner FP32 CG on the correction: A32 d = r32
        const double tin = std::min(0.5, std::max(inner_tol, 0.3 * tol / worst));
        CGResult cr{};
        int rc = cg_solve_t<float>(A32, A32op, prec_kind, w.diag32, none, k, max_inner, 1.0, w.r32, k, 0.0, w.d32, k,
                                   tin, 0, w.inner, w.inner_bytes, &cr, s);
        if (rc) return rc;
        res->inner_total += cr.iterations;
        hipLaunchKernelGGL(mx_update_kernel, dim3(grid_m(N)), dim3(kBlock), 0, s, N, (const float*)w.d32, w.x64);
        TG_LAUNCH_CHECK();
        ++outer;
    }
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
                                 int32_t nrhs, int32_t max_outer, int32_t max_inner, double inner_tol, double a_coef,
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
    if (spmm_may_slab(A, nrhs)) A.band = csr_band(A, as_stream(stream));
    MixedResult res{};
    const int rc = cg_mixed(A, a32, prec_kind, prec_diag, nrhs, max_outer, max_inner, inner_tol, a_coef, B, ldB,
                            b_coef, X, ldX, tol, use_guess, workspace, workspace_bytes, &res, as_stream(stream));
    if (rc) return rc;
    if (succeeded) *succeeded = res.succeeded;
    if (outer_iterations) *outer_iterations = res.outer;
    if (inner_iterations) *inner_iterations = res.inner_total;
    if (rel_residuals)
        for (int c = 0; c < nrhs; ++c) rel_residuals[c] = res.rel_res[c];
    return TG_OK;
}
