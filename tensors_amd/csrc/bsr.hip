// Block-CSR SpMM: the product of Sparse::KernelMatrixCSR<Kernel_T>::Dot
// (KernelMatrixCSR.hpp:153-227) with the dense block kernels
// (BlockKernels/DenseBlockKernel_Tiny.hpp:160-180, BlockKernels/BlockKernel_Tiny.hpp:
// 156-196).  A pattern of mb x nb blocks (rp, ci) carries one R x C dense block
// per nonzero; for every block row i
//     Y_i <- alpha * sum_{k in row i} A_k X_{ci[k]} + beta * Y_i
// with X_j the C x nrhs chunk at X + j*C*nrhs and Y_i the R x nrhs chunk at
// Y + i*R*nrhs.  Storage orders as the kernel's template flags: A_k row-major
// (a_RM) or column-major; X / Y chunks row-major (x_RM / y_RM: entry (c, h) at
// c*nrhs + h) or transposed (h*C + c).  An empty block row gets Y_i <- beta Y_i
// (WriteYZero, :185-190); alpha == 0 or no nonzeros scales all of Y by beta
// (KernelMatrixCSR.hpp:137-151, 158-162).
//
// GPU layout: a lane group of L lanes (L = R*nrhs rounded up to a power of two,
// at most 64) owns a block row; lane t computes output entries t, t+L, ... of the
// chunk, so lanes of one group read the same A entries (broadcast) and
// neighbouring X entries (coalesced).  The per-lane sum runs over the row's
// blocks in order and over c in order: deterministic.  CSR SpMM bytes per call
// apply with the block sizes: nnzb*(R*C*s + 4) + 4*(mb+1) + nb*C*k*s +
// mb*R*k*s*(1 + [beta != 0]); the gathered X chunks are C*k*s contiguous bytes.
// Memory-bound at every block size the reference instantiates (R, C <= 8:
// 2RCk flops per RCs + Cks bytes stays below 2 flop/byte), so no MFMA.

#include <string>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

enum BsrBeta : int { kBsrBetaZero = 0, kBsrBetaOne = 1, kBsrBetaGeneric = 2 };

// 1: row-major X / Y chunks use bsr_spmm_vec_kernel (lane = block row x V
// right-hand sides); 0: every storage order uses bsr_spmm_kernel (lane = entry).
static int g_bsr_vector = 1;

int bsr_tune_set(const char* key, int32_t value, bool* known) {
    *known = std::string(key) == "bsr_vector";
    if (!*known) return TG_OK;
    if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: bsr_vector must be 0 or 1");
    g_bsr_vector = value;
    return TG_OK;
}
int32_t bsr_tune_get(const char* key) { return std::string(key) == "bsr_vector" ? g_bsr_vector : -1; }

template <typename T>
__device__ __forceinline__ T bsr_out(T acc, const T* yp, T alpha, int alpha_one, T beta, int beta_mode) {
    T v = alpha_one ? acc : alpha * acc;
    if (beta_mode == kBsrBetaOne)
        v = v + *yp;
    else if (beta_mode == kBsrBetaGeneric)
        v = v + beta * *yp;
    return v;
}

// CT > 0: C fixed at compile time (loads of a block's C entries issued together).
template <typename T, typename RP, int CT, bool A_RM, bool X_RM, bool Y_RM>
__global__ __launch_bounds__(kBlock) void bsr_spmm_kernel(int mb, int R, int Cr, int k, const RP* __restrict__ rp,
                                                          const int32_t* __restrict__ ci, const T* __restrict__ A,
                                                          T alpha, const T* __restrict__ X, T beta,
                                                          T* __restrict__ Y, int L, int alpha_one, int beta_mode) {
    const int C = CT > 0 ? CT : Cr;
    const int groups = kBlock / L;
    const int t = threadIdx.x % L;
    const int i = blockIdx.x * groups + threadIdx.x / L;
    if (i >= mb) return;
    const RP lo = rp[i], hi = rp[i + 1];
    const int E = R * k;  // entries of the output chunk
    const int64_t RC = (int64_t)R * C, CK = (int64_t)C * k;
    T* __restrict__ Yi = Y + (int64_t)i * E;
    for (int e = t; e < E; e += L) {
        // entry e of the chunk in its storage order -> (r, h)
        const int r = Y_RM ? e / k : e % R;
        const int h = Y_RM ? e % k : e / R;
        T* yp = Yi + e;
        if (hi == lo) {  // WriteYZero: Y_i <- beta Y_i
            *yp = beta_mode == kBsrBetaZero ? T(0) : beta_mode == kBsrBetaOne ? *yp : beta * *yp;
            continue;
        }
        T acc = T(0);
        // two blocks per step (the second clamped to the row's last block and
        // masked), so their index, A and X loads are in flight together
        for (RP kb = lo; kb < hi; kb += 2) {
            const bool two = kb + 1 < hi;
            const RP kb1 = two ? kb + 1 : kb;
            const int64_t j0 = ci[kb], j1 = ci[kb1];
            const T* a0 = A + (int64_t)kb * RC;
            const T* a1 = A + (int64_t)kb1 * RC;
            const T* x0 = X + j0 * CK;
            const T* x1 = X + j1 * CK;
            T s0 = T(0), s1 = T(0);
#pragma unroll
            for (int c = 0; c < (CT > 0 ? CT : 1); ++c) {
                if constexpr (CT > 0) {
                    const int ia = A_RM ? r * C + c : c * R + r;
                    const int ix = X_RM ? c * k + h : h * C + c;
                    s0 += a0[ia] * x0[ix];
                    s1 += a1[ia] * x1[ix];
                }
            }
            if constexpr (CT == 0) {
                for (int c = 0; c < C; ++c) {
                    const int ia = A_RM ? r * C + c : c * R + r;
                    const int ix = X_RM ? c * k + h : h * C + c;
                    s0 += a0[ia] * x0[ix];
                    s1 += a1[ia] * x1[ix];
                }
            }
            acc += s0;
            acc += two ? s1 : T(0);
        }
        *yp = bsr_out(acc, yp, alpha, alpha_one, beta, beta_mode);
    }
}

// Row-major X and Y chunks (the default storage): lane t of a group owns block
// row r and V consecutive right-hand sides h0 .. h0+V-1 (entries t, t+L, ... of
// the R x k/V grid of such tiles), so each block costs it C loads of A and C
// vector loads of V X entries.  k % V == 0 and V-aligned X / Y are checked by the
// host.
template <typename T, typename RP, int CT, int V, bool A_RM>
__global__ __launch_bounds__(kBlock) void bsr_spmm_vec_kernel(int mb, int R, int Cr, int k, const RP* __restrict__ rp,
                                                              const int32_t* __restrict__ ci, const T* __restrict__ A,
                                                              T alpha, const T* __restrict__ X, T beta,
                                                              T* __restrict__ Y, int L, int alpha_one, int beta_mode) {
    const int C = CT > 0 ? CT : Cr;
    const int groups = kBlock / L;
    const int t = threadIdx.x % L;
    const int i = blockIdx.x * groups + threadIdx.x / L;
    if (i >= mb) return;
    const RP lo = rp[i], hi = rp[i + 1];
    const int kv = k / V;
    const int E = R * kv;  // V-wide tiles of the output chunk
    const int64_t RC = (int64_t)R * C, CK = (int64_t)C * k;
    T* __restrict__ Yi = Y + (int64_t)i * R * k;
    for (int e = t; e < E; e += L) {
        const int r = e / kv, h0 = (e % kv) * V;
        T* yp = Yi + (int64_t)r * k + h0;
        T acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = T(0);
        if (hi == lo) {  // WriteYZero: Y_i <- beta Y_i
            T y[V];
            if (beta_mode != kBsrBetaZero) load_vec<T, V>(yp, y);
#pragma unroll
            for (int v = 0; v < V; ++v)
                y[v] = beta_mode == kBsrBetaZero ? T(0) : beta_mode == kBsrBetaOne ? y[v] : beta * y[v];
            store_vec<T, V>(yp, y);
            continue;
        }
        for (RP kb = lo; kb < hi; kb += 2) {
            const bool two = kb + 1 < hi;
            const RP kb1 = two ? kb + 1 : kb;
            const int64_t j0 = ci[kb], j1 = ci[kb1];
            const T* a0 = A + (int64_t)kb * RC;
            const T* a1 = A + (int64_t)kb1 * RC;
            const T* x0 = X + j0 * CK + h0;
            const T* x1 = X + j1 * CK + h0;
            T s0[V], s1[V];
#pragma unroll
            for (int v = 0; v < V; ++v) s0[v] = s1[v] = T(0);
            auto step = [&](int c) {
                const int ia = A_RM ? r * C + c : c * R + r;
                const T av0 = a0[ia], av1 = a1[ia];
                T xv0[V], xv1[V];
                load_vec<T, V>(x0 + (int64_t)c * k, xv0);
                load_vec<T, V>(x1 + (int64_t)c * k, xv1);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    s0[v] += av0 * xv0[v];
                    s1[v] += av1 * xv1[v];
                }
            };
            if constexpr (CT > 0) {
#pragma unroll
                for (int c = 0; c < CT; ++c) step(c);
            } else {
                for (int c = 0; c < C; ++c) step(c);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                acc[v] += s0[v];
                acc[v] += two ? s1[v] : T(0);
            }
        }
        T y[V];
        if (beta_mode != kBsrBetaZero) load_vec<T, V>(yp, y);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            T out = alpha_one ? acc[v] : alpha * acc[v];
            if (beta_mode == kBsrBetaOne)
                out = out + y[v];
            else if (beta_mode == kBsrBetaGeneric)
                out = out + beta * y[v];
            y[v] = out;
        }
        store_vec<T, V>(yp, y);
    }
}

// Y <- beta Y over `count` entries (beta == 0: zero, NaN-free).
template <typename T>
__global__ __launch_bounds__(kBlock) void bsr_scale_kernel(int64_t count, T beta, int beta_mode, T* __restrict__ Y) {
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < count; e += (int64_t)gridDim.x * kBlock)
        Y[e] = beta_mode == kBsrBetaZero ? T(0) : beta * Y[e];
}

// FillLowerTriangleFromUpperTriangle (KernelMatrixCSR.hpp:71-128): for every block
// (i, j) with j < i, the block (j, i) of the upper triangle, transposed.  One
// thread per block row; the upper block is found by binary search in row j from
// its diagonal position (the first column >= j).  Square blocks only.  A lower
// block whose mirror is absent from the pattern is left as it is.
template <typename T, typename RP>
__global__ __launch_bounds__(kBlock) void bsr_fill_lower_kernel(int mb, int R, const RP* __restrict__ rp,
                                                                const int32_t* __restrict__ ci, T* A) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= mb) return;
    const int64_t RR = (int64_t)R * R;
    for (RP kb = rp[i]; kb < rp[i + 1] && ci[kb] < i; ++kb) {
        const int j = ci[kb];
        RP lo = rp[j];
        RP hi = rp[j + 1];
        while (lo < hi && ci[lo] < j) ++lo;  // diagonal position of row j (rows are short)
        // last position in [lo, hi) with column <= i
        RP L = lo, Rr = hi - 1;
        while (L < Rr) {
            const RP M = Rr - (Rr - L) / 2;
            if (ci[M] > i)
                Rr = M - 1;
            else
                L = M;
        }
        if (L >= hi || ci[L] != i) continue;
        const T* from = A + (int64_t)L * RR;
        T* to = A + (int64_t)kb * RR;
        for (int a = 0; a < R; ++a)
            for (int b = 0; b < R; ++b) to[(int64_t)a * R + b] = from[(int64_t)b * R + a];
    }
}

template <typename T, typename RP, int CT>
static void launch_bsr_ct(int a_rm, int x_rm, int y_rm, dim3 grid, hipStream_t s, int mb, int R, int C, int k,
                          const RP* rp, const int32_t* ci, const T* A, T alpha, const T* X, T beta, T* Y, int L,
                          int alpha_one, int beta_mode) {
#define TG_BSR(AR, XR, YR)                                                                                    \
    hipLaunchKernelGGL((bsr_spmm_kernel<T, RP, CT, AR, XR, YR>), grid, dim3(kBlock), 0, s, mb, R, C, k, rp, ci, A, \
                       alpha, X, beta, Y, L, alpha_one, beta_mode)
    const int code = (a_rm ? 4 : 0) | (x_rm ? 2 : 0) | (y_rm ? 1 : 0);
    switch (code) {
        case 0: TG_BSR(false, false, false); break;
        case 1: TG_BSR(false, false, true); break;
        case 2: TG_BSR(false, true, false); break;
        case 3: TG_BSR(false, true, true); break;
        case 4: TG_BSR(true, false, false); break;
        case 5: TG_BSR(true, false, true); break;
        case 6: TG_BSR(true, true, false); break;
        default: TG_BSR(true, true, true); break;
    }
#undef TG_BSR
}

template <typename T, typename RP>
static int bsr_typed(int32_t mb, int64_t nnzb, int32_t R, int32_t C, const RP* rp, const int32_t* ci, const T* A,
                     int a_rm, double alpha_d, const T* X, int x_rm, double beta_d, T* Y, int y_rm, int32_t k,
                     hipStream_t s) {
    const T alpha = (T)alpha_d, beta = (T)beta_d;
    const int beta_mode = beta_d == 0.0 ? kBsrBetaZero : beta_d == 1.0 ? kBsrBetaOne : kBsrBetaGeneric;
    const int64_t total = (int64_t)mb * R * k;
    if (alpha_d == 0.0 || nnzb <= 0) {  // Scale (KernelMatrixCSR.hpp:137-151)
        if (beta_mode == kBsrBetaOne || total == 0) return TG_OK;
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock, 8192));
        hipLaunchKernelGGL((bsr_scale_kernel<T>), dim3(grid), dim3(kBlock), 0, s, total, beta, beta_mode, Y);
        TG_LAUNCH_CHECK();
        return TG_OK;
    }
    const int alpha_one = alpha_d == 1.0;
    if (x_rm && y_rm && g_bsr_vector) {  // lane = (block row r, V right-hand sides)
        const auto aligned = [&](int bytes) {
            return (uintptr_t)X % bytes == 0 && (uintptr_t)Y % bytes == 0;
        };
        // V * C X values of two blocks stay in registers: V * C <= 16 (FP64) / 32 (FP32)
        const int cap = (int)(128 / sizeof(T));
        int V = 1;
        if (k % 4 == 0 && aligned(4 * (int)sizeof(T)) && 4 * C <= cap)
            V = 4;
        else if (k % 2 == 0 && aligned(2 * (int)sizeof(T)) && 2 * C <= cap)
            V = 2;
        int L = 1;
        while (L < R * (k / V) && L < kWave) L *= 2;
        const int groups = kBlock / L;
        const dim3 grid((unsigned)((mb + groups - 1) / groups));
#define TG_BSRV(CC, VV)                                                                                            \
    do {                                                                                                           \
        if (a_rm)                                                                                                  \
            hipLaunchKernelGGL((bsr_spmm_vec_kernel<T, RP, CC, VV, true>), grid, dim3(kBlock), 0, s, mb, R, C, k, rp, \
                               ci, A, alpha, X, beta, Y, L, alpha_one, beta_mode);                                 \
        else                                                                                                       \
            hipLaunchKernelGGL((bsr_spmm_vec_kernel<T, RP, CC, VV, false>), grid, dim3(kBlock), 0, s, mb, R, C, k,     \
                               rp, ci, A, alpha, X, beta, Y, L, alpha_one, beta_mode);                             \
    } while (0)
#define TG_BSRVC(CC)             \
    case CC:                     \
        if (V == 4)              \
            TG_BSRV(CC, 4);      \
        else if (V == 2)         \
            TG_BSRV(CC, 2);      \
        else                     \
            TG_BSRV(CC, 1);      \
        break;
        switch (C) {
            TG_BSRVC(1)
            TG_BSRVC(2)
            TG_BSRVC(3)
            TG_BSRVC(4)
            TG_BSRVC(6)
            TG_BSRVC(8)
            default:
                if (V == 4)
                    TG_BSRV(0, 4);
                else if (V == 2)
                    TG_BSRV(0, 2);
                else
                    TG_BSRV(0, 1);
        }
#undef TG_BSRVC
#undef TG_BSRV
        TG_LAUNCH_CHECK();
        return TG_OK;
    }
    int L = 1;
    while (L < R * k && L < kWave) L *= 2;
    const int groups = kBlock / L;
    const dim3 grid((unsigned)((mb + groups - 1) / groups));
    switch (C) {
#define TG_BSRC(CC)                                                                                                 \
    case CC:                                                                                                        \
        launch_bsr_ct<T, RP, CC>(a_rm, x_rm, y_rm, grid, s, mb, R, C, k, rp, ci, A, alpha, X, beta, Y, L, alpha_one, \
                                 beta_mode);                                                                        \
        break;
        TG_BSRC(1)
        TG_BSRC(2)
        TG_BSRC(3)
        TG_BSRC(4)
        TG_BSRC(6)
        TG_BSRC(8)
#undef TG_BSRC
        default:
            launch_bsr_ct<T, RP, 0>(a_rm, x_rm, y_rm, grid, s, mb, R, C, k, rp, ci, A, alpha, X, beta, Y, L,
                                    alpha_one, beta_mode);
    }
    TG_LAUNCH_CHECK();
    return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" int tg_bsr_spmm(int32_t mb, int32_t nb, int64_t nnzb, int32_t R, int32_t C, const void* rp,
                           tg_dtype rp_type, const int32_t* ci, const void* a, tg_dtype val_type, int32_t a_row_major,
                           double alpha, const void* X, int32_t x_row_major, double beta, void* Y,
                           int32_t y_row_major, int32_t nrhs, void* stream) {
    if (mb < 0 || nb < 0 || nnzb < 0 || R < 1 || C < 1 || nrhs < 0 || R > 64 || C > 64)
        return fail(TG_ERR_INVALID, "tg_bsr_spmm: bad size (block sizes must be in [1, 64])");
    if (val_type != TG_F64 && val_type != TG_F32) return fail(TG_ERR_DTYPE, "tg_bsr_spmm: val_type must be F64/F32");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_bsr_spmm: rp_type must be I32/I64");
    if (mb == 0 || nrhs == 0) return TG_OK;
    if (!Y) return fail(TG_ERR_INVALID, "tg_bsr_spmm: Y is null");
    const bool compute = alpha != 0.0 && nnzb > 0;
    if (compute) {
        if (!rp || !ci || !a || !X) return fail(TG_ERR_INVALID, "tg_bsr_spmm: rp, ci, values or X is null");
        if (X == Y) return fail(TG_ERR_ALIAS, "tg_bsr_spmm: input and output pointer coincide. Doing nothing.");
    }
    hipStream_t s = as_stream(stream);
    const int ar = a_row_major != 0, xr = x_row_major != 0, yr = y_row_major != 0;
    if (val_type == TG_F64) {
        if (rp_type == TG_I64)
            return bsr_typed<double, int64_t>(mb, nnzb, R, C, (const int64_t*)rp, ci, (const double*)a, ar, alpha,
                                              (const double*)X, xr, beta, (double*)Y, yr, nrhs, s);
        return bsr_typed<double, int32_t>(mb, nnzb, R, C, (const int32_t*)rp, ci, (const double*)a, ar, alpha,
                                          (const double*)X, xr, beta, (double*)Y, yr, nrhs, s);
    }
    if (rp_type == TG_I64)
        return bsr_typed<float, int64_t>(mb, nnzb, R, C, (const int64_t*)rp, ci, (const float*)a, ar, alpha,
                                         (const float*)X, xr, beta, (float*)Y, yr, nrhs, s);
    return bsr_typed<float, int32_t>(mb, nnzb, R, C, (const int32_t*)rp, ci, (const float*)a, ar, alpha,
                                     (const float*)X, xr, beta, (float*)Y, yr, nrhs, s);
}

extern "C" int tg_bsr_fill_lower(int32_t mb, const void* rp, tg_dtype rp_type, const int32_t* ci, int32_t R, void* a,
                                 tg_dtype val_type, void* stream) {
    if (mb < 0 || R < 1 || !rp || !a) return fail(TG_ERR_INVALID, "tg_bsr_fill_lower: bad argument");
    if (val_type != TG_F64 && val_type != TG_F32)
        return fail(TG_ERR_DTYPE, "tg_bsr_fill_lower: val_type must be F64/F32");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_bsr_fill_lower: rp_type must be I32/I64");
    if (mb == 0) return TG_OK;
    hipStream_t s = as_stream(stream);
    const dim3 grid((unsigned)((mb + kBlock - 1) / kBlock));
#define TG_FL(TT, RR) \
    hipLaunchKernelGGL((bsr_fill_lower_kernel<TT, RR>), grid, dim3(kBlock), 0, s, mb, R, (const RR*)rp, ci, (TT*)a)
    if (val_type == TG_F64) {
        if (rp_type == TG_I64)
            TG_FL(double, int64_t);
        else
            TG_FL(double, int32_t);
    } else {
        if (rp_type == TG_I64)
            TG_FL(float, int64_t);
        else
            TG_FL(float, int32_t);
    }
#undef TG_FL
    TG_LAUNCH_CHECK();
    return TG_OK;
}
