// CSR SpMV / SpMM for gfx950:  Y <- alpha * A X + beta * Y.
//
// Reference semantics: SparseBLAS::Multiply_DenseMatrix routing
// (src/SparseBLAS/Multiply_DenseMatrix.hpp:9-45), SpMV (src/SparseBLAS/SpMV.hpp:
// 3-288), SpMM (src/SparseBLAS/SpMM.hpp:3-569).  The reference's compile-time
// flag dispatch (a_flag x alpha_flag x beta_flag, SpMM.hpp:28-131) becomes
// wave-uniform runtime branches here: they cost nothing in a bandwidth-bound
// kernel and keep the instantiation count down.
//
// Two kernel families, both HBM-bound (no MFMA: ~0.1-5 flop/byte):
//  * spmv_team<TEAM>: nrhs == 1.  A team of TEAM lanes (power of two, picked
//    from the mean row length) owns one row; lanes stride through the row so
//    the int32 column indices and values stream coalesced; the row sum is a
//    butterfly over the team.  64/TEAM rows per wave.
//  * spmm_wave<G,V>: nrhs >= 2.  One wavefront owns one row.  G lanes cover a
//    tile of G*V right-hand-side columns (V-wide vector loads of the row-major
//    X rows), and the 64/G lane groups of the wave split the row's nonzeros
//    between them.  The wave loads 64 (column, value) pairs coalesced and
//    hands each group its pair through a cross-lane read; partial sums of the
//    groups are combined by butterfly at the end of the row.
// Rows are distributed over workgroups either evenly (rows_per_block) or by a
// device partition (job_ptr) that balances nonzeros + rows per block.

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>

#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

enum BetaMode : int { kBetaZero = 0, kBetaOne = 1, kBetaGeneric = 2 };

// Row epilogue for V consecutive entries of one Y row: the combine_scalars /
// combine_buffers contract y = (alpha * z) + (beta * y) (SpMV.hpp:224-225,
// SpMM.hpp:388-393) and the empty-row rules (SpMM.hpp:395-410).
// The rounding is fixed, not left to the compiler's contraction: alpha * z rounded
// on its own, beta * y + (alpha * z) as one fused multiply-add, y + (alpha * z) as
// an add; an empty row with a general beta gets beta * y rounded.
template <typename T>
__device__ __forceinline__ T fma_rn(T a, T b, T c) {
    if constexpr (sizeof(T) == 8)
        return __builtin_fma(a, b, c);
    else
        return __builtin_fmaf(a, b, c);
}

template <typename T, int V>
__device__ __forceinline__ void epilogue(T* __restrict__ yp, const T (&z)[V], bool empty, T alpha,
                                         int alpha_one, T beta, int beta_mode) {
#pragma clang fp contract(off)
    T out[V];
    if (empty) {
        if (beta_mode == kBetaOne) return;
        if (beta_mode == kBetaZero) {
#pragma unroll
            for (int i = 0; i < V; ++i) out[i] = T(0);
        } else {
            load_vec<T, V>(yp, out);
#pragma unroll
            for (int i = 0; i < V; ++i) out[i] = beta * out[i];
        }
        store_vec<T, V>(yp, out);
        return;
    }
#pragma unroll
    for (int i = 0; i < V; ++i) out[i] = alpha_one ? z[i] : alpha * z[i];
    if (beta_mode != kBetaZero) {
        T yv[V];
        load_vec<T, V>(yp, yv);
        if (beta_mode == kBetaOne) {
#pragma unroll
            for (int i = 0; i < V; ++i) out[i] = out[i] + yv[i];
        } else {
#pragma unroll
            for (int i = 0; i < V; ++i) out[i] = fma_rn(beta, yv[i], out[i]);
        }
    }
    store_vec<T, V>(yp, out);
}

__device__ __forceinline__ void block_rows(int m, const int32_t* __restrict__ job_ptr, int rows_per_block,
                                           int& r0, int& r1) {
    const int b = xcd_swizzle(blockIdx.x, gridDim.x);
    if (job_ptr) {
        r0 = job_ptr[b];
        r1 = job_ptr[b + 1];
    } else {
        r0 = b * rows_per_block;
        r1 = min(m, r0 + rows_per_block);
    }
}

// ---------------------------------------------------------------------------
// nrhs == 1
// ---------------------------------------------------------------------------
template <typename T, typename RP, int TEAM>
__global__ __launch_bounds__(kBlock) void spmv_team_kernel(
    int m, const RP* __restrict__ rp, const RP* __restrict__ rpe, const int32_t* __restrict__ ci, const T* __restrict__ a, T alpha,
    const T* __restrict__ x, int64_t ldx, T beta, T* __restrict__ y, int64_t ldy, int alpha_one,
    int beta_mode, const int32_t* __restrict__ job_ptr, int rows_per_block, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int TEAMS = kBlock / TEAM;
    const int lt = threadIdx.x % TEAM;
    const int team = threadIdx.x / TEAM;
    int r0, r1;
    block_rows(m, job_ptr, rows_per_block, r0, r1);
    for (int row = r0 + team; row < r1; row += TEAMS) {
        const RP lb = rp[row], le = rpe[row];
        T z = T(0);
        if (a) {
            for (RP l = lb + lt; l < le; l += TEAM) {
                const int32_t c = __builtin_nontemporal_load(ci + l);
                const T av = __builtin_nontemporal_load(a + l);
                z += av * x[(int64_t)c * ldx];
            }
        } else {
            for (RP l = lb + lt; l < le; l += TEAM) {
                const int32_t c = __builtin_nontemporal_load(ci + l);
                z += x[(int64_t)c * ldx];
            }
        }
        z = group_sum<TEAM>(z);
        if (lt == 0) {
            T zz[1] = {z};
            epilogue<T, 1>(y + (int64_t)row * ldy, zz, le == lb, alpha, alpha_one, beta, beta_mode);
        }
    }
}

// nrhs == 1, streaming form: the workgroup walks its nonzeros in chunks of
// kBlock * ELEMS.  Every lane loads ELEMS (column, value) pairs of the chunk
// with coalesced loads and gathers the x entries -- all independent, so a wave
// keeps 2*ELEMS loads in flight instead of one dependent rp -> ci -> x chain
// per row -- and leaves the products in LDS.  Teams of TEAM lanes then sum each
// row's segment out of LDS (strided, butterfly); rows that span chunks
// accumulate in LDS.  Rows are taken RMAX at a time.
template <typename T, typename RP, int ELEMS, int TEAM>
__global__ __launch_bounds__(kBlock) void spmv_stream_kernel(
    int m, const RP* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ a, T alpha,
    const T* __restrict__ x, int64_t ldx, T beta, T* __restrict__ y, int64_t ldy, int alpha_one, int beta_mode,
    const int32_t* __restrict__ job_ptr, int rows_per_block, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int CAP = kBlock * ELEMS;
    constexpr int RMAX = 1024;
    constexpr int TEAMS = kBlock / TEAM;
    __shared__ T prod[CAP];
    __shared__ RP srp[RMAX + 1];
    __shared__ T racc[RMAX];
    int r0, r1;
    block_rows(m, job_ptr, rows_per_block, r0, r1);
    const int lt = threadIdx.x % TEAM, team = threadIdx.x / TEAM;
    for (int rs = r0; rs < r1; rs += RMAX) {
        const int nr = min(RMAX, r1 - rs);
        for (int i = threadIdx.x; i <= nr; i += kBlock) srp[i] = rp[rs + i];
        for (int i = threadIdx.x; i < nr; i += kBlock) racc[i] = T(0);
        __syncthreads();
        const RP L0 = srp[0], L1 = srp[nr];
        for (RP cs = L0; cs < L1; cs += CAP) {
            const int cnt = (int)((L1 - cs) < (RP)CAP ? (L1 - cs) : (RP)CAP);
            int32_t c[ELEMS];
            T av[ELEMS];
#pragma unroll
            for (int j = 0; j < ELEMS; ++j) {
                const int e = threadIdx.x + j * kBlock;
                c[j] = 0;
                av[j] = T(0);
                if (e < cnt) {
                    c[j] = __builtin_nontemporal_load(ci + cs + e);
                    av[j] = a ? __builtin_nontemporal_load(a + cs + e) : T(1);
                }
            }
            // gathers unconditionally (c = 0 for the lanes past the chunk, a valid
            // entry) so that all ELEMS are in flight together; only the store is masked
            T xv[ELEMS];
#pragma unroll
            for (int j = 0; j < ELEMS; ++j) xv[j] = x[(int64_t)c[j] * ldx];
#pragma unroll
            for (int j = 0; j < ELEMS; ++j) {
                const int e = threadIdx.x + j * kBlock;
                if (e < cnt) prod[e] = av[j] * xv[j];
            }
            __syncthreads();
            for (int i = team; i < nr; i += TEAMS) {
                const RP lb = srp[i] > cs ? srp[i] : cs;
                const RP le = srp[i + 1] < cs + cnt ? srp[i + 1] : cs + cnt;
                T sum = T(0);
                for (RP l = lb + lt; l < le; l += TEAM) sum += prod[l - cs];
                sum = group_sum<TEAM>(sum);
                if (lt == 0 && le > lb) racc[i] += sum;
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < nr; i += kBlock) {
            T zz[1] = {racc[i]};
            epilogue<T, 1>(y + (int64_t)(rs + i) * ldy, zz, srp[i + 1] == srp[i], alpha, alpha_one, beta, beta_mode);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// nrhs >= 2
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T lane_read(T v, int src) {
    return __shfl(v, src, kWave);
}

// Broadcast of lane `src` (wave-uniform) through the scalar unit.
__device__ __forceinline__ int32_t lane_read_uniform(int32_t v, int src) {
    return __builtin_amdgcn_readlane(v, src);
}
__device__ __forceinline__ float lane_read_uniform(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_read_uniform(double v, int src) {
    const int64_t b = __double_as_longlong(v);
    const int32_t lo = __builtin_amdgcn_readlane((int32_t)(b & 0xffffffff), src);
    const int32_t hi = __builtin_amdgcn_readlane((int32_t)(b >> 32), src);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

template <typename T, typename RP, int G, int V>
__global__ __launch_bounds__(kBlock) void spmm_wave_kernel(
    int m, const RP* __restrict__ rp, const RP* __restrict__ rpe, const int32_t* __restrict__ ci, const T* __restrict__ a, T alpha,
    const T* __restrict__ X, int64_t ldX, T beta, T* __restrict__ Y, int64_t ldY, int nrhs, int alpha_one,
    int beta_mode, const int32_t* __restrict__ job_ptr, int rows_per_block, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int S = kWave / G;  // lane groups splitting a row's nonzeros
    constexpr int KT = G * V;     // columns per tile
    constexpr int WAVES = kBlock / kWave;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int s = lane / G, g = lane % G;
    int r0, r1;
    block_rows(m, job_ptr, rows_per_block, r0, r1);
    for (int row = r0 + wave; row < r1; row += WAVES) {
        const RP lb = rp[row], le = rpe[row];
        T* __restrict__ yrow = Y + (int64_t)row * ldY;
        for (int c0 = 0; c0 < nrhs; c0 += KT) {
            const int col = c0 + g * V;
            const bool colok = col < nrhs;
            T acc[V];
#pragma unroll
            for (int i = 0; i < V; ++i) acc[i] = T(0);
            for (RP base = lb; base < le; base += kWave) {
                const RP e = base + lane;
                int32_t myc = 0;
                T mya = T(0);
                if (e < le) {
                    myc = __builtin_nontemporal_load(ci + e);
                    mya = a ? __builtin_nontemporal_load(a + e) : T(1);
                }
                const int cnt = (int)((le - base) < (RP)kWave ? (le - base) : (RP)kWave);
                // G steps of S nonzeros cover the 64 loaded pairs.  Steps run in
                // groups of U without branches inside a group, so the X loads of
                // a group are all in flight together; the wave-uniform test
                // between groups skips the tail of short rows.
                constexpr int U = G < 4 ? G : 4;
                for (int st0 = 0; st0 < G; st0 += U) {
                    if (st0 * S >= cnt) break;
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int j = (st0 + u) * S + s;
                        int32_t c;
                        T av;
                        if constexpr (S == 1) {
                            c = lane_read_uniform(myc, st0 + u);
                            av = lane_read_uniform(mya, st0 + u);
                        } else {
                            c = lane_read(myc, j);
                            av = lane_read(mya, j);
                        }
                        const bool live = (j < cnt) && colok;
                        T xv[V];
                        load_vec<T, V>(X + (int64_t)(live ? c : 0) * ldX + (live ? col : 0), xv);
#pragma unroll
                        for (int i = 0; i < V; ++i) acc[i] += live ? av * xv[i] : T(0);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < V; ++i) {
#pragma unroll
                for (int off = G; off < kWave; off <<= 1) acc[i] += shfl_xor(acc[i], off);
            }
            if (s == 0 && colok) epilogue<T, V>(yrow + col, acc, le == lb, alpha, alpha_one, beta, beta_mode);
        }
    }
}

// spmm_rows<G,V>: short rows (grid and FEM matrices, a few nonzeros per row) with
// many right-hand sides.  Splitting one such row over a whole wave leaves most
// lanes idle and makes every row a chain of dependent loads; here each of the
// 64/G lane groups of a wave owns its own row: the group loads G (column, value)
// pairs of the row at once, hands them out with cross-lane reads, and gathers the
// X rows U at a time, so a wave keeps 64/G rows and 4 gathers per lane in flight.
// A row's sum stays in its group (no butterfly); the summation order is the
// nonzero order.
template <typename T, typename RP, int G, int V>
__global__ __launch_bounds__(kBlock) void spmm_rows_kernel(
    int m, const RP* __restrict__ rp, const RP* __restrict__ rpe, const int32_t* __restrict__ ci, const T* __restrict__ a, T alpha,
    const T* __restrict__ X, int64_t ldX, T beta, T* __restrict__ Y, int64_t ldY, int nrhs, int alpha_one,
    int beta_mode, const int32_t* __restrict__ job_ptr, int rows_per_block, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int S = kWave / G;  // rows in flight per wave
    constexpr int KT = G * V;     // columns per tile
    constexpr int WAVES = kBlock / kWave;
    constexpr int U = 4;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int s = lane / G, g = lane % G;
    int r0, r1;
    block_rows(m, job_ptr, rows_per_block, r0, r1);
    for (int rbase = r0 + wave * S; rbase < r1; rbase += WAVES * S) {
        const int row = rbase + s;
        const bool rok = row < r1;
        const RP lb = rok ? rp[row] : RP(0), le = rok ? rpe[row] : RP(0);
        const int len = (int)(le - lb);
        int wmax = len;  // longest row of the wave (uniform loop bound)
#pragma unroll
        for (int off = G; off < kWave; off <<= 1) wmax = max(wmax, __shfl_xor(wmax, off, kWave));
        for (int c0 = 0; c0 < nrhs; c0 += KT) {
            const int col = c0 + g * V;
            const bool colok = col < nrhs;
            T acc[V];
#pragma unroll
            for (int i = 0; i < V; ++i) acc[i] = T(0);
            for (int base = 0; base < wmax; base += G) {
                const int e = base + g;
                int32_t myc = 0;
                T mya = T(0);
                if (e < len) {
                    myc = __builtin_nontemporal_load(ci + lb + e);
                    mya = a ? __builtin_nontemporal_load(a + lb + e) : T(1);
                }
                const int cnt = min(G, len - base);  // this group's pairs in the chunk
                const int wcnt = min(G, wmax - base);
                for (int u0 = 0; u0 < wcnt; u0 += U) {
#pragma unroll
                    for (int uu = 0; uu < U; ++uu) {
                        const int j = u0 + uu;
                        const int src = s * G + (j < G ? j : 0);
                        const int32_t c = __shfl(myc, src, kWave);
                        const T av = __shfl(mya, src, kWave);
                        const bool live = (j < cnt) && colok;
                        T xv[V];
                        load_vec<T, V>(X + (int64_t)(live ? c : 0) * ldX + (live ? col : 0), xv);
#pragma unroll
                        for (int i = 0; i < V; ++i) acc[i] += live ? av * xv[i] : T(0);
                    }
                }
            }
            if (rok && colok) epilogue<T, V>(Y + (int64_t)row * ldY + col, acc, le == lb, alpha, alpha_one, beta, beta_mode);
        }
    }
}

// Y <- beta * Y over m x nrhs (alpha == 0 path of SpMM.hpp:18-25 / SpMV.hpp:14-29)
// and the general axpby used by the solvers.
template <typename T>
__global__ __launch_bounds__(kBlock) void axpby_kernel(int m, int nrhs, T alpha, const T* __restrict__ X,
                                                       int64_t ldX, T beta, T* __restrict__ Y, int64_t ldY,
                                                       int alpha_mode, int beta_mode) {
    const int64_t total = (int64_t)m * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / nrhs;
        const int k = (int)(e - i * nrhs);
        T* yp = Y + i * ldY + k;
        T out = T(0);
        if (alpha_mode != kBetaZero) {
            const T xv = X[i * ldX + k];
            out = (alpha_mode == kBetaOne) ? xv : alpha * xv;
        }
        if (beta_mode == kBetaOne)
            out = out + *yp;
        else if (beta_mode == kBetaGeneric)
            out = out + beta * *yp;
        *yp = out;
    }
}

template <typename RP>
__global__ void job_pointers_kernel(int m, const RP* __restrict__ rp, int blocks, int32_t* __restrict__ jp) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > blocks) return;
    if (b == 0) {
        jp[0] = 0;
        return;
    }
    if (b == blocks) {
        jp[blocks] = m;
        return;
    }
    const int64_t total = (int64_t)rp[m] - (int64_t)rp[0] + m;
    const int64_t target = (int64_t)(((__int128)total * b) / blocks) + (int64_t)rp[0];
    // first i in [0, m] with rp[i] + i >= target (rp[i] + i is strictly increasing)
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if ((int64_t)rp[mid] + mid >= target)
            hi = mid;
        else
            lo = mid + 1;
    }
    jp[b] = lo;
}

// ---------------------------------------------------------------------------
// Host-side dispatch.
// ---------------------------------------------------------------------------
static int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

constexpr int kRowsPerTeam = 8;

// Shape overrides (0 = automatic), set through tg_tune_set by the autotuner:
// lanes per row of the nrhs == 1 kernels, and lanes per RHS tile of spmm_wave.
static int g_spmv_team = 0;
static int g_spmm_group = 0;
static int g_spmm_rows = 0;  // nrhs >= 2 kernel: 0 = by mean row length, 1 = row per wave, 2 = row per lane group
// planned SpMM: 0 = by X size (slice_count, the default), 1 = never sliced,
// 2/4/8 = that many slices
static int g_spmm_slices = 0;
// planned SpMM with values registered by tg_csr_plan_set_values: 1 = the re-laid
// layout where it pays (default), 0 = never
static int g_spmm_split = 1;
// rows longer than this are split off by plans created afterwards (0: never)
static int g_spmm_long_row = 8192;

// Lanes per row of the nrhs == 1 kernels.  Team-per-row: the row length rounded up
// (4..64).  Streaming (the products are formed with coalesced loads first; the team
// only adds a row's products from LDS): a quarter of the row length rounded down
// (1..32) -- on 5- and 7-point stencils one lane per row measured 31 % and 23 %
// faster than 8, on the random 32-per-row matrix every width from 1 to 32 the same,
// and long rows keep wide teams so that no lane adds a whole chunk alone.
static int spmv_team_width(int32_t m, int64_t nnz, bool stream) {
    if (g_spmv_team) return g_spmv_team;
    const double avg = m > 0 ? (double)nnz / m : 0.0;
    if (stream) {
        int t = 1;
        while (t < 32 && 2.0 * t <= avg / 4.0) t *= 2;
        return t;
    }
    int t = pow2ceil((int)std::ceil(avg));
    return std::min(64, std::max(4, t));
}

static int stream_rows_per_block(int32_t m, int64_t nnz);
static int spmv_variant();

static int rows_per_block_for(int32_t m, int64_t nnz, int32_t nrhs) {
    if (nrhs == 1)
        return spmv_variant() == 0 ? stream_rows_per_block(m, nnz)
                                   : (kBlock / std::max(4, spmv_team_width(m, nnz, false))) * kRowsPerTeam;
    return (kBlock / kWave) * kRowsPerTeam;
}

constexpr int kStreamElems = 8;
constexpr int kStreamCap = kBlock * kStreamElems;

// nrhs == 1 kernel choice: 0 = streaming (default), 1 = team-per-row.  Set
// through tg_tune_set("spmv_kernel", v); the environment variable
// TG_SPMV_KERNEL ("stream" / "team") gives the initial value.
static int g_spmv_variant = -1;
static int spmv_variant() {
    if (g_spmv_variant < 0) {
        const char* e = std::getenv("TG_SPMV_KERNEL");
        g_spmv_variant = (e && std::string(e) == "team") ? 1 : 0;
    }
    return g_spmv_variant;
}

static int stream_rows_per_block(int32_t m, int64_t nnz) {
    const double avg = m > 0 ? (double)nnz / m : 0.0;
    int r = (int)(kStreamCap / (avg + 1.0));
    return std::max(1, std::min(1024, r));
}

template <typename T, typename RP>
static int launch_spmv(int32_t m, int64_t nnz, const RP* rp, const RP* rpe, const int32_t* ci, const T* a, T alpha,
                       const T* X, int64_t ldX, T beta, T* Y, int64_t ldY, int alpha_one, int beta_mode,
                       const int32_t* job_ptr, int32_t blocks, const int* skip, hipStream_t st) {
    // the streaming kernel walks whole nonzero ranges: with split-off long rows (row
    // ends other than rp + 1) the team-per-row kernel, which reads each row's end
    const bool stream = spmv_variant() == 0 && rpe == rp + 1;
    const int team = spmv_team_width(m, nnz, stream);
    if (stream) {
        const int rpb = stream_rows_per_block(m, nnz);
        const int grid = job_ptr ? blocks : (m + rpb - 1) / rpb;
        if (grid <= 0) return TG_OK;
        const int steam = std::min(team, 32);
#define TG_SPMVS(TM)                                                                                          \
    case TM:                                                                                                  \
        hipLaunchKernelGGL((spmv_stream_kernel<T, RP, kStreamElems, TM>), dim3(grid), dim3(kBlock), 0, st, m, rp, \
                           ci, a, alpha, X, ldX, beta, Y, ldY, alpha_one, beta_mode, job_ptr, rpb, skip);      \
        break;
        switch (steam) {
            TG_SPMVS(1)
            TG_SPMVS(2)
            TG_SPMVS(4)
            TG_SPMVS(8)
            TG_SPMVS(16)
            TG_SPMVS(32)
            default: return fail(TG_ERR_INVALID, "internal: bad team width %d", steam);
        }
#undef TG_SPMVS
        TG_LAUNCH_CHECK();
        return TG_OK;
    }
    const int tteam = std::max(4, team);  // the team-per-row kernel has teams of 4..64 lanes
    const int rpb = (kBlock / tteam) * kRowsPerTeam;
    const int grid = job_ptr ? blocks : (m + rpb - 1) / rpb;
    if (grid <= 0) return TG_OK;
#define TG_SPMV(TM)                                                                                     \
    case TM:                                                                                            \
        hipLaunchKernelGGL((spmv_team_kernel<T, RP, TM>), dim3(grid), dim3(kBlock), 0, st, m, rp, rpe, ci, a, \
                           alpha, X, ldX, beta, Y, ldY, alpha_one, beta_mode, job_ptr, rpb, skip);      \
        break;
    switch (tteam) {
        TG_SPMV(4)
        TG_SPMV(8)
        TG_SPMV(16)
        TG_SPMV(32)
        TG_SPMV(64)
        default: return fail(TG_ERR_INVALID, "internal: bad team width %d", team);
    }
#undef TG_SPMV
    TG_LAUNCH_CHECK();
    return TG_OK;
}

template <typename T, typename RP, int V>
static int launch_spmm_v(int G, bool rows_mode, int32_t m, const RP* rp, const RP* rpe, const int32_t* ci, const T* a,
                         T alpha,
                         const T* X, int64_t ldX, T beta, T* Y, int64_t ldY, int32_t nrhs, int alpha_one,
                         int beta_mode, const int32_t* job_ptr, int32_t blocks, const int* skip, hipStream_t st) {
    if (rows_mode && G < kWave) {
        // 64/G rows per wave at a time; kRowsPerTeam rounds per wave per block
        const int rpb = (kBlock / G) * kRowsPerTeam;
        const int grid = job_ptr ? blocks : (m + rpb - 1) / rpb;
        if (grid <= 0) return TG_OK;
#define TG_SPMMR(GG)                                                                                          \
    case GG:                                                                                                  \
        hipLaunchKernelGGL((spmm_rows_kernel<T, RP, GG, V>), dim3(grid), dim3(kBlock), 0, st, m, rp, rpe, ci, a, alpha, \
                           X, ldX, beta, Y, ldY, nrhs, alpha_one, beta_mode, job_ptr, rpb, skip);             \
        break;
        switch (G) {
            TG_SPMMR(1)
            TG_SPMMR(2)
            TG_SPMMR(4)
            TG_SPMMR(8)
            TG_SPMMR(16)
            TG_SPMMR(32)
            default: return fail(TG_ERR_INVALID, "internal: bad lane group %d", G);
        }
#undef TG_SPMMR
        TG_LAUNCH_CHECK();
        return TG_OK;
    }
    const int rpb = (kBlock / kWave) * kRowsPerTeam;
    const int grid = job_ptr ? blocks : (m + rpb - 1) / rpb;
    if (grid <= 0) return TG_OK;
#define TG_SPMM(GG)                                                                                      \
    case GG:                                                                                             \
        hipLaunchKernelGGL((spmm_wave_kernel<T, RP, GG, V>), dim3(grid), dim3(kBlock), 0, st, m, rp, rpe, ci, a, \
                           alpha, X, ldX, beta, Y, ldY, nrhs, alpha_one, beta_mode, job_ptr, rpb, skip); \
        break;
    switch (G) {
        TG_SPMM(1)
        TG_SPMM(2)
        TG_SPMM(4)
        TG_SPMM(8)
        TG_SPMM(16)
        TG_SPMM(32)
        TG_SPMM(64)
        default: return fail(TG_ERR_INVALID, "internal: bad lane group %d", G);
    }
#undef TG_SPMM
    TG_LAUNCH_CHECK();
    return TG_OK;
}

template <typename T, typename RP>
static int spmm_typed(int32_t m, int64_t nnz, const RP* rp, const int32_t* ci, const T* a, double alpha_d,
                      const T* X, int64_t ldX, double beta_d, T* Y, int64_t ldY, int32_t nrhs,
                      const int32_t* job_ptr, int32_t blocks, const int* skip, hipStream_t st,
                      const RP* rpe = nullptr) {
    if (!rpe) rpe = rp + 1;  // row i ends at rpe[i]
    const T alpha = (T)alpha_d, beta = (T)beta_d;
    const int alpha_one = (alpha_d == 1.0);
    const int beta_mode = (beta_d == 0.0) ? kBetaZero : (beta_d == 1.0) ? kBetaOne : kBetaGeneric;
    if (nrhs == 1)
        return launch_spmv<T, RP>(m, nnz, rp, rpe, ci, a, alpha, X, ldX, beta, Y, ldY, alpha_one, beta_mode,
                                  job_ptr, blocks, skip, st);
    // Vector width: 16-byte loads when X and Y rows allow it.
    const auto aligned = [&](int bytes) {
        return ((uintptr_t)X % bytes == 0) && ((uintptr_t)Y % bytes == 0) &&
               ((ldX * (int64_t)sizeof(T)) % bytes == 0) && ((ldY * (int64_t)sizeof(T)) % bytes == 0);
    };
    int V = 1;
    if (sizeof(T) == 8) {
        if (nrhs % 2 == 0 && aligned(16)) V = 2;
    } else {
        if (nrhs % 4 == 0 && aligned(16))
            V = 4;
        else if (nrhs % 2 == 0 && aligned(8))
            V = 2;
    }
    const int G = g_spmm_group ? g_spmm_group : std::min(64, pow2ceil((nrhs + V - 1) / V));
    // short rows: one row per lane group (auto: mean row length up to 16)
    const double avg = m > 0 ? (double)nnz / m : 0.0;
    const bool rows_mode = g_spmm_rows == 2 || (g_spmm_rows == 0 && avg <= 16.0);
    switch (V) {
        case 1:
            return launch_spmm_v<T, RP, 1>(G, rows_mode, m, rp, rpe, ci, a, alpha, X, ldX, beta, Y, ldY, nrhs, alpha_one,
                                           beta_mode, job_ptr, blocks, skip, st);
        case 2:
            return launch_spmm_v<T, RP, 2>(G, rows_mode, m, rp, rpe, ci, a, alpha, X, ldX, beta, Y, ldY, nrhs, alpha_one,
                                           beta_mode, job_ptr, blocks, skip, st);
        default:
            if constexpr (sizeof(T) == 4)
                return launch_spmm_v<T, RP, 4>(G, rows_mode, m, rp, rpe, ci, a, alpha, X, ldX, beta, Y, ldY, nrhs, alpha_one,
                                               beta_mode, job_ptr, blocks, skip, st);
            return fail(TG_ERR_INVALID, "internal: bad vector width");
    }
}

template <typename T>
static int axpby_typed(int32_t m, int32_t nrhs, double alpha_d, const T* X, int64_t ldX, double beta_d, T* Y,
                       int64_t ldY, hipStream_t st) {
    const int alpha_mode = (alpha_d == 0.0) ? kBetaZero : (alpha_d == 1.0) ? kBetaOne : kBetaGeneric;
    const int beta_mode = (beta_d == 0.0) ? kBetaZero : (beta_d == 1.0) ? kBetaOne : kBetaGeneric;
    if (alpha_mode == kBetaZero && beta_mode == kBetaOne) return TG_OK;  // nothing to do
    const int64_t total = (int64_t)m * nrhs;
    if (total <= 0) return TG_OK;
    const int grid = (int)std::min<int64_t>((total + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL((axpby_kernel<T>), dim3(grid), dim3(kBlock), 0, st, m, nrhs, (T)alpha_d, X, ldX,
                       (T)beta_d, Y, ldY, alpha_mode, beta_mode);
    TG_LAUNCH_CHECK();
    return TG_OK;
}


// ---------------------------------------------------------------------------
// Column-sliced SpMM for dense blocks a few L2s large (nrhs <= 4).
//
// Each XCD has its own 4 MB L2, and a random-column gather from an X of 8-32 MB
// misses it almost every time (a 128-byte line fill per nonzero).  Here the
// columns are cut into S equal ranges (S = 2, 4, 8), and the workgroups that
// the hardware places on XCD x (blockIdx % 8 == x) only gather from range
// x % S, so each XCD's L2 holds its slice of X.  Columns are sorted within a
// row, so a row's nonzeros of one range form one run; the plan stores the run
// boundaries of the 8-way cut (7 int32 per row; S = 4 and 2 use every second
// / fourth).  The slices' partial sums go to S partial blocks that a second
// kernel adds in slice order, with the alpha/beta epilogue.  Extra traffic:
// S partial blocks written and read, 7 cut arrays read.
// ---------------------------------------------------------------------------
constexpr int kCuts = kXcds - 1;

template <typename RP>
__global__ __launch_bounds__(kBlock) void slice_plan_kernel(int m, int n, const RP* __restrict__ rp,
                                                            const int32_t* __restrict__ ci,
                                                            int32_t* __restrict__ cuts, int* __restrict__ unsorted) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const int64_t lb = rp[i], le = rp[i + 1];
        bool bad = false;
        for (int64_t l = lb + 1; l < le; ++l) bad |= ci[l] < ci[l - 1];
        if (bad) *unsorted = 1;
        for (int s = 1; s <= kCuts; ++s) {
            const int32_t bound = (int32_t)((int64_t)n * s / kXcds);
            int64_t lo = lb, hi = le;  // first l with ci[l] >= bound
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (ci[mid] < bound)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            cuts[(int64_t)(s - 1) * m + i] = (int32_t)(lo - lb);  // offset within the row
        }
    }
}

// Run lengths of the re-laid layout: len[s * m + i] = nonzeros of row i in column
// slice s of the 8-way cut (len[8m] = 0, so the exclusive scan ends with nnz).
template <typename RP>
__global__ __launch_bounds__(kBlock) void split_len_kernel(int m, const RP* __restrict__ rp,
                                                           const int32_t* __restrict__ cuts, int32_t* __restrict__ len) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const int32_t total = (int32_t)(rp[i + 1] - rp[i]);
        int32_t prev = 0;
        for (int s = 0; s < kXcds; ++s) {
            const int32_t cut = s == kXcds - 1 ? total : cuts[(int64_t)s * m + i];
            len[(int64_t)s * m + i] = cut - prev;
            prev = cut;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) len[(int64_t)kXcds * m] = 0;
}

// dst[srp[s*m + i] + j] = src[rp[i] + (start of slice s in row i) + j]: ci or values
// in the slice-major order of the re-laid layout (a row per thread; built once).
template <typename RP, typename V>
__global__ __launch_bounds__(kBlock) void split_scatter_kernel(int m, const RP* __restrict__ rp,
                                                               const int32_t* __restrict__ cuts,
                                                               const int32_t* __restrict__ srp,
                                                               const V* __restrict__ src, V* __restrict__ dst) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const RP base = rp[i];
        int32_t from = 0;
        for (int s = 0; s < kXcds; ++s) {
            const int32_t to = s == kXcds - 1 ? (int32_t)(rp[i + 1] - base) : cuts[(int64_t)s * m + i];
            const int32_t d0 = srp[(int64_t)s * m + i];
            for (int32_t j = from; j < to; ++j) dst[d0 + (j - from)] = src[base + j];
            from = to;
        }
    }
}

// Exclusive scan of v[0..n) in place into v[0..n] (v[n] = total), n <= 4 * kBlock,
// by one workgroup: 4 entries per thread, a wave scan of the thread totals, then
// the wave totals.  All threads must call it.
__device__ __forceinline__ void block_exclusive_scan(int* v, int n) {
    __shared__ int wsum[kBlock / kWave];
    const int t = threadIdx.x, base = 4 * t;
    int loc[4], tot = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        loc[q] = base + q < n ? v[base + q] : 0;
        tot += loc[q];
    }
    // inclusive scan of the thread totals within the wave
    const int lane = t % kWave, w = t / kWave;
    int inc = tot;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int o = __shfl_up(inc, off, kWave);
        if (lane >= off) inc += o;
    }
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    int wpre = 0;
    for (int q = 0; q < w; ++q) wpre += wsum[q];
    int run = wpre + inc - tot;  // exclusive prefix of this thread
    __syncthreads();             // everyone has read v before it is overwritten
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (base + q < n) v[base + q] = run;
        run += loc[q];
    }
    if (t == kBlock - 1) v[n] = wpre + inc;  // last thread holds the block total
    __syncthreads();
}

// Streaming form of the sliced product: the workgroup's rows' runs in its column
// range are laid end to end (exclusive scan of the run lengths in LDS); every lane
// takes ELEMS positions of that sequence (consecutive lanes, consecutive positions,
// so loads of a run coalesce), finds its row by binary search in LDS, gathers the
// X row and leaves the K products in LDS; teams of TEAM lanes then add each row's
// products (rows spanning chunks accumulate in LDS).
// SPLIT: the re-laid layout of the plan -- every slice's runs stored as a CSR of
// their own, slice after slice (srp: the 8m + 1 row pointers of all slices, ci and a
// the re-laid column indices and values), so a workgroup's runs are one contiguous
// range and no XCD touches another slice's cache lines of A.  S is 8 then.
template <typename T, typename RP, int ELEMS, int TEAM, int K, bool SPLIT>
__global__ __launch_bounds__(kBlock) void spmm_slice_stream_kernel(int m, int S, const RP* __restrict__ rp,
                                                                   const int32_t* __restrict__ cuts,
                                                                   const int32_t* __restrict__ srp,
                                                                   const int32_t* __restrict__ ci,
                                                                   const T* __restrict__ a, const T* __restrict__ X,
                                                                   int64_t ldX, T* __restrict__ part,
                                                                   int rows_per_block) {
    constexpr int CAP = kBlock * ELEMS;
    constexpr int RMAX = 1024 / K;
    constexpr int TEAMS = kBlock / TEAM;
    __shared__ T prod[CAP * K];
    __shared__ int soff[RMAX + 1];
    __shared__ RP slo[RMAX];
    __shared__ T racc[RMAX * K];
    const int b = blockIdx.x, xcd = b % kXcds;
    const int s = xcd % S, per = kXcds / S, step = kXcds / S;
    const int rb = (b / kXcds) * per + xcd / S;
    const int r0 = rb * rows_per_block, r1 = min(m, r0 + rows_per_block);
    const int lt = threadIdx.x % TEAM, team = threadIdx.x / TEAM;
    for (int rs = r0; rs < r1; rs += RMAX) {
        const int nr = min(RMAX, r1 - rs);
        for (int i = threadIdx.x; i < nr; i += kBlock) {
            const int row = rs + i;
            RP lb, le;
            if constexpr (SPLIT) {
                lb = srp[(int64_t)s * m + row];
                le = srp[(int64_t)s * m + row + 1];
            } else {
                const RP base = rp[row];
                lb = s == 0 ? base : base + cuts[(int64_t)(s * step - 1) * m + row];
                le = s == S - 1 ? rp[row + 1] : base + cuts[(int64_t)((s + 1) * step - 1) * m + row];
            }
            slo[i] = lb;
            soff[i] = (int)(le - lb);
        }
        for (int i = threadIdx.x; i < nr * K; i += kBlock) racc[i] = T(0);
        __syncthreads();
        block_exclusive_scan(soff, nr);
        const int total = soff[nr];
        for (int cs = 0; cs < total; cs += CAP) {
            const int cnt = min(CAP, total - cs);
            int32_t c[ELEMS];
            T av[ELEMS];
#pragma unroll
            for (int j = 0; j < ELEMS; ++j) {
                const int e = threadIdx.x + j * kBlock;
                c[j] = 0;
                av[j] = T(0);
                if (e < cnt) {
                    const int v = cs + e;
                    int lo = 0, hi = nr - 1;  // last row with soff <= v
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (soff[mid] <= v)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    const RP l = slo[lo] + (v - soff[lo]);
                    c[j] = __builtin_nontemporal_load(ci + l);
                    av[j] = a ? __builtin_nontemporal_load(a + l) : T(1);
                }
            }
            T xv[ELEMS][K];  // gathered unconditionally (c = 0 past the chunk) so they overlap
#pragma unroll
            for (int j = 0; j < ELEMS; ++j)
#pragma unroll
                for (int q = 0; q < K; ++q) xv[j][q] = X[(int64_t)c[j] * ldX + q];
#pragma unroll
            for (int j = 0; j < ELEMS; ++j) {
                const int e = threadIdx.x + j * kBlock;
                if (e < cnt)
#pragma unroll
                    for (int q = 0; q < K; ++q) prod[e * K + q] = av[j] * xv[j][q];
            }
            __syncthreads();
            for (int i = team; i < nr; i += TEAMS) {
                const int lb = max(soff[i], cs), le = min(soff[i + 1], cs + cnt);
                if (le <= lb) continue;  // team-uniform
                T sum[K];
#pragma unroll
                for (int q = 0; q < K; ++q) sum[q] = T(0);
                for (int l = lb + lt; l < le; l += TEAM)
#pragma unroll
                    for (int q = 0; q < K; ++q) sum[q] += prod[(l - cs) * K + q];
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    sum[q] = group_sum<TEAM>(sum[q]);
                    if (lt == 0) racc[i * K + q] += sum[q];
                }
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < nr * K; i += kBlock) part[((int64_t)s * m + rs) * K + i] = racc[i];
        __syncthreads();
    }
}

// Y[i, :] = alpha * sum_s part_s[i, :] + beta * Y[i, :] (slices added in order)
template <typename T, typename RP, int K>
__global__ __launch_bounds__(kBlock) void slice_reduce_kernel(int m, int S, const RP* __restrict__ rp,
                                                              const T* __restrict__ part, T alpha, T* __restrict__ Y,
                                                              int64_t ldY, T beta, int alpha_one, int beta_mode) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        T z[K];
#pragma unroll
        for (int q = 0; q < K; ++q) z[q] = part[(int64_t)i * K + q];
        for (int s = 1; s < S; ++s)
#pragma unroll
            for (int q = 0; q < K; ++q) z[q] += part[((int64_t)s * m + i) * K + q];
        const bool empty = rp[i + 1] == rp[i];
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const T zq[1] = {z[q]};
            epilogue<T, 1>(Y + (int64_t)i * ldY + q, zq, empty, alpha, alpha_one, beta, beta_mode);
        }
    }
}

// Slices for an X of `xbytes`: 0 = not worth it (X fits an L2 already, or is so
// large that the partial blocks cost more than the misses they save).
// Slices for an X block of xbytes: none up to 6 MB (X mostly stays in the L2s
// anyway) or above 32 MB (not measured to pay); otherwise 2 or 4 slices of at
// most 4 MB each where possible.  Measured on a random 1e6-row CSR with 32
// nonzeros per row: 2 slices for the 8 MB FP64 SpMV, 4 for the 16 MB FP32 and
// 32 MB FP64 nrhs = 4 blocks were the fastest counts; 8 slices were slower.
// The re-laid layout (8 slices, one per XCD): X blocks of more than 2 MB and at
// most 64 MB, i.e. slices of 0.25-8 MB.
static bool split_pays(int64_t xbytes) {
    constexpr int64_t kMB = 1ll << 20;
    return xbytes > 2 * kMB && xbytes <= 64 * kMB;
}

static int slice_count(int64_t xbytes) {
    constexpr int64_t kMB = 1ll << 20;
    constexpr int64_t kSliceBytes = 4 * kMB;
    constexpr int kMaxSlices = 4;
    if (xbytes <= 6 * kMB || xbytes > 32 * kMB) return 0;
    int S = 2;
    while (S < kMaxSlices && xbytes / S > kSliceBytes) S *= 2;
    return S;
}

template <typename T, typename RP, int K, bool SPLIT>
static void launch_slice_stream(int S, int team, int grid, int m, const RP* rp, const int32_t* cuts,
                                const int32_t* srp, const int32_t* ci, const T* a, const T* X, int64_t ldX, T* part,
                                int rpb, hipStream_t st) {
    // ELEMS keeps the product stage at 2048 entries (16 KB in FP64)
    constexpr int E = K == 1 ? 8 : K == 2 ? 4 : 2;
#define TG_SLS(TM)                                                                                                \
    case TM:                                                                                                      \
        hipLaunchKernelGGL((spmm_slice_stream_kernel<T, RP, E, TM, K, SPLIT>), dim3(grid), dim3(kBlock), 0, st, m, \
                           S, rp, cuts, srp, ci, a, X, ldX, part, rpb);                                           \
        break;
    switch (team) {
        TG_SLS(2)
        TG_SLS(4)
        TG_SLS(8)
        default:
            TG_SLS(16)
    }
#undef TG_SLS
}

template <typename T, typename RP>
static int spmm_sliced_typed(int S, int32_t m, int64_t nnz, const RP* rp, const int32_t* cuts, const int32_t* srp,
                             const int32_t* ci, const T* a, double alpha_d, const T* X, int64_t ldX, double beta_d,
                             T* Y, int64_t ldY, int32_t K, T* part, hipStream_t st) {
    const double avg = m > 0 ? (double)nnz / m / S : 0.0;  // mean run length per row and slice
    const int team = std::min(16, std::max(2, pow2ceil((int)std::ceil(avg))));
    const int cap = kBlock * (K == 1 ? 8 : K == 2 ? 4 : 2);
    const int rmax = 1024 / (K == 3 ? 4 : K);
    const int rpb = std::max(1, std::min(rmax, (int)(cap / (avg + 1.0))));
    const int row_blocks = (m + rpb - 1) / rpb;
    const int per = kXcds / S;
    const int grid = kXcds * ((row_blocks + per - 1) / per);
#define TG_SLK(KK)                                                                                              \
    if (srp)                                                                                                    \
        launch_slice_stream<T, RP, KK, true>(S, team, grid, m, rp, cuts, srp, ci, a, X, ldX, part, rpb, st);    \
    else                                                                                                        \
        launch_slice_stream<T, RP, KK, false>(S, team, grid, m, rp, cuts, srp, ci, a, X, ldX, part, rpb, st);
    switch (K) {
        case 1: TG_SLK(1) break;
        case 2: TG_SLK(2) break;
        case 3: TG_SLK(3) break;
        default: TG_SLK(4) break;
    }
#undef TG_SLK
    const int alpha_one = (alpha_d == 1.0);
    const int beta_mode = (beta_d == 0.0) ? kBetaZero : (beta_d == 1.0) ? kBetaOne : kBetaGeneric;
    const int rgrid = (int)std::min<int64_t>((m + kBlock - 1) / kBlock, 65536);
#define TG_SRED(KK)                                                                                       \
    case KK:                                                                                              \
        hipLaunchKernelGGL((slice_reduce_kernel<T, RP, KK>), dim3(rgrid), dim3(kBlock), 0, st, m, S, rp, \
                           (const T*)part, (T)alpha_d, Y, ldY, (T)beta_d, alpha_one, beta_mode);          \
        break;
    switch (K) {
        TG_SRED(1)
        TG_SRED(2)
        TG_SRED(3)
        TG_SRED(4)
    }
#undef TG_SRED
    TG_LAUNCH_CHECK();
    return TG_OK;
}

}  // namespace tg

using namespace tg;

static bool pow2_in(int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; }

extern "C" int tg_tune_set(const char* key, int32_t value) {
    if (!key) return fail(TG_ERR_INVALID, "tg_tune_set: null key");
    const std::string k(key);
    if (k == "spmv_kernel") {
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: spmv_kernel must be 0 or 1");
        g_spmv_variant = value;
        return TG_OK;
    }
    if (k == "spmv_team") {
        if (value != 0 && !pow2_in(value, 1, 64))
            return fail(TG_ERR_INVALID, "tg_tune_set: spmv_team must be 0 (auto) or a power of two in [1, 64]");
        g_spmv_team = value;
        return TG_OK;
    }
    if (k == "spmm_group") {
        if (value != 0 && !pow2_in(value, 1, 64))
            return fail(TG_ERR_INVALID, "tg_tune_set: spmm_group must be 0 (auto) or a power of two in [1, 64]");
        g_spmm_group = value;
        return TG_OK;
    }
    if (k == "spmm_rows") {
        if (value < 0 || value > 2) return fail(TG_ERR_INVALID, "tg_tune_set: spmm_rows must be 0 (auto), 1 or 2");
        g_spmm_rows = value;
        return TG_OK;
    }
    if (k == "spmm_long_row") {
        if (value < 0) return fail(TG_ERR_INVALID, "tg_tune_set: spmm_long_row must be >= 0");
        g_spmm_long_row = value;
        return TG_OK;
    }
    if (k == "spmm_split") {
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: spmm_split must be 0 or 1");
        g_spmm_split = value;
        return TG_OK;
    }
    if (k == "spmm_slices") {
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
            return fail(TG_ERR_INVALID, "tg_tune_set: spmm_slices must be 0 (auto), 1 (off), 2, 4 or 8");
        g_spmm_slices = value;
        return TG_OK;
    }
    bool known = false;
    int rc = batch_tune_set(key, value, &known);
    if (known) return rc;
    rc = cg_tune_set(key, value, &known);
    if (known) return rc;
    rc = bsr_tune_set(key, value, &known);
    if (known) return rc;
    rc = gmres_tune_set(key, value, &known);
    if (known) return rc;
    rc = gbatch_tune_set(key, value, &known);
    if (known) return rc;
    rc = serve_tune_set(key, value, &known);
    if (known) return rc;
    return fail(TG_ERR_INVALID, "tg_tune_set: unknown key '%s'", key);
}

extern "C" int32_t tg_tune_get(const char* key) {
    if (!key) return -1;
    const std::string k(key);
    if (k == "spmv_kernel") return spmv_variant();
    if (k == "spmv_team") return g_spmv_team;
    if (k == "spmm_group") return g_spmm_group;
    if (k == "spmm_slices") return g_spmm_slices;
    if (k == "spmm_split") return g_spmm_split;
    if (k == "spmm_long_row") return g_spmm_long_row;
    if (k == "spmm_rows") return g_spmm_rows;
    // the other modules' keys: each getter returns -1 for a key it does not know (no
    // knob takes a negative value)
    for (auto get : {cg_tune_get, bsr_tune_get, gmres_tune_get, gbatch_tune_get, serve_tune_get, batch_tune_get}) {
        const int32_t v = get(key);
        if (v != -1) return v;
    }
    return -1;
}

extern "C" int32_t tg_csr_default_blocks(int32_t m, int64_t nnz, int32_t nrhs) {
    if (m <= 0) return 0;
    const int rpb = rows_per_block_for(m, nnz, nrhs < 1 ? 1 : nrhs);
    return (m + rpb - 1) / rpb;
}

extern "C" int tg_csr_job_pointers(int32_t m, const void* rp, tg_dtype rp_type, int32_t blocks, int32_t* job_ptr,
                                   void* stream) {
    if (m < 0 || blocks < 1 || !rp || !job_ptr) return fail(TG_ERR_INVALID, "tg_csr_job_pointers: bad argument");
    const int threads = 256, grid = (blocks + 1 + threads - 1) / threads;
    hipStream_t st = as_stream(stream);
    if (rp_type == TG_I32)
        hipLaunchKernelGGL((job_pointers_kernel<int32_t>), dim3(grid), dim3(threads), 0, st, m,
                           (const int32_t*)rp, blocks, job_ptr);
    else if (rp_type == TG_I64)
        hipLaunchKernelGGL((job_pointers_kernel<int64_t>), dim3(grid), dim3(threads), 0, st, m,
                           (const int64_t*)rp, blocks, job_ptr);
    else
        return fail(TG_ERR_DTYPE, "tg_csr_job_pointers: rp_type must be TG_I32 or TG_I64");
    TG_LAUNCH_CHECK();
    return TG_OK;
}

extern "C" int tg_axpby(int32_t m, int32_t nrhs, tg_dtype val_type, double alpha, const void* X, int64_t ldX,
                        double beta, void* Y, int64_t ldY, void* stream) {
    if (m < 0 || nrhs < 0 || !Y || ldY < nrhs || (alpha != 0.0 && (!X || ldX < nrhs)))
        return fail(TG_ERR_INVALID, "tg_axpby: bad argument");
    hipStream_t st = as_stream(stream);
    if (val_type == TG_F64) return axpby_typed<double>(m, nrhs, alpha, (const double*)X, ldX, beta, (double*)Y, ldY, st);
    if (val_type == TG_F32) return axpby_typed<float>(m, nrhs, alpha, (const float*)X, ldX, beta, (float*)Y, ldY, st);
    return fail(TG_ERR_DTYPE, "tg_axpby: val_type must be TG_F64 or TG_F32");
}

namespace tg {

// Random gathers of X rows only stay cheap while X fits the 256 MB Infinity
// Cache; wider blocks are processed in column tiles of at most this many
// bytes of X (A is re-streamed once per tile).
constexpr int64_t kXTileBytes = 128ll << 20;

// Per-block maxima of |ci - row| (no atomics; the host takes the max of the blocks).
template <typename RP>
__global__ __launch_bounds__(kBlock) void band_kernel(int m, const RP* __restrict__ rp, const int32_t* __restrict__ ci,
                                                      int32_t* __restrict__ block_max) {
    __shared__ int32_t sh[kBlock];
    int32_t b = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const RP lb = rp[i], le = rp[i + 1];
        if (le > lb) {  // columns need not be sorted: look at all of them
            for (RP l = lb; l < le; ++l) b = max(b, abs(ci[l] - i));
        }
    }
    sh[threadIdx.x] = b;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = sh[0];
}

static int band_grid(const CsrView& A) { return (int)std::min<int64_t>((A.m + kBlock - 1) / kBlock, kBandInts); }

// The band pass with a caller's device buffer of kBandInts int32 (no allocation).
int64_t csr_band_into(const CsrView& A, int32_t* d, hipStream_t s) {
    if (A.m <= 0 || A.nnz <= 0 || !A.ci) return 0;
    const int grid = band_grid(A);
    if (A.rp_type == TG_I32)
        hipLaunchKernelGGL((band_kernel<int32_t>), dim3(grid), dim3(kBlock), 0, s, A.m, (const int32_t*)A.rp, A.ci, d);
    else
        hipLaunchKernelGGL((band_kernel<int64_t>), dim3(grid), dim3(kBlock), 0, s, A.m, (const int64_t*)A.rp, A.ci, d);
    std::vector<int32_t> h(grid, 0);
    const bool ok = hipMemcpyAsync(h.data(), d, grid * sizeof(int32_t), hipMemcpyDeviceToHost, s) == hipSuccess;
    if (!ok || hipStreamSynchronize(s) != hipSuccess) return -1;
    int64_t b = 0;
    for (int32_t v : h) b = std::max<int64_t>(b, v);
    return b;
}

int64_t csr_band(const CsrView& A, hipStream_t s) {
    if (A.m <= 0 || A.nnz <= 0 || !A.ci) return 0;
    int32_t* d = nullptr;
    if (hipMallocAsync((void**)&d, band_grid(A) * sizeof(int32_t), s) != hipSuccess) return -1;
    const int64_t b = csr_band_into(A, d, s);  // synchronises s
    (void)hipFreeAsync(d, s);
    return b;
}

// ---------------------------------------------------------------------------
// Long rows (plan path): a row of hundreds of thousands of nonzeros would keep one
// wave or one workgroup busy long after the rest of the matrix is done.  The row
// kernels see such rows as empty (row ends from LongRows::rpe) and apply the
// empty-row rule; these two kernels then add the rows' products.
// ---------------------------------------------------------------------------

// One workgroup per segment of a long row: part[seg * k + c] = sum over the segment's
// nonzeros of a_l X[ci_l, c].  Columns in tiles of CT lanes; the segment's nonzeros are
// dealt over the kBlock / CT row lanes and summed by a fixed tree: deterministic.
template <typename T>
__global__ __launch_bounds__(kBlock) void lr_segment_kernel(const int64_t* __restrict__ seg_lb,
                                                            const int64_t* __restrict__ seg_le,
                                                            const int32_t* __restrict__ ci, const T* __restrict__ a,
                                                            const T* __restrict__ X, int64_t ldX, int k, int CT,
                                                            T* __restrict__ part) {
    __shared__ T sh[kBlock];
    const int sg = blockIdx.x;
    const int64_t lb = seg_lb[sg], le = seg_le[sg];
    const int lc = threadIdx.x % CT, r = threadIdx.x / CT, R = kBlock / CT;
    constexpr int U = 8;  // nonzeros per thread in flight: index loads first, then the gathers
    for (int c0 = 0; c0 < k; c0 += CT) {
        const int c = c0 + lc;
        const int cs = c < k ? c : 0;
        T acc = T(0);
        for (int64_t l0 = lb + r; l0 < le; l0 += (int64_t)R * U) {
            int32_t col[U];
            T av[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t l = l0 + (int64_t)u * R;
                const int64_t ls = l < le ? l : lb;  // a valid entry; masked below
                col[u] = ci[ls];
                av[u] = (l < le && c < k) ? (a ? a[ls] : T(1)) : T(0);
            }
            T xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = X[(int64_t)col[u] * ldX + cs];
#pragma unroll
            for (int u = 0; u < U; ++u) acc += av[u] * xv[u];  // nonzero order, as one at a time
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int h = R / 2; h > 0; h >>= 1) {
            if (r < h) sh[threadIdx.x] += sh[threadIdx.x + h * CT];
            __syncthreads();
        }
        if (r == 0 && c < k) part[(int64_t)sg * k + c] = sh[threadIdx.x];
        __syncthreads();
    }
}

// One workgroup per long row: z = the row's segment sums in segment order, then
// y = alpha z + y over the value the row kernels left there for an empty row (0 for
// beta = 0, y for beta = 1, beta y otherwise).
template <typename T>
__global__ __launch_bounds__(kBlock) void lr_combine_kernel(const int32_t* __restrict__ rows,
                                                            const int32_t* __restrict__ seg_ptr,
                                                            const T* __restrict__ part, int k, T alpha, int alpha_one,
                                                            int beta_mode, T* __restrict__ Y, int64_t ldY) {
#pragma clang fp contract(off)
    const int i = blockIdx.x;
    const int64_t row = rows[i];
    const int s0 = seg_ptr[i], s1 = seg_ptr[i + 1];
    for (int c = threadIdx.x; c < k; c += kBlock) {
        T z = T(0);
        for (int sg = s0; sg < s1; ++sg) z += part[(int64_t)sg * k + c];
        const T az = alpha_one ? z : alpha * z;
        T* y = Y + row * ldY + c;
        *y = beta_mode == kBetaZero ? az : az + *y;
    }
}

template <typename T>
static int long_rows_apply(const CsrView& A, double alpha_d, const T* X, int64_t ldX, double beta_d, T* Y,
                           int64_t ldY, int32_t nrhs, hipStream_t st) {
    const LongRows& L = *A.longr;
    if (L.nrows == 0) return TG_OK;
    int CT = 1;
    while (CT < nrhs && CT < 64) CT *= 2;
    const int beta_mode = (beta_d == 0.0) ? kBetaZero : (beta_d == 1.0) ? kBetaOne : kBetaGeneric;
    hipLaunchKernelGGL((lr_segment_kernel<T>), dim3(L.nseg), dim3(kBlock), 0, st, L.seg_lb, L.seg_le, A.ci,
                       (const T*)A.val, X, ldX, nrhs, CT, (T*)L.part);
    hipLaunchKernelGGL((lr_combine_kernel<T>), dim3(L.nrows), dim3(kBlock), 0, st, L.rows, L.seg_ptr,
                       (const T*)L.part, nrhs, (T)alpha_d, alpha_d == 1.0 ? 1 : 0, beta_mode, Y, ldY);
    TG_LAUNCH_CHECK();
    return TG_OK;
}

bool spmm_may_slab(const CsrView& A, int32_t nrhs) {
    const int64_t vs = A.val_type == TG_F64 ? 8 : 4;
    return (int64_t)A.n * nrhs * vs > kXTileBytes && nrhs > 8;
}

// Column slabs keep a random-column X resident in the 256 MB Infinity Cache; a
// banded matrix (known band) only touches a window of X around the current
// rows, so there they would just re-read A once per slab.
bool spmm_uses_slabs(const CsrView& A, int32_t nrhs) {
    const int64_t vs = A.val_type == TG_F64 ? 8 : 4;
    const bool banded = A.band >= 0 && (2 * A.band + 1) * (int64_t)nrhs * vs <= kXTileBytes / 2;
    return spmm_may_slab(A, nrhs) && !banded;
}

int spmm_dispatch(const CsrView& A, double alpha, const void* X, int64_t ldX, double beta, void* Y, int64_t ldY,
                  int32_t nrhs, const int32_t* job_ptr, int32_t blocks, const int* skip, hipStream_t st) {
    const int32_t m = A.m, n = A.n;
    if (m < 0 || n < 0 || A.nnz < 0 || nrhs < 0) return fail(TG_ERR_INVALID, "spmm: negative size");
    if (A.val_type != TG_F64 && A.val_type != TG_F32) return fail(TG_ERR_DTYPE, "spmm: val_type must be F64/F32");
    if (A.rp_type != TG_I32 && A.rp_type != TG_I64) return fail(TG_ERR_DTYPE, "spmm: rp_type must be I32/I64");
    if (m == 0 || nrhs == 0) return TG_OK;
    if (!Y || ldY < nrhs) return fail(TG_ERR_INVALID, "spmm: Y is null or ldY < nrhs");
    if (job_ptr && blocks < 1) return fail(TG_ERR_INVALID, "spmm: job_ptr given with blocks < 1");
    const bool f64 = A.val_type == TG_F64;
    if (alpha == 0.0) {  // SpMM.hpp:18-25, SpMV.hpp:14-29
        if (f64) return axpby_typed<double>(m, nrhs, 0.0, nullptr, 0, beta, (double*)Y, ldY, st);
        return axpby_typed<float>(m, nrhs, 0.0, nullptr, 0, beta, (float*)Y, ldY, st);
    }
    if (X == Y) return fail(TG_ERR_ALIAS, "spmm: input and output pointer coincide. Doing nothing.");
    if (!A.rp || !X || ldX < nrhs) return fail(TG_ERR_INVALID, "spmm: rp/X null or ldX < nrhs");
    if (A.nnz > 0 && !A.ci) return fail(TG_ERR_INVALID, "spmm: ci is null");
    const size_t vs = f64 ? 8 : 4;
    const bool long_rows = A.longr && A.longr->nrows > 0 && !job_ptr && !skip;
    const void* rpe = long_rows ? A.longr->rpe : nullptr;
    int32_t kt = nrhs;
    if (spmm_uses_slabs(A, nrhs)) {
        kt = 8;
        while ((int64_t)n * (2 * kt) * (int64_t)vs <= kXTileBytes && 2 * kt < nrhs) kt *= 2;
    }
    for (int32_t c0 = 0; c0 < nrhs; c0 += kt) {
        const int32_t kc = std::min(kt, nrhs - c0);
        int rc;
        if (f64) {
            const double* Xc = (const double*)X + c0;
            double* Yc = (double*)Y + c0;
            if (A.rp_type == TG_I32)
                rc = spmm_typed<double, int32_t>(m, A.nnz, (const int32_t*)A.rp, A.ci, (const double*)A.val, alpha, Xc,
                                                 ldX, beta, Yc, ldY, kc, job_ptr, blocks, skip, st,
                                                 (const int32_t*)rpe);
            else
                rc = spmm_typed<double, int64_t>(m, A.nnz, (const int64_t*)A.rp, A.ci, (const double*)A.val, alpha, Xc,
                                                 ldX, beta, Yc, ldY, kc, job_ptr, blocks, skip, st,
                                                 (const int64_t*)rpe);
        } else {
            const float* Xc = (const float*)X + c0;
            float* Yc = (float*)Y + c0;
            if (A.rp_type == TG_I32)
                rc = spmm_typed<float, int32_t>(m, A.nnz, (const int32_t*)A.rp, A.ci, (const float*)A.val, alpha, Xc,
                                                ldX, beta, Yc, ldY, kc, job_ptr, blocks, skip, st,
                                                (const int32_t*)rpe);
            else
                rc = spmm_typed<float, int64_t>(m, A.nnz, (const int64_t*)A.rp, A.ci, (const float*)A.val, alpha, Xc,
                                                ldX, beta, Yc, ldY, kc, job_ptr, blocks, skip, st,
                                                (const int64_t*)rpe);
        }
        if (rc) return rc;
    }
    if (long_rows) {
        if (f64) return long_rows_apply<double>(A, alpha, (const double*)X, ldX, beta, (double*)Y, ldY, nrhs, st);
        return long_rows_apply<float>(A, alpha, (const float*)X, ldX, beta, (float*)Y, ldY, nrhs, st);
    }
    return TG_OK;
}

int op_apply(const OpRef& op, const void* x, void* y, int k, const int* skip, hipStream_t s) {
    if (op.csr) return spmm_dispatch(*op.csr, 1.0, x, k, 0.0, y, k, k, nullptr, 0, skip, s);
    const int rc = op.fn(op.ctx, x, y, k, (void*)s);
    if (rc) return fail(TG_ERR_INVALID, "operator callback returned %d", rc);
    return TG_OK;
}

}  // namespace tg

struct tg_csr_plan_s {
    int32_t m = 0, n = 0;
    int64_t nnz = 0;
    const void* rp = nullptr;
    tg_dtype rp_type = TG_I32;
    const int32_t* ci = nullptr;
    int32_t* cuts = nullptr;  // kCuts x m row offsets of the 8-way column cut
    int sorted = 0;
    int64_t band = -1;        // max |column - row|
    // Re-laid layout (tg_csr_plan_set_values): the 8 slices' runs as CSRs of their own,
    // slice after slice -- row pointers of all slices (8m + 1), column indices, and the
    // values of one registered pointer per value type (0 FP64, 1 FP32).
    int32_t* srp = nullptr;
    int32_t* sci = nullptr;
    void* sval[2] = {nullptr, nullptr};
    const void* sval_src[2] = {nullptr, nullptr};
    void* part = nullptr;     // partial blocks of the sliced path (grown on demand)
    size_t part_bytes = 0;
    // long rows (more than spmm_long_row nonzeros when the plan was made): row ends
    // with those rows empty, the rows (ascending), their segments' first index and
    // the segments' nonzero ranges (lb then le); the segment sums use `part`
    void* rpe = nullptr;
    int32_t* lrows = nullptr;
    int32_t* lseg_ptr = nullptr;
    int64_t* lseg = nullptr;
    int32_t nlong = 0, nseg = 0;
    // The partial blocks are shared by every sliced call on the plan: a call on another
    // stream than the previous one first waits for it (event `used`), and host threads
    // take turns.
    hipEvent_t used = nullptr;
    hipStream_t used_on = nullptr;
    std::mutex mu;
};

// Row ends with the long rows (more than `thr` nonzeros) empty, and the long rows'
// ids appended to `list` (in no particular order; the host sorts them).
template <typename RP>
__global__ __launch_bounds__(kBlock) void long_rows_kernel(int m, const RP* __restrict__ rp, int64_t thr,
                                                           RP* __restrict__ rpe, int32_t* __restrict__ list,
                                                           int* __restrict__ count, int cap) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const RP lb = rp[i], le = rp[i + 1];
        const bool lng = (int64_t)(le - lb) > thr;
        rpe[i] = lng ? lb : le;
        if (lng) {
            const int p = atomicAdd(count, 1);
            if (p < cap) list[p] = i;
        }
    }
}

extern "C" int tg_csr_plan_destroy(tg_csr_plan_t plan);

// Finds the plan's long rows and lays out their segments (once, at plan creation).
static int plan_long_rows(tg_csr_plan_t P, hipStream_t st) {
    const int64_t thr = g_spmm_long_row;
    if (thr <= 0 || P->m == 0 || P->nnz <= thr) return TG_OK;
    const size_t rs = P->rp_type == TG_I32 ? 4 : 8;
    const int cap = (int)std::min<int64_t>(P->nnz / (thr + 1) + 1, P->m);
    int32_t* list = nullptr;
    int* count = nullptr;
    TG_HIP(hipMallocAsync(&P->rpe, (size_t)P->m * rs, st));
    TG_HIP(hipMallocAsync((void**)&list, (size_t)cap * sizeof(int32_t), st));
    TG_HIP(hipMallocAsync((void**)&count, sizeof(int), st));
    TG_HIP(hipMemsetAsync(count, 0, sizeof(int), st));
    const int grid = (int)std::min<int64_t>((P->m + kBlock - 1) / kBlock, 65536);
    if (P->rp_type == TG_I32)
        hipLaunchKernelGGL((long_rows_kernel<int32_t>), dim3(grid), dim3(kBlock), 0, st, P->m, (const int32_t*)P->rp,
                           thr, (int32_t*)P->rpe, list, count, cap);
    else
        hipLaunchKernelGGL((long_rows_kernel<int64_t>), dim3(grid), dim3(kBlock), 0, st, P->m, (const int64_t*)P->rp,
                           thr, (int64_t*)P->rpe, list, count, cap);
    TG_LAUNCH_CHECK();
    int nl = 0;
    TG_HIP(hipMemcpyAsync(&nl, count, sizeof(int), hipMemcpyDeviceToHost, st));
    TG_HIP(hipStreamSynchronize(st));
    nl = std::min(nl, cap);
    std::vector<int32_t> rows(nl);
    if (nl > 0) TG_HIP(hipMemcpyAsync(rows.data(), list, (size_t)nl * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    TG_HIP(hipFreeAsync(list, st));
    TG_HIP(hipFreeAsync(count, st));
    if (nl == 0) {
        TG_HIP(hipFreeAsync(P->rpe, st));
        P->rpe = nullptr;
        TG_HIP(hipStreamSynchronize(st));
        return TG_OK;
    }
    std::vector<int64_t> rph((size_t)P->m + 1);
    if (P->rp_type == TG_I32) {
        std::vector<int32_t> r32((size_t)P->m + 1);
        TG_HIP(hipMemcpyAsync(r32.data(), P->rp, r32.size() * 4, hipMemcpyDeviceToHost, st));
        TG_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < r32.size(); ++i) rph[i] = r32[i];
    } else {
        TG_HIP(hipMemcpyAsync(rph.data(), P->rp, rph.size() * 8, hipMemcpyDeviceToHost, st));
        TG_HIP(hipStreamSynchronize(st));
    }
    std::sort(rows.begin(), rows.end());
    std::vector<int32_t> sptr(1, 0);
    std::vector<int64_t> lb, le;
    for (int32_t r : rows) {
        for (int64_t l = rph[r]; l < rph[r + 1]; l += kLongSeg) {
            lb.push_back(l);
            le.push_back(std::min<int64_t>(rph[r + 1], l + kLongSeg));
        }
        sptr.push_back((int32_t)lb.size());
    }
    const size_t ns = lb.size();
    TG_HIP(hipMallocAsync((void**)&P->lrows, rows.size() * sizeof(int32_t), st));
    TG_HIP(hipMallocAsync((void**)&P->lseg_ptr, sptr.size() * sizeof(int32_t), st));
    TG_HIP(hipMallocAsync((void**)&P->lseg, 2 * ns * sizeof(int64_t), st));
    TG_HIP(hipMemcpyAsync(P->lrows, rows.data(), rows.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    TG_HIP(hipMemcpyAsync(P->lseg_ptr, sptr.data(), sptr.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    TG_HIP(hipMemcpyAsync(P->lseg, lb.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
    TG_HIP(hipMemcpyAsync(P->lseg + ns, le.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
    TG_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    P->nlong = (int32_t)rows.size();
    P->nseg = (int32_t)ns;
    return TG_OK;
}

extern "C" int tg_csr_plan_create(int32_t m, int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type,
                                  const int32_t* ci, tg_csr_plan_t* out, void* stream) {
    if (!out || m < 0 || n < 0 || nnz < 0 || !rp || (nnz > 0 && !ci))
        return fail(TG_ERR_INVALID, "tg_csr_plan_create: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_csr_plan_create: rp_type must be I32/I64");
    hipStream_t st = as_stream(stream);
    auto* P = new tg_csr_plan_s();
    P->m = m;
    P->n = n;
    P->nnz = nnz;
    P->rp = rp;
    P->rp_type = rp_type;
    P->ci = ci;
    int* flag = nullptr;
    if (m > 0 && n > 0) {
        if (hipMallocAsync((void**)&P->cuts, (size_t)kCuts * m * sizeof(int32_t), st) != hipSuccess ||
            hipMallocAsync((void**)&flag, sizeof(int), st) != hipSuccess) {
            delete P;
            return fail(TG_ERR_HIP, "tg_csr_plan_create: allocation failed");
        }
        int zero = 0;
        (void)hipMemcpyAsync(flag, &zero, sizeof(int), hipMemcpyHostToDevice, st);
        const int grid = (int)std::min<int64_t>((m + kBlock - 1) / kBlock, 65536);
        if (rp_type == TG_I32)
            hipLaunchKernelGGL((slice_plan_kernel<int32_t>), dim3(grid), dim3(kBlock), 0, st, m, n, (const int32_t*)rp,
                               ci, P->cuts, flag);
        else
            hipLaunchKernelGGL((slice_plan_kernel<int64_t>), dim3(grid), dim3(kBlock), 0, st, m, n, (const int64_t*)rp,
                               ci, P->cuts, flag);
        int unsorted = 1;
        (void)hipMemcpyAsync(&unsorted, flag, sizeof(int), hipMemcpyDeviceToHost, st);
        (void)hipFreeAsync(flag, st);
        if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) {
            (void)hipFree(P->cuts);
            delete P;
            return fail(TG_ERR_HIP, "tg_csr_plan_create: plan kernel failed");
        }
        P->sorted = unsorted ? 0 : 1;
        CsrView V;
        V.m = m;
        V.n = n;
        V.nnz = nnz;
        V.rp = rp;
        V.rp_type = rp_type;
        V.ci = ci;
        P->band = csr_band(V, st);
        const int rc = plan_long_rows(P, st);
        if (rc) {
            (void)tg_csr_plan_destroy(P);
            return rc;
        }
    }
    *out = P;
    return TG_OK;
}

extern "C" int tg_csr_plan_destroy(tg_csr_plan_t plan) {
    if (!plan) return TG_OK;
    if (plan->srp) (void)hipFree(plan->srp);
    if (plan->sci) (void)hipFree(plan->sci);
    for (void* v : plan->sval)
        if (v) (void)hipFree(v);
    if (plan->cuts) (void)hipFree(plan->cuts);
    if (plan->part) (void)hipFree(plan->part);
    if (plan->rpe) (void)hipFree(plan->rpe);
    if (plan->lrows) (void)hipFree(plan->lrows);
    if (plan->lseg_ptr) (void)hipFree(plan->lseg_ptr);
    if (plan->lseg) (void)hipFree(plan->lseg);
    if (plan->used) (void)hipEventDestroy(plan->used);
    delete plan;
    return TG_OK;
}

extern "C" int tg_csr_plan_set_values(tg_csr_plan_t plan, const void* a, tg_dtype val_type, void* stream) {
    if (!plan || (val_type != TG_F64 && val_type != TG_F32))
        return fail(TG_ERR_INVALID, "tg_csr_plan_set_values: bad argument");
    const int vi = val_type == TG_F64 ? 0 : 1;
    std::lock_guard<std::mutex> lk(plan->mu);
    plan->sval_src[vi] = nullptr;
    if (!a || !plan->sorted || plan->m == 0 || plan->n == 0 || plan->nnz == 0) return TG_OK;
    if (plan->nnz >= (1ll << 31)) return TG_OK;  // int32 offsets of the layout
    hipStream_t st = as_stream(stream);
    const int32_t m = plan->m;
    const int grid = (int)std::min<int64_t>((m + kBlock - 1) / kBlock, 65536);
    if (!plan->srp) {  // the layout of the pattern, once
        const size_t cnt = (size_t)kXcds * m + 1;
        int32_t* len = nullptr;
        size_t tmp_bytes = 0;
        TG_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, (const int32_t*)nullptr, (int32_t*)nullptr, (int32_t)0, cnt,
                                       rocprim::plus<int32_t>(), st));
        void* tmp = nullptr;
        TG_HIP(hipMallocAsync((void**)&len, cnt * sizeof(int32_t), st));
        TG_HIP(hipMallocAsync(&tmp, std::max<size_t>(tmp_bytes, 8), st));
        TG_HIP(hipMallocAsync((void**)&plan->srp, cnt * sizeof(int32_t), st));
        TG_HIP(hipMallocAsync((void**)&plan->sci, (size_t)plan->nnz * sizeof(int32_t), st));
        if (plan->rp_type == TG_I32)
            hipLaunchKernelGGL((split_len_kernel<int32_t>), dim3(grid), dim3(kBlock), 0, st, m,
                               (const int32_t*)plan->rp, (const int32_t*)plan->cuts, len);
        else
            hipLaunchKernelGGL((split_len_kernel<int64_t>), dim3(grid), dim3(kBlock), 0, st, m,
                               (const int64_t*)plan->rp, (const int32_t*)plan->cuts, len);
        TG_HIP(rocprim::exclusive_scan(tmp, tmp_bytes, (const int32_t*)len, plan->srp, (int32_t)0, cnt,
                                       rocprim::plus<int32_t>(), st));
        if (plan->rp_type == TG_I32)
            hipLaunchKernelGGL((split_scatter_kernel<int32_t, int32_t>), dim3(grid), dim3(kBlock), 0, st, m,
                               (const int32_t*)plan->rp, (const int32_t*)plan->cuts, (const int32_t*)plan->srp,
                               plan->ci, plan->sci);
        else
            hipLaunchKernelGGL((split_scatter_kernel<int64_t, int32_t>), dim3(grid), dim3(kBlock), 0, st, m,
                               (const int64_t*)plan->rp, (const int32_t*)plan->cuts, (const int32_t*)plan->srp,
                               plan->ci, plan->sci);
        TG_LAUNCH_CHECK();
        TG_HIP(hipFreeAsync(len, st));
        TG_HIP(hipFreeAsync(tmp, st));
    }
    const size_t vs = val_type == TG_F64 ? 8 : 4;
    if (!plan->sval[vi]) TG_HIP(hipMallocAsync(&plan->sval[vi], (size_t)plan->nnz * vs, st));
#define TG_SPLIT_VALS(RPT, VT)                                                                                   \
    hipLaunchKernelGGL((split_scatter_kernel<RPT, VT>), dim3(grid), dim3(kBlock), 0, st, m, (const RPT*)plan->rp, \
                       (const int32_t*)plan->cuts, (const int32_t*)plan->srp, (const VT*)a, (VT*)plan->sval[vi])
    if (plan->rp_type == TG_I32) {
        if (vi == 0)
            TG_SPLIT_VALS(int32_t, double);
        else
            TG_SPLIT_VALS(int32_t, float);
    } else {
        if (vi == 0)
            TG_SPLIT_VALS(int64_t, double);
        else
            TG_SPLIT_VALS(int64_t, float);
    }
#undef TG_SPLIT_VALS
    TG_LAUNCH_CHECK();
    TG_HIP(hipStreamSynchronize(st));
    plan->sval_src[vi] = a;
    return TG_OK;
}

// Slices of the planned product and whether the re-laid layout serves it (the values
// `a` are the ones registered for the value type and X is a few L2s large).
static int plan_slices(tg_csr_plan_t plan, tg_dtype val_type, int32_t nrhs, const void* a, bool* split) {
    *split = false;
    if (!plan || !plan->sorted || nrhs < 1 || nrhs > 4 || plan->m == 0 || plan->n == 0) return 0;
    if (g_spmm_slices == 1) return 0;
    const int64_t vs = val_type == TG_F64 ? 8 : 4;
    const int64_t xbytes = (int64_t)plan->n * nrhs * vs;
    const int vi = val_type == TG_F64 ? 0 : 1;
    if (g_spmm_split && a && plan->srp && plan->sval_src[vi] == a && (g_spmm_slices == 0 || g_spmm_slices == kXcds) &&
        (g_spmm_slices == kXcds || split_pays(xbytes))) {
        *split = true;
        return kXcds;
    }
    if (g_spmm_slices > 1) return g_spmm_slices;
    return slice_count(xbytes);
}

extern "C" int32_t tg_csr_plan_slices(tg_csr_plan_t plan, tg_dtype val_type, int32_t nrhs) {
    bool split = false;
    return plan_slices(plan, val_type, nrhs, nullptr, &split);
}

extern "C" int32_t tg_csr_plan_long_rows(tg_csr_plan_t plan) { return plan ? plan->nlong : 0; }

extern "C" int32_t tg_csr_plan_uses_split(tg_csr_plan_t plan, const void* a, tg_dtype val_type, int32_t nrhs) {
    bool split = false;
    if (!plan || (val_type != TG_F64 && val_type != TG_F32)) return 0;
    (void)plan_slices(plan, val_type, nrhs, a, &split);
    return split ? 1 : 0;
}

extern "C" int tg_csr_spmm_plan(tg_csr_plan_t plan, const void* a, tg_dtype val_type, double alpha, const void* X,
                                int64_t ldX, double beta, void* Y, int64_t ldY, int32_t nrhs, int32_t NRHS,
                                void* stream) {
    if (!plan) return fail(TG_ERR_INVALID, "tg_csr_spmm_plan: null plan");
    if (NRHS > 0 && nrhs != NRHS)
        return fail(TG_ERR_NRHS, "tg_csr_spmm: nrhs (%d) != NRHS (%d). Doing nothing.", nrhs, NRHS);
    CsrView A;
    A.m = plan->m;
    A.n = plan->n;
    A.nnz = plan->nnz;
    A.rp = plan->rp;
    A.rp_type = plan->rp_type;
    A.ci = plan->ci;
    A.val = a;
    A.val_type = val_type;
    A.band = plan->band;
    hipStream_t st = as_stream(stream);
    if (plan->nlong > 0 && (val_type == TG_F64 || val_type == TG_F32) && alpha != 0.0 && X && Y && X != Y &&
        ldX >= nrhs && ldY >= nrhs && nrhs > 0) {
        // long rows split off: the row kernels on the rest, then the rows' segments
        const size_t need = (size_t)plan->nseg * nrhs * (val_type == TG_F64 ? 8 : 4);
        std::lock_guard<std::mutex> lk(plan->mu);
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        TG_HIP(hipStreamIsCapturing(st, &cap));
        const bool track = cap == hipStreamCaptureStatusNone;
        if (track && plan->used && plan->used_on != st) TG_HIP(hipStreamWaitEvent(st, plan->used, 0));
        if (plan->part_bytes < need && !track)
            return fail(TG_ERR_WORKSPACE,
                        "tg_csr_spmm_plan: the plan's partial blocks (%zu bytes) must be sized by one call of this shape "
                        "outside stream capture before the call is captured",
                        need);
        if (plan->part_bytes < need) {
            if (plan->part) TG_HIP(hipFreeAsync(plan->part, st));
            plan->part = nullptr;
            plan->part_bytes = 0;
            TG_HIP(hipMallocAsync(&plan->part, need, st));
            plan->part_bytes = need;
        }
        LongRows L;
        L.nrows = plan->nlong;
        L.nseg = plan->nseg;
        L.rpe = plan->rpe;
        L.rows = plan->lrows;
        L.seg_ptr = plan->lseg_ptr;
        L.seg_lb = plan->lseg;
        L.seg_le = plan->lseg + plan->nseg;
        L.part = plan->part;
        A.longr = &L;
        const int rc = spmm_dispatch(A, alpha, X, ldX, beta, Y, ldY, nrhs, nullptr, 0, nullptr, st);
        if (track) {
            if (!plan->used) TG_HIP(hipEventCreateWithFlags(&plan->used, hipEventDisableTiming));
            TG_HIP(hipEventRecord(plan->used, st));
            plan->used_on = st;
        }
        return rc;
    }
    bool split = false;
    const int S = (val_type == TG_F64 || val_type == TG_F32) ? plan_slices(plan, val_type, nrhs, a, &split) : 0;
    // the plain path covers every special case (alpha = 0, aliasing, bad arguments)
    if (S == 0 || alpha == 0.0 || X == Y || !X || !Y || ldX < nrhs || ldY < nrhs)
        return spmm_dispatch(A, alpha, X, ldX, beta, Y, ldY, nrhs, nullptr, 0, nullptr, st);
    const size_t vs = val_type == TG_F64 ? 8 : 4;
    const size_t need = (size_t)S * plan->m * nrhs * vs;
    std::lock_guard<std::mutex> lk(plan->mu);
    // under stream capture the caller orders the uses of one plan; otherwise a call on
    // another stream waits for the previous user of the partial blocks
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    TG_HIP(hipStreamIsCapturing(st, &cap));
    const bool track = cap == hipStreamCaptureStatusNone;
    if (track && plan->used && plan->used_on != st) TG_HIP(hipStreamWaitEvent(st, plan->used, 0));
    if (plan->part_bytes < need && !track)
        // growing the blocks under capture would record an allocation node whose memory
        // belongs to the graph; the plan must be sized by an eager call of this shape first
        return fail(TG_ERR_WORKSPACE,
                    "tg_csr_spmm_plan: the plan's partial blocks (%zu bytes) must be sized by one call of this shape "
                    "outside stream capture before the call is captured",
                    need);
    if (plan->part_bytes < need) {
        if (plan->part) TG_HIP(hipFreeAsync(plan->part, st));
        plan->part = nullptr;
        plan->part_bytes = 0;
        TG_HIP(hipMallocAsync(&plan->part, need, st));
        plan->part_bytes = need;
    }
    const int32_t m = plan->m;
    const int vi = val_type == TG_F64 ? 0 : 1;
    // the re-laid layout reads its own column indices and values (a binary matrix has none)
    const int32_t* srp = split ? plan->srp : nullptr;
    const int32_t* ci_use = split ? plan->sci : plan->ci;
    const void* a_use = split ? plan->sval[vi] : a;
    int rc;
    if (val_type == TG_F64) {
        if (plan->rp_type == TG_I32)
            rc = spmm_sliced_typed<double, int32_t>(S, m, plan->nnz, (const int32_t*)plan->rp, plan->cuts, srp, ci_use,
                                                    (const double*)a_use, alpha, (const double*)X, ldX, beta,
                                                    (double*)Y, ldY, nrhs, (double*)plan->part, st);
        else
            rc = spmm_sliced_typed<double, int64_t>(S, m, plan->nnz, (const int64_t*)plan->rp, plan->cuts, srp, ci_use,
                                                    (const double*)a_use, alpha, (const double*)X, ldX, beta,
                                                    (double*)Y, ldY, nrhs, (double*)plan->part, st);
    } else {
        if (plan->rp_type == TG_I32)
            rc = spmm_sliced_typed<float, int32_t>(S, m, plan->nnz, (const int32_t*)plan->rp, plan->cuts, srp, ci_use,
                                                   (const float*)a_use, alpha, (const float*)X, ldX, beta, (float*)Y,
                                                   ldY, nrhs, (float*)plan->part, st);
        else
            rc = spmm_sliced_typed<float, int64_t>(S, m, plan->nnz, (const int64_t*)plan->rp, plan->cuts, srp, ci_use,
                                                   (const float*)a_use, alpha, (const float*)X, ldX, beta, (float*)Y,
                                                   ldY, nrhs, (float*)plan->part, st);
    }
    if (track) {
        if (!plan->used) TG_HIP(hipEventCreateWithFlags(&plan->used, hipEventDisableTiming));
        TG_HIP(hipEventRecord(plan->used, st));
        plan->used_on = st;
    }
    return rc;
}

extern "C" int tg_csr_spmm(int32_t m, int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                           const void* a, tg_dtype val_type, double alpha, const void* X, int64_t ldX, double beta,
                           void* Y, int64_t ldY, int32_t nrhs, int32_t NRHS, const int32_t* job_ptr, int32_t blocks,
                           void* stream) {
    if (NRHS > 0 && nrhs != NRHS)
        return fail(TG_ERR_NRHS, "tg_csr_spmm: nrhs (%d) != NRHS (%d). Doing nothing.", nrhs, NRHS);
    CsrView A;
    A.m = m;
    A.n = n;
    A.nnz = nnz;
    A.rp = rp;
    A.rp_type = rp_type;
    A.ci = ci;
    A.val = a;
    A.val_type = val_type;
    return spmm_dispatch(A, alpha, X, ldX, beta, Y, ldY, nrhs, job_ptr, blocks, nullptr, as_stream(stream));
}
