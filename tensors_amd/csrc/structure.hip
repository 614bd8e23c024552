// Sparse structure work on the GPU: triples -> CSR, transpose, sparse x sparse.
//
// triples -> CSR follows MatrixCSR::FromTriples (MatrixCSR/FromTriples.hpp:3-209):
// entries are bucketed by row, each row is ordered by column (SortInner,
// MatrixCSR/SortInner.hpp:3-49), duplicates (i,j) are merged by summing their
// values when compressQ (PatternCSR/Compress.hpp:81-106), and symmetrizeQ adds
// (j,i,a) for every off-diagonal (i,j,a) (FromTriples.hpp:131-145).  Here the
// bucketing + row ordering is one stable LSD radix sort of the 64-bit keys
// (row << 32 | column) with the triple index as payload (rocPRIM), so rows,
// columns and nonzero counts are exactly determined and duplicates are summed
// in input order (deterministic).
//
// transpose follows MatrixCSR::Op<Trans> (MatrixCSR.hpp:445-490): the
// transposed triples go through the same sort, giving sorted rows.
// sparse x sparse follows MatrixCSR::Dot(MatrixCSR) / PatternCSR::DotBinary_
// (MatrixCSR.hpp:835-921, PatternCSR.hpp:731-839): expand every a_ij * b_jk
// into a triple (i, k), then sort and compress.

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

constexpr uint64_t kSentinel = ~0ull;

__global__ void st_expand_kernel(int64_t count, const int32_t* __restrict__ I, const int32_t* __restrict__ J,
                                 int32_t m, int32_t n, int symmetrize, uint64_t* __restrict__ keys,
                                 uint32_t* __restrict__ idx, int* __restrict__ bad) {
    const int64_t total = symmetrize ? 2 * count : count;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = e < count ? e : e - count;
        int32_t i = I[t], j = J[t];
        if (i < 0 || i >= m || j < 0 || j >= n) atomicOr(bad, 1);
        uint64_t key;
        if (e >= count) {  // transposed copy; the diagonal is not duplicated
            key = (i == j) ? kSentinel : (((uint64_t)(uint32_t)j << 32) | (uint32_t)i);
        } else {
            key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
        }
        keys[e] = key;
        idx[e] = (uint32_t)e;
    }
}

// head[e] = 1 where a new output entry starts (every entry when not compressing);
// sentinels (dropped diagonal copies) sort last and are never heads.
__global__ void st_heads_kernel(int64_t total, const uint64_t* __restrict__ keys, int compress,
                                int64_t* __restrict__ head) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[e];
        int64_t h = 0;
        if (k != kSentinel) h = (!compress || e == 0 || keys[e - 1] != k) ? 1 : 0;
        head[e] = h;
    }
}

// Writes column indices and (summed) values; counts entries per row.
template <typename T>
__global__ void st_write_kernel(int64_t total, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                const int64_t* __restrict__ head, const int64_t* __restrict__ pos, int64_t count,
                                const T* __restrict__ V, int compress, int32_t* __restrict__ ci, T* __restrict__ val,
                                int64_t* __restrict__ rowcnt) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (!head[e]) continue;
        const uint64_t k = keys[e];
        const int64_t p = pos[e];
        ci[p] = (int32_t)(uint32_t)(k & 0xffffffffull);
        atomicAdd((unsigned long long*)&rowcnt[(uint32_t)(k >> 32)], 1ull);
        if (V) {
            const uint32_t t0 = idx[e];
            T s = V[t0 < count ? t0 : t0 - count];
            if (compress)
                for (int64_t f = e + 1; f < total && keys[f] == k; ++f) {
                    const uint32_t t = idx[f];
                    s += V[t < count ? t : t - count];
                }
            val[p] = s;
        }
    }
}

__global__ void st_nnz_kernel(int64_t total, const int64_t* __restrict__ head, const int64_t* __restrict__ pos,
                              int64_t* __restrict__ nnz) {
    if (threadIdx.x == 0 && blockIdx.x == 0) nnz[0] = total > 0 ? pos[total - 1] + head[total - 1] : 0;
}

template <typename RP>
__global__ void st_rp_kernel(int32_t m, const int64_t* __restrict__ scanned, RP* __restrict__ rp) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= m; i += (int64_t)gridDim.x * blockDim.x)
        rp[i] = (RP)scanned[i];
}

__global__ void st_zero64_kernel(int64_t count, int64_t* __restrict__ a) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

static int grid_of(int64_t count) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 16384));
}

// Workspace layout of from_triples.
struct TriplesWs {
    uint64_t* keys[2];
    uint32_t* idx[2];
    int64_t* head;
    int64_t* pos;
    int64_t* rowcnt;  // m + 1
    int64_t* rowscan; // m + 1
    int64_t* nnz;
    int* bad;
    void* temp;
    size_t temp_bytes;
    size_t total_bytes;
};

static size_t triples_ws(int64_t count, int32_t m, int symmetrize, char* base, TriplesWs* out) {
    const int64_t total = symmetrize ? 2 * count : count;
    size_t sort_tmp = 0, scan_tmp = 0, scan_tmp2 = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_tmp, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                              (uint32_t*)nullptr, (size_t)std::max<int64_t>(total, 1), 0, 64);
    (void)rocprim::exclusive_scan(nullptr, scan_tmp, (const int64_t*)nullptr, (int64_t*)nullptr, (int64_t)0,
                            (size_t)std::max<int64_t>(total, 1), rocprim::plus<int64_t>());
    (void)rocprim::exclusive_scan(nullptr, scan_tmp2, (const int64_t*)nullptr, (int64_t*)nullptr, (int64_t)0,
                            (size_t)m + 1, rocprim::plus<int64_t>());
    const size_t tmp = std::max(sort_tmp, std::max(scan_tmp, scan_tmp2));
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t o = off;
        off += align256(std::max<size_t>(b, 8));
        return base ? (void*)(base + o) : nullptr;
    };
    TriplesWs w;
    w.keys[0] = (uint64_t*)take(8 * (size_t)total);
    w.keys[1] = (uint64_t*)take(8 * (size_t)total);
    w.idx[0] = (uint32_t*)take(4 * (size_t)total);
    w.idx[1] = (uint32_t*)take(4 * (size_t)total);
    w.head = (int64_t*)take(8 * (size_t)total);
    w.pos = (int64_t*)take(8 * (size_t)total);
    w.rowcnt = (int64_t*)take(8 * ((size_t)m + 1));
    w.rowscan = (int64_t*)take(8 * ((size_t)m + 1));
    w.nnz = (int64_t*)take(8);
    w.bad = (int*)take(8);
    w.temp = take(tmp);
    w.temp_bytes = tmp;
    w.total_bytes = off;
    if (out) *out = w;
    return off;
}

template <typename T>
static int from_triples_t(int64_t count, const int32_t* I, const int32_t* J, const T* V, int32_t m, int32_t n,
                          int compress, int symmetrize, void* rp_out, tg_dtype rp_type, int32_t* ci_out, T* val_out,
                          int64_t cap, void* ws, size_t ws_bytes, int64_t* nnz_host, hipStream_t s) {
    const int64_t total = symmetrize ? 2 * count : count;
    if (cap < total) return fail(TG_ERR_INVALID, "from_triples: output capacity %lld < %lld", (long long)cap,
                                 (long long)total);
    if (total > (int64_t)0xffffffffll) return fail(TG_ERR_INVALID, "from_triples: too many triples");
    TriplesWs w;
    if (!ws || ws_bytes < triples_ws(count, m, symmetrize, nullptr, nullptr))
        return fail(TG_ERR_WORKSPACE, "from_triples: workspace missing or too small");
    triples_ws(count, m, symmetrize, (char*)ws, &w);
    hipLaunchKernelGGL(st_zero64_kernel, dim3(grid_of(m + 1)), dim3(256), 0, s, (int64_t)m + 1, w.rowcnt);
    hipLaunchKernelGGL(st_zero64_kernel, dim3(1), dim3(256), 0, s, (int64_t)1, (int64_t*)w.bad);
    if (total > 0) {
        hipLaunchKernelGGL(st_expand_kernel, dim3(grid_of(total)), dim3(256), 0, s, count, I, J, m, n, symmetrize,
                           w.keys[0], w.idx[0], w.bad);
        size_t tb = w.temp_bytes;
        TG_HIP(rocprim::radix_sort_pairs(w.temp, tb, w.keys[0], w.keys[1], w.idx[0], w.idx[1], (size_t)total, 0, 64,
                                         s));
        hipLaunchKernelGGL(st_heads_kernel, dim3(grid_of(total)), dim3(256), 0, s, total, (const uint64_t*)w.keys[1],
                           compress, w.head);
        tb = w.temp_bytes;
        TG_HIP(rocprim::exclusive_scan(w.temp, tb, (const int64_t*)w.head, w.pos, (int64_t)0, (size_t)total,
                                       rocprim::plus<int64_t>(), s));
        hipLaunchKernelGGL((st_write_kernel<T>), dim3(grid_of(total)), dim3(256), 0, s, total,
                           (const uint64_t*)w.keys[1], (const uint32_t*)w.idx[1], (const int64_t*)w.head,
                           (const int64_t*)w.pos, count, V, compress, ci_out, val_out, w.rowcnt);
    }
    hipLaunchKernelGGL(st_nnz_kernel, dim3(1), dim3(64), 0, s, total, (const int64_t*)w.head, (const int64_t*)w.pos,
                       w.nnz);
    size_t tb = w.temp_bytes;
    TG_HIP(rocprim::exclusive_scan(w.temp, tb, (const int64_t*)w.rowcnt, w.rowscan, (int64_t)0, (size_t)m + 1,
                                   rocprim::plus<int64_t>(), s));
    if (rp_type == TG_I64)
        hipLaunchKernelGGL((st_rp_kernel<int64_t>), dim3(grid_of(m + 1)), dim3(256), 0, s, m,
                           (const int64_t*)w.rowscan, (int64_t*)rp_out);
    else
        hipLaunchKernelGGL((st_rp_kernel<int32_t>), dim3(grid_of(m + 1)), dim3(256), 0, s, m,
                           (const int64_t*)w.rowscan, (int32_t*)rp_out);
    TG_LAUNCH_CHECK();
    int64_t res[2] = {0, 0};
    TG_HIP(hipMemcpyAsync(&res[0], w.nnz, 8, hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(&res[1], w.bad, 8, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    if ((int)res[1]) return fail(TG_ERR_INVALID, "from_triples: a row or column index is out of range");
    if (rp_type == TG_I32 && res[0] > 0x7fffffffll)
        return fail(TG_ERR_DTYPE, "from_triples: nnz does not fit int32 row pointers");
    if (nnz_host) *nnz_host = res[0];
    return TG_OK;
}

// ---- transpose: triples (ci[e], row(e), val[e]) ---------------------------
template <typename RP>
__global__ void st_rows_kernel(int32_t m, const RP* __restrict__ rp, int32_t* __restrict__ rows) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        for (int64_t e = rp[i]; e < (int64_t)rp[i + 1]; ++e) rows[e] = (int32_t)i;
}

// ---- sparse x sparse expansion ---------------------------------------------
template <typename RP>
__global__ void st_expand_count_kernel(int32_t m, const RP* __restrict__ Arp, const int32_t* __restrict__ Aci,
                                       const RP* __restrict__ Brp, int64_t* __restrict__ cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = 0;
        for (int64_t e = Arp[i]; e < (int64_t)Arp[i + 1]; ++e) {
            const int32_t j = Aci[e];
            c += (int64_t)Brp[j + 1] - (int64_t)Brp[j];
        }
        cnt[i] = c;
    }
}

template <typename T, typename RP>
__global__ void st_expand_pairs_kernel(int32_t m, const RP* __restrict__ Arp, const int32_t* __restrict__ Aci,
                                       const T* __restrict__ Aval, const RP* __restrict__ Brp,
                                       const int32_t* __restrict__ Bci, const T* __restrict__ Bval,
                                       const int64_t* __restrict__ off, int32_t* __restrict__ I, int32_t* __restrict__ J,
                                       T* __restrict__ V) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = off[i];
        for (int64_t e = Arp[i]; e < (int64_t)Arp[i + 1]; ++e) {
            const int32_t j = Aci[e];
            const T a = Aval ? Aval[e] : T(1);
            for (int64_t f = Brp[j]; f < (int64_t)Brp[j + 1]; ++f, ++p) {
                I[p] = (int32_t)i;
                J[p] = Bci[f];
                if (V) V[p] = a * Bval[f];
            }
        }
    }
}

}  // namespace tg

using namespace tg;

extern "C" size_t tg_csr_from_triples_workspace_bytes(int64_t count, int32_t m, int32_t symmetrize) {
    if (count < 0 || m < 0) return 0;
    return triples_ws(count, m, symmetrize ? 1 : 0, nullptr, nullptr);
}

extern "C" int tg_csr_from_triples(int64_t count, const int32_t* I, const int32_t* J, const void* V,
                                   tg_dtype val_type, int32_t m, int32_t n, int32_t compress, int32_t symmetrize,
                                   void* rp_out, tg_dtype rp_type, int32_t* ci_out, void* val_out, int64_t capacity,
                                   void* workspace, size_t workspace_bytes, int64_t* nnz_out, void* stream) {
    if (nnz_out) *nnz_out = 0;
    if (count < 0 || m < 0 || n < 0 || !rp_out || (count > 0 && (!I || !J || !ci_out)) || (V && !val_out))
        return fail(TG_ERR_INVALID, "tg_csr_from_triples: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_csr_from_triples: rp_type must be I32/I64");
    hipStream_t s = as_stream(stream);
    if (val_type == TG_F64)
        return from_triples_t<double>(count, I, J, (const double*)V, m, n, compress, symmetrize, rp_out, rp_type, ci_out,
                                      (double*)val_out, capacity, workspace, workspace_bytes, nnz_out, s);
    if (val_type == TG_F32)
        return from_triples_t<float>(count, I, J, (const float*)V, m, n, compress, symmetrize, rp_out, rp_type, ci_out,
                                     (float*)val_out, capacity, workspace, workspace_bytes, nnz_out, s);
    return fail(TG_ERR_DTYPE, "tg_csr_from_triples: val_type must be F64/F32");
}

extern "C" size_t tg_csr_transpose_workspace_bytes(int32_t m, int32_t n, int64_t nnz) {
    if (m < 0 || n < 0 || nnz < 0) return 0;
    return align256((size_t)std::max<int64_t>(nnz, 1) * 4) + triples_ws(nnz, n, 0, nullptr, nullptr);
}

extern "C" int tg_csr_transpose(int32_t m, int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type,
                                const int32_t* ci, const void* val, tg_dtype val_type, void* out_rp,
                                int32_t* out_ci, void* out_val, void* workspace, size_t workspace_bytes,
                                void* stream) {
    if (m < 0 || n < 0 || nnz < 0 || !rp || !out_rp || (nnz > 0 && (!ci || !out_ci)) || (val && !out_val))
        return fail(TG_ERR_INVALID, "tg_csr_transpose: bad argument");
    if (!workspace || workspace_bytes < tg_csr_transpose_workspace_bytes(m, n, nnz))
        return fail(TG_ERR_WORKSPACE, "tg_csr_transpose: workspace missing or too small");
    hipStream_t s = as_stream(stream);
    int32_t* rows = (int32_t*)workspace;
    void* rest = (char*)workspace + align256((size_t)std::max<int64_t>(nnz, 1) * 4);
    const size_t rest_bytes = workspace_bytes - align256((size_t)std::max<int64_t>(nnz, 1) * 4);
    if (rp_type == TG_I64)
        hipLaunchKernelGGL((st_rows_kernel<int64_t>), dim3(grid_of(m)), dim3(256), 0, s, m, (const int64_t*)rp, rows);
    else if (rp_type == TG_I32)
        hipLaunchKernelGGL((st_rows_kernel<int32_t>), dim3(grid_of(m)), dim3(256), 0, s, m, (const int32_t*)rp, rows);
    else
        return fail(TG_ERR_DTYPE, "tg_csr_transpose: rp_type must be I32/I64");
    TG_LAUNCH_CHECK();
    int64_t got = 0;
    // (i, j, a) of A becomes (j, i, a): rows of A^T are the columns of A
    int rc;
    if (val_type == TG_F32)
        rc = from_triples_t<float>(nnz, ci, rows, (const float*)val, n, m, 0, 0, out_rp, rp_type, out_ci,
                                   (float*)out_val, nnz, rest, rest_bytes, &got, s);
    else
        rc = from_triples_t<double>(nnz, ci, rows, (const double*)val, n, m, 0, 0, out_rp, rp_type, out_ci,
                                    (double*)out_val, nnz, rest, rest_bytes, &got, s);
    return rc;
}

// Expanded (pre-compression) nonzero count of A * B; synchronises `stream`.
extern "C" int tg_csr_spgemm_expanded_count(int32_t m, const void* Arp, const int32_t* Aci, const void* Brp,
                                            tg_dtype rp_type, int64_t* count_out, void* workspace,
                                            size_t workspace_bytes, void* stream) {
    if (m < 0 || !Arp || !Brp || !count_out) return fail(TG_ERR_INVALID, "tg_csr_spgemm_expanded_count: bad argument");
    size_t scan_tmp = 0;
    (void)rocprim::exclusive_scan(nullptr, scan_tmp, (const int64_t*)nullptr, (int64_t*)nullptr, (int64_t)0, (size_t)m + 1,
                            rocprim::plus<int64_t>());
    const size_t need = 2 * align256(8 * ((size_t)m + 1)) + align256(scan_tmp);
    if (!workspace || workspace_bytes < need)
        return fail(TG_ERR_WORKSPACE, "tg_csr_spgemm_expanded_count: workspace must hold %zu bytes", need);
    hipStream_t s = as_stream(stream);
    int64_t* cnt = (int64_t*)workspace;
    int64_t* off = (int64_t*)((char*)workspace + align256(8 * ((size_t)m + 1)));
    void* tmp = (char*)workspace + 2 * align256(8 * ((size_t)m + 1));
    hipLaunchKernelGGL(st_zero64_kernel, dim3(grid_of(m + 1)), dim3(256), 0, s, (int64_t)m + 1, cnt);
    if (rp_type == TG_I64)
        hipLaunchKernelGGL((st_expand_count_kernel<int64_t>), dim3(grid_of(m)), dim3(256), 0, s, m, (const int64_t*)Arp,
                           Aci, (const int64_t*)Brp, cnt);
    else
        hipLaunchKernelGGL((st_expand_count_kernel<int32_t>), dim3(grid_of(m)), dim3(256), 0, s, m, (const int32_t*)Arp,
                           Aci, (const int32_t*)Brp, cnt);
    TG_LAUNCH_CHECK();
    TG_HIP(rocprim::exclusive_scan(tmp, scan_tmp, (const int64_t*)cnt, off, (int64_t)0, (size_t)m + 1,
                                   rocprim::plus<int64_t>(), s));
    TG_HIP(hipMemcpyAsync(count_out, off + m, 8, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    return TG_OK;
}

extern "C" size_t tg_csr_spgemm_workspace_bytes(int32_t m, int32_t n, int64_t expanded, tg_dtype val_type) {
    size_t scan_tmp = 0;
    (void)rocprim::exclusive_scan(nullptr, scan_tmp, (const int64_t*)nullptr, (int64_t*)nullptr, (int64_t)0, (size_t)m + 1,
                            rocprim::plus<int64_t>());
    const size_t vs = val_type == TG_F32 ? 4 : 8;
    return 2 * align256(8 * ((size_t)m + 1)) + align256(scan_tmp) + 2 * align256(4 * (size_t)std::max<int64_t>(expanded, 1)) +
           align256(vs * (size_t)std::max<int64_t>(expanded, 1)) + triples_ws(expanded, m, 0, nullptr, nullptr);
}

// C = A * B (A: m x k, B: k x n).  Outputs need capacity `expanded` (from
// tg_csr_spgemm_expanded_count); *nnz_out receives the compressed count.
// Values null on both A and B -> pattern product (DotBinary_).
extern "C" int tg_csr_spgemm(int32_t m, int32_t k, int32_t n, const void* Arp, const int32_t* Aci, const void* Aval,
                             const void* Brp, const int32_t* Bci, const void* Bval, tg_dtype rp_type,
                             tg_dtype val_type, int64_t expanded, void* Crp, int32_t* Cci, void* Cval,
                             void* workspace, size_t workspace_bytes, int64_t* nnz_out, void* stream) {
    if (m < 0 || k < 0 || n < 0 || expanded < 0 || !Arp || !Brp || !Crp || ((Aval == nullptr) != (Bval == nullptr)))
        return fail(TG_ERR_INVALID, "tg_csr_spgemm: bad argument");
    if (!workspace || workspace_bytes < tg_csr_spgemm_workspace_bytes(m, n, expanded, val_type))
        return fail(TG_ERR_WORKSPACE, "tg_csr_spgemm: workspace missing or too small");
    hipStream_t s = as_stream(stream);
    size_t scan_tmp = 0;
    (void)rocprim::exclusive_scan(nullptr, scan_tmp, (const int64_t*)nullptr, (int64_t*)nullptr, (int64_t)0, (size_t)m + 1,
                            rocprim::plus<int64_t>());
    char* w = (char*)workspace;
    int64_t* cnt = (int64_t*)w;
    w += align256(8 * ((size_t)m + 1));
    int64_t* off = (int64_t*)w;
    w += align256(8 * ((size_t)m + 1));
    void* tmp = w;
    w += align256(scan_tmp);
    const size_t E = (size_t)std::max<int64_t>(expanded, 1);
    int32_t* I = (int32_t*)w;
    w += align256(4 * E);
    int32_t* J = (int32_t*)w;
    w += align256(4 * E);
    void* V = w;
    w += align256((val_type == TG_F32 ? 4 : 8) * E);
    const size_t rest_bytes = workspace_bytes - (size_t)(w - (char*)workspace);
    hipLaunchKernelGGL(st_zero64_kernel, dim3(grid_of(m + 1)), dim3(256), 0, s, (int64_t)m + 1, cnt);
    const bool i64 = rp_type == TG_I64;
    if (i64)
        hipLaunchKernelGGL((st_expand_count_kernel<int64_t>), dim3(grid_of(m)), dim3(256), 0, s, m, (const int64_t*)Arp,
                           Aci, (const int64_t*)Brp, cnt);
    else
        hipLaunchKernelGGL((st_expand_count_kernel<int32_t>), dim3(grid_of(m)), dim3(256), 0, s, m, (const int32_t*)Arp,
                           Aci, (const int32_t*)Brp, cnt);
    TG_HIP(rocprim::exclusive_scan(tmp, scan_tmp, (const int64_t*)cnt, off, (int64_t)0, (size_t)m + 1,
                                   rocprim::plus<int64_t>(), s));
    const bool vals = Aval != nullptr;
#define TG_EXPAND(TT, RPT)                                                                                          \
    hipLaunchKernelGGL((st_expand_pairs_kernel<TT, RPT>), dim3(grid_of(m)), dim3(256), 0, s, m, (const RPT*)Arp, Aci, \
                       (const TT*)Aval, (const RPT*)Brp, Bci, (const TT*)Bval, (const int64_t*)off, I, J,             \
                       vals ? (TT*)V : (TT*)nullptr)
    if (val_type == TG_F32) {
        if (i64) TG_EXPAND(float, int64_t); else TG_EXPAND(float, int32_t);
    } else {
        if (i64) TG_EXPAND(double, int64_t); else TG_EXPAND(double, int32_t);
    }
#undef TG_EXPAND
    TG_LAUNCH_CHECK();
    if (val_type == TG_F32)
        return from_triples_t<float>(expanded, I, J, vals ? (const float*)V : nullptr, m, n, 1, 0, Crp, rp_type, Cci,
                                     (float*)Cval, expanded, w, rest_bytes, nnz_out, s);
    return from_triples_t<double>(expanded, I, J, vals ? (const double*)V : nullptr, m, n, 1, 0, Crp, rp_type, Cci,
                                  (double*)Cval, expanded, w, rest_bytes, nnz_out, s);
}
