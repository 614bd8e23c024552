// tg_cg_solve_noalloc: the block PCG of tg_cg_solve with its workspace in a
// file-scope device array instead of caller- or library-allocated memory, for
// callers that may not allocate device memory (serving loops that must not
// call hipMalloc between requests, harnesses that forbid allocation).
//
// The array `scratch` (TG_CG_NOALLOC_SCRATCH_BYTES) lives in the library's code
// object, so every GPU the library runs on holds one copy until the process
// ends.  Host code never uses the array's name as an address: the device
// address of the current GPU's copy comes from hipGetSymbolAddress, taken once
// per device on that device's first call and kept in scratch_ptr[device].
//
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two host threads, or two streams whose work overlaps) race on it.  The array
// keeps the last call's contents; the solver initialises every part of its
// workspace that it reads (state block by a copy, vectors and partial sums by
// its setup kernels), and the band pass writes every entry of its buffer that it
// reads, so nothing has to start at zero.
//
// Layout: the solver's workspace (at most TG_CG_NOALLOC_SCRATCH_BYTES) at the
// start, then the band pass's kBandInts int32 in the last 4 KiB.

#include <mutex>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

constexpr size_t kWorkspaceBytes = TG_CG_NOALLOC_SCRATCH_BYTES;
constexpr size_t kBandBytes = 4096;
static_assert(kBandInts * sizeof(int32_t) <= kBandBytes, "band buffer");
constexpr size_t kScratchBytes = kWorkspaceBytes + kBandBytes;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex scratch_mu;
static void* scratch_ptr[kMaxDevices];  // device address of `scratch` on each GPU (null: not set up)

// Device address of the current GPU's `scratch` (set up on the first call there).
static int scratch_address(void** out) {
    int dev = 0;
    TG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TG_ERR_INVALID, "noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(TG_ERR_HIP, "noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return TG_OK;
}

template <typename T>
static int noalloc_entry(const CsrView& A, int32_t prec_kind, const void* prec_diag, int32_t nrhs, int32_t max_iter,
                         double a_coef, const void* B, int64_t ldB, double b_coef, void* X, int64_t ldX, double tol,
                         int32_t use_guess, void* ws, int32_t* succeeded, int32_t* iterations, double* rel_residuals,
                         hipStream_t s) {
    OpRef Aop;
    Aop.csr = &A;
    OpRef Pop;
    CGResult res{};
    const int rc = cg_solve_t<T>(A, Aop, prec_kind, (const T*)prec_diag, Pop, nrhs, max_iter, a_coef, (const T*)B,
                                 ldB, b_coef, (T*)X, ldX, tol, use_guess, ws, kWorkspaceBytes, &res, s, nullptr,
                                 /*allow_graphs=*/false);
    if (rc) return rc;
    if (succeeded) *succeeded = res.succeeded;
    if (iterations) *iterations = res.iterations;
    if (rel_residuals)
        for (int c = 0; c < nrhs; ++c) rel_residuals[c] = res.rel_res[c];
    return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" int tg_cg_solve_noalloc(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                                   const void* a, tg_dtype val_type, int32_t prec_kind, const void* prec_diag,
                                   int32_t nrhs, int32_t max_iter, double a_coef, const void* B, int64_t ldB,
                                   double b_coef, void* X, int64_t ldX, double tol, int32_t use_guess,
                                   int32_t* succeeded, int32_t* iterations, double* rel_residuals, void* stream) {
    if (succeeded) *succeeded = 0;
    if (iterations) *iterations = 0;
    if (n < 0 || nrhs < 1 || !B || !X || ldB < nrhs || ldX < nrhs || (prec_kind != 0 && prec_kind != 1))
        return fail(TG_ERR_INVALID, "tg_cg_solve_noalloc: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_cg_solve_noalloc: rp_type must be I32/I64");
    if (val_type != TG_F64 && val_type != TG_F32) return fail(TG_ERR_DTYPE, "tg_cg_solve_noalloc: val_type must be F64/F32");
    const size_t need = cg_workspace_bytes(n, nrhs, val_type == TG_F32 ? 4 : 8, prec_kind ? 1 : 0);
    if (need > kWorkspaceBytes)
        return fail(TG_ERR_WORKSPACE,
                    "tg_cg_solve_noalloc: n = %d with %d right-hand sides needs %zu bytes of workspace; the static "
                    "scratch holds %zu (use tg_cg_solve with a workspace instead)",
                    n, nrhs, need, kWorkspaceBytes);
    void* ws = nullptr;
    const int rs = scratch_address(&ws);
    if (rs) return rs;
    CsrView A;
    A.m = n;
    A.n = n;
    A.nnz = nnz;
    A.rp = rp;
    A.rp_type = rp_type;
    A.ci = ci;
    A.val = a;
    A.val_type = val_type;
    // The band decides the SpMM's row order as in tg_cg_solve (column slabs never apply:
    // X is at most a fifth of the workspace); it is measured into the array's last 4 KiB.
    hipStream_t s = as_stream(stream);
    if (cg_panel_wants_band(n)) {
        A.band = csr_band_into(A, (int32_t*)((char*)ws + kWorkspaceBytes), s);
        if (A.band < 0) return fail(TG_ERR_HIP, "tg_cg_solve_noalloc: band pass failed");
    }
    if (val_type == TG_F64)
        return noalloc_entry<double>(A, prec_kind, prec_diag, nrhs, max_iter, a_coef, B, ldB, b_coef, X, ldX, tol,
                                     use_guess, ws, succeeded, iterations, rel_residuals, s);
    return noalloc_entry<float>(A, prec_kind, prec_diag, nrhs, max_iter, a_coef, B, ldB, b_coef, X, ldX, tol,
                                use_guess, ws, succeeded, iterations, rel_residuals, s);
}
