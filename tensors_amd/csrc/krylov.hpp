// Internal interfaces shared by the SpMM and Krylov translation units.
#pragma once

#include <functional>
#include <vector>

#include "common.hpp"

namespace tg {

// Device CSR operand of the solvers (all pointers device-resident).
// Long rows split off the per-row SpMM kernels (a plan's rows of more than
// spmm_long_row nonzeros): the row kernels read each row's end from `rpe`, where a
// long row is empty, and the long rows' nonzeros are added in segments of at most
// kLongSeg by a workgroup each, then combined per row in segment order.
constexpr int kLongSeg = 4096;
struct LongRows {
    int32_t nrows = 0, nseg = 0;
    const void* rpe = nullptr;         // m row ends, the row pointers' type
    const int32_t* rows = nullptr;     // [nrows] ascending
    const int32_t* seg_ptr = nullptr;  // [nrows + 1] first segment of each long row
    const int64_t* seg_lb = nullptr;   // [nseg] nonzero range of each segment
    const int64_t* seg_le = nullptr;
    void* part = nullptr;              // [nseg * nrhs] segment sums (sized by the caller)
};

struct CsrView {
    int32_t m = 0, n = 0;
    int64_t nnz = 0;
    const void* rp = nullptr;
    tg_dtype rp_type = TG_I32;
    const int32_t* ci = nullptr;
    const void* val = nullptr;  // null: binary matrix
    tg_dtype val_type = TG_F64;
    int64_t band = -1;          // max |column - row| if known (csr_band), else -1
    const LongRows* longr = nullptr;  // split-off long rows (plan path), or none
};

// Bandwidth max |ci - row| of a device CSR (one pass over ci; synchronises s).
// Returns -1 if it could not be computed.
int64_t csr_band(const CsrView& A, hipStream_t s);
// csr_band into a caller's device buffer of kBandInts int32 (allocates nothing).
constexpr int kBandInts = 1024;
int64_t csr_band_into(const CsrView& A, int32_t* buf, hipStream_t s);
// Whether tg_cg_solve's panel-ordered SpMM could apply to an n-row system, so that
// the caller measures the band (cg.hip).
bool cg_panel_wants_band(int64_t n);
// Whether spmm_dispatch would cut X into column slabs for this shape, so that
// knowing the band can matter.
bool spmm_may_slab(const CsrView& A, int32_t nrhs);

// Whether spmm_dispatch cuts X into column slabs (uses A.band when known).
bool spmm_uses_slabs(const CsrView& A, int32_t nrhs);

// Tuning keys of the batched engine and of the CG driver, forwarded by
// tg_tune_set / tg_tune_get.
int batch_tune_set(const char* key, int32_t value, bool* known);
int32_t batch_tune_get(const char* key);
int cg_tune_set(const char* key, int32_t value, bool* known);
int32_t cg_tune_get(const char* key);
int bsr_tune_set(const char* key, int32_t value, bool* known);
int32_t bsr_tune_get(const char* key);
int gmres_tune_set(const char* key, int32_t value, bool* known);
int32_t gmres_tune_get(const char* key);
int gbatch_tune_set(const char* key, int32_t value, bool* known);
int32_t gbatch_tune_get(const char* key);
int serve_tune_set(const char* key, int32_t value, bool* known);
int32_t serve_tune_get(const char* key);

// Y <- alpha A X + beta Y with the SpMM kernels; when `skip` is non-null and
// *skip != 0 at kernel start, the kernels return without touching Y.
int spmm_dispatch(const CsrView& A, double alpha, const void* X, int64_t ldX, double beta, void* Y, int64_t ldY,
                  int32_t nrhs, const int32_t* job_ptr, int32_t blocks, const int* skip, hipStream_t s);

// An operator y <- Op(x) on n x k row-major device blocks: either a device CSR
// matrix or a host callback (tg_apply_fn) that the library calls with device
// pointers.
struct OpRef {
    const CsrView* csr = nullptr;
    tg_apply_fn fn = nullptr;
    void* ctx = nullptr;
    bool is_callback() const { return csr == nullptr; }
};

int op_apply(const OpRef& op, const void* x, void* y, int k, const int* skip, hipStream_t s);

template <typename T>
inline int csr_apply(const CsrView& A, const T* X, T* Y, int k, hipStream_t s) {
    return spmm_dispatch(A, 1.0, X, k, 0.0, Y, k, k, nullptr, 0, nullptr, s);
}
template <typename T>
inline int csr_apply_guarded(const CsrView& A, const T* X, T* Y, int k, const int* skip, hipStream_t s) {
    return spmm_dispatch(A, 1.0, X, k, 0.0, Y, k, k, nullptr, 0, skip, s);
}

constexpr int kSolverBlocks = 1024;  // workgroups of the column-blocked vector kernels
constexpr int kFusedBlocks = 2048;   // workgroups of the fused CG direction + SpMM kernel
constexpr int kMaxCols = 256;        // nrhs limit of the solvers

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device-resident CG state (one per solve).
struct CGState {
    int iter;
    int done;
    int succeeded;
    int max_iter;
    int use_guess;
    int guard;  // alpha = 0 where p.u <= 0 (set by the mixed solver for its FP32 iteration)
    // Deferred x update (form 1): the p of iteration t sits in ring slot t % xdefer and its
    // alpha in the alpha history; x holds the updates of the iterations before xfrom.  The
    // p step of an iteration with xcnt > 0 adds those of [xfirst, xfirst + xcnt) to x.
    int xdefer;
    int xfrom;
    int xfirst;
    int xcnt;
    int pad[6];
    double rho[kMaxCols];
    double rho_old[kMaxCols];
    double alpha[kMaxCols];
    double beta[kMaxCols];
    double TOL[kMaxCols];
    double bsq[kMaxCols];
};

struct CGResult {
    int succeeded;
    int iterations;
    double rel_res[kMaxCols];
};

struct GmresResult {
    int succeeded;
    int iterations;
    int restarts;
    double rel_res[kMaxCols];
};

size_t gmres_workspace_bytes(int32_t n, int32_t max_iter, int32_t k, size_t vs);

// State a caller keeps across several gmres_solve_t calls on the same operator,
// workspace, stream and shapes (the mixed solver's one-cycle inner solves): the
// captured restart cycle, reused from the first call's first cycle on, and `checked`,
// which says the right-hand side is known finite and non-zero, so the call neither
// reads its norms back before the cycle nor handles those special cases.
struct GmresReuse {
    hipStream_t cs = nullptr;
    hipGraphExec_t ex = nullptr;
    bool checked = false;
    GmresReuse() = default;
    GmresReuse(const GmresReuse&) = delete;
    GmresReuse& operator=(const GmresReuse&) = delete;
    ~GmresReuse() {
        if (ex) (void)hipGraphExecDestroy(ex);
        if (cs) (void)hipStreamDestroy(cs);
    }
};

template <typename T>
int gmres_solve_t(const CsrView& A, const OpRef& Aop, int prec_kind, const T* d, const OpRef& Pop, int side,
                  int32_t k, int32_t max_iter, double a, const T* Bm, int64_t ldB, double b, T* X, int64_t ldX,
                  double tol, int32_t max_restarts, int use_guess, void* ws, size_t ws_bytes, GmresResult* res,
                  hipStream_t s, GmresReuse* reuse = nullptr);

// Ring slots of p for the CG's deferred x update: cg_workspace_bytes with `slots` = 2 is
// the least a solve accepts; with kCgRingSlots (what tg_cg_workspace_bytes reports) the
// x update is deferred over up to kCgRingSlots iterations.
constexpr int kCgRingSlots = 4;
size_t cg_workspace_bytes(int32_t n, int32_t k, size_t vs, int prec_kind, int slots = kCgRingSlots);

// Host hook of cg_solve_t: with a hook the host polls every `every` iterations and
// calls fn with the device state copied to the host (hs), the device buffers of the
// iteration (r, z, x, the workgroup partials, the device state).  fn may change
// them (e.g. replace the residual) and sets *stop to end the solve; if the device
// stopped (hs.done) and fn neither stops nor clears `done`, the solve ends.
struct CGHook {
    int every = 16;
    int guard = 0;  // 1: alpha = 0 for a column whose p.u is not positive (its p has vanished)
    // fn(state, r, z, x, p, partials, device state, stop): p is null where the direction
    // is double-buffered (form 2)
    std::function<int(const CGState& hs, void* r, void* z, void* x, void* p, double* partial, CGState* st,
                      bool* stop)>
        fn;
};

// allow_graphs = false keeps every iteration an eager launch: no stream or graph is
// created (the allocation-free entry point).
template <typename T>
int cg_solve_t(const CsrView& A, const OpRef& Aop, int prec_kind, const T* d, const OpRef& Pop, int32_t k, int32_t max_iter, double a, const T* Bm,
               int64_t ldB, double b, T* X, int64_t ldX, double tol, int use_guess, void* ws, size_t ws_bytes,
               CGResult* res, hipStream_t s, const CGHook* hook = nullptr, bool allow_graphs = true);

}  // namespace tg
