/*
 * tensors_amd.h -- C ABI of the MI355X (gfx950) sparse hot path.
 *
 * Plain pointers and sizes only.  Every pointer to matrix or vector data is a
 * DEVICE pointer (hipMalloc'd or otherwise GPU-resident) unless a comment says
 * "host".  Every call that takes a `stream` (a hipStream_t passed as void*,
 * NULL = the null stream) is stream-ordered: it enqueues work and returns
 * without synchronising, except where a comment says it synchronises.
 *
 * Return values are tg_status codes; tg_last_error() gives a message for the
 * calling thread's most recent non-zero status.
 *
 * Index types: column indices are int32 (`Int` of the reference); row
 * pointers (`LInt`) are int32 or int64, named by an rp_type argument.
 * Values / dense blocks are FP64 or FP32 (val_type); A, X and Y share it.
 * Dense blocks are row-major with a leading dimension (ld >= nrhs), as in the
 * reference's Tensor2 (src/Tensor2.hpp:9-30).
 */
#ifndef TENSORS_AMD_H
#define TENSORS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tg_dtype { TG_F64 = 0, TG_F32 = 1, TG_I32 = 2, TG_I64 = 3 } tg_dtype;

/* Operator callback used by the *_op solver entry points: y <- Op(x) on
 * n x nrhs row-major DEVICE blocks (ld = nrhs).  The library calls it from the
 * host thread that called the solver; the callback must enqueue its work on
 * `stream` (or finish it before returning).  Non-zero return = failure. */
typedef int (*tg_apply_fn)(void* ctx, const void* x, void* y, int32_t nrhs, void* stream);

typedef enum tg_status {
    TG_OK = 0,
    TG_ERR_INVALID = -1,  /* bad argument (negative size, null pointer, ld < nrhs, ...) */
    TG_ERR_ALIAS = -2,    /* X == Y: the reference prints and does nothing (SpMM.hpp:191-198) */
    TG_ERR_NRHS = -3,     /* fixed NRHS != nrhs: reference no-op (Multiply_DenseMatrix.hpp:19-26) */
    TG_ERR_DTYPE = -4,    /* unsupported index / value type */
    TG_ERR_HIP = -5,      /* a HIP runtime call failed */
    TG_ERR_WORKSPACE = -6,/* workspace missing or too small */
    TG_ERR_SHAPE = -7     /* shape mismatch (MatrixCSR.hpp:999-1002,1018-1021) */
} tg_status;

/* Message for the calling thread's last non-zero status ("" if none). */
const char* tg_last_error(void);
/* Library version string. */
const char* tg_version(void);

/* ------------------------------------------------------------------------
 * Row partition -- replaces Tools' JobPointers<Int>(m, outer, T, false) that
 * PatternCSR::RequireJobPtr builds and caches (src/Sparse/PatternCSR.hpp:389-398).
 * Writes job_ptr[0..blocks] (device, int32): block b owns rows
 * [job_ptr[b], job_ptr[b+1]), cut where rp[i] + i first reaches
 * b * (nnz + m) / blocks, so every block carries about the same nonzeros + rows.
 * ------------------------------------------------------------------------ */
int tg_csr_job_pointers(int32_t m, const void* rp, tg_dtype rp_type, int32_t blocks, int32_t* job_ptr,
                        void* stream);

/* Number of partition blocks tg_csr_spmm uses for a matrix of m rows and nnz
 * nonzeros multiplied against nrhs columns (host arithmetic only). */
int32_t tg_csr_default_blocks(int32_t m, int64_t nnz, int32_t nrhs);

/* ------------------------------------------------------------------------
 * Y <- alpha * A X + beta * Y   (A: m x n CSR; X: n x nrhs, ldX; Y: m x nrhs, ldY)
 *
 * Replaces SparseBLAS<Scal,Int,LInt>::Multiply_DenseMatrix<NRHS>
 * (src/SparseBLAS/Multiply_DenseMatrix.hpp:9-45), i.e. the SpMV
 * (src/SparseBLAS/SpMV.hpp:3-288) and SpMM (src/SparseBLAS/SpMM.hpp:3-569)
 * kernels behind Sparse::MatrixCSR::Dot (src/Sparse/MatrixCSR.hpp:965-1086)
 * and PatternCSR::Dot_ (src/Sparse/PatternCSR.hpp:844-890).
 *
 *  a == NULL   : binary matrix, every stored entry is 1 (a_flag Plus).
 *  NRHS > 0    : the fixed-NRHS instantiation; NRHS != nrhs -> TG_ERR_NRHS, no-op.
 *  alpha == 0  : Y <- beta * Y only (beta 0: zero, 1: untouched).
 *  beta == 0   : Y is written without being read (NaNs in Y do not propagate).
 *  X == Y      : TG_ERR_ALIAS, no-op.
 *  empty rows  : Y row <- beta * Y row (beta 0: zero, 1: untouched).
 *  nnz         : rp[m] - rp[0], known to the caller (MatrixCSR::NonzeroCount);
 *                it picks the lane layout, it is not read from the device.
 *  job_ptr     : optional device partition from tg_csr_job_pointers with
 *                `blocks` blocks; NULL = rows split evenly inside the kernel.
 * ------------------------------------------------------------------------ */
int tg_csr_spmm(int32_t m, int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci, const void* a,
                tg_dtype val_type, double alpha, const void* X, int64_t ldX, double beta, void* Y,
                int64_t ldY, int32_t nrhs, int32_t NRHS, const int32_t* job_ptr, int32_t blocks,
                void* stream);

/* Column-sliced execution plan of one CSR pattern (no reference counterpart;
 * cached per matrix like PatternCSR::RequireJobPtr, PatternCSR.hpp:389-398).
 * For dense blocks of 6-32 MB (nrhs <= 4) the columns are cut into 2-8 ranges,
 * one per group of XCDs, so every XCD gathers from a slice of X that stays in
 * its L2; the slices' partial sums are added in slice order.  The plan keeps
 * the rp/ci pointers (they must stay valid) and 7 int32 per row; it needs
 * sorted columns within rows (checked when built, else the plain kernels
 * run).  tg_csr_spmm_plan has tg_csr_spmm's semantics and falls back to it
 * whenever slicing does not pay.  Calls on one plan may come from several host
 * threads and streams: they share the plan's partial blocks, so each sliced call
 * first waits for the previous one on the device (outside stream capture; under
 * capture the caller orders them).  The partial blocks grow on the first call of
 * a larger shape; a captured call never allocates: it returns TG_ERR_WORKSPACE
 * (nothing enqueued) unless one call of that shape (same nrhs and value type, or
 * a larger one) ran on the plan outside capture first.
 * tg_csr_plan_slices returns the slice count it would use (0 = plain) for values
 * that are not registered (below).
 *
 * tg_csr_plan_set_values registers the value array `a` of one value type (NULL
 * drops the registration): the plan then keeps a re-laid copy of the pattern and
 * of those values in which each of 8 column slices (one per XCD) is a CSR of its
 * own, stored slice after slice, so the workgroups of one XCD stream only their
 * slice's nonzeros instead of touching every cache line of A for a short run per
 * row.  Calls of tg_csr_spmm_plan with exactly that pointer and an X block of
 * 2-64 MB (nrhs <= 4) use it (8 slices; the same run sums in the same order as
 * the 8-slice cut path); any other pointer takes the paths above.  The plan
 * copies the values when the call is made: if the values at `a` change, call
 * tg_csr_plan_set_values again (the drop-in MatrixCSR does, since changing its
 * values drops its plan).  Extra device memory: 4 (8m + 1) + 4 nnz bytes for
 * the pattern and nnz * sizeof(value) per registered type.  Synchronises
 * `stream`.  tg_csr_plan_uses_split tells whether a call with `a` would use it.
 *
 * Long rows: a plan also finds the rows of more than spmm_long_row nonzeros
 * (tg_tune_set knob, default 8192, 0 = never; read when the plan is made).  A
 * single such row would keep one wave busy long after the rest of the matrix is
 * done (one row of 1e6 nonzeros made an FP64 SpMV 150x slower).  tg_csr_spmm_plan
 * then runs the row kernels with those rows seen as empty, adds each long row's
 * nonzeros in segments of 4096 with a workgroup per segment, and combines each
 * row's segment sums in order (deterministic; y = alpha z + beta y rounded as
 * alpha z + (beta y)); slicing is not used for such a plan.  Extra device memory:
 * m row ends of rp's type, 20 bytes per segment, and the segment sums (in the
 * partial blocks).  tg_csr_plan_long_rows returns the number of long rows. */
typedef struct tg_csr_plan_s* tg_csr_plan_t;
int tg_csr_plan_create(int32_t m, int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                       tg_csr_plan_t* out, void* stream);
int tg_csr_plan_destroy(tg_csr_plan_t plan);
int32_t tg_csr_plan_slices(tg_csr_plan_t plan, tg_dtype val_type, int32_t nrhs);
int tg_csr_plan_set_values(tg_csr_plan_t plan, const void* a, tg_dtype val_type, void* stream);
int32_t tg_csr_plan_uses_split(tg_csr_plan_t plan, const void* a, tg_dtype val_type, int32_t nrhs);
int32_t tg_csr_plan_long_rows(tg_csr_plan_t plan);
int tg_csr_spmm_plan(tg_csr_plan_t plan, const void* a, tg_dtype val_type, double alpha, const void* X, int64_t ldX,
                     double beta, void* Y, int64_t ldY, int32_t nrhs, int32_t NRHS, void* stream);

/* Kernel-shape knobs read by tg_csr_spmm (no reference counterpart; the
 * reference picks its kernel by template arguments).  Keys:
 *   "spmv_kernel": nrhs == 1 kernel, 0 = streaming (default), 1 = team-per-row.
 *   "spmv_team":   lanes per row for nrhs == 1, 0 = from the mean row length
 *                  (streaming: a quarter of it, 1..32; team-per-row: all of it,
 *                  4..64), else a power of two in [1, 64] (the streaming kernel
 *                  caps it at 32, the team-per-row kernel raises it to 4).
 *   "spmm_group":  lanes per RHS tile of the nrhs >= 2 kernel (a wave's 64/G lane
 *                  groups split each row's nonzeros), 0 = enough lanes for all
 *                  columns, else a power of two in [1, 64].
 *   "spmm_rows":   nrhs >= 2 kernel, 0 = from the mean row length (up to 16:
 *                  row per lane group), 1 = the wave's lane groups split one row,
 *                  2 = each lane group owns a row.
 *   "spmm_slices": column slices of tg_csr_spmm_plan, 0 = from the size of X
 *                  (default: the re-laid 8 slices for registered values and X of
 *                  2-64 MB, else 2 or 4 slices for X blocks of 6-32 MB, else none),
 *                  1 = never, 2 / 4 / 8 = always that many (nrhs <= 4).
 *   "spmm_long_row": rows of more than this many nonzeros are split off by
 *                  plans made afterwards (default 8192, 0 = never; see above).
 *   "spmm_split":  1 = registered values use the re-laid layout (default),
 *                  0 = never.
 *   "batch_tile_rows": rows per workgroup tile of systems added to a batched
 *                  engine from now on, a power of two in [256, 8192] (1024).
 *   "batch_stream": SpMM kernel of the batched engine's fused form, 0 = a row
 *                  per lane (default), 1 = streaming chunks (measured slower).
 *   "batch_split": batched iteration, 1 = direction pass + SpMM of p + update
 *                  (default), 0 = fused direction and SpMM (bitwise the same).
 *                  The split form defers x += alpha p to the next direction pass:
 *                  change this knob only while no batch has systems in flight.
 *   "batch_xcd_order": 0 = each XCD works through a contiguous run of the
 *                  batched engine's system-major tile list (default: a tile's stencil
 *                  neighbours are in flight on its XCD), 1 = each XCD's tiles spread
 *                  over the whole list (measured 8-9 % slower).  Bitwise the same.
 *   "batch_graph": 1 = each tg_batch_step runs its iterations as one captured
 *                  hipGraph (re-captured when the set of active systems changes;
 *                  default), 0 = eager launches.  Bitwise the same results.
 *   "batch_idx16": 1 = 16-bit column offsets (ci - row) in the batched SpMM for
 *                  systems whose offsets fit (default), 0 = int32 indices.
 *   "batch_spmm":  SpMM kernel of the split batched iteration, 2 = the offsets and
 *                  values of 256-row segments staged through LDS with coalesced
 *                  loads (default), 1 = row loop with the next row's indices
 *                  loaded during this row's gathers, 0 = row loop.  Bitwise the
 *                  same results.
 *   "batch_replace_drop_pct", "batch_check_every": residual replacement of
 *                  batched modes 1 / 2 when the estimate fell to this percent
 *                  (10) at checkpoints every this many iterations (16).
 *   "cg_fused":    iteration of tg_cg_solve / tg_cg_solve_mixed for CSR operators
 *                  with P = I or Jacobi and mean row length <= 16: 0 = SpMM then
 *                  four vector kernels, 1 = p.u in the SpMM epilogue, 2 = p
 *                  update, SpMM and p.u in one kernel plus one vector kernel,
 *                  3 = default (form 2 for blocks of up to 2^20 entries
 *                  n * nrhs, form 1 above).  Results agree to rounding.
 *   "cg_panel":    rows of that iteration's SpMM kernel, 1 = panel order where
 *                  the band allows (default: the rows cut into panels of one band,
 *                  every workgroup taking the same offsets of each panel, so a row's
 *                  neighbours one band away are reused from the L2), 0 = each
 *                  workgroup a contiguous chunk.  Results differ by rounding only.
 *   "cg_graph":    tg_cg_solve and the mixed CG with device operators, 1 = after
 *                  the first batch, blocks of 16 iterations run as one captured
 *                  hipGraph (default; the CSR iteration forms), 0 = eager launches.
 *                  Bitwise the same results.  tg_cg_solve_noalloc always launches
 *                  eagerly (it creates no stream or graph).
 *   "cg_xdefer":   form 1's x += alpha p deferred over 1, 2 or 4 (default)
 *                  iterations and applied in one pass from a ring of p, so three p
 *                  steps in four neither read nor write x.  Bitwise the same
 *                  results (the same fused multiply-adds in the same order).
 *   "gmres_graph": tg_gmres_solve with device operators, 1 = every restart cycle
 *                  after the first runs as one captured hipGraph (default), 0 =
 *                  eager launches.  Bitwise the same results.
 *   "gmres_delayed": Gram-Schmidt of tg_gmres_solve (and of the FP32 cycles of
 *                  tg_gmres_solve_mixed), 1 = CGS2 with each vector's second pass
 *                  delayed to the next step and fused with its projection (default
 *                  for restart lengths up to 32 and (2m + 2) * nrhs <= 2048: two
 *                  passes over the basis and five launches per step), 0 = CGS2
 *                  (three passes, eight launches).  Results agree to rounding.
 *   "gmres_split_dots": delayed steps with one right-hand side and more than 16
 *                  basis vectors (FP32: the cycles of tg_gmres_solve_mixed; FP64:
 *                  systems below 2^19 rows): 1 = the dots pass with its u-side and
 *                  v-side dot products on alternate lanes (default: half the
 *                  accumulators per lane), and in FP32 the update pass with its two
 *                  output vectors on alternate lanes; 0 = both on every lane.
 *                  Results agree to rounding.
 *   "bsr_vector":  tg_bsr_spmm with row-major X / Y chunks, 1 = a lane per
 *                  (block row, 1-4 right-hand sides) with vector loads (default),
 *                  0 = a lane per output entry (the mapping of the other
 *                  storage orders).
 *   "gbatch_big_tiles": systems added to a batched GMRES engine with more than
 *                  this many 1024-row tiles (default 64) get each reduction's totals
 *                  from one workgroup per system instead of every tile workgroup
 *                  adding all tiles' partials.  Bitwise the same results.
 *   "gbatch_delayed": batched GMRES engines created afterwards, 1 = the second
 *                  Gram-Schmidt pass delayed to the next step (default: two
 *                  launches and two passes over the basis per step, every
 *                  workgroup computing the system's scalars from the totals), 0 =
 *                  CGS2 (four launches, three passes).  Results agree to rounding;
 *                  within one engine a system's result is bitwise the same alone
 *                  and in any mix in both forms.
 *   "serve_gmres_batch", "serve_gmres_batch_rows": GMRES requests of servers
 *                  created afterwards; 1 (default) = requests of up to
 *                  serve_gmres_batch_rows rows (default 131072) join one batched
 *                  GMRES engine (tg_gbatch_*), larger ones run on the worker
 *                  threads; 0 = every request on a worker thread.
 *   "serve_gmres_mixed": GMRES requests submitted afterwards that run on a worker
 *                  thread: 1 = tg_gmres_solve_mixed (default: FP32 cycles, stop test
 *                  on the FP64 true residual; the result's iterations count the
 *                  FP32 Arnoldi steps of all cycles), 0 = tg_gmres_solve.
 * Process-wide; not thread-safe against concurrent SpMM calls. */
int tg_tune_set(const char* key, int32_t value);
int32_t tg_tune_get(const char* key);

/* Y <- alpha * X + beta * Y over an m x nrhs block (flag shortcuts of
 * combine_matrices_auto: beta 0 -> Y not read; alpha 0 -> X not read).
 * Replaces Tools' combine_matrices_auto / scale_matrix as used by
 * ConjugateGradient.hpp:342-346 and GMRES.hpp:835-845. */
int tg_axpby(int32_t m, int32_t nrhs, tg_dtype val_type, double alpha, const void* X, int64_t ldX,
             double beta, void* Y, int64_t ldY, void* stream);

/* ------------------------------------------------------------------------
 * Block preconditioned conjugate gradient -- replaces
 * ConjugateGradient<NRHS,Scal,Int>::operator()(A, P, a, B, ldB, b, X, ldX,
 * tol, use_initial_guessQ) (ConjugateGradient.hpp:105-353) for the case where
 * A is a CSR matrix and P is the identity (prec_kind 0) or a diagonal scaling
 * z_i = prec_diag[i] * r_i (prec_kind 1, e.g. a Jacobi / lumped-mass inverse).
 *
 *   X <- a_coef * A^{-1} B + b_coef * X        (n x nrhs blocks, row-major)
 *
 * Same recurrences, per-column scalars and stopping rule as the reference:
 * all columns satisfy |r.z|_k <= tol^2 |b.Pb|_k, or max_iter iterations.
 * Special cases as the reference: a_coef == 0 or B == 0 -> X <- b_coef X,
 * succeeded; NaN in B -> nothing written, not succeeded.  x starts at 0.
 * workspace: device memory of tg_cg_workspace_bytes() bytes (not zeroed); a
 * workspace of two n x nrhs blocks less (aligned to 256 bytes) is accepted too,
 * with the x update then deferred over at most 2 iterations ("cg_xdefer").
 * Synchronises `stream` (the host polls convergence between batches of
 * iterations); *succeeded, *iterations and rel_residuals[nrhs] (sqrt(|r.z| /
 * |b.Pb|) per column, as ConjugateGradient::RelativeResiduals) are host
 * outputs.  nrhs <= 256.
 * ------------------------------------------------------------------------ */
size_t tg_cg_workspace_bytes(int32_t n, int32_t nrhs, tg_dtype val_type, int32_t prec_kind);
/* The generic-operator form: A and P are callables, as in the reference's
 * A(cptr x, mptr y) / P(cptr x, mptr y) contract (ConjugateGradient.hpp:
 * 120-128), here called with device pointers (apply_P NULL = identity).
 * Workspace: tg_cg_workspace_bytes(n, nrhs, val_type, apply_P != NULL). */
int tg_cg_solve_op(int32_t n, tg_apply_fn apply_A, void* ctx_A, tg_apply_fn apply_P, void* ctx_P,
                   tg_dtype val_type, int32_t nrhs, int32_t max_iter, double a_coef, const void* B, int64_t ldB,
                   double b_coef, void* X, int64_t ldX, double tol, int32_t use_guess, void* workspace,
                   size_t workspace_bytes, int32_t* succeeded, int32_t* iterations, double* rel_residuals,
                   void* stream);
int tg_cg_solve(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci, const void* a,
                tg_dtype val_type, int32_t prec_kind, const void* prec_diag, int32_t nrhs, int32_t max_iter,
                double a_coef, const void* B, int64_t ldB, double b_coef, void* X, int64_t ldX, double tol,
                int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                int32_t* iterations, double* rel_residuals, void* stream);
/* After a tg_cg_solve / tg_cg_solve_op on `workspace` that reached its
 * iteration loop: host copies of sqrt|r.z| per column -- replaces
 * ConjugateGradient::Residuals (ConjugateGradient.hpp:447-456) -- and of
 * |b.Pb| (the member b_squared_norms, :45) into residuals[nrhs] /
 * b_squared_norms[nrhs] (either may be NULL).  Synchronises `stream`. */
int tg_cg_residuals(const void* workspace, size_t workspace_bytes, int32_t nrhs, double* residuals,
                    double* b_squared_norms, void* stream);
/* tg_cg_solve without any allocation.  No counterpart in the reference: it
 * serves callers that cannot allocate device memory (serving loops that must not
 * call hipMalloc between requests, harnesses that forbid allocation).  Same
 * arguments, kernels and results as tg_cg_solve, minus the workspace, which is a
 * static device array of TG_CG_NOALLOC_SCRATCH_BYTES (plus 4 KiB for the band
 * pass) that every GPU the library runs on holds until the process ends.  A system whose
 * tg_cg_workspace_bytes(n, nrhs, val_type, prec_kind) exceeds it is refused with
 * TG_ERR_WORKSPACE before any device work.  All calls on one GPU share that
 * GPU's array: calls that overlap on one GPU (several host threads, or streams
 * whose work overlaps) race.  Synchronises `stream`. */
#define TG_CG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tg_cg_solve_noalloc(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci, const void* a,
                        tg_dtype val_type, int32_t prec_kind, const void* prec_diag, int32_t nrhs, int32_t max_iter,
                        double a_coef, const void* B, int64_t ldB, double b_coef, void* X, int64_t ldX, double tol,
                        int32_t use_guess, int32_t* succeeded, int32_t* iterations, double* rel_residuals,
                        void* stream);

/* ------------------------------------------------------------------------
 * Mixed-precision CG: the recurrences of tg_cg_solve in FP32 (FP32 values a32
 * on the same pattern, FP32 Krylov vectors, FP64 reductions) with FP64
 * residual replacement: every 16 iterations, once the FP32 residual estimate
 * has fallen by the factor replace_drop (0 = 0.1) since the last replacement,
 * or when the FP32 iteration meets its own test, the FP32 correction is folded
 * into the FP64 solution and the residual is recomputed as b - A x in FP64; the
 * search direction is kept.  No reference counterpart (the reference CG is
 * precision-agnostic); it solves what tg_cg_solve solves:
 *
 *   X <- a_coef * A^{-1} B + b_coef * X      (FP64 in and out)
 *
 * Stops when every column has ||B - A x||_2 / ||B||_2 <= tol (true FP64
 * residual), after max_outer replacements or max_inner FP32 iterations.
 * a32 may be NULL (the FP64 values are rounded into the workspace; pass
 * with_a32 = 0 to the workspace query then).  rel_residuals[nrhs] are the final
 * true relative residuals; outer_iterations counts the replacements.
 * Synchronises `stream`.
 * ------------------------------------------------------------------------ */
size_t tg_cg_mixed_workspace_bytes(int32_t n, int64_t nnz, int32_t nrhs, int32_t prec_kind, int32_t with_a32);
int tg_cg_solve_mixed(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                      const double* a64, const float* a32, int32_t prec_kind, const double* prec_diag,
                      int32_t nrhs, int32_t max_outer, int32_t max_inner, double replace_drop, double a_coef,
                      const double* B, int64_t ldB, double b_coef, double* X, int64_t ldX, double tol,
                      int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                      int32_t* outer_iterations, int32_t* inner_iterations, double* rel_residuals, void* stream);

/* ------------------------------------------------------------------------
 * Mixed-precision restarted GMRES(m): every cycle's Arnoldi process runs in
 * FP32 (FP32 values a32 on the same pattern, FP32 basis and Krylov vectors)
 * while the Hessenberg matrix, Givens rotations, beta and least squares of the
 * cycle are FP64, and every restart recomputes the residual b - A x in FP64
 * with the solution x kept and updated in FP64 (iterative refinement, one
 * FP32 GMRES(m) cycle per refinement step).  No reference counterpart (the
 * reference GMRES is precision-agnostic); it solves what tg_gmres_solve solves
 * (GMRES.hpp:127-260):
 *
 *   X <- a_coef * A^{-1} B + b_coef * X      (FP64 in and out)
 *
 * Stops when every column has ||P (B - A x)|| <= tol ||P B|| (P = the diagonal
 * for side 0 with prec_kind 1, else the identity; the reference's test on the
 * true FP64 residual), or after max_restarts + 1 cycles.  a32 may be NULL (the
 * FP64 values are rounded into the workspace; pass with_a32 = 0 to the
 * workspace query then).  m = min(max_iter, n + 1) <= 63, nrhs <= 32.
 * *restarts counts the cycles after the first, *inner_iterations the FP32
 * Arnoldi steps over all cycles, rel_residuals[nrhs] the final true relative
 * residuals.  Synchronises `stream`.
 * ------------------------------------------------------------------------ */
size_t tg_gmres_mixed_workspace_bytes(int32_t n, int64_t nnz, int32_t max_iter, int32_t nrhs, int32_t prec_kind,
                                      int32_t with_a32);
int tg_gmres_solve_mixed(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                         const double* a64, const float* a32, int32_t prec_kind, const double* prec_diag,
                         int32_t side, int32_t nrhs, int32_t max_iter, int32_t max_restarts, double a_coef,
                         const double* B, int64_t ldB, double b_coef, double* X, int64_t ldX, double tol,
                         int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                         int32_t* restarts, int32_t* inner_iterations, double* rel_residuals, void* stream);

/* ------------------------------------------------------------------------
 * Restarted GMRES(m) -- replaces GMRES<NRHS,Scal,Int,Side>::operator()(A, P,
 * a, B, ldB, b, X, ldX, tol, max_restarts, use_initial_guessQ)
 * (GMRES.hpp:127-260) for a CSR operator and P = identity (prec_kind 0) or a
 * diagonal scaling (prec_kind 1); side 0 = Side::Left (default), 1 = Right.
 *
 *   X <- a_coef * A^{-1} B + b_coef * X        (n x nrhs blocks, row-major)
 *
 * m = min(max_iter, n + 1) (GMRES.hpp:89), m <= 63; nrhs <= 32.  Stops when
 * every column has |beta_iter|_k <= tol ||P b||_k; restarts while not
 * converged and restarts <= max_restarts.  Orthogonalisation is classical
 * Gram-Schmidt applied twice (the reference applies modified Gram-Schmidt
 * twice: same projection in exact arithmetic); Hessenberg, Givens rotations
 * and the triangular solve are FP64 on the device.  Special cases as the
 * reference: B == 0 -> X <- b X (zeroed if b == 0); NaN in B -> nothing
 * written.  Synchronises `stream`.  Host outputs: *succeeded, *iterations
 * (of the last cycle, as GMRES::IterationCount), *restarts
 * (GMRES::RestartCount) and rel_residuals[nrhs] = |beta_iter| / ||P b||.
 * ------------------------------------------------------------------------ */
size_t tg_gmres_workspace_bytes(int32_t n, int32_t max_iter, int32_t nrhs, tg_dtype val_type);
/* Generic-operator form (callables A and P with device pointers, apply_P
 * NULL = identity), as the reference's A / P contract (GMRES.hpp:143-151). */
int tg_gmres_solve_op(int32_t n, tg_apply_fn apply_A, void* ctx_A, tg_apply_fn apply_P, void* ctx_P,
                      tg_dtype val_type, int32_t side, int32_t nrhs, int32_t max_iter, double a_coef, const void* B,
                      int64_t ldB, double b_coef, void* X, int64_t ldX, double tol, int32_t max_restarts,
                      int32_t use_guess, void* workspace, size_t workspace_bytes, int32_t* succeeded,
                      int32_t* iterations, int32_t* restarts, double* rel_residuals, void* stream);
int tg_gmres_solve(int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci, const void* a,
                   tg_dtype val_type, int32_t prec_kind, const void* prec_diag, int32_t side, int32_t nrhs,
                   int32_t max_iter, double a_coef, const void* B, int64_t ldB, double b_coef, void* X,
                   int64_t ldX, double tol, int32_t max_restarts, int32_t use_guess, void* workspace,
                   size_t workspace_bytes, int32_t* succeeded, int32_t* iterations, int32_t* restarts,
                   double* rel_residuals, void* stream);
/* Device views into a GMRES workspace after a solve with the same n, max_iter,
 * nrhs and val_type -- the members behind GMRES::GetBasis, GetHessenbergMatrix,
 * Solution and Residuals (GMRES.hpp:779-789, 820-833).  m = min(max_iter, n+1)
 * (at least 1).  *Q: (m+1) x n x nrhs basis of the last cycle (val_type);
 * *H: (m+1) x m x nrhs FP64, the Hessenberg matrix after the Givens rotations,
 * entry (i, j, k) at (i*m + j)*nrhs + k, as the reference's Tensor3 H;
 * *beta: (m+1) x nrhs FP64 (|beta[iter][k]| = GMRES::Residuals()[k]);
 * *solution: n x nrhs (val_type) before the a/b combination.  Any output may
 * be NULL.  Host arithmetic only. */
int tg_gmres_workspace_parts(void* workspace, size_t workspace_bytes, int32_t n, int32_t max_iter, int32_t nrhs,
                             tg_dtype val_type, void** Q, double** H, double** beta, void** solution);

/* ------------------------------------------------------------------------
 * Structure work on the GPU (integer results are exact and deterministic).
 *
 * Triples -> CSR: replaces MatrixCSR::FromTriples (MatrixCSR/FromTriples.hpp:
 * 3-209) with SortInner (MatrixCSR/SortInner.hpp:3-49) and Compress
 * (PatternCSR/Compress.hpp:3-214): rows bucketed, each row sorted by column,
 * duplicates summed in input order when `compress`, (j,i,a) added for every
 * off-diagonal (i,j,a) when `symmetrize`.  V may be NULL (pattern only).
 * Outputs: rp_out[m+1] (rp_type), ci_out / val_out of `capacity` >=
 * count * (symmetrize ? 2 : 1) entries; *nnz_out (host) = the nonzero count.
 * Out-of-range indices -> TG_ERR_INVALID.  Synchronises `stream`.
 * ------------------------------------------------------------------------ */
size_t tg_csr_from_triples_workspace_bytes(int64_t count, int32_t m, int32_t symmetrize);
int tg_csr_from_triples(int64_t count, const int32_t* I, const int32_t* J, const void* V, tg_dtype val_type,
                        int32_t m, int32_t n, int32_t compress, int32_t symmetrize, void* rp_out, tg_dtype rp_type,
                        int32_t* ci_out, void* val_out, int64_t capacity, void* workspace, size_t workspace_bytes,
                        int64_t* nnz_out, void* stream);

/* Transpose: replaces MatrixCSR::Op<Op::Trans> / Transpose (MatrixCSR.hpp:
 * 389-495, CreateTransposeCounters PatternCSR.hpp:642-684).  out_rp[n+1],
 * out_ci / out_val[nnz] (val may be NULL).  Synchronises `stream`. */
size_t tg_csr_transpose_workspace_bytes(int32_t m, int32_t n, int64_t nnz);
int tg_csr_transpose(int32_t m, int32_t n, int64_t nnz, const void* rp, tg_dtype rp_type, const int32_t* ci,
                     const void* val, tg_dtype val_type, void* out_rp, int32_t* out_ci, void* out_val,
                     void* workspace, size_t workspace_bytes, void* stream);

/* Sparse x sparse C = A B (A m x k, B k x n): replaces MatrixCSR::Dot(MatrixCSR)
 * (MatrixCSR.hpp:835-921) and, with both value arrays NULL, the pattern product
 * PatternCSR::DotBinary_ (PatternCSR.hpp:731-839).  Step 1 gives the expanded
 * count E (workspace >= 2*align256(8(m+1)) + scan scratch; 1 MiB + 16(m+1)
 * bytes is enough); step 2 writes C with capacity E and returns its nnz.
 * Both synchronise `stream`. */
int tg_csr_spgemm_expanded_count(int32_t m, const void* Arp, const int32_t* Aci, const void* Brp, tg_dtype rp_type,
                                 int64_t* count_out, void* workspace, size_t workspace_bytes, void* stream);
size_t tg_csr_spgemm_workspace_bytes(int32_t m, int32_t n, int64_t expanded, tg_dtype val_type);
int tg_csr_spgemm(int32_t m, int32_t k, int32_t n, const void* Arp, const int32_t* Aci, const void* Aval,
                  const void* Brp, const int32_t* Bci, const void* Bval, tg_dtype rp_type, tg_dtype val_type,
                  int64_t expanded, void* Crp, int32_t* Cci, void* Cval, void* workspace, size_t workspace_bytes,
                  int64_t* nnz_out, void* stream);

/* ------------------------------------------------------------------------
 * Batched, continuously fed CG (no single reference counterpart: it runs many
 * ConjugateGradient::operator() solves, ConjugateGradient.hpp:105-353, at
 * once).  Each system: n x n CSR (int32 rp/ci, FP64 values, device), optional
 * Jacobi diagonal `diag` (z = diag * r), nrhs in 1..8, B and X n x nrhs
 * row-major device blocks (ld = nrhs), X <- A^{-1} B.
 *   mixed == 0: FP64 CG with the reference stopping rule |r.z| <= tol^2 |b.Pb|.
 *   mixed == 1: CG in FP32 (FP32 copy of A, FP32 vectors, FP64 reductions)
 *               with FP64 residual replacement: every 16 iterations the host
 *               may replace the recursive residual by b - A x computed in FP64
 *               (x accumulated in FP64), keeping the search direction; done
 *               when the true ||B - A X|| / ||B|| <= tol in every column
 *               (max_iter bounds the total iterations).  `outer` in
 *               tg_batch_result counts the replacements.
 *   mixed == 2: FP64 CG with the same residual replacement and the same
 *               true-residual stop as mode 1 (the FP64 comparison at equal
 *               accuracy).
 * All active systems advance together (three launches per CG iteration cover
 * every system's row tiles); a system's arithmetic does not depend on the
 * others, so its result is bitwise the same alone and in any mix.  Systems may be added
 * between steps; tg_batch_add_cg only enqueues the system's setup (it does not wait
 * for the GPU), and n * nrhs * 8 bytes must be below 4 GiB (the engine addresses its
 * vectors by 32-bit offsets; larger systems belong to tg_cg_solve).  The caller's
 * arrays must stay valid until the system has finished; tg_batch_release frees the
 * system's workspace.  tg_batch_step synchronises the batch's stream.
 * ------------------------------------------------------------------------ */
typedef struct tg_batch_s* tg_batch_t;
int tg_batch_create(tg_batch_t* out, void* stream);
int tg_batch_destroy(tg_batch_t batch);
int tg_batch_add_cg(tg_batch_t batch, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                    const double* val, const double* diag, int32_t nrhs, const double* B, double* X, double tol,
                    int32_t max_iter, int32_t mixed, int64_t* id_out);
/* Advances every active system by up to `iterations` CG iterations. */
int tg_batch_step(tg_batch_t batch, int32_t iterations);
/* Ids of systems finished since the last call (at most cap; *count = number). */
int tg_batch_finished(tg_batch_t batch, int64_t* ids, int32_t cap, int32_t* count);
int tg_batch_result(tg_batch_t batch, int64_t id, int32_t* finished, int32_t* succeeded, int32_t* iterations,
                    int32_t* outer, double* rel_residuals);
int tg_batch_release(tg_batch_t batch, int64_t id);
int tg_batch_active(tg_batch_t batch, int32_t* systems, int64_t* bytes_in_use);
/* Device bytes the engine itself allocates for one such system. */
size_t tg_batch_request_bytes(int32_t n, int64_t nnz, int32_t nrhs, int32_t has_diag, int32_t mixed);

/* ------------------------------------------------------------------------
 * Batched, continuously fed GMRES(m) (no single reference counterpart: it runs
 * many GMRES::operator() solves, GMRES.hpp:127-260, at once, with a = 1, b = 0,
 * no initial guess, Side::Left).  Each system: n x n CSR (int32 rp/ci, FP64
 * values, device), optional left diagonal preconditioner `inv_diag` (z = d r),
 * nrhs in 1..8, restart length m = min(restart, n + 1) <= 63, B and X n x nrhs
 * row-major device blocks (ld = nrhs); X <- A^{-1} B.  The same recurrences,
 * stop test (|beta_iter|_k <= tol ||P b||_k) and restart rule as tg_gmres_solve
 * (classical Gram-Schmidt applied twice; Hessenberg, Givens and back
 * substitution in FP64 on the device).
 * All active systems run restart cycles in lock step: a cycle is 4 launches per
 * Arnoldi step plus 5 for ALL systems (captured as one hipGraph), and a
 * system's arithmetic does not depend on the others, so its result is bitwise
 * the same alone and in any mix.  tg_gbatch_add only enqueues the system's setup
 * (no host wait); the system joins at the next cycle.  tg_gbatch_step runs up to
 * `cycles` cycles and synchronises the batch's stream.  Results: succeeded,
 * iterations of the last cycle, restarts (cycles after the first) and
 * rel_residuals[nrhs] = |beta_iter| / ||P b||.  A NaN right-hand side finishes
 * at once with nothing written; an all-zero one with X = 0, succeeded.  The
 * caller's arrays must stay valid until the system has finished.
 * ------------------------------------------------------------------------ */
typedef struct tg_gbatch_s* tg_gbatch_t;
int tg_gbatch_create(tg_gbatch_t* out, void* stream);
int tg_gbatch_destroy(tg_gbatch_t batch);
int tg_gbatch_add(tg_gbatch_t batch, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const double* val,
                  const double* inv_diag, int32_t nrhs, const double* B, double* X, double tol, int32_t restart,
                  int32_t max_restarts, int64_t* id_out);
int tg_gbatch_step(tg_gbatch_t batch, int32_t cycles);
int tg_gbatch_finished(tg_gbatch_t batch, int64_t* ids, int32_t cap, int32_t* count);
int tg_gbatch_result(tg_gbatch_t batch, int64_t id, int32_t* finished, int32_t* succeeded, int32_t* iterations,
                     int32_t* restarts, double* rel_residuals);
int tg_gbatch_release(tg_gbatch_t batch, int64_t id);
int tg_gbatch_active(tg_gbatch_t batch, int32_t* systems, int64_t* bytes_in_use);
/* Device bytes the engine allocates for one such system. */
size_t tg_gbatch_request_bytes(int32_t n, int32_t nrhs, int32_t restart);

/* ------------------------------------------------------------------------
 * Serving loop (no reference counterpart).  Requests carry HOST arrays (int32
 * rp/ci, FP64 values, B and X n x nrhs row-major with ld = nrhs); the caller
 * keeps them valid until the request has finished.  A request is admitted, in
 * arrival order, once its device bytes (matrix, vectors, solver workspace) fit
 * in the budget next to the requests already admitted; it is then uploaded and
 * solved while other requests run:
 *   CG requests join the batched CG engine (tg_batch_*), which advances all of
 *   them together on one stream (Jacobi `inv_diag` optional, mixed as there);
 *   GMRES(restart) requests (P = I, left side) of up to serve_gmres_batch_rows
 *   rows join one batched GMRES engine (tg_gbatch_*) advanced by a thread of its
 *   own; larger ones (and all, with serve_gmres_batch = 0) run
 *   tg_gmres_solve_mixed (tg_gmres_solve with serve_gmres_mixed = 0) on one of
 *   `gmres_workers` host threads, each with its own stream.
 * On completion X is downloaded and all of the request's device memory freed.
 * Results do not depend on what else is in flight.  Times are seconds since
 * tg_server_create on the host steady clock.
 * ------------------------------------------------------------------------ */
typedef struct tg_server_s* tg_server_t;
int tg_server_create(int64_t budget_bytes, int32_t gmres_workers, int32_t cg_iterations_per_poll, tg_server_t* out);
/* Waits for every admitted and queued request, then frees the server. */
int tg_server_destroy(tg_server_t server);
int tg_server_submit_cg(tg_server_t server, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                        const double* val, const double* inv_diag, int32_t nrhs, const double* B, double* X,
                        double tol, int32_t max_iter, int32_t mixed, int64_t* id_out);
int tg_server_submit_gmres(tg_server_t server, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                           const double* val, int32_t nrhs, const double* B, double* X, double tol, int32_t restart,
                           int32_t max_restarts, int64_t* id_out);
/* Ids finished since the last call (at most cap), waiting up to timeout_s for one. */
int tg_server_wait(tg_server_t server, int64_t* ids, int32_t cap, int32_t* count, double timeout_s);
/* status: tg_status of the request; times[3] = submit, admit, done. */
int tg_server_result(tg_server_t server, int64_t id, int32_t* status, int32_t* succeeded, int32_t* iterations,
                     int32_t* restarts, double* rel_residuals, double* times);
int tg_server_release(tg_server_t server, int64_t id);
int tg_server_stats(tg_server_t server, int64_t* bytes_in_use, int64_t* peak_bytes, int32_t* queued,
                    int32_t* running);
/* Bytes a request submitted now is charged against the budget (mixed: the CG
 * request's batched-engine mode; GMRES follows the serve_gmres_* knobs). */
size_t tg_server_request_bytes(int32_t kind, int32_t n, int64_t nnz, int32_t nrhs, int32_t has_diag, int32_t mixed,
                               int32_t restart);

/* Streaming probes (no reference counterpart): dst <- src over `bytes`
 * (multiple of 16), and a read-only pass over src (sink: one device float that
 * is practically never written).  They measure the achievable HBM bandwidth
 * that the SpMM roofline fraction is quoted against. */
int tg_probe_copy(const void* src, void* dst, int64_t bytes, void* stream);
int tg_probe_read(const void* src, int64_t bytes, float* sink, void* stream);

/* ------------------------------------------------------------------------
 * Synthetic workload generators (HOST buffers; no device work).  No
 * reference counterpart: the reference draws its test inputs from
 * std::random_device (src/Tensor_Common.hpp:141-211), which cannot be
 * reproduced; these are counter-based (SplitMix64 of seed, stream, index) so
 * every run and every thread count yields the same inputs.  Row pointers are
 * int64, column indices int32.  Size-returning functions return the nonzero
 * count or a negative tg_status.
 * ------------------------------------------------------------------------ */
int tg_gen_uniform(uint64_t seed, uint64_t offset, int64_t count, tg_dtype dtype, void* host_out);
int64_t tg_gen_random_csr_rowptr(int64_t m, double mean, int32_t lo, int32_t hi, uint64_t seed, int64_t* rp);
int tg_gen_random_csr_fill(int64_t m, int64_t n, const int64_t* rp, uint64_t seed, int32_t* ci, double* val64,
                           float* val32);
int64_t tg_gen_laplacian_3d(int32_t g, double shift, int64_t* rp, int32_t* ci, double* val);
int64_t tg_gen_weighted_grid(int32_t dims, int32_t s0, int32_t s1, int32_t s2, double shift, uint64_t seed,
                             int64_t* rp, int32_t* ci, double* val);
int64_t tg_gen_convection_diffusion_2d(int32_t g, double peclet, double theta, int64_t* rp, int32_t* ci,
                                       double* val);

/* ------------------------------------------------------------------------
 * Block-CSR SpMM: replaces Sparse::KernelMatrixCSR<Kernel_T>::Dot
 * (KernelMatrixCSR.hpp:153-227) with the dense block kernels
 * (BlockKernels/DenseBlockKernel_Tiny.hpp:160-180, BlockKernel_Tiny.hpp:156-196).
 * A pattern of mb x nb blocks (rp[mb+1], ci[nnzb]) with one R x C block per
 * nonzero (a + k*R*C; row-major if a_row_major, else column-major).  For every
 * block row i:  Y_i <- alpha * sum_k A_k X_{ci[k]} + beta * Y_i, where X_j is
 * the C x nrhs chunk at X + j*C*nrhs and Y_i the R x nrhs chunk at
 * Y + i*R*nrhs, each row-major (x_row_major / y_row_major: entry (c, h) at
 * c*nrhs + h) or transposed (h*C + c).  Empty block rows: Y_i <- beta Y_i.
 * alpha == 0 or nnzb == 0: Y <- beta Y (beta == 0 writes zeros).  X == Y ->
 * TG_ERR_ALIAS.  Block sizes 1..64.  Device pointers; stream-ordered.
 *
 * tg_bsr_fill_lower: FillLowerTriangleFromUpperTriangle (KernelMatrixCSR.hpp:
 * 71-128) for square R x R blocks and sorted block columns: every block
 * (i, j), j < i, becomes the transpose of block (j, i) (left as it is where
 * the pattern has no (j, i)).
 * ------------------------------------------------------------------------ */
int tg_bsr_spmm(int32_t mb, int32_t nb, int64_t nnzb, int32_t R, int32_t C, const void* rp, tg_dtype rp_type,
                const int32_t* ci, const void* a, tg_dtype val_type, int32_t a_row_major, double alpha,
                const void* X, int32_t x_row_major, double beta, void* Y, int32_t y_row_major, int32_t nrhs,
                void* stream);
int tg_bsr_fill_lower(int32_t mb, const void* rp, tg_dtype rp_type, const int32_t* ci, int32_t R, void* a,
                      tg_dtype val_type, void* stream);

/* ------------------------------------------------------------------------
 * MatrixMarket coordinate files (HOST buffers; no device work).  Replace
 * MatrixCSR::LoadFromMatrixMarket / WriteToMatrixMarket
 * (MatrixCSR/MatrixMarket.hpp:3-306); loading = tg_mm_info + tg_mm_read, then
 * tg_csr_from_triples with compress = 1 and symmetrize = the file's symmetric
 * flag (the reference's MatrixCSR(nnz, i, j, a, m, n, threads, true,
 * symmetrizeQ), :190-196).
 *
 * tg_mm_info: banner and size line.  Field "complex" -> TG_ERR_DTYPE;
 * symmetry other than general / symmetric -> TG_ERR_INVALID.
 * tg_mm_read: the first `entries` (= the announced count) entry lines, in
 * file order, as zero-based I, J (int32) and V (double; 1 for "pattern"
 * files; V may be NULL).  A malformed line or an index outside [1, m] x
 * [1, n] -> TG_ERR_INVALID naming the entry.  threads <= 0: all cores.
 * tg_mm_write: "coordinate real general" (or "pattern" when vals is NULL),
 * one "i j value" line per nonzero in row order, %.17E for F64, %.8E for F32
 * (enough digits that every float reads back exactly).  rp / ci / vals are
 * host arrays.
 * ------------------------------------------------------------------------ */
typedef enum tg_mm_field { TG_MM_REAL = 0, TG_MM_INTEGER = 1, TG_MM_PATTERN = 2 } tg_mm_field;
int tg_mm_info(const char* path, int32_t* m, int32_t* n, int64_t* entries, int32_t* field, int32_t* symmetric);
int tg_mm_read(const char* path, int64_t entries, int32_t* I, int32_t* J, double* V, int32_t threads);
int tg_mm_write(const char* path, int32_t m, int32_t n, const void* rp, tg_dtype rp_type, const int32_t* ci,
                const void* vals, tg_dtype val_type, int32_t threads);

#ifdef __cplusplus
}
#endif

#endif /* TENSORS_AMD_H */
