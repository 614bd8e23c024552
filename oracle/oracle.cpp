// =============================================================================
// oracle.cpp -- CPU restatement of the Tensors sparse hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load the library built from it (liboracle.so for parity checks,
// liboracle_fast.so -- the same source under the reference's release flags --
// for the timed CPU baseline).  Nothing under tensors_amd/ or include/ links,
// loads or calls it.
//
// Every function restates the reference algorithm it cites (paths relative to
// the HenrikSchumacher/Tensors tree).  The reference itself cannot be compiled
// here: Base.hpp:10 includes submodules/Tools/Tools.hpp and that submodule is
// empty (.gitmodules:1-3, no pinned commit), so the Tools pieces the path uses
// (JobPointers, combine_buffers, Sort, ParallelDo) are restated from their
// call sites.  Pinning of each function (see tests/test_oracle_*.py):
//   job pointers (uniform / cost-balanced)   invariants + brute force
//   spmv / spmm + alpha/beta flag epilogue   dense brute force (numpy),
//                                            exact-integer cases, flag table
//   grid_laplacian (+ parallel assembly)     closed form 5-point graph
//                                            Laplacian; sequential == parallel
//                                            (Experiment_SpMM/main.cpp:55-57)
//   from_triples (sort, compress, symmetrize) scipy.sparse coo->csr
//   transpose / spgemm / pattern spgemm      scipy.sparse, boolean products
//   conjugate_gradient                       dense solve (numpy), finite
//                                            termination on diagonal systems
//   gmres (left/right, restarts)             dense solve; GMRES(n) exactness
//   upper back substitution                  numpy triangular solve
//   block CSR product (all storage orders)   the scalar CSR of the expanded
//                                            blocks (spmm above), dense numpy
//   fill lower blocks from upper             dense symmetric closed form
// Bit-level float parity with the reference is unpinned everywhere (Tools'
// reduction splits, Sort stability and -ffast-math evaluation order are not
// available); structure (rp, ci) is pinned bit-exact.
// =============================================================================

#include <pthread.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <thread>
#include <vector>

namespace orc {

using i32 = int32_t;
using i64 = int64_t;

// ---------------------------------------------------------------------------
// Fork-join over `threads` workers: the role of Tools' ParallelDo
// (used e.g. at src/SparseBLAS/SpMM.hpp:286-414).
// ---------------------------------------------------------------------------
// With pinning on (orc_set_pinning, used by the timed CPU baseline) worker t runs
// on the t-th CPU of the process's affinity mask, taken once when pinning is
// switched on (a pinned thread's own mask is one CPU, and threads it starts would
// inherit that); the calling thread gets its own mask back after the join.  The
// results do not depend on it.
static std::atomic<int> g_pin{0};
static cpu_set_t g_proc_mask;
static std::atomic<int> g_proc_cpus{0};

static void take_process_mask() {
    cpu_set_t mask;
    CPU_ZERO(&mask);
    if (sched_getaffinity(0, sizeof(mask), &mask) != 0) return;
    g_proc_mask = mask;
    g_proc_cpus.store(CPU_COUNT(&mask));
}

static void pin_to(int t) {
    const int cpus = g_proc_cpus.load();
    if (cpus <= 0) return;
    int want = t % cpus, seen = 0;
    for (int c = 0; c < CPU_SETSIZE; ++c)
        if (CPU_ISSET(c, &g_proc_mask) && seen++ == want) {
            cpu_set_t one;
            CPU_ZERO(&one);
            CPU_SET(c, &one);
            (void)pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
            return;
        }
}

template <class F>
static void parallel_do(int threads, F&& f) {
    if (threads <= 1) {
        f(0);
        return;
    }
    const bool pin = g_pin.load() != 0;
    cpu_set_t caller;
    const bool restore = pin && pthread_getaffinity_np(pthread_self(), sizeof(caller), &caller) == 0;
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    for (int t = 1; t < threads; ++t)
        pool.emplace_back([&f, t, pin] {
            if (pin) pin_to(t);
            f(t);
        });
    if (pin) pin_to(0);
    f(0);
    for (auto& th : pool) th.join();
    if (restore) (void)pthread_setaffinity_np(pthread_self(), sizeof(caller), &caller);
}

// Uniform split of [0,n) into T blocks: the formula AssemblyCounters.hpp:87-95
// uses for its own chunking, which Tools' JobPointer(n, T, t) shares.
static i64 job_pointer(i64 n, i64 T, i64 t) {
    const i64 step = n / T, corr = n % T;
    return step * t + (corr * t) / T;
}

// Cost-balanced split (Tools' JobPointers(m, acc_costs, T, false), called at
// src/Sparse/PatternCSR.hpp:395): rows [0,m) cut where the accumulated cost
// first reaches t/T of the total.
static std::vector<i64> job_pointers_costs(i64 m, const i64* acc, i64 T) {
    std::vector<i64> jp(T + 1, 0);
    const i64 total = acc[m] - acc[0];
    for (i64 t = 1; t < T; ++t) {
        const i64 target = acc[0] + (i64)(((__int128)total * t) / T);
        i64 i = std::lower_bound(acc, acc + m + 1, target) - acc;
        jp[t] = std::min(std::max(i, jp[t - 1]), m);
    }
    jp[T] = m;
    return jp;
}

// ---------------------------------------------------------------------------
// alpha/beta flag semantics (Scalar::Flag, SpMM.hpp:28-131): Zero means the
// term is dropped without reading memory; Plus means multiplication by one is
// skipped; Generic multiplies.
// ---------------------------------------------------------------------------
enum Flag { Zero = 0, Plus = 1, Generic = 2 };

template <class T>
static inline T combine(Flag af, T alpha, T z, Flag bf, T beta, T y) {
    // combine_scalars / combine_buffers: y = (alpha * z) + (beta * y)
    // (SpMV.hpp:224-225, UnitTests/main.cpp:23-34).
    const T az = (af == Plus) ? z : alpha * z;
    if (bf == Zero) return az;
    if (bf == Plus) return az + y;
    return az + beta * y;
}

enum Status { OK = 0, ERR_INVALID = -1, ERR_ALIAS = -2, ERR_NRHS = -3 };

// SparseBLAS::SpMV (SpMV.hpp:3-131) + SpMV_impl (SpMV.hpp:135-288).
template <class T>
static int spmv(const i64* rp, const i32* ci, const T* a, i32 m, T alpha, const T* x, T beta, T* y,
                const std::vector<i64>& jp) {
    const int threads = (int)jp.size() - 1;
    if (alpha == T(0)) {  // SpMV.hpp:15-30
        if (beta == T(0))
            std::fill(y, y + m, T(0));
        else if (beta != T(1))
            for (i32 i = 0; i < m; ++i) y[i] *= beta;
        return OK;
    }
    if ((const void*)x == (const void*)y) return ERR_ALIAS;  // SpMV.hpp:163-170
    const Flag af = (alpha == T(1)) ? Plus : Generic;
    const Flag bf = (beta == T(0)) ? Zero : (beta == T(1)) ? Plus : Generic;
    parallel_do(threads, [&](int t) {
        for (i64 i = jp[t]; i < jp[t + 1]; ++i) {
            T z = T(0);  // SpMV.hpp:261
            for (i64 l = rp[i]; l < rp[i + 1]; ++l) z += (a ? a[l] : T(1)) * x[ci[l]];
            y[i] = combine(af, alpha, z, bf, beta, y[i]);  // SpMV.hpp:283
        }
    });
    return OK;
}

// SparseBLAS::SpMM (SpMM.hpp:3-131) + SpMM_impl (SpMM.hpp:134-415).  K > 0 is
// the fixed-NRHS instantiation that SpMM_vec (SpMM.hpp:419-569) provides, with
// the same summation order; K == 0 is the runtime-nrhs path.
template <class T, int K>
static int spmm(const i64* rp, const i32* ci, const T* a, i32 m, T alpha, const T* X, i64 ldX, T beta,
                T* Y, i64 ldY, i32 nrhs_rt, const std::vector<i64>& jp) {
    const int threads = (int)jp.size() - 1;
    const i32 nrhs = K > 0 ? K : nrhs_rt;
    if (alpha == T(0)) {  // modify_matrix_auto, SpMM.hpp:18-25
        if (beta == T(0)) {
            for (i32 i = 0; i < m; ++i) std::fill(Y + ldY * i, Y + ldY * i + nrhs, T(0));
        } else if (beta != T(1)) {
            for (i32 i = 0; i < m; ++i)
                for (i32 k = 0; k < nrhs; ++k) Y[ldY * i + k] *= beta;
        }
        return OK;
    }
    if ((const void*)X == (const void*)Y) return ERR_ALIAS;  // SpMM.hpp:191-198
    const Flag af = (alpha == T(1)) ? Plus : Generic;
    const Flag bf = (beta == T(0)) ? Zero : (beta == T(1)) ? Plus : Generic;
    parallel_do(threads, [&](int t) {
        std::vector<T> zbuf(K > 0 ? 1 : nrhs);
        T zfix[K > 0 ? K : 1];
        T* z = K > 0 ? zfix : zbuf.data();
        for (i64 i = jp[t]; i < jp[t + 1]; ++i) {
            T* y = Y + ldY * i;
            const i64 lb = rp[i], le = rp[i + 1];
            if (le > lb) {
                {  // first entry overwrites z (SpMM.hpp:322-352)
                    const T av = a ? a[lb] : T(1);
                    const T* x = X + ldX * (i64)ci[lb];
                    for (i32 k = 0; k < nrhs; ++k) z[k] = av * x[k];
                }
                for (i64 l = lb + 1; l < le; ++l) {  // add-in (SpMM.hpp:355-386)
                    const T av = a ? a[l] : T(1);
                    const T* x = X + ldX * (i64)ci[l];
                    for (i32 k = 0; k < nrhs; ++k) z[k] += av * x[k];
                }
                for (i32 k = 0; k < nrhs; ++k) y[k] = combine(af, alpha, z[k], bf, beta, y[k]);
            } else {  // empty row (SpMM.hpp:395-410)
                if (bf == Zero)
                    for (i32 k = 0; k < nrhs; ++k) y[k] = T(0);
                else if (bf == Generic)
                    for (i32 k = 0; k < nrhs; ++k) y[k] *= beta;
            }
        }
    });
    return OK;
}

// SparseBLAS::Multiply_DenseMatrix (src/SparseBLAS/Multiply_DenseMatrix.hpp:9-45)
// with the job pointers PatternCSR::RequireJobPtr builds (PatternCSR.hpp:389-398).
template <class T>
static int multiply_dense(const i64* rp, const i32* ci, const T* a, i32 m, i32 n, T alpha, const T* X,
                          i64 ldX, T beta, T* Y, i64 ldY, i32 nrhs, i32 NRHS, i32 threads) {
    (void)n;
    if (m < 0 || nrhs < 0 || threads < 1) return ERR_INVALID;
    if (NRHS > 0 && nrhs != NRHS) return ERR_NRHS;  // :19-26, "Doing nothing."
    std::vector<i64> jp = job_pointers_costs(m, rp, threads);
    if (nrhs == 1 && ldX == 1 && ldY == 1) return spmv(rp, ci, a, m, alpha, X, beta, Y, jp);
    switch (NRHS) {
        case 1: return spmm<T, 1>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
        case 2: return spmm<T, 2>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
        case 4: return spmm<T, 4>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
        case 8: return spmm<T, 8>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
        case 16: return spmm<T, 16>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
        case 32: return spmm<T, 32>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
        case 64: return spmm<T, 64>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
        default: return spmm<T, 0>(rp, ci, a, m, alpha, X, ldX, beta, Y, ldY, nrhs, jp);
    }
}

// ---------------------------------------------------------------------------
// Structure work.
// ---------------------------------------------------------------------------

// AccumulateAssemblyCounters (AssemblyCounters.hpp:36-157): 2D prefix sum over
// (row, list) in row-major-by-row order; after it counters[t][i] is one past
// the last slot list t may write for row i.  The parallel chunked form
// produces the same numbers as the sequential loop (AssemblyCounters.hpp:63-75).
static void accumulate_counters(std::vector<std::vector<i64>>& c, i64 m) {
    const i64 T = (i64)c.size();
    if (m <= 0 || T <= 0) return;
    for (i64 t = 1; t < T; ++t) c[t][0] += c[t - 1][0];
    for (i64 i = 1; i < m; ++i) {
        c[0][i] += c[T - 1][i - 1];
        for (i64 t = 1; t < T; ++t) c[t][i] += c[t - 1][i];
    }
}

struct CSR {
    i64 m = 0, n = 0;
    std::vector<i64> rp;
    std::vector<i32> ci;
    std::vector<double> val;
};

// PatternCSR::WellFormedQ (PatternCSR.hpp:1146-1167).
static bool well_formed(const CSR& A) { return A.rp.size() > 1 && A.rp.back() > 0; }

// SortInner_impl (MatrixCSR/SortInner.hpp:3-49) with TwoArraySort
// (Sorting/TwoArraySort.hpp:28-60): per row, order (column, value) pairs by
// column.  A stable sort is used: Tools::Sort's tie order is unpinned.
static void sort_inner(CSR& A, bool values) {
    if (!well_formed(A)) return;
    std::vector<i64> perm;
    std::vector<i32> kc;
    std::vector<double> kv;
    for (i64 i = 0; i < A.m; ++i) {
        const i64 b = A.rp[i], e = A.rp[i + 1], len = e - b;
        if (len <= 1) continue;
        perm.resize(len);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(),
                         [&](i64 p, i64 q) { return A.ci[b + p] < A.ci[b + q]; });
        kc.assign(A.ci.begin() + b, A.ci.begin() + e);
        if (values) kv.assign(A.val.begin() + b, A.val.begin() + e);
        for (i64 k = 0; k < len; ++k) {
            A.ci[b + k] = kc[perm[k]];
            if (values) A.val[b + k] = kv[perm[k]];
        }
    }
}

// Compress_impl (PatternCSR/Compress.hpp:3-214): merge runs of equal column
// indices inside each (sorted) row, summing their values in order.
static void compress(CSR& A, bool values) {
    if (!well_formed(A)) return;
    sort_inner(A, values);
    std::vector<i64> rp(A.m + 1, 0);
    std::vector<i32> ci;
    std::vector<double> val;
    ci.reserve(A.ci.size());
    if (values) val.reserve(A.val.size());
    for (i64 i = 0; i < A.m; ++i) {
        i64 k = A.rp[i];
        const i64 ke = A.rp[i + 1];
        i64 cnt = 0;
        while (k < ke) {
            const i32 j = A.ci[k];
            double s = values ? A.val[k] : 0.0;
            ++k;
            while (k < ke && A.ci[k] == j) {
                if (values) s += A.val[k];
                ++k;
            }
            ci.push_back(j);
            if (values) val.push_back(s);
            ++cnt;
        }
        rp[i + 1] = rp[i] + cnt;
    }
    A.rp.swap(rp);
    A.ci.swap(ci);
    if (values) A.val.swap(val);
}

// MatrixCSR::FromTriples (MatrixCSR/FromTriples.hpp:3-209) with
// AssemblyCounters (AssemblyCounters.hpp:162-225).
static CSR from_triples(const std::vector<const i32*>& I, const std::vector<const i32*>& J,
                        const std::vector<const double*>& V, const std::vector<i64>& counts, i64 m,
                        i64 n, bool compressQ, bool symmetrize) {
    CSR A;
    A.m = m;
    A.n = n;
    A.rp.assign(m + 1, 0);
    const i64 L = (i64)counts.size();
    i64 triple_count = 0;
    for (i64 c : counts) triple_count += c;
    if (triple_count <= 0 || L <= 0 || m <= 0) return A;
    std::vector<std::vector<i64>> c(L, std::vector<i64>(m, 0));
    for (i64 t = 0; t < L; ++t)
        for (i64 k = 0; k < counts[t]; ++k) {
            const i32 i = I[t][k], j = J[t][k];
            ++c[t][i];
            if (symmetrize && i != j) ++c[t][j];
        }
    accumulate_counters(c, m);
    const i64 nnz = c[L - 1][m - 1];
    if (nnz <= 0) return A;
    A.ci.assign(nnz, 0);
    A.val.assign(nnz, 0.0);
    for (i64 i = 0; i < m; ++i) A.rp[i + 1] = c[L - 1][i];
    // Reverse scatter keeps each row's entries in input order (:89,114).
    for (i64 t = 0; t < L; ++t)
        for (i64 k = counts[t]; k-- > 0;) {
            const i32 i = I[t][k], j = J[t][k];
            const double a = V[t][k];
            i64 pos = --c[t][i];
            A.ci[pos] = j;
            A.val[pos] = a;
            if (symmetrize && i != j) {  // :131-145
                pos = --c[t][j];
                A.ci[pos] = i;
                A.val[pos] = a;
            }
        }
    sort_inner(A, true);
    if (compressQ) compress(A, true);
    return A;
}

// GridLaplacian (src/Sparse/GridLaplacian.hpp:11-77): graph Laplacian plus
// mass on a g x g grid, vertex V = g*i + j, assembled from one triple list.
static CSR grid_laplacian(i32 g, double mass) {
    std::vector<i32> I, J;
    std::vector<double> V;
    auto push = [&](i32 i, i32 j, double v) {
        I.push_back(i);
        J.push_back(j);
        V.push_back(v);
    };
    for (i32 i = 0; i < g - 1; ++i)
        for (i32 j = 0; j < g - 1; ++j) {
            const i32 v = g * i + j, l = v + g, r = v + 1;
            push(v, v, 1);
            push(v, r, -1);
            push(r, v, -1);
            push(r, r, 1);
            push(v, v, 1);
            push(v, l, -1);
            push(l, v, -1);
            push(l, l, 1);
        }
    for (i32 i = 0; i < g - 1; ++i) {
        const i32 v = g * i + (g - 1), l = v + g;
        push(v, v, 1);
        push(v, l, -1);
        push(l, v, -1);
        push(l, l, 1);
    }
    for (i32 j = 0; j < g - 1; ++j) {
        const i32 v = g * (g - 1) + j, r = v + 1;
        push(v, v, 1);
        push(v, r, -1);
        push(r, v, -1);
        push(r, r, 1);
    }
    for (i32 i = 0; i < g; ++i)
        for (i32 j = 0; j < g; ++j) push(g * i + j, g * i + j, mass);
    return from_triples({I.data()}, {J.data()}, {V.data()}, {(i64)I.size()}, (i64)g * g, (i64)g * g,
                        true, false);
}

// GridLaplacian_Parallel (GridLaplacian.hpp:85-158): T triple lists holding the
// upper triangle only, assembled with symmetrizeQ = true.
static CSR grid_laplacian_parallel(i32 g, double mass, i32 T) {
    std::vector<std::vector<i32>> I(T), J(T);
    std::vector<std::vector<double>> V(T);
    for (i32 t = 0; t < T; ++t) {
        auto push = [&](i32 i, i32 j, double v) {
            I[t].push_back(i);
            J[t].push_back(j);
            V[t].push_back(v);
        };
        const i32 ib = (i32)job_pointer(g - 1, T, t), ie = (i32)job_pointer(g - 1, T, t + 1);
        for (i32 i = ib; i < ie; ++i) {
            for (i32 j = 0; j < g - 1; ++j) {
                const i32 v = g * i + j, l = v + g, r = v + 1;
                push(v, v, 2.0 + mass);
                push(v, r, -1.0);
                push(r, r, 1.0);
                push(v, l, -1.0);
                push(l, l, 1.0);
            }
            const i32 v = g * i + (g - 1), l = v + g;
            push(v, v, 1.0 + mass);
            push(v, l, -1.0);
            push(l, l, 1.0);
        }
        if (t + 1 == T) {
            for (i32 j = 0; j < g - 1; ++j) {
                const i32 v = g * (g - 1) + j, r = v + 1;
                push(v, v, 1.0 + mass);
                push(v, r, -1.0);
                push(r, r, 1.0);
            }
            push(g * g - 1, g * g - 1, mass);
        }
    }
    std::vector<const i32*> pi, pj;
    std::vector<const double*> pv;
    std::vector<i64> cnt;
    for (i32 t = 0; t < T; ++t) {
        pi.push_back(I[t].data());
        pj.push_back(J[t].data());
        pv.push_back(V[t].data());
        cnt.push_back((i64)I[t].size());
    }
    return from_triples(pi, pj, pv, cnt, (i64)g * g, (i64)g * g, true, true);
}

// MatrixCSR::Op<Trans> (MatrixCSR.hpp:445-490) with CreateTransposeCounters
// (PatternCSR.hpp:642-684): counting sort over column indices, rows visited in
// reverse inside each thread block, then SortInner.
static CSR transpose(const CSR& A, i32 T, bool values) {
    CSR B;
    B.m = A.n;
    B.n = A.m;
    B.rp.assign(A.n + 1, 0);
    if (!well_formed(A)) return B;
    std::vector<i64> jp = job_pointers_costs(A.m, A.rp.data(), T);
    std::vector<std::vector<i64>> c(T, std::vector<i64>(A.n, 0));
    for (i32 t = 0; t < T; ++t)
        for (i64 i = jp[t]; i < jp[t + 1]; ++i)
            for (i64 k = A.rp[i]; k < A.rp[i + 1]; ++k) ++c[t][A.ci[k]];
    accumulate_counters(c, A.n);
    const i64 nnz = A.rp[A.m];
    B.ci.assign(nnz, 0);
    if (values) B.val.assign(nnz, 0.0);
    for (i64 j = 0; j < A.n; ++j) B.rp[j + 1] = c[T - 1][j];
    for (i32 t = 0; t < T; ++t)
        for (i64 i = jp[t]; i < jp[t + 1]; ++i)
            for (i64 k = A.rp[i + 1]; k-- > A.rp[i];) {
                const i64 pos = --c[t][A.ci[k]];
                B.ci[pos] = (i32)i;
                if (values) B.val[pos] = A.val[k];
            }
    sort_inner(B, values);
    return B;
}

// MatrixCSR::Dot(MatrixCSR) (MatrixCSR.hpp:835-921) and the pattern-only
// PatternCSR::DotBinary_ (PatternCSR.hpp:731-839): expansion by counting sort
// (B's row entries visited in reverse), then Compress.
static CSR spgemm(const CSR& A, const CSR& B, i32 T, bool values) {
    CSR C;
    C.m = A.m;
    C.n = B.n;
    C.rp.assign(A.m + 1, 0);
    if (!well_formed(A) || !well_formed(B)) return C;
    std::vector<i64> jp = job_pointers_costs(A.m, A.rp.data(), T);
    std::vector<std::vector<i64>> c(T, std::vector<i64>(A.m, 0));
    for (i32 t = 0; t < T; ++t)
        for (i64 i = jp[t]; i < jp[t + 1]; ++i) {
            i64 s = 0;
            for (i64 jj = A.rp[i]; jj < A.rp[i + 1]; ++jj) {
                const i32 j = A.ci[jj];
                s += B.rp[j + 1] - B.rp[j];
            }
            c[t][i] = s;
        }
    accumulate_counters(c, A.m);
    const i64 nnz = c[T - 1][A.m - 1];
    for (i64 i = 0; i < A.m; ++i) C.rp[i + 1] = c[T - 1][i];
    C.ci.assign(nnz, 0);
    if (values) C.val.assign(nnz, 0.0);
    for (i32 t = 0; t < T; ++t)
        for (i64 i = jp[t]; i < jp[t + 1]; ++i)
            for (i64 jj = A.rp[i]; jj < A.rp[i + 1]; ++jj) {
                const i32 j = A.ci[jj];
                for (i64 kk = B.rp[j + 1]; kk-- > B.rp[j];) {
                    const i64 pos = --c[t][i];
                    C.ci[pos] = B.ci[kk];
                    if (values) C.val[pos] = A.val[jj] * B.val[kk];
                }
            }
    compress(C, values);
    return C;
}

// ---------------------------------------------------------------------------
// Block CSR: KernelMatrixCSR<DenseBlockKernel_Tiny<ROWS, COLS, NRHS, ...>>::Dot
// (KernelMatrixCSR.hpp:153-227).  Per block row i: the kernel's y = 0
// (CleanseY, BlockKernel_Tiny.hpp:170-174), for each block k in row order
// ReadX(ci[k]) and ReadA(k) then y += a x (DenseBlockKernel_Tiny.hpp:150-180:
// y(r,h) += sum_c a(r,c) x(c,h)), then WriteY: Y_i = alpha y + beta Y_i
// (BlockKernel_Tiny.hpp:176-183, flags as combine_buffers); an empty row
// gets WriteYZero (:185-190): Y_i = beta Y_i.  alpha == 0 or no nonzeros:
// Scale (KernelMatrixCSR.hpp:137-151).  Storage orders: a_RM (a(r,c) at
// r*COLS + c, else c*ROWS + r), x_RM (x(c,h) at c*NRHS + h, else h*COLS + c),
// y_RM likewise with ROWS.
// ---------------------------------------------------------------------------
template <typename T>
void bsr_spmm(i64 mb, const i64* rp, const i32* ci, int R, int C, const T* A, bool a_rm, T alpha, const T* X,
              bool x_rm, T beta, T* Y, bool y_rm, int k) {
    const i64 E = (i64)R * k;
    if (alpha == T(0) || rp[mb] - rp[0] <= 0) {
        for (i64 e = 0; e < mb * E; ++e) Y[e] = (beta == T(0)) ? T(0) : beta * Y[e];
        return;
    }
    std::vector<T> y((size_t)E);
    for (i64 i = 0; i < mb; ++i) {
        T* Yi = Y + i * E;
        if (rp[i + 1] == rp[i]) {
            for (i64 e = 0; e < E; ++e) Yi[e] = (beta == T(0)) ? T(0) : (beta == T(1)) ? Yi[e] : beta * Yi[e];
            continue;
        }
        std::fill(y.begin(), y.end(), T(0));
        for (i64 kb = rp[i]; kb < rp[i + 1]; ++kb) {
            const T* a = A + kb * R * C;
            const T* x = X + (i64)ci[kb] * C * k;
            for (int r = 0; r < R; ++r)
                for (int h = 0; h < k; ++h) {
                    T z = T(0);
                    for (int c = 0; c < C; ++c)
                        z += a[a_rm ? r * C + c : c * R + r] * x[x_rm ? c * k + h : h * C + c];
                    y[(size_t)r * k + h] += z;
                }
        }
        for (int r = 0; r < R; ++r)
            for (int h = 0; h < k; ++h) {
                T& out = Yi[y_rm ? r * k + h : h * R + r];
                const T v = (alpha == T(1)) ? y[(size_t)r * k + h] : alpha * y[(size_t)r * k + h];
                out = (beta == T(0)) ? v : (beta == T(1)) ? v + out : v + beta * out;
            }
    }
}

// FillLowerTriangleFromUpperTriangle (KernelMatrixCSR.hpp:71-128) with the
// diagonal positions of PatternCSR::RequireDiag (PatternCSR.hpp:408-445: first
// position with column >= row) and TransposeBlock (DenseBlockKernel_Tiny.hpp:
// 112-148).  A lower block whose mirror (j, i) is not in the pattern is left
// unchanged (the reference's search then lands on another block).
template <typename T>
void bsr_fill_lower(i64 mb, const i64* rp, const i32* ci, int R, T* A) {
    std::vector<i64> diag((size_t)mb);
    for (i64 i = 0; i < mb; ++i) {
        i64 kk = rp[i];
        while (kk < rp[i + 1] && ci[kk] < i) ++kk;
        diag[(size_t)i] = kk;
    }
    const i64 RR = (i64)R * R;
    for (i64 i = 0; i < mb; ++i)
        for (i64 kb = rp[i]; kb < diag[(size_t)i]; ++kb) {
            const i64 j = ci[kb];
            i64 L = diag[(size_t)j], Rr = rp[j + 1] - 1;
            while (L < Rr) {
                const i64 M = Rr - (Rr - L) / 2;
                if (ci[M] > i)
                    Rr = M - 1;
                else
                    L = M;
            }
            if (L >= rp[j + 1] || ci[L] != i) continue;
            for (int a = 0; a < R; ++a)
                for (int b = 0; b < R; ++b) A[kb * RR + (i64)a * R + b] = A[L * RR + (i64)b * R + a];
        }
}

// ---------------------------------------------------------------------------
// Krylov solvers.  The operator is y = A x on n x nrhs row-major blocks.
// Preconditioner kinds: 0 = identity (copy), 1 = row scaling z_i = d_i * r_i
// (the lumped-mass-inverse / Jacobi diagonal the reference's CG tests use,
// Mathematica/Test_CG.nb).
// ---------------------------------------------------------------------------
struct Op {
    const i64* rp;
    const i32* ci;
    const double* a;
    i32 n;
    i32 threads;
    void apply(const double* x, double* y, i32 nrhs) const {
        multiply_dense<double>(rp, ci, a, n, n, 1.0, x, nrhs, 0.0, y, nrhs, nrhs, 0, threads);
    }
};
// z = P r, rows split over `threads` uniform blocks (elementwise: the split does not
// change the result).
struct Prec {
    int kind;
    const double* d;
    i32 threads = 1;
    void apply(const double* r, double* z, i64 n, i32 nrhs) const {
        parallel_do(threads, [&](int t) {
            const i64 i0 = job_pointer(n, threads, t), i1 = job_pointer(n, threads, t + 1);
            if (kind == 1) {
                for (i64 i = i0; i < i1; ++i)
                    for (i32 k = 0; k < nrhs; ++k) z[i * nrhs + k] = d[i] * r[i * nrhs + k];
            } else {
                std::memcpy(z + i0 * nrhs, r + i0 * nrhs, sizeof(double) * (i1 - i0) * nrhs);
            }
        });
    }
};

// Per-column dot products: each thread sums its uniform row block
// (JobPointers(n, T), ConjugateGradient.hpp:87) in row order, then the
// thread partials are added in thread order (ThreadTensor2::AddReduce,
// ThreadTensor2.hpp:348-367).  ConjugateGradient.hpp:395-438,
// GMRES.hpp:660-729.
static void dots(const double* v, const double* w, i64 n, i32 nrhs, i32 T, double* out) {
    std::vector<std::vector<double>> part(T, std::vector<double>(nrhs, 0.0));
    parallel_do(T, [&](int t) {
        auto& s = part[t];
        for (i64 i = job_pointer(n, T, t); i < job_pointer(n, T, t + 1); ++i)
            for (i32 k = 0; k < nrhs; ++k) s[k] += v[i * nrhs + k] * w[i * nrhs + k];
    });
    for (i32 k = 0; k < nrhs; ++k) out[k] = part[0][k];
    for (i32 t = 1; t < T; ++t)
        for (i32 k = 0; k < nrhs; ++k) out[k] += part[t][k];
}

// X <- a * x + b * X with the flag shortcuts of combine_matrices_auto
// (ConjugateGradient.hpp:342-346, GMRES.hpp:835-845).
static void write_solution(double a, const double* x, double b, double* X, i64 ldX, i64 n, i32 nrhs) {
    const Flag af = (a == 1.0) ? Plus : Generic;
    const Flag bf = (b == 0.0) ? Zero : (b == 1.0) ? Plus : Generic;
    for (i64 i = 0; i < n; ++i)
        for (i32 k = 0; k < nrhs; ++k) {
            double& y = X[i * ldX + k];
            y = combine(af, a, x[i * nrhs + k], bf, b, y);
        }
}

struct CGResult {
    int succeeded;
    i32 iter;
};

// ConjugateGradient::operator() (ConjugateGradient.hpp:105-353).
// Departure: the reference leaves x uninitialised when the loop body never
// runs without an initial guess (:214-250, 342-346); here x starts at 0.
static CGResult conjugate_gradient(const Op& A, const Prec& P, i32 n, i32 nrhs, i32 max_iter, i32 T,
                                   double a, const double* B, i64 ldB, double b, double* X, i64 ldX,
                                   double tol, bool use_guess, double* rel_res, double* abs_res = nullptr) {
    CGResult res{0, 0};
    max_iter = std::max(0, max_iter);
    const i64 N = (i64)n * nrhs;
    if (a == 0.0) {  // :147-163
        for (i64 i = 0; i < n; ++i)
            for (i32 k = 0; k < nrhs; ++k) X[i * ldX + k] *= b;
        res.succeeded = 1;
        return res;
    }
    std::vector<double> r(N), u(N), p(N), x(N, 0.0), z(N);
    std::vector<double> rho(nrhs), rho_old(nrhs), alpha(nrhs), beta(nrhs), TOL(nrhs), bsq(nrhs);
    for (i64 i = 0; i < n; ++i)
        for (i32 k = 0; k < nrhs; ++k) r[i * nrhs + k] = B[i * ldB + k];
    P.apply(r.data(), z.data(), n, nrhs);
    dots(r.data(), z.data(), n, nrhs, T, rho.data());
    for (i32 k = 0; k < nrhs; ++k) {
        bsq[k] = std::fabs(rho[k]);
        TOL[k] = bsq[k] * tol * tol;
    }
    for (i32 k = 0; k < nrhs; ++k)
        if (std::isnan(bsq[k])) return res;  // :185-195
    double bmax = 0.0;
    for (i32 k = 0; k < nrhs; ++k) bmax = std::max(bmax, bsq[k]);
    if (!(bmax > 0.0)) {  // :197-211
        for (i64 i = 0; i < n; ++i)
            for (i32 k = 0; k < nrhs; ++k) X[i * ldX + k] *= b;
        res.succeeded = 1;
        return res;
    }
    if (use_guess) {  // :214-238
        for (i64 i = 0; i < n; ++i)
            for (i32 k = 0; k < nrhs; ++k) x[i * nrhs + k] = X[i * ldX + k];
        A.apply(x.data(), u.data(), nrhs);
        for (i64 e = 0; e < N; ++e) r[e] -= u[e];
    }
    P.apply(r.data(), z.data(), n, nrhs);
    p = z;
    dots(r.data(), z.data(), n, nrhs, T, rho.data());
    auto check = [&] {
        bool ok = true;
        for (i32 k = 0; k < nrhs; ++k) ok = ok && (std::fabs(rho[k]) <= TOL[k]);
        return ok;
    };
    bool succeeded = check();
    i32 iter = 0;
    while (!succeeded && iter < max_iter) {  // :250-339
        A.apply(p.data(), u.data(), nrhs);
        dots(p.data(), u.data(), n, nrhs, T, alpha.data());
        for (i32 k = 0; k < nrhs; ++k) alpha[k] = rho[k] / alpha[k];
        const bool first = (iter == 0) && !use_guess;
        // x and r updates over the threads' row blocks (ConjugateGradient.hpp:262-303)
        parallel_do(T, [&](int t) {
            for (i64 i = job_pointer(n, T, t); i < job_pointer(n, T, t + 1); ++i)
                for (i32 k = 0; k < nrhs; ++k) {
                    const i64 e = i * nrhs + k;
                    x[e] = first ? alpha[k] * p[e] : x[e] + alpha[k] * p[e];
                    r[e] -= alpha[k] * u[e];
                }
        });
        P.apply(r.data(), z.data(), n, nrhs);
        std::swap(rho_old, rho);
        dots(r.data(), z.data(), n, nrhs, T, rho.data());
        for (i32 k = 0; k < nrhs; ++k) beta[k] = rho[k] / rho_old[k];
        // p = z + beta p (:323-335)
        parallel_do(T, [&](int t) {
            for (i64 i = job_pointer(n, T, t); i < job_pointer(n, T, t + 1); ++i)
                for (i32 k = 0; k < nrhs; ++k) {
                    const i64 e = i * nrhs + k;
                    p[e] = z[e] + beta[k] * p[e];
                }
        });
        succeeded = check();
        ++iter;
    }
    write_solution(a, x.data(), b, X, ldX, n, nrhs);
    if (rel_res)
        for (i32 k = 0; k < nrhs; ++k) rel_res[k] = std::sqrt(std::fabs(rho[k]) / bsq[k]);
    if (abs_res)  // ConjugateGradient::Residuals (ConjugateGradient.hpp:447-456)
        for (i32 k = 0; k < nrhs; ++k) abs_res[k] = std::sqrt(std::fabs(rho[k]));
    res.succeeded = succeeded ? 1 : 0;
    res.iter = iter;
    return res;
}

// Upper-triangular, non-unit, row-major back substitution: what
// BLAS::trsv<RowMajor,Upper,Id,NonUnit> (src/BLAS/trsv.hpp:7-48) asks CBLAS
// for at GMRES.hpp:404-406.
static void trsv_upper(i32 k, const double* H, i32 ldH, double* b) {
    for (i32 i = k - 1; i >= 0; --i) {
        double s = b[i];
        for (i32 j = i + 1; j < k; ++j) s -= H[i * ldH + j] * b[j];
        b[i] = s / H[i * ldH + i];
    }
}

struct GMRESResult {
    int succeeded;
    i32 iter;
    i32 restarts;
};

// GMRES::operator() / Solve / ArnoldiStep / ApplyGivensRotations
// (GMRES.hpp:127-260, 264-485, 526-601, 603-658).  side: 0 = Left, 1 = Right.
struct GMRES {
    const Op& A;
    const Prec& P;
    i32 n, nrhs, m, T, side;
    std::vector<double> Q, H, cs, sn, beta, x, z, solution, TOL, b_norms, r_norms, q_norms, h;
    i32 iter = 0;

    GMRES(const Op& A_, const Prec& P_, i32 n_, i32 max_iter, i32 nrhs_, i32 T_, i32 side_)
        : A(A_), P(P_), n(n_), nrhs(nrhs_), m(std::min(max_iter, n_ + 1)), T(T_), side(side_) {
        Q.assign((i64)(m + 1) * n * nrhs, 0.0);
        H.assign((i64)(m + 1) * m * nrhs, 0.0);
        cs.assign((i64)m * nrhs, 0.0);
        sn.assign((i64)m * nrhs, 0.0);
        beta.assign((i64)(m + 1) * nrhs, 0.0);
        x.assign((i64)n * nrhs, 0.0);
        z.assign((i64)n * nrhs, 0.0);
        solution.assign((i64)n * nrhs, 0.0);
        TOL.assign(nrhs, 0.0);
        b_norms.assign(nrhs, 0.0);
        r_norms.assign(nrhs, 0.0);
        q_norms.assign(nrhs, 0.0);
        h.assign(nrhs, 0.0);
    }
    double* q(i32 j) { return Q.data() + (i64)j * n * nrhs; }
    double& Hm(i32 i, i32 j, i32 k) { return H[((i64)i * m + j) * nrhs + k]; }
    void norms(const double* v, std::vector<double>& out) {
        dots(v, v, n, nrhs, T, out.data());
        for (auto& o : out) o = std::sqrt(o);
    }
    void inverse_scale(double* v, const std::vector<double>& f) {
        for (i64 i = 0; i < n; ++i)
            for (i32 k = 0; k < nrhs; ++k) v[i * nrhs + k] *= 1.0 / f[k];
    }
    bool check() {
        bool ok = true;
        for (i32 k = 0; k < nrhs; ++k) ok = ok && (std::fabs(beta[(i64)iter * nrhs + k]) <= TOL[k]);
        return ok;
    }
    void arnoldi() {
        double* qp = q(iter + 1);
        if (side == 0) {
            A.apply(q(iter), z.data(), nrhs);
            P.apply(z.data(), qp, n, nrhs);
        } else {
            P.apply(q(iter), z.data(), n, nrhs);
            A.apply(z.data(), qp, nrhs);
        }
        for (int gs = 0; gs < 2; ++gs)  // gram_schmidt_counts = 2 (GMRES.hpp:31)
            for (i32 i = 0; i < iter + 1; ++i) {
                double* qi = q(i);
                dots(qi, qp, n, nrhs, T, h.data());
                for (i32 k = 0; k < nrhs; ++k) Hm(i, iter, k) += h[k];
                for (i64 e = 0; e < (i64)n; ++e)
                    for (i32 k = 0; k < nrhs; ++k) qp[e * nrhs + k] -= qi[e * nrhs + k] * h[k];
            }
        norms(qp, q_norms);
        for (i32 k = 0; k < nrhs; ++k) Hm(iter + 1, iter, k) = q_norms[k];
        inverse_scale(qp, q_norms);
    }
    void givens() {
        for (i32 i = 0; i < iter; ++i)
            for (i32 k = 0; k < nrhs; ++k) {
                const double xi = Hm(i, iter, k), eta = Hm(i + 1, iter, k);
                const double c = cs[(i64)i * nrhs + k], s = sn[(i64)i * nrhs + k];
                Hm(i, iter, k) = c * xi + s * eta;
                Hm(i + 1, iter, k) = -s * xi + c * eta;
            }
        for (i32 k = 0; k < nrhs; ++k) {
            const double xi = Hm(iter, iter, k), eta = Hm(iter + 1, iter, k);
            double c, s;
            const double r = std::sqrt(xi * xi + eta * eta);
            if (std::fabs(xi) <= 2.220446049250313e-16 * r) {
                c = 0.0;
                s = 1.0;
            } else {
                const double rinv = 1.0 / r;
                c = std::fabs(xi) * rinv;
                s = (xi / std::fabs(xi)) * eta * rinv;
            }
            cs[(i64)iter * nrhs + k] = c;
            sn[(i64)iter * nrhs + k] = s;
            Hm(iter, iter, k) = c * xi + s * eta;
            Hm(iter + 1, iter, k) = 0.0;
            const double bi = beta[(i64)iter * nrhs + k];
            beta[(i64)(iter + 1) * nrhs + k] = -s * bi;
            beta[(i64)iter * nrhs + k] = c * bi;
        }
    }
    bool solve(const double* B, i64 ldB, bool implicit_zero) {
        std::fill(h.begin(), h.end(), 0.0);
        if (!implicit_zero) x = solution;
        const i64 N = (i64)n * nrhs;
        double* q0 = q(0);
        double* target = (side == 0) ? z.data() : q0;
        if (implicit_zero) {
            for (i64 i = 0; i < n; ++i)
                for (i32 k = 0; k < nrhs; ++k) target[i * nrhs + k] = -B[i * ldB + k];
        } else {
            A.apply(x.data(), target, nrhs);
            for (i64 i = 0; i < n; ++i)
                for (i32 k = 0; k < nrhs; ++k) target[i * nrhs + k] += -B[i * ldB + k];
        }
        if (side == 0) P.apply(z.data(), q0, n, nrhs);
        norms(q0, r_norms);
        inverse_scale(q0, r_norms);
        std::fill(beta.begin(), beta.end(), 0.0);
        for (i32 k = 0; k < nrhs; ++k) beta[k] = r_norms[k];
        std::fill(H.begin(), H.end(), 0.0);
        iter = 0;
        bool succ = check();
        if (succ) return succ;  // GMRES.hpp:371-375
        while (!succ && iter < m) {
            arnoldi();
            givens();
            ++iter;
            succ = check();
        }
        std::vector<double> y((i64)iter * nrhs), Hk((i64)iter * iter), bk(iter);
        for (i32 k = 0; k < nrhs; ++k) {
            for (i32 i = 0; i < iter; ++i) {
                bk[i] = beta[(i64)i * nrhs + k];
                for (i32 j = 0; j < iter; ++j) Hk[(i64)i * iter + j] = Hm(i, j, k);
            }
            trsv_upper(iter, Hk.data(), iter, bk.data());
            for (i32 j = 0; j < iter; ++j) y[(i64)j * nrhs + k] = bk[j];
        }
        std::fill(z.begin(), z.end(), 0.0);
        for (i64 i = 0; i < n; ++i)
            for (i32 j = 0; j < iter; ++j) {
                const double* qj = q(j);
                for (i32 k = 0; k < nrhs; ++k) z[i * nrhs + k] += y[(i64)j * nrhs + k] * qj[i * nrhs + k];
            }
        if (side == 0)
            std::swap(x, z);
        else
            P.apply(z.data(), x.data(), n, nrhs);
        for (i64 e = 0; e < N; ++e) solution[e] = implicit_zero ? -x[e] : solution[e] - x[e];
        return succ;
    }
};

// Optional state outputs: the members behind GetBasis / GetHessenbergMatrix /
// Solution / Residuals (GMRES.hpp:779-789, 820-833) after the last cycle.
struct GMRESState {
    double* Q = nullptr;         // (m+1) x n x nrhs
    double* H = nullptr;         // (m+1) x m x nrhs
    double* beta = nullptr;      // (m+1) x nrhs
    double* solution = nullptr;  // n x nrhs
};

static GMRESResult gmres(const Op& A, const Prec& P, i32 n, i32 nrhs, i32 max_iter, i32 T, i32 side,
                         double a, const double* B, i64 ldB, double b, double* X, i64 ldX, double tol,
                         i32 max_restarts, bool use_guess, double* rel_res, const GMRESState* out = nullptr) {
    GMRESResult res{0, 0, 0};
    GMRES g(A, P, n, max_iter, nrhs, T, side);
    const i64 N = (i64)n * nrhs;
    for (i64 i = 0; i < n; ++i)
        for (i32 k = 0; k < nrhs; ++k) g.x[i * nrhs + k] = B[i * ldB + k];
    if (side == 0)
        P.apply(g.x.data(), g.z.data(), n, nrhs);
    else
        std::swap(g.x, g.z);
    g.norms(g.z.data(), g.b_norms);
    for (i32 k = 0; k < nrhs; ++k) g.TOL[k] = g.b_norms[k] * tol;
    for (i32 k = 0; k < nrhs; ++k)
        if (std::isnan(g.b_norms[k])) return res;  // GMRES.hpp:193-204
    double bmax = 0.0;
    for (double v : g.b_norms) bmax = std::max(bmax, v);
    if (!(bmax > 0.0)) {  // :206-229
        for (i64 i = 0; i < n; ++i)
            for (i32 k = 0; k < nrhs; ++k) X[i * ldX + k] = (b == 0.0) ? 0.0 : X[i * ldX + k] * b;
        res.succeeded = 1;
        return res;
    }
    (void)N;
    if (use_guess)
        for (i64 i = 0; i < n; ++i)
            for (i32 k = 0; k < nrhs; ++k) g.solution[i * nrhs + k] = X[i * ldX + k];
    i32 restarts = 0;
    bool succeeded = g.solve(B, ldB, !use_guess);
    ++restarts;
    while (!succeeded && restarts <= max_restarts) {  // :244-250
        succeeded = g.solve(B, ldB, false);
        ++restarts;
    }
    --restarts;
    write_solution(a, g.solution.data(), b, X, ldX, n, nrhs);
    if (rel_res)
        for (i32 k = 0; k < nrhs; ++k)
            rel_res[k] = std::fabs(g.beta[(i64)g.iter * nrhs + k]) / g.b_norms[k];
    if (out) {
        if (out->Q) std::copy(g.Q.begin(), g.Q.end(), out->Q);
        if (out->H) std::copy(g.H.begin(), g.H.end(), out->H);
        if (out->beta) std::copy(g.beta.begin(), g.beta.end(), out->beta);
        if (out->solution) std::copy(g.solution.begin(), g.solution.end(), out->solution);
    }
    res.succeeded = succeeded ? 1 : 0;
    res.iter = g.iter;
    res.restarts = restarts;
    return res;
}

static void copy_out(const CSR& C, i64* rp, i32* ci, double* val) {
    std::copy(C.rp.begin(), C.rp.end(), rp);
    std::copy(C.ci.begin(), C.ci.end(), ci);
    if (val && !C.val.empty()) std::copy(C.val.begin(), C.val.end(), val);
}

static CSR make_csr(i64 m, i64 n, const i64* rp, const i32* ci, const double* val) {
    CSR A;
    A.m = m;
    A.n = n;
    A.rp.assign(rp, rp + m + 1);
    A.ci.assign(ci, ci + rp[m]);
    if (val) A.val.assign(val, val + rp[m]);
    return A;
}

}  // namespace orc

// =============================================================================
// C interface (ctypes).  Row pointers are int64, column indices int32.
// =============================================================================
extern "C" {

// Pins the worker threads of parallel regions to CPUs (1) or not (0, the default).
void orc_set_pinning(int32_t on) {
    if (on && orc::g_proc_cpus.load() == 0) orc::take_process_mask();
    orc::g_pin.store(on ? 1 : 0);
}

int64_t orc_job_pointer(int64_t n, int64_t T, int64_t t) { return orc::job_pointer(n, T, t); }

void orc_job_pointers_costs(int64_t m, const int64_t* acc, int64_t T, int64_t* out) {
    auto jp = orc::job_pointers_costs(m, acc, T);
    std::copy(jp.begin(), jp.end(), out);
}

int orc_bsr_spmm_f64(int64_t mb, const int64_t* rp, const int32_t* ci, int32_t R, int32_t C, const double* A,
                     int32_t a_rm, double alpha, const double* X, int32_t x_rm, double beta, double* Y,
                     int32_t y_rm, int32_t k) {
    orc::bsr_spmm<double>(mb, rp, ci, R, C, A, a_rm != 0, alpha, X, x_rm != 0, beta, Y, y_rm != 0, k);
    return 0;
}

int orc_bsr_spmm_f32(int64_t mb, const int64_t* rp, const int32_t* ci, int32_t R, int32_t C, const float* A,
                     int32_t a_rm, float alpha, const float* X, int32_t x_rm, float beta, float* Y, int32_t y_rm,
                     int32_t k) {
    orc::bsr_spmm<float>(mb, rp, ci, R, C, A, a_rm != 0, alpha, X, x_rm != 0, beta, Y, y_rm != 0, k);
    return 0;
}

void orc_bsr_fill_lower(int64_t mb, const int64_t* rp, const int32_t* ci, int32_t R, double* A) {
    orc::bsr_fill_lower<double>(mb, rp, ci, R, A);
}

int orc_multiply_dense_f64(const int64_t* rp, const int32_t* ci, const double* a, int32_t m, int32_t n,
                           double alpha, const double* X, int64_t ldX, double beta, double* Y, int64_t ldY,
                           int32_t nrhs, int32_t NRHS, int32_t threads) {
    return orc::multiply_dense<double>(rp, ci, a, m, n, alpha, X, ldX, beta, Y, ldY, nrhs, NRHS, threads);
}

int orc_multiply_dense_f32(const int64_t* rp, const int32_t* ci, const float* a, int32_t m, int32_t n,
                           float alpha, const float* X, int64_t ldX, float beta, float* Y, int64_t ldY,
                           int32_t nrhs, int32_t NRHS, int32_t threads) {
    return orc::multiply_dense<float>(rp, ci, a, m, n, alpha, X, ldX, beta, Y, ldY, nrhs, NRHS, threads);
}

// Triples -> CSR.  The triples are cut into `lists` contiguous lists with the
// uniform split (MatrixCSR.hpp:203-235).  Outputs need room for m+1 / cap.
int64_t orc_from_triples(int64_t count, const int32_t* I, const int32_t* J, const double* V, int32_t lists,
                         int64_t m, int64_t n, int compressQ, int symmetrizeQ, int64_t* rp, int32_t* ci,
                         double* val, int64_t cap) {
    if (lists < 1) return -1;
    std::vector<const int32_t*> pi, pj;
    std::vector<const double*> pv;
    std::vector<int64_t> cnt;
    for (int32_t t = 0; t < lists; ++t) {
        const int64_t b = orc::job_pointer(count, lists, t), e = orc::job_pointer(count, lists, t + 1);
        pi.push_back(I + b);
        pj.push_back(J + b);
        pv.push_back(V + b);
        cnt.push_back(e - b);
    }
    orc::CSR A = orc::from_triples(pi, pj, pv, cnt, m, n, compressQ != 0, symmetrizeQ != 0);
    const int64_t nnz = A.rp[m];
    if (nnz > cap) return -2;
    orc::copy_out(A, rp, ci, val);
    return nnz;
}

int64_t orc_grid_laplacian(int32_t g, double mass, int32_t parallel_threads, int64_t* rp, int32_t* ci,
                           double* val, int64_t cap) {
    orc::CSR A = parallel_threads > 0 ? orc::grid_laplacian_parallel(g, mass, parallel_threads)
                                      : orc::grid_laplacian(g, mass);
    const int64_t nnz = A.rp[(int64_t)g * g];
    if (nnz > cap) return -2;
    orc::copy_out(A, rp, ci, val);
    return nnz;
}

// Transpose: outputs rp (n+1), ci (nnz), val (nnz or null).
int orc_transpose(int64_t m, int64_t n, const int64_t* rp, const int32_t* ci, const double* val, int32_t T,
                  int64_t* out_rp, int32_t* out_ci, double* out_val) {
    orc::CSR A = orc::make_csr(m, n, rp, ci, val);
    orc::CSR B = orc::transpose(A, T, val != nullptr);
    orc::copy_out(B, out_rp, out_ci, out_val);
    return 0;
}

// Upper bound on the expanded nonzero count of A*B.
int64_t orc_spgemm_expanded_count(int64_t m, const int64_t* Arp, const int32_t* Aci, const int64_t* Brp) {
    int64_t s = 0;
    for (int64_t i = 0; i < m; ++i)
        for (int64_t jj = Arp[i]; jj < Arp[i + 1]; ++jj) s += Brp[Aci[jj] + 1] - Brp[Aci[jj]];
    return s;
}

int64_t orc_spgemm(int64_t m, int64_t k, int64_t n, const int64_t* Arp, const int32_t* Aci, const double* Aval,
                   const int64_t* Brp, const int32_t* Bci, const double* Bval, int32_t T, int64_t* Crp,
                   int32_t* Cci, double* Cval, int64_t cap) {
    const bool values = Aval && Bval;
    orc::CSR A = orc::make_csr(m, k, Arp, Aci, values ? Aval : nullptr);
    orc::CSR B = orc::make_csr(k, n, Brp, Bci, values ? Bval : nullptr);
    orc::CSR C = orc::spgemm(A, B, T, values);
    const int64_t nnz = C.rp[m];
    if (nnz > cap) return -2;
    orc::copy_out(C, Crp, Cci, values ? Cval : nullptr);
    return nnz;
}

void orc_trsv_upper(int32_t k, const double* H, int32_t ldH, double* b) { orc::trsv_upper(k, H, ldH, b); }

// Returns succeeded (0/1); iterations in *iter; relative residuals sqrt(|rho|/|b.Pb|).
int orc_cg(int32_t n, const int64_t* rp, const int32_t* ci, const double* val, int32_t prec_kind,
           const double* prec_diag, int32_t nrhs, int32_t max_iter, int32_t threads, double a, const double* B,
           int64_t ldB, double b, double* X, int64_t ldX, double tol, int32_t use_guess, int32_t* iter,
           double* rel_res) {
    orc::Op A{rp, ci, val, n, threads};
    orc::Prec P{prec_kind, prec_diag, threads};
    auto r = orc::conjugate_gradient(A, P, n, nrhs, max_iter, threads, a, B, ldB, b, X, ldX, tol,
                                     use_guess != 0, rel_res);
    if (iter) *iter = r.iter;
    return r.succeeded;
}

// orc_cg plus abs_res[nrhs] = sqrt|rho| (ConjugateGradient::Residuals).
int orc_cg_ex(int32_t n, const int64_t* rp, const int32_t* ci, const double* val, int32_t prec_kind,
              const double* prec_diag, int32_t nrhs, int32_t max_iter, int32_t threads, double a, const double* B,
              int64_t ldB, double b, double* X, int64_t ldX, double tol, int32_t use_guess, int32_t* iter,
              double* rel_res, double* abs_res) {
    orc::Op A{rp, ci, val, n, threads};
    orc::Prec P{prec_kind, prec_diag, threads};
    auto r = orc::conjugate_gradient(A, P, n, nrhs, max_iter, threads, a, B, ldB, b, X, ldX, tol,
                                     use_guess != 0, rel_res, abs_res);
    if (iter) *iter = r.iter;
    return r.succeeded;
}

int orc_gmres(int32_t n, const int64_t* rp, const int32_t* ci, const double* val, int32_t prec_kind,
              const double* prec_diag, int32_t side, int32_t nrhs, int32_t max_iter, int32_t threads, double a,
              const double* B, int64_t ldB, double b, double* X, int64_t ldX, double tol, int32_t max_restarts,
              int32_t use_guess, int32_t* iter, int32_t* restarts, double* rel_res) {
    orc::Op A{rp, ci, val, n, threads};
    orc::Prec P{prec_kind, prec_diag};
    auto r = orc::gmres(A, P, n, nrhs, max_iter, threads, side, a, B, ldB, b, X, ldX, tol, max_restarts,
                        use_guess != 0, rel_res);
    if (iter) *iter = r.iter;
    if (restarts) *restarts = r.restarts;
    return r.succeeded;
}

// orc_gmres plus the final-cycle state: Q ((m+1) x n x nrhs), H ((m+1) x m x nrhs),
// beta ((m+1) x nrhs), solution (n x nrhs) with m = min(max_iter, n + 1); any may be null.
int orc_gmres_ex(int32_t n, const int64_t* rp, const int32_t* ci, const double* val, int32_t prec_kind,
                 const double* prec_diag, int32_t side, int32_t nrhs, int32_t max_iter, int32_t threads, double a,
                 const double* B, int64_t ldB, double b, double* X, int64_t ldX, double tol, int32_t max_restarts,
                 int32_t use_guess, int32_t* iter, int32_t* restarts, double* rel_res, double* Q, double* H,
                 double* beta, double* solution) {
    orc::Op A{rp, ci, val, n, threads};
    orc::Prec P{prec_kind, prec_diag};
    orc::GMRESState st{Q, H, beta, solution};
    auto r = orc::gmres(A, P, n, nrhs, max_iter, threads, side, a, B, ldB, b, X, ldX, tol, max_restarts,
                        use_guess != 0, rel_res, &st);
    if (iter) *iter = r.iter;
    if (restarts) *restarts = r.restarts;
    return r.succeeded;
}

}  // extern "C"
