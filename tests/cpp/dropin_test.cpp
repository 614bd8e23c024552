// Exercises the header-only drop-in (include/Tensors) the way code written
// against the reference uses it: GridLaplacian, MatrixCSR::Dot overloads
// (host and device pointers, Tensor1/Tensor2, strided), the reference's no-op
// error cases, Transpose, sparse x sparse, ConjugateGradient and GMRES with CSR
// operators, host lambdas and device lambdas.  Prints one line per check and
// exits non-zero on the first failure.  argv[1] (optional): file that receives
// the GridLaplacian(16) structure for the pytest to compare with the oracle;
// argv[2] (optional): directory that receives the solver accessors' values.
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <thread>

#include "Tensors/Tensors.hpp"

using namespace Tensors;
using Real = double;
using Int = int32_t;
using LInt = int64_t;
using CSR = Sparse::MatrixCSR<Real, Int, LInt, Parallel>;

static int failures = 0;
static void check(bool ok, const char* what, double err = 0.0) {
    std::printf("%s %-40s %.3e\n", ok ? "OK  " : "FAIL", what, err);
    if (!ok) ++failures;
}

static void brute(const CSR& A, Real alpha, const Real* X, Int ldX, Real beta, Real* Y, Int ldY, Int k) {
    for (Int i = 0; i < A.RowCount(); ++i)
        for (Int c = 0; c < k; ++c) {
            long double z = 0;
            for (LInt l = A.Outer()[i]; l < A.Outer()[i + 1]; ++l)
                z += (long double)A.Values()[l] * X[(LInt)A.Inner()[l] * ldX + c];
            Real& y = Y[(LInt)i * ldY + c];
            y = (Real)(alpha * z) + (beta == 0 ? 0.0 : beta * y);
        }
}

static double relerr(const Real* a, const Real* b, size_t n) {
    double e = 0, m = 0;
    for (size_t i = 0; i < n; ++i) {
        e = std::max(e, std::abs(a[i] - b[i]));
        m = std::max(m, std::abs(b[i]));
    }
    return m > 0 ? e / m : e;
}

int main(int argc, char** argv) {
    // --- GridLaplacian (GridLaplacian.hpp:11-77) -------------------------------
    const Int g = 16;
    CSR L = Sparse::GridLaplacian<Real, Int, LInt>(g, Real(1), 4);
    check(L.NonzeroCount() == 5 * g * g - 4 * g, "GridLaplacian nnz = 5g^2 - 4g");
    bool sorted = true, values_ok = true;
    for (Int i = 0; i < L.RowCount(); ++i) {
        Int deg = 0;
        for (LInt l = L.Outer()[i]; l < L.Outer()[i + 1]; ++l) {
            if (l > L.Outer()[i] && L.Inner()[l] <= L.Inner()[l - 1]) sorted = false;
            if (L.Inner()[l] != i) {
                ++deg;
                values_ok &= L.Values()[l] == -1.0;
            }
        }
        for (LInt l = L.Outer()[i]; l < L.Outer()[i + 1]; ++l)
            if (L.Inner()[l] == i) values_ok &= L.Values()[l] == deg + 1.0;
    }
    check(sorted, "GridLaplacian rows strictly sorted");
    check(values_ok, "GridLaplacian diag = degree + mass, off = -1");
    CSR Lp = Sparse::GridLaplacian_Parallel<Real, Int, LInt>(g, Real(1), 3);
    check(Lp.Outer() == L.Outer() && Lp.Inner() == L.Inner() && Lp.Values() == L.Values(),
          "GridLaplacian == GridLaplacian_Parallel");
    if (argc > 1) {
        FILE* f = std::fopen(argv[1], "wb");
        const int64_t m = L.RowCount(), nnz = L.NonzeroCount();
        std::fwrite(&m, 8, 1, f);
        std::fwrite(&nnz, 8, 1, f);
        std::fwrite(L.Outer().data(), sizeof(LInt), m + 1, f);
        std::fwrite(L.Inner().data(), sizeof(Int), nnz, f);
        std::fwrite(L.Values().data(), sizeof(Real), nnz, f);
        std::fclose(f);
    }

    // --- MatrixMarket round trip (MatrixCSR/MatrixMarket.hpp:3-306) ---------------
    {
        const auto path = std::filesystem::temp_directory_path() / ("tensors_amd_dropin_" +
                                                                    std::to_string((long)getpid()) + ".mtx");
        L.WriteToMatrixMarket(path);
        CSR M;
        M.LoadFromMatrixMarket(path, 2);
        check(M.RowCount() == L.RowCount() && M.Outer() == L.Outer() && M.Inner() == L.Inner() &&
                  M.Values() == L.Values(),
              "WriteToMatrixMarket -> LoadFromMatrixMarket reproduces the matrix");
        CSR keep = M;
        M.LoadFromMatrixMarket(path.string() + ".missing", 2);  // prints, leaves M unchanged
        check(M.Outer() == keep.Outer() && M.Values() == keep.Values(), "LoadFromMatrixMarket(missing) is a no-op");
        std::filesystem::remove(path);
    }

    // --- KernelMatrixCSR with dense 3 x 3 blocks (KernelMatrixCSR.hpp:153-227) ----
    {
        using Ker = DenseBlockKernel_Tiny<3, 3, 2, Real, Real, Real, Int, LInt, Scalar::Flag::Generic,
                                          Scalar::Flag::Generic, true, false, true, true, true, true>;
        Sparse::KernelMatrixCSR<Ker> K(L);  // the grid pattern, one 3 x 3 block per nonzero
        const Int mb = L.RowCount();
        const LInt nnzb = L.NonzeroCount();
        check(K.RowCount() == 3 * mb && K.NonzeroCount() == 9 * nnzb, "KernelMatrixCSR sizes");
        Tensor1<Real, LInt> A(9 * nnzb);
        Tensor2<Real, Int> X(3 * mb, 2), Y(3 * mb, 2), Yr(3 * mb, 2);
        A.Random();
        X.Random();
        Y.Random();
        Yr = Y;
        K.Dot(A.data(), Real(2), X.data(), Real(-0.5), Y.data(), 2);
        for (Int i = 0; i < mb; ++i)
            for (int r = 0; r < 3; ++r)
                for (int h = 0; h < 2; ++h) {
                    double z = 0;
                    for (LInt l = L.Outer()[i]; l < L.Outer()[i + 1]; ++l)
                        for (int c = 0; c < 3; ++c) z += A[9 * l + 3 * r + c] * X(3 * L.Inner()[l] + c, h);
                    Yr(3 * i + r, h) = 2.0 * z - 0.5 * Yr(3 * i + r, h);
                }
        const double e = relerr(Y.data(), Yr.data(), (size_t)6 * mb);
        check(e < 1e-13, "KernelMatrixCSR::Dot (host pointers) matches the block product", e);
        K.FillLowerTriangleFromUpperTriangle(A.data());
        bool sym = true;
        for (Int i = 0; i < mb; ++i)
            for (LInt l = L.Outer()[i]; l < L.Outer()[i + 1]; ++l) {
                const Int j = L.Inner()[l];
                if (j >= i) continue;
                for (LInt u = L.Outer()[j]; u < L.Outer()[j + 1]; ++u)
                    if (L.Inner()[u] == i)
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b) sym &= A[9 * l + 3 * a + b] == A[9 * u + 3 * b + a];
            }
        check(sym, "KernelMatrixCSR::FillLowerTriangleFromUpperTriangle mirrors the upper blocks");
    }

    // --- Dot overloads (MatrixCSR.hpp:965-1086) --------------------------------
    const Int n = L.RowCount();
    {
        Tensor2<Real, Int> X(n, 4), Y(n, 4), Yr(n, 4);
        X.Random();
        Y.Random();
        Yr = Y;
        L.Dot(Real(2), X, Real(-1), Y);
        brute(L, 2, X.data(), 4, -1, Yr.data(), 4, 4);
        const double e = relerr(Y.data(), Yr.data(), (size_t)n * 4);
        check(e <= 1e-12, "Dot(Tensor2) alpha=2 beta=-1", e);
    }
    {
        Tensor1<Real, Int> x(n), y(n), yr(n);
        x.Random();
        L.Dot(Real(1), x, Real(0), y);
        brute(L, 1, x.data(), 1, 0, yr.data(), 1, 1);
        const double e = relerr(y.data(), yr.data(), n);
        check(e <= 1e-12, "Dot(Tensor1)", e);
    }
    {
        const Int k = 3, ld = 5;
        Tensor2<Real, Int> X(n, ld), Y(n, ld), Yr(n, ld);
        X.Random();
        Y.Random();
        Yr = Y;
        L.Dot(Real(0.5), X.data(), ld, Real(3), Y.data(), ld, k);
        brute(L, 0.5, X.data(), ld, 3, Yr.data(), ld, k);
        const double e = relerr(Y.data(), Yr.data(), (size_t)n * ld);
        check(e <= 1e-12, "Dot(ptr, ld=5, nrhs=3) with untouched tail", e);
    }
    {
        const Int k = 2;
        Tensor2<Real, Int> X(n, k), Yh(n, k), Yr(n, k);
        X.Random();
        Device::Buffer<Real> dX, dY((Size_T)n * k);
        dX.Upload(X.data(), (Size_T)n * k);
        L.Dot(Real(1), dX.data(), Real(0), dY.data(), k);
        dY.Download(Yh.data(), (Size_T)n * k);
        brute(L, 1, X.data(), k, 0, Yr.data(), k, k);
        const double e = relerr(Yh.data(), Yr.data(), (size_t)n * k);
        check(e <= 1e-12, "Dot on device pointers", e);
    }
    {
        // Dot from several host threads on one fresh matrix: the first calls race
        // for the lazy device upload and all share the staging buffers
        CSR F = L;
        const int T = 6;
        std::vector<Tensor2<Real, Int>> Xs, Ys, Yr;
        for (int t = 0; t < T; ++t) {
            const Int k = 1 + t % 4;
            Xs.emplace_back(n, k);
            Ys.emplace_back(n, k);
            Yr.emplace_back(n, k);
            Xs[t].Random();
            brute(L, 1, Xs[t].data(), k, 0, Yr[t].data(), k, k);
        }
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (int rep = 0; rep < 4; ++rep) F.Dot(Real(1), Xs[t], Real(0), Ys[t]);
            });
        for (auto& x : th) x.join();
        double e = 0;
        for (int t = 0; t < T; ++t)
            e = std::max(e, relerr(Ys[t].data(), Yr[t].data(), (size_t)n * (1 + t % 4)));
        check(e <= 1e-12, "Dot from 6 host threads on one matrix (first use inside the race)", e);
    }
    {
        Tensor2<Real, Int> X(n, 2), snap(n, 2);
        X.Random();
        snap = X;
        L.Dot(Real(1), X.data(), Real(0), X.data(), 2);  // reference: message, no-op
        check(X == snap, "Dot X == Y is a no-op");
        Tensor2<Real, Int> Yb(n + 1, 2, Real(7));
        L.Dot(Real(1), X, Real(0), Yb);  // shape mismatch: message, no-op
        bool untouched = true;
        for (Int i = 0; i < n + 1; ++i) untouched &= Yb(i, 0) == 7.0 && Yb(i, 1) == 7.0;
        check(untouched, "Dot shape mismatch is a no-op");
        Tensor2<Real, Int> Y4(n, 2, Real(7));
        L.Dot<4>(Real(1), X.data(), Real(0), Y4.data(), 2);  // NRHS != nrhs: message, no-op
        check(Y4(0, 0) == 7.0, "Dot NRHS != nrhs is a no-op");
    }

    // --- Transpose and sparse x sparse ---------------------------------------------
    {
        // a nonsymmetric matrix from triples: A(i, (3i+1) % n) = i + 1, A(i,i) = 2
        std::vector<Int> I, J;
        std::vector<Real> V;
        for (Int i = 0; i < n; ++i) {
            I.push_back(i); J.push_back((3 * i + 1) % n); V.push_back(i + 1.0);
            I.push_back(i); J.push_back(i); V.push_back(2.0);
            I.push_back(i); J.push_back(i); V.push_back(0.5);  // duplicate: summed
        }
        CSR A((LInt)I.size(), I.data(), J.data(), V.data(), n, n, 2);
        std::map<std::pair<Int, Int>, double> dense;
        for (size_t t = 0; t < I.size(); ++t) dense[{I[t], J[t]}] += V[t];
        bool ok = (size_t)A.NonzeroCount() == dense.size();
        for (Int i = 0; i < n && ok; ++i)
            for (LInt l = A.Outer()[i]; l < A.Outer()[i + 1]; ++l) ok &= dense[{i, A.Inner()[l]}] == A.Values()[l];
        check(ok, "MatrixCSR from triples (compress duplicates)");
        CSR At = A.Transpose();
        bool tok = At.NonzeroCount() == A.NonzeroCount();
        for (Int i = 0; i < n && tok; ++i)
            for (LInt l = At.Outer()[i]; l < At.Outer()[i + 1]; ++l) {
                tok &= dense[{At.Inner()[l], i}] == At.Values()[l];
                if (l > At.Outer()[i]) tok &= At.Inner()[l] > At.Inner()[l - 1];
            }
        check(tok, "Transpose (structure exact, sorted)");
        CSR C = A.Dot(At);  // A * A^T
        std::map<std::pair<Int, Int>, double> cd;
        for (auto& [ij, a] : dense)
            for (auto& [kl, b] : dense)
                if (ij.second == kl.second) cd[{ij.first, kl.first}] += a * b;
        bool cok = (size_t)C.NonzeroCount() == cd.size();
        double cerr = 0;
        for (Int i = 0; i < n && cok; ++i)
            for (LInt l = C.Outer()[i]; l < C.Outer()[i + 1]; ++l)
                cerr = std::max(cerr, std::abs(cd[{i, C.Inner()[l]}] - C.Values()[l]));
        check(cok && cerr <= 1e-12, "sparse x sparse A * A^T", cerr);
        CSR C2 = Sparse::Dot<Op::Id, Op::Trans>(A, A);  // Sparse/Dot.hpp:280-323
        check(C2.Outer() == C.Outer() && C2.Inner() == C.Inner() && C2.Values() == C.Values(),
              "Sparse::Dot<Id, Trans>(A, A) == A.Dot(A.Transpose())");
        CSR At2 = A.Op<Op::Trans>();
        check(At2.Outer() == At.Outer() && At2.Inner() == At.Inner(), "MatrixCSR::Op<Trans> == Transpose");
    }

    // --- ConjugateGradient (ConjugateGradient.hpp:105-353) ---------------------
    {
        const Int k = 2;
        Tensor2<Real, Int> B(n, k), X(n, k, Real(0)), R(n, k);
        B.Random();
        IdentityPreconditioner P;
        ConjugateGradient<0, Real, Int> cg(n, 1000, k, 4);
        const bool ok = cg(L, P, Real(1), B.data(), k, Real(0), X.data(), k, Real(1e-10));
        R = B;
        brute(L, -1, X.data(), k, 1, R.data(), k, k);  // R = B - A X
        const double res = R.MaxNorm() / B.MaxNorm();
        check(ok && res < 1e-8, "CG(MatrixCSR, Identity) residual", res);
        // the reference's lambda idiom (Mathematica/Dev_CG.nb): host pointers, staged
        auto A_host = [&](const Real* x, Real* y) { L.Dot(Real(1), x, Real(0), y, k); };
        Tensor2<Real, Int> X2(n, k, Real(0));
        const bool ok2 = cg(A_host, P, Real(1), B.data(), k, Real(0), X2.data(), k, Real(1e-10));
        // a MatrixCSR takes the fused iteration (p.u summed over another workgroup
        // partition), a lambda the four-kernel one: equal up to rounding
        const double e2 = relerr(X2.data(), X.data(), (size_t)n * k);
        check(ok2 && e2 < 1e-9, "CG(host lambda) matches CG(MatrixCSR)", e2);
        auto A_dev = DeviceOperator([&](const Real* x, Real* y) { L.Dot(Real(1), x, Real(0), y, k); });
        Tensor2<Real, Int> X3(n, k, Real(0));
        const bool ok3 = cg(A_dev, P, Real(1), B.data(), k, Real(0), X3.data(), k, Real(1e-10));
        check(ok3 && X3 == X2, "CG(device lambda) == CG(host lambda) bitwise");
        Tensor1<Real, Int> diag(n);
        for (Int i = 0; i < n; ++i)
            for (LInt l = L.Outer()[i]; l < L.Outer()[i + 1]; ++l)
                if (L.Inner()[l] == i) diag[i] = 1.0 / L.Values()[l];
        DiagonalPreconditioner<Real> D(diag);
        Tensor2<Real, Int> X4(n, k, Real(0));
        const bool ok4 = cg(L, D, Real(1), B.data(), k, Real(0), X4.data(), k, Real(1e-10));
        const double e4 = relerr(X4.data(), X.data(), (size_t)n * k);
        check(ok4 && e4 < 1e-8, "CG(MatrixCSR, Jacobi) matches", e4);
    }

    // --- GMRES (GMRES.hpp:127-260) -------------------------------------------------
    {
        const Int k = 1;
        Tensor2<Real, Int> B(n, k), X(n, k, Real(0)), R(n, k);
        B.Random();
        IdentityPreconditioner P;
        GMRES<0, Real, Int, Side::Left> gm(n, 30, k, 4);
        const bool ok = gm(L, P, Real(1), B.data(), k, Real(0), X.data(), k, Real(1e-10), 50);
        R = B;
        brute(L, -1, X.data(), k, 1, R.data(), k, k);
        const double res = R.MaxNorm() / B.MaxNorm();
        check(ok && res < 1e-8, "GMRES(MatrixCSR) residual", res);
        auto A_host = [&](const Real* x, Real* y) { L.Dot(Real(1), x, Real(0), y, k); };
        Tensor2<Real, Int> X2(n, k, Real(0));
        const bool ok2 = gm(A_host, P, Real(1), B.data(), k, Real(0), X2.data(), k, Real(1e-10), 50);
        const double e2 = relerr(X2.data(), X.data(), n);
        check(ok2 && e2 < 1e-12, "GMRES(host lambda) matches", e2);
    }

    // --- Solver accessors at a fixed iteration count -----------------------------
    // Residuals (ConjugateGradient.hpp:447-456; GMRES.hpp:779-789), GetBasis,
    // GetHessenbergMatrix and Solution (GMRES.hpp:820-833) after tol = 0 runs; argv[2]
    // names a directory that receives them (raw float64) for the pytest to compare with
    // the oracle's values on the same inputs.
    {
        const Int k = 2, steps = 20, m = 10;
        Tensor2<Real, Int> B(n, k), X(n, k, Real(0));
        B.Random();
        IdentityPreconditioner P;
        ConjugateGradient<0, Real, Int> cg(n, steps, k, 4);
        const bool ok = cg(L, P, Real(1), B.data(), k, Real(0), X.data(), k, Real(0));
        const auto res = cg.Residuals();
        const auto rel = cg.RelativeResiduals();
        bool pos = true;
        for (Int c = 0; c < k; ++c) pos &= res[c] > 0 && rel[c] > 0 && res[c] > rel[c];
        check(!ok && cg.IterationCount() == steps && pos, "CG Residuals() after 20 steps");
        GMRES<0, Real, Int, Side::Left> gm(n, m, k, 4);
        Tensor2<Real, Int> Xg(n, k, Real(0));
        const bool okg = gm(L, P, Real(1), B.data(), k, Real(0), Xg.data(), k, Real(0), 1);
        const auto& Q = gm.GetBasis();
        const auto& H = gm.GetHessenbergMatrix();
        const auto& S = gm.Solution();
        double orth = 0;  // columns of the last cycle's basis are orthonormal
        for (Int c = 0; c < k; ++c)
            for (Int i = 0; i <= m; ++i)
                for (Int j = 0; j <= m; ++j) {
                    double d = 0;
                    for (Int r = 0; r < n; ++r) d += Q(i, r, c) * Q(j, r, c);
                    orth = std::max(orth, std::abs(d - (i == j ? 1.0 : 0.0)));
                }
        check(!okg && gm.IterationCount() == m && gm.RestartCount() == 1 && Q.Dim(0) == m + 1 && Q.Dim(1) == n &&
                  H.Dim(0) == m + 1 && H.Dim(1) == m && orth < 1e-12,
              "GMRES GetBasis / GetHessenbergMatrix", orth);
        check(relerr(S.data(), Xg.data(), (size_t)n * k) == 0.0, "GMRES Solution() == X (a = 1, b = 0)");
        if (argc > 2) {
            const std::string dir = argv[2];
            auto dump = [&](const char* name, const double* p, size_t count) {
                FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
                if (f) {
                    std::fwrite(p, sizeof(double), count, f);
                    std::fclose(f);
                }
            };
            const auto gres = gm.Residuals();
            const auto grel = gm.RelativeResiduals();
            dump("B.bin", B.data(), (size_t)n * k);
            dump("cg_X.bin", X.data(), (size_t)n * k);
            dump("cg_res.bin", res.data(), (size_t)k);
            dump("cg_rel.bin", rel.data(), (size_t)k);
            dump("gm_X.bin", Xg.data(), (size_t)n * k);
            dump("gm_res.bin", gres.data(), (size_t)k);
            dump("gm_rel.bin", grel.data(), (size_t)k);
            dump("gm_Q.bin", Q.data(), (size_t)Q.Size());
            dump("gm_H.bin", H.data(), (size_t)H.Size());
            dump("gm_sol.bin", S.data(), (size_t)S.Size());
        }
    }

    std::printf(failures ? "FAILED %d\n" : "ALL OK\n", failures);
    return failures ? 1 : 0;
}
