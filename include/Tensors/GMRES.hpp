// GMRES drop-in: the class template, constructor, operator() and accessors of
// the reference (GMRES.hpp:18-879), solving on the GPU through tg_gmres_solve
// (CSR operator + identity/diagonal preconditioner) or tg_gmres_solve_op (any
// other operator / preconditioner, see Operators.hpp).
//
//   X_inout <- a * A^{-1} B_in + b * X_inout      (n x nrhs, row-major, ld)
//
// Restart length m = min(max_iter, n + 1) (GMRES.hpp:89).  B_in / X_inout may
// be host or device pointers.
#pragma once

#include "Operators.hpp"

namespace Tensors {

template <Size_T NRHS_, typename Scal_, typename Int_, Side side = Side::Left, bool A_verboseQ = true,
          bool P_verboseQ = true>
class GMRES final {
  public:
    using Scal = Scal_;
    using Real = Scal_;
    using Int = Int_;
    static constexpr Int NRHS = (Int)NRHS_;
    using RealVector_T = Tensor1<Real, Int>;
    static_assert(std::is_same_v<Scal, double> || std::is_same_v<Scal, float>, "real double/float only");

    GMRES(const Int n_, const Int max_iter_, const Size_T nrhs_ = NRHS_, const Size_T thread_count_ = 1)
        : n(n_), max_iter(std::min<Int>(max_iter_, n_ + 1)), nrhs(NRHS > 0 ? NRHS : (Int)nrhs_),
          thread_count((Int)thread_count_), rel64(nrhs), res64(nrhs) {
        ws.Resize(tg_gmres_workspace_bytes((int32_t)n, (int32_t)max_iter, (int32_t)nrhs, Device::DType<Scal>()));
    }

    template <typename Operator_T, typename Preconditioner_T, typename a_T, typename B_T, typename b_T, typename X_T>
    bool operator()(Operator_T& A, Preconditioner_T& P, const a_T a, cptr<B_T> B_in, const Int ldB, const b_T b,
                    mptr<X_T> X_inout, const Int ldX, const Real relative_tolerance_, const Int max_restarts_,
                    const bool use_initial_guessQ_ = false) {
        static_assert(std::is_same_v<B_T, Scal> && std::is_same_v<X_T, Scal>, "B and X must have the scalar type");
        relative_tolerance = relative_tolerance_;
        max_restarts = max_restarts_;
        use_initial_guessQ = use_initial_guessQ_;
        if (use_initial_guessQ && (double)b != 0.0)
            wprint(ClassName() + ": use_initial_guessQ == true and b != 0. Typically, this does not make sense.");
        detail::Staged<Scal> Bd(B_in, ldB, (Size_T)n, (Size_T)nrhs, true);
        detail::Staged<Scal> Xd(X_inout, ldX, (Size_T)n, (Size_T)nrhs, (double)b != 0.0 || use_initial_guessQ);
        int32_t ok = 0, it = 0, rs = 0;
        const int32_t sd = side == Side::Left ? 0 : 1;
        using AO = std::decay_t<Operator_T>;
        using PO = std::decay_t<Preconditioner_T>;
        int st;
        if constexpr (detail::is_csr<AO>::value &&
                      (std::is_same_v<PO, IdentityPreconditioner> || detail::is_diag<PO>::value)) {
            const void* dptr = nullptr;
            int32_t pk = 0;
            if constexpr (detail::is_diag<PO>::value) {
                dptr = P.DeviceData();
                pk = 1;
            }
            st = tg_gmres_solve((int32_t)n, (int64_t)A.NonzeroCount(), A.DeviceRowPointers(),
                                A.DeviceRowPointerType(), A.DeviceColumnIndices(), A.DeviceValues(),
                                Device::DType<Scal>(), pk, dptr, sd, (int32_t)nrhs, (int32_t)max_iter, (double)a,
                                Bd.ptr, Bd.ld, (double)b, Xd.ptr, Xd.ld, (double)relative_tolerance,
                                (int32_t)max_restarts, use_initial_guessQ ? 1 : 0, ws.data(), ws.size(), &ok, &it, &rs,
                                rel64.data(), nullptr);
        } else {
            detail::OpAdapter<Scal, Operator_T> fa(&A, (Size_T)n);
            detail::OpAdapter<Scal, Preconditioner_T> fp(&P, (Size_T)n);
            const bool identity = std::is_same_v<PO, IdentityPreconditioner>;
            st = tg_gmres_solve_op((int32_t)n, &detail::OpAdapter<Scal, Operator_T>::call, &fa,
                                   identity ? nullptr : &detail::OpAdapter<Scal, Preconditioner_T>::call,
                                   identity ? nullptr : (void*)&fp, Device::DType<Scal>(), sd, (int32_t)nrhs,
                                   (int32_t)max_iter, (double)a, Bd.ptr, Bd.ld, (double)b, Xd.ptr, Xd.ld,
                                   (double)relative_tolerance, (int32_t)max_restarts, use_initial_guessQ ? 1 : 0,
                                   ws.data(), ws.size(), &ok, &it, &rs, rel64.data(), nullptr);
        }
        Device::Status(st, ClassName() + "::operator()");
        if (Xd.host) Device::Copy2D(X_inout, (Size_T)ldX, Xd.ptr, (Size_T)nrhs, (Size_T)n, (Size_T)nrhs,
                                    hipMemcpyDeviceToHost);
        iter = it;
        restarts = rs;
        succeeded = ok != 0;
        // |beta[iter][k]| for Residuals() (GMRES.hpp:779-789)
        double* beta_d = nullptr;
        Device::Status(tg_gmres_workspace_parts(ws.data(), ws.size(), (int32_t)n, (int32_t)max_iter, (int32_t)nrhs,
                                                Device::DType<Scal>(), nullptr, nullptr, &beta_d, nullptr),
                       ClassName() + "::Residuals");
        Device::Copy(res64.data(), beta_d + (Size_T)iter * (Size_T)nrhs, (Size_T)nrhs, hipMemcpyDeviceToHost);
        for (Int k = 0; k < nrhs; ++k) res64[k] = std::abs(res64[k]);
        return succeeded;
    }

    Int ThreadCount() const { return thread_count; }
    Int IterationCount() const { return iter; }
    Int RestartCount() const { return restarts; }
    // |beta[iter][k]| per column (GMRES.hpp:779-789).
    RealVector_T Residuals() const {
        RealVector_T r(nrhs);
        for (Int k = 0; k < nrhs; ++k) r[k] = (Real)res64[k];
        return r;
    }
    RealVector_T RelativeResiduals() const {
        RealVector_T r(nrhs);
        for (Int k = 0; k < nrhs; ++k) r[k] = (Real)rel64[k];
        return r;
    }

    // Host copies of the device workspace (GMRES.hpp:820-833): the basis Q of the last
    // cycle ((m+1) x n x nrhs), the Hessenberg matrix H after the Givens rotations
    // ((m+1) x m x nrhs) and the solution before the a/b combination (n x nrhs).
    const Tensor3<Scal, Int>& GetBasis() const {
        void* q = nullptr;
        Parts(&q, nullptr, nullptr);
        Q_host = Tensor3<Scal, Int>(max_iter_ws() + 1, n, nrhs);
        Device::Copy(Q_host.data(), (const Scal*)q, (Size_T)Q_host.Size(), hipMemcpyDeviceToHost);
        return Q_host;
    }
    const Tensor3<Scal, Int>& GetHessenbergMatrix() const {
        double* h = nullptr;
        Parts(nullptr, &h, nullptr);
        const Int m = max_iter_ws();
        std::vector<double> tmp((Size_T)(m + 1) * (Size_T)m * (Size_T)nrhs);
        Device::Copy(tmp.data(), (const double*)h, tmp.size(), hipMemcpyDeviceToHost);
        H_host = Tensor3<Scal, Int>(m + 1, m, nrhs);
        H_host.Read(tmp.data());
        return H_host;
    }
    const Tensor2<Scal, Int>& Solution() const {
        void* x = nullptr;
        Parts(nullptr, nullptr, &x);
        sol_host = Tensor2<Scal, Int>(n, nrhs);
        Device::Copy(sol_host.data(), (const Scal*)x, (Size_T)sol_host.Size(), hipMemcpyDeviceToHost);
        return sol_host;
    }

    std::string Stats() const {
        return "\n==== " + ClassName() + " Stats ====\n\n succeeded          = " +
               std::string(succeeded ? "true" : "false") + "\n n                  = " + ToString(n) +
               "\n nrhs               = " + ToString(nrhs) + "\n restarts           = " + ToString(restarts) +
               "\n max_restarts       = " + ToString(max_restarts) + "\n iter               = " + ToString(iter) +
               "\n max_iter           = " + ToString(max_iter) + "\n";
    }

    std::string ClassName() const {
        return std::string("GMRES<") + (NRHS <= 0 ? std::string("VarSize") : ToString(NRHS)) + "," +
               TypeName<Scal>() + "," + TypeName<Int>() + "," + (side == Side::Left ? "Left" : "Right") + ">(" +
               ToString(nrhs) + ")";
    }

  private:
    // restart length of the workspace: m = min(max_iter, n + 1), at least 1
    Int max_iter_ws() const { return std::max<Int>(Int(1), max_iter); }
    void Parts(void** q, double** h, void** x) const {
        Device::Status(tg_gmres_workspace_parts((void*)ws.data(), ws.size(), (int32_t)n, (int32_t)max_iter,
                                                (int32_t)nrhs, Device::DType<Scal>(), q, h, nullptr, x),
                       ClassName() + "::workspace");
    }

    const Int n;
    const Int max_iter;
    const Int nrhs;
    const Int thread_count;
    Tensor1<double, Int> rel64;
    Tensor1<double, Int> res64;
    mutable Tensor3<Scal, Int> Q_host;
    mutable Tensor3<Scal, Int> H_host;
    mutable Tensor2<Scal, Int> sol_host;
    Device::Buffer<char> ws;
    Int iter = 0;
    Int restarts = 0;
    Int max_restarts = 30;
    bool succeeded = false;
    Real relative_tolerance = Real(0.00001);
    bool use_initial_guessQ = false;
};

}  // namespace Tensors
