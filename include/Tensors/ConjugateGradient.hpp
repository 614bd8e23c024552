// ConjugateGradient drop-in: the class template, constructor, operator() and
// accessors of the reference (ConjugateGradient.hpp:16-509), solving on the GPU
// through tg_cg_solve (CSR operator + identity/diagonal preconditioner) or
// tg_cg_solve_op (any other operator / preconditioner, see Operators.hpp).
//
//   X_inout <- a * A^{-1} B_in + b * X_inout      (n x nrhs, row-major, ld)
//
// B_in / X_inout may be host or device pointers.  The device workspace
// (r, u, p, x, z and reduction buffers) is allocated once in the constructor,
// as the reference allocates its Tensor2 workspaces (:38-52, 64-88).
#pragma once

#include "Operators.hpp"

namespace Tensors {

template <Size_T NRHS_, typename Scal_, typename Int_, bool A_verboseQ = true, bool P_verboseQ = true>
class ConjugateGradient final {
  public:
    using Scal = Scal_;
    using Real = Scal_;
    using Int = Int_;
    static constexpr Int NRHS = (Int)NRHS_;
    using RealVector_T = Tensor1<Real, Int>;
    static_assert(std::is_same_v<Scal, double> || std::is_same_v<Scal, float>, "real double/float only");

    ConjugateGradient(const Int n_, const Int max_iter_, const Size_T eq_count_ = NRHS_,
                      const Size_T thread_count_ = 1)
        : n(n_), max_iter(std::max<Int>(0, max_iter_)), nrhs(NRHS > 0 ? NRHS : (Int)eq_count_),
          thread_count((Int)thread_count_), rel64(nrhs), res64(nrhs) {
        ws.Resize(tg_cg_workspace_bytes((int32_t)n, (int32_t)nrhs, Device::DType<Scal>(), 1));
    }

    template <typename Operator_T, typename Preconditioner_T, typename a_T, typename B_T, typename b_T, typename X_T>
    bool operator()(Operator_T& A, Preconditioner_T& P, const a_T a, cptr<B_T> B_in, const Int ldB, const b_T b,
                    mptr<X_T> X_inout, const Int ldX, const Real relative_tolerance_,
                    const bool use_initial_guessQ_ = false) {
        static_assert(std::is_same_v<B_T, Scal> && std::is_same_v<X_T, Scal>, "B and X must have the scalar type");
        relative_tolerance = relative_tolerance_;
        use_initial_guessQ = use_initial_guessQ_;
        if (use_initial_guessQ && (double)b != 0.0)
            wprint(ClassName() + ": use_initial_guessQ == true and b != 0. Typically, this does not make sense.");
        detail::Staged<Scal> Bd(B_in, ldB, (Size_T)n, (Size_T)nrhs, true);
        detail::Staged<Scal> Xd(X_inout, ldX, (Size_T)n, (Size_T)nrhs, (double)b != 0.0 || use_initial_guessQ);
        int32_t ok = 0, it = 0;
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
            st = tg_cg_solve((int32_t)n, (int64_t)A.NonzeroCount(), A.DeviceRowPointers(), A.DeviceRowPointerType(),
                             A.DeviceColumnIndices(), A.DeviceValues(), Device::DType<Scal>(), pk, dptr,
                             (int32_t)nrhs, (int32_t)max_iter, (double)a, Bd.ptr, Bd.ld, (double)b, Xd.ptr, Xd.ld,
                             (double)relative_tolerance, use_initial_guessQ ? 1 : 0, ws.data(), ws.size(), &ok, &it,
                             rel64.data(), nullptr);
        } else {
            detail::OpAdapter<Scal, Operator_T> fa(&A, (Size_T)n);
            detail::OpAdapter<Scal, Preconditioner_T> fp(&P, (Size_T)n);
            const bool identity = std::is_same_v<PO, IdentityPreconditioner>;
            st = tg_cg_solve_op((int32_t)n, &detail::OpAdapter<Scal, Operator_T>::call, &fa,
                                identity ? nullptr : &detail::OpAdapter<Scal, Preconditioner_T>::call,
                                identity ? nullptr : (void*)&fp, Device::DType<Scal>(), (int32_t)nrhs,
                                (int32_t)max_iter, (double)a, Bd.ptr, Bd.ld, (double)b, Xd.ptr, Xd.ld,
                                (double)relative_tolerance, use_initial_guessQ ? 1 : 0, ws.data(), ws.size(), &ok, &it,
                                rel64.data(), nullptr);
        }
        Device::Status(st, ClassName() + "::operator()");
        // sqrt|r.z| per column for Residuals() (the solver's final rho); with a == 0 the
        // reference returns before touching rho (:147-163), so the last values stay
        if ((double)a != 0.0)
            Device::Status(tg_cg_residuals(ws.data(), ws.size(), (int32_t)nrhs, res64.data(), nullptr, nullptr),
                           ClassName() + "::Residuals");
        if (Xd.host) Device::Copy2D(X_inout, (Size_T)ldX, Xd.ptr, (Size_T)nrhs, (Size_T)n, (Size_T)nrhs,
                                    hipMemcpyDeviceToHost);
        iter = it;
        succeeded = ok != 0;
        return succeeded;
    }

    Int IterationCount() const { return iter; }
    // sqrt|rho_k| per column (ConjugateGradient.hpp:447-456).
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
    bool CheckResiduals() const { return succeeded; }

    std::string Stats() const {
        return "\n==== " + ClassName() + " Stats ====\n\n n                  = " + ToString(n) +
               "\n nrhs               = " + ToString(nrhs) + "\n iter               = " + ToString(iter) +
               "\n max_iter           = " + ToString(max_iter) + "\n relative_tolerance = " +
               std::to_string(relative_tolerance) + "\n";
    }

    std::string ClassName() const {
        return std::string("ConjugateGradient<") + (NRHS <= 0 ? std::string("VarSize") : ToString(NRHS)) + "," +
               TypeName<Scal>() + "," + TypeName<Int>() + ">(" + ToString(nrhs) + ")";
    }

  private:
    const Int n;
    const Int max_iter;
    const Int nrhs;
    const Int thread_count;
    Tensor1<double, Int> rel64;
    Tensor1<double, Int> res64;
    Device::Buffer<char> ws;
    Int iter = 0;
    bool succeeded = false;
    Real relative_tolerance = Real(0.0001);
    bool use_initial_guessQ = false;
};

}  // namespace Tensors
