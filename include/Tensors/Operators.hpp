// Operator adapters for the drop-in solvers.
//
// The reference solvers take any callables A(cptr<Scal> x, mptr<Scal> y) and
// P(cptr<Scal> x, mptr<Scal> y) on n x nrhs row-major host blocks
// (ConjugateGradient.hpp:120-128, GMRES.hpp:143-151).  Here the Krylov vectors
// live on the GPU, so an operator is adapted to a tg_apply_fn callback that the
// library calls with device pointers:
//  * Sparse::MatrixCSR          -- runs the CSR kernels directly (the solver even
//                                 skips the callback and uses the built-in CSR path);
//  * IdentityPreconditioner      -- copy (or the built-in identity path);
//  * DiagonalPreconditioner      -- z_i = d_i r_i (built-in path for CSR operators);
//  * DeviceOperator(f)           -- f is called with the device pointers;
//  * any other callable          -- staged: x is copied to host memory, f runs on
//                                 host pointers, y is copied back (correct, slow).
#pragma once

#include "Sparse/MatrixCSR.hpp"

namespace Tensors {

struct IdentityPreconditioner {
    // host form, for code that calls the preconditioner itself
    template <typename T> void operator()(const T*, T*) const {}
};

template <typename Scal> class DiagonalPreconditioner {
  public:
    DiagonalPreconditioner() = default;
    template <typename S, IntQ I> DiagonalPreconditioner(const S* diag, I n) : n_((Size_T)n) {
        std::vector<Scal> h(n_);
        for (Size_T i = 0; i < n_; ++i) h[i] = (Scal)diag[i];
        d_.Upload(h.data(), n_);
    }
    template <typename S, typename I> explicit DiagonalPreconditioner(const Tensor1<S, I>& diag)
        : DiagonalPreconditioner(diag.data(), diag.Size()) {}
    const Scal* DeviceData() const { return d_.data(); }
    Size_T Size() const { return n_; }

  private:
    Size_T n_ = 0;
    Device::Buffer<Scal> d_;
};

template <typename F> struct DeviceOperator_T {
    F f;
};
// Marks a callable that accepts DEVICE pointers (x, y), e.g. a lambda calling
// MatrixCSR::Dot on them; no host staging is done for it.
template <typename F> DeviceOperator_T<std::decay_t<F>> DeviceOperator(F&& f) { return {std::forward<F>(f)}; }

namespace detail {

template <typename T> struct is_csr : std::false_type {};
template <typename S, typename I, typename L, Parallel_T p> struct is_csr<Sparse::MatrixCSR<S, I, L, p>> : std::true_type {};
template <typename T> struct is_diag : std::false_type {};
template <typename S> struct is_diag<DiagonalPreconditioner<S>> : std::true_type {};
template <typename T> struct is_devop : std::false_type {};
template <typename F> struct is_devop<DeviceOperator_T<F>> : std::true_type {};

template <typename Scal, typename Op> struct OpAdapter {
    Op* op;
    Size_T n;
    std::vector<Scal> hx, hy;
    OpAdapter(Op* o, Size_T n_) : op(o), n(n_) {}

    static int call(void* ctx, const void* x, void* y, int32_t k, void* stream) {
        auto* self = static_cast<OpAdapter*>(ctx);
        return self->apply((const Scal*)x, (Scal*)y, k, (hipStream_t)stream);
    }

    int apply(const Scal* x, Scal* y, int32_t k, hipStream_t s) {
        using O = std::decay_t<Op>;
        const Size_T N = n * (Size_T)k;
        if constexpr (is_csr<O>::value) {
            const auto& A = *op;
            return tg_csr_spmm((int32_t)A.RowCount(), (int32_t)A.ColCount(), (int64_t)A.NonzeroCount(),
                               A.DeviceRowPointers(), A.DeviceRowPointerType(), A.DeviceColumnIndices(),
                               A.DeviceValues(), Device::DType<Scal>(), 1.0, x, k, 0.0, y, k, k, 0, nullptr, 0, s);
        } else if constexpr (std::is_same_v<O, IdentityPreconditioner>) {
            return hipMemcpyAsync(y, x, N * sizeof(Scal), hipMemcpyDeviceToDevice, s) == hipSuccess ? 0 : 1;
        } else if constexpr (is_diag<O>::value) {
            // y = d .* x per row: a diagonal CSR with one entry per row
            if (rp.size() < n + 1) {
                std::vector<int32_t> r(n + 1), c(n);
                for (Size_T i = 0; i <= n; ++i) r[i] = (int32_t)i;
                for (Size_T i = 0; i < n; ++i) c[i] = (int32_t)i;
                rp.Upload(r.data(), r.size());
                ci.Upload(c.data(), c.size());
            }
            return tg_csr_spmm((int32_t)n, (int32_t)n, (int64_t)n, rp.data(), TG_I32, ci.data(), op->DeviceData(),
                               Device::DType<Scal>(), 1.0, x, k, 0.0, y, k, k, 0, nullptr, 0, s);
        } else if constexpr (is_devop<O>::value) {
            (void)op->f(x, y);
            return 0;
        } else {
            hx.resize(N);
            hy.resize(N);
            if (hipMemcpyAsync(hx.data(), x, N * sizeof(Scal), hipMemcpyDeviceToHost, s) != hipSuccess) return 1;
            if (hipStreamSynchronize(s) != hipSuccess) return 1;
            (void)(*op)(hx.data(), hy.data());
            if (hipMemcpyAsync(y, hy.data(), N * sizeof(Scal), hipMemcpyHostToDevice, s) != hipSuccess) return 1;
            return hipStreamSynchronize(s) == hipSuccess ? 0 : 1;
        }
    }
    Device::Buffer<int32_t> rp, ci;
};

// Stages a host n x k block (leading dimension ld) into a device block with
// ld = k, or passes a device pointer through.
template <typename T> struct Staged {
    T* ptr = nullptr;
    int64_t ld = 0;
    bool host = false;
    Device::Buffer<T> buf;
    Staged(const T* p, int64_t ld_, Size_T rows, Size_T k, bool copy_in) {
        if (Device::IsDevicePointer(p)) {
            ptr = const_cast<T*>(p);
            ld = ld_;
        } else {
            host = true;
            buf.Resize(rows * k);
            ptr = buf.data();
            ld = (int64_t)k;
            if (copy_in) Device::Copy2D(ptr, k, p, (Size_T)ld_, rows, k, hipMemcpyHostToDevice);
        }
    }
};

}  // namespace detail
}  // namespace Tensors
