// Sparse::KernelMatrixCSR<Kernel_T> drop-in (KernelMatrixCSR.hpp:7-227): a block
// sparse matrix whose pattern is a PatternCSR of block positions and whose
// nonzeros are Kernel_T::ROWS x Kernel_T::COLS dense blocks, the values passed
// to each call.  Dot and FillLowerTriangleFromUpperTriangle run on the GPU
// (tg_bsr_spmm / tg_bsr_fill_lower); host pointers are staged through device
// buffers, device pointers are used in place.
//
// Differences from the reference: the pattern is held by value (the reference
// keeps a const reference, so it must outlive the matrix); Dot uses the
// kernel's compile-time NRHS for the chunk sizes, as the reference's Tiny
// kernels do, and the runtime nrhs only for the alpha == 0 scaling
// (KernelMatrixCSR.hpp:137-162); X == Y prints and does nothing.
#pragma once

#include <vector>

#include "../BlockKernels.hpp"
#include "PatternCSR.hpp"

namespace Tensors {
namespace Sparse {

template <typename Kernel_T> class KernelMatrixCSR final {
  public:
    using Scal = typename Kernel_T::Scal;
    using Int = typename Kernel_T::Int;
    using LInt = typename Kernel_T::LInt;
    using Scal_in = typename Kernel_T::Scal_in;
    using Scal_out = typename Kernel_T::Scal_out;
    static constexpr Parallel_T parQ = Parallel;
    using Pattern_T = Sparse::PatternCSR<Int, LInt, parQ>;

    explicit KernelMatrixCSR(const Pattern_T& pattern_) : pattern(pattern_) {}
    // From any CSR matrix with Outer() / Inner() (the reference's MatrixCSR is a PatternCSR).
    template <typename CSR_T>
    explicit KernelMatrixCSR(const CSR_T& A) : pattern(A) {}

    KernelMatrixCSR() = delete;

    Int RowCount() const { return pattern.RowCount() * Kernel_T::RowCount(); }
    Int ColCount() const { return pattern.ColCount() * Kernel_T::ColCount(); }
    LInt NonzeroCount() const { return pattern.NonzeroCount() * Kernel_T::BLOCK_NNZ; }
    const Pattern_T& Pattern() const { return pattern; }

    // Y <- beta Y over RowCount() * nrhs entries (KernelMatrixCSR.hpp:137-151).
    void Scale(mptr<Scal_out> Y, cref<Scal_out> beta, const Int nrhs) const {
        Dot(nullptr, Scal_out(0), nullptr, beta, Y, nrhs);
    }

    // Y <- alpha * A X + beta * Y (KernelMatrixCSR.hpp:153-227).
    void Dot(cptr<Scal> A, cref<Scal_out> alpha, cptr<Scal_in> X, cref<Scal_out> beta, mptr<Scal_out> Y,
             const Int nrhs) const {
        const std::string where = ClassName() + "::Dot";
        const bool scale_only = alpha == Scal_out(0) || NonzeroCount() <= LInt(0);
        const Int k = scale_only ? nrhs : Kernel_T::NRHS;
        if (!scale_only && X == Y) {
            eprint(where + ": Input and output pointer coincide. Doing nothing.");
            return;
        }
        const Size_T nA = (Size_T)NonzeroCount(), nX = (Size_T)ColCount() * (Size_T)k,
                     nY = (Size_T)RowCount() * (Size_T)k;
        const bool ad = !A || Device::IsDevicePointer(A), xd = !X || Device::IsDevicePointer(X),
                   yd = Device::IsDevicePointer(Y);
        Device::Buffer<Scal> dA, dX, dY;
        const Scal* Ad = A;
        const Scal* Xd = X;
        Scal* Yd = Y;
        if (!scale_only) {
            if (!ad) {
                dA.Upload(A, nA);
                Ad = dA.data();
            }
            if (!xd) {
                dX.Upload(X, nX);
                Xd = dX.data();
            }
        }
        if (!yd) {
            dY.Resize(nY);
            if (beta != Scal_out(0)) dY.Upload(Y, nY);
            Yd = dY.data();
        }
        Device::Status(tg_bsr_spmm((int32_t)pattern.RowCount(), (int32_t)pattern.ColCount(),
                                   (int64_t)pattern.NonzeroCount(), (int32_t)Kernel_T::ROWS, (int32_t)Kernel_T::COLS,
                                   pattern.DeviceRowPointers(), pattern.DeviceRowPointerType(),
                                   pattern.DeviceColumnIndices(), Ad, Device::DType<Scal>(), Kernel_T::A_ROW_MAJOR,
                                   (double)alpha, Xd, Kernel_T::X_ROW_MAJOR, (double)beta, Yd,
                                   Kernel_T::Y_ROW_MAJOR, (int32_t)k, nullptr),
                       where);
        if (!yd)
            dY.Download(Y, nY);
        else
            Device::Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    }

    // FillLowerTriangleFromUpperTriangle (KernelMatrixCSR.hpp:71-128): every block
    // (i, j), j < i, becomes the transpose of block (j, i).  Square blocks.
    void FillLowerTriangleFromUpperTriangle(mptr<Scal> values) const {
        static_assert(Kernel_T::ROWS == Kernel_T::COLS, "FillLowerTriangleFromUpperTriangle needs square blocks");
        if (!pattern.WellFormedQ() || pattern.RowCount() < pattern.ColCount()) return;
        const Size_T nA = (Size_T)NonzeroCount();
        const bool vd = Device::IsDevicePointer(values);
        Device::Buffer<Scal> dA;
        Scal* Ad = values;
        if (!vd) {
            dA.Upload(values, nA);
            Ad = dA.data();
        }
        Device::Status(tg_bsr_fill_lower((int32_t)pattern.RowCount(), pattern.DeviceRowPointers(),
                                         pattern.DeviceRowPointerType(), pattern.DeviceColumnIndices(),
                                         (int32_t)Kernel_T::ROWS, Ad, Device::DType<Scal>(), nullptr),
                       ClassName() + "::FillLowerTriangleFromUpperTriangle");
        if (!vd)
            dA.Download(values, nA);
        else
            Device::Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    }

    std::string ClassName() const { return "Sparse::KernelMatrixCSR<" + Kernel_T::ClassName() + ">"; }

  private:
    Pattern_T pattern;
};

}  // namespace Sparse
}  // namespace Tensors
