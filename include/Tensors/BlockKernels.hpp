// Block kernels of Sparse::KernelMatrixCSR, drop-in form.
//
// The reference's kernel classes (BlockKernels/DenseBlockKernel_Tiny.hpp:24-206,
// BlockKernels/BlockKernel_Tiny.hpp:4-217) carry the block shape, the right-hand
// side count, the scalar types, the alpha/beta flags and the storage orders as
// template parameters, and apply one block at a time on the CPU.  Here the same
// template parameter list describes the product; KernelMatrixCSR hands it to one
// GPU launch (tg_bsr_spmm).  Only Scal == Scal_in == Scal_out (double or float)
// is on this path.  a_intRM, a_copy and x_prefetch only tune the CPU kernel and
// are accepted and ignored, as the reference itself ignores a_intRM and a_copy
// (DenseBlockKernel_Tiny.hpp:40-43).
#pragma once

#include "Base.hpp"

namespace Tensors {

namespace Scalar {
// The alpha/beta shortcuts of combine_buffers (Tools): Zero drops the term,
// Plus skips the multiplication by one, Generic multiplies.  The GPU path picks
// its shortcuts from the runtime values (alpha == 1, beta == 0 / 1), which
// agree with the flags whenever the flags describe the values.
enum class Flag : signed char { Minus = -1, Zero = 0, Plus = 1, Generic = 2 };
}  // namespace Scalar

template <int ROWS_, int COLS_, int NRHS_, typename Scal_, typename Scal_in_, typename Scal_out_, IntQ Int_,
          IntQ LInt_, Scalar::Flag alpha_flag, Scalar::Flag beta_flag, bool a_RM, bool a_intRM, bool a_copy,
          bool x_RM, bool x_prefetch, bool y_RM>
class DenseBlockKernel_Tiny {
    static_assert(std::is_same_v<Scal_, Scal_in_> && std::is_same_v<Scal_, Scal_out_>,
                  "the GPU block product needs one scalar type for A, X and Y");
    static_assert(std::is_same_v<Scal_, double> || std::is_same_v<Scal_, float>, "double or float");
    static_assert(ROWS_ >= 1 && COLS_ >= 1 && ROWS_ <= 64 && COLS_ <= 64, "block sizes 1..64");

  public:
    using Scal = Scal_;
    using Scal_in = Scal_in_;
    using Scal_out = Scal_out_;
    using Int = Int_;
    using LInt = LInt_;

    static constexpr Int ROWS = ROWS_;
    static constexpr Int COLS = COLS_;
    static constexpr Int NRHS = NRHS_;
    static constexpr Int MAX_NRHS = NRHS_;
    static constexpr Int ROWS_SIZE = ROWS_ * NRHS_;
    static constexpr Int COLS_SIZE = COLS_ * NRHS_;
    static constexpr LInt BLOCK_NNZ = (LInt)ROWS_ * (LInt)COLS_;
    static constexpr bool A_ROW_MAJOR = a_RM;
    static constexpr bool X_ROW_MAJOR = x_RM;
    static constexpr bool Y_ROW_MAJOR = y_RM;

    static constexpr Int RowCount() { return ROWS; }
    static constexpr Int ColCount() { return COLS; }

    static std::string ClassName() {
        return "DenseBlockKernel_Tiny<" + ToString(ROWS) + "," + ToString(COLS) + "," + ToString(NRHS) + "," +
               TypeName<Scal>() + "," + TypeName<Int>() + "," + TypeName<LInt>() + ">";
    }
};

}  // namespace Tensors
