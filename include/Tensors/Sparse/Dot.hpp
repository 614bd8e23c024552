// Sparse::Dot<opA, opB>(A, B) drop-in (Sparse/Dot.hpp:280-323): the product of
// op_A(A) and op_B(B) for real CSR matrices, each operand transposed on the GPU
// where its tag asks for it (MatrixCSR::Op) and multiplied by the GPU SpGEMM
// (MatrixCSR::Dot(MatrixCSR)).
#pragma once

#include "MatrixCSR.hpp"

namespace Tensors {
namespace Sparse {

template <Op opA, Op opB, typename Scal, IntQ Int, IntQ LInt, Parallel_T parQ>
MatrixCSR<Scal, Int, LInt, parQ> Dot(const MatrixCSR<Scal, Int, LInt, parQ>& A,
                                     const MatrixCSR<Scal, Int, LInt, parQ>& B) {
    if constexpr (opA == Op::Id && opB == Op::Id)
        return A.Dot(B);
    else if constexpr (opA == Op::Id)
        return A.Dot(B.template Op<opB>());
    else if constexpr (opB == Op::Id)
        return A.template Op<opA>().Dot(B);
    else
        return A.template Op<opA>().Dot(B.template Op<opB>());
}

}  // namespace Sparse
}  // namespace Tensors
