// Umbrella header of the MI355X drop-in for the Tensors sparse hot path
// (the include surface of the reference's Tensors.hpp:5-51 restricted to the
// CSR SpMV/SpMM + CG/GMRES path, block CSR, MatrixMarket files).  Link with libtensors_amd.so.
#pragma once

#include "Base.hpp"
#include "Tensor.hpp"
#include "Sparse/MatrixCSR.hpp"
#include "Sparse/GridLaplacian.hpp"
#include "Sparse/Dot.hpp"
#include "BlockKernels.hpp"
#include "Sparse/PatternCSR.hpp"
#include "Sparse/KernelMatrixCSR.hpp"
#include "Operators.hpp"
#include "ConjugateGradient.hpp"
#include "GMRES.hpp"
