"""tensors_amd: the Tensors sparse hot path (CSR SpMV/SpMM behind MatrixCSR::Dot and the
ConjugateGradient / GMRES solvers) as hand-written HIP kernels for MI355X (gfx950).

The product is libtensors_amd.so (C ABI: include/tensors_amd.h) and the header-only
C++ drop-in (include/Tensors/).  This Python package binds the C ABI with ctypes for
tests and benchmarks.
"""
from . import _capi
from ._capi import TensorsError, lib, last_error

__all__ = ["TensorsError", "lib", "last_error"]
