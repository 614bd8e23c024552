// GridLaplacian drop-in: the same triple push sequence as the reference
// (src/Sparse/GridLaplacian.hpp:11-158) -- graph Laplacian plus mass on a
// grid_size x grid_size grid, vertex V = grid_size * i + j -- assembled into CSR
// by the GPU triples -> CSR path of MatrixCSR.
#pragma once

#include "MatrixCSR.hpp"

namespace Tensors {
namespace Sparse {

template <typename Scal, typename Int, typename LInt, Parallel_T parQ = Parallel>
MatrixCSR<Scal, Int, LInt, parQ> GridLaplacian(const Int grid_size, const Scal mass, const Int thread_count) {
    TripleAggregator<Int, Int, Scal, LInt> triples;
    for (Int i = 0; i < grid_size - 1; ++i)
        for (Int j = 0; j < grid_size - 1; ++j) {
            const Int V = grid_size * i + j, L = V + grid_size, R = V + 1;
            triples.Push(V, V, 1);
            triples.Push(V, R, -1);
            triples.Push(R, V, -1);
            triples.Push(R, R, 1);
            triples.Push(V, V, 1);
            triples.Push(V, L, -1);
            triples.Push(L, V, -1);
            triples.Push(L, L, 1);
        }
    for (Int i = 0; i < grid_size - 1; ++i) {
        const Int V = grid_size * i + (grid_size - 1), L = V + grid_size;
        triples.Push(V, V, 1);
        triples.Push(V, L, -1);
        triples.Push(L, V, -1);
        triples.Push(L, L, 1);
    }
    for (Int j = 0; j < grid_size - 1; ++j) {
        const Int V = grid_size * (grid_size - 1) + j, R = V + 1;
        triples.Push(V, V, 1);
        triples.Push(V, R, -1);
        triples.Push(R, V, -1);
        triples.Push(R, R, 1);
    }
    for (Int i = 0; i < grid_size; ++i)
        for (Int j = 0; j < grid_size; ++j) triples.Push(grid_size * i + j, grid_size * i + j, mass);
    return MatrixCSR<Scal, Int, LInt, parQ>(triples, grid_size * grid_size, grid_size * grid_size, thread_count, true,
                                            0, false);
}

// The upper-triangle assembly with symmetrizeQ (GridLaplacian.hpp:85-158); it
// yields the same matrix as GridLaplacian (Experiment_SpMM/main.cpp:55-57).
template <typename Scal, typename Int, typename LInt, Parallel_T parQ = Parallel>
MatrixCSR<Scal, Int, LInt, parQ> GridLaplacian_Parallel(const Int grid_size, const Scal mass, const Int thread_count) {
    const Int T = std::max<Int>(1, thread_count);
    std::vector<TripleAggregator<Int, Int, Scal, LInt>> lists((Size_T)T);
    auto jp = [](Int n, Int T_, Int t) { return (n / T_) * t + ((n % T_) * t) / T_; };
    for (Int t = 0; t < T; ++t) {
        auto& tr = lists[(Size_T)t];
        for (Int i = jp(grid_size - 1, T, t); i < jp(grid_size - 1, T, t + 1); ++i) {
            for (Int j = 0; j < grid_size - 1; ++j) {
                const Int V = grid_size * i + j, L = V + grid_size, R = V + 1;
                tr.Push(V, V, Scal(2) + mass);
                tr.Push(V, R, Scal(-1));
                tr.Push(R, R, Scal(1));
                tr.Push(V, L, Scal(-1));
                tr.Push(L, L, Scal(1));
            }
            const Int V = grid_size * i + grid_size - 1, L = V + grid_size;
            tr.Push(V, V, Scal(1) + mass);
            tr.Push(V, L, Scal(-1));
            tr.Push(L, L, Scal(1));
        }
        if (t + 1 == T) {
            for (Int j = 0; j < grid_size - 1; ++j) {
                const Int V = grid_size * (grid_size - 1) + j, R = V + 1;
                tr.Push(V, V, Scal(1) + mass);
                tr.Push(V, R, Scal(-1));
                tr.Push(R, R, Scal(1));
            }
            tr.Push(grid_size * grid_size - 1, grid_size * grid_size - 1, mass);
        }
    }
    return MatrixCSR<Scal, Int, LInt, parQ>(lists, grid_size * grid_size, grid_size * grid_size, thread_count, true, 1);
}

}  // namespace Sparse
}  // namespace Tensors
