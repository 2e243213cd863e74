// Exact FP64 coarsest-level solve: setup here (coarse_solve.hip); the apply is side work of the
// second bank-apply launch (mas_kernels.hip).
#pragma once

#include "kernels.hpp"

#include <vector>

namespace psmas
{
    /// Largest number of active coarsest nodes handled by the dense solve (3x that many rows;
    /// at 512 nodes and block size 3 the FP64 inverse is 18.9 MB read once per apply).
    constexpr int DENSE_COARSE_MAX_NODES = 512;

    /// Full Galerkin matrix R A R^T over the active coarsest nodes into dense (nd x nd, zeroed).
    void launch_dense_galerkin(int bd, int npad, int level, const SellView<double> &A, const int *row_len,
                               const int *block_of_slot, const int4 *cmap, const int *active_index, int nd,
                               double *dense, hipStream_t s);

    /// In-place inverse of a dense SPD matrix via Cholesky (A = L L^T, A^-1 = L^-T L^-1).
    /// Returns false when the matrix is not numerically positive definite.
    bool cholesky_inverse(int n, std::vector<double> &A);
} // namespace psmas
