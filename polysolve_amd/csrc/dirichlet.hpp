// FEM Dirichlet elimination on the GPU from a column-compressed matrix in device memory.
#pragma once

#include "common.hpp"

namespace psmas
{
    /// The elimination of eliminate_dirichlet (host_setup.hpp; FEMSolver.cpp:123-160 and
    /// 269-306), bit for bit, on device arrays.  A′ keeps a_ij when neither i nor j is fixed, has
    /// 1 on a fixed dof's diagonal (its only entry in row and column) and a_kk, or an explicit
    /// zero, on a free one; rows ascend in every column.  keep[k] is the position in A′ of A's k-th
    /// non-zero or -1.  The couplings a_ij (i free, j fixed) are grouped by row: crow[t] with
    /// entries cptr[t] .. cptr[t+1] holding ccol (ascending) and cidx (index among A's non-zeros).
    struct DirichletDevice
    {
        long long nnz = 0; ///< non-zeros of A′
        int nrows = 0;     ///< rows with couplings
        long long ncoupl = 0;
        DeviceBuffer<int> colptr, rowidx, keep;
        DeviceBuffer<double> vals;
        DeviceBuffer<int> crow, cptr, ccol, cidx;
    };

    /// A is n x n with nnz non-zeros (d_colptr, d_rowidx, d_vals, validated); d_nodes holds the
    /// nfix fixed dofs (in [0, n), duplicates allowed).
    void eliminate_dirichlet_device(int n, long long nnz, const int *d_colptr, const int *d_rowidx,
                                    const double *d_vals, const int *d_nodes, int nfix, hipStream_t s,
                                    DirichletDevice &out);
} // namespace psmas
