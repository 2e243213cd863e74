// Graph aggregation into <=32-node domains on the GPU (no coordinates needed).
#pragma once

#include "host_setup.hpp"

#include <hip/hip_runtime.h>

namespace psmas
{
    /// Domains from the block graph of the n x n CSC matrix in device memory (d_colptr n+1,
    /// d_rowidx and d_vals nnz entries) with block size bd, each at most max_size nodes and
    /// connected.  Edge weight of blocks {u, v}: the sum over the scalars of blocks (u, v) and
    /// (v, u) of |a| / max|a| quantised to 2^20 (the reference's adjacency weight,
    /// BSRAdjacency.cu:60-84, quantised as there so that sums are exact).  Rounds of handshake
    /// heavy-edge matching under the size cap (each unmatched domain proposes to its heaviest
    /// neighbour that fits; mutual proposals merge; four proposal passes per round), then the
    /// domain graph is contracted (radix sort of domain-pair keys, integer sums); until a round
    /// merges nothing.  Deterministic.  Replaces the reference's KaMinPar call
    /// (GraphPartition.cpp:176-289) and its host round trip of the adjacency
    /// (BSRAdjacency.cu:451-457).
    Partition gpu_aggregate(int n, long long nnz, int bd, const int *d_colptr, const int *d_rowidx,
                            const double *d_vals, int max_size, hipStream_t s);
} // namespace psmas
