// Lattice restriction of the learned coarse correction, per 256-node block: shared by the
// standalone restriction kernel and the residual update that produces r (no counterpart in the
// reference).
#pragma once

#include <hip/hip_runtime.h>

namespace psmas
{
    constexpr int LCC_TMAX = 64; ///< distinct lattice points one node block may touch

    /// Trilinear weight of corner (bit 0: x, bit 1: y, bit 2: z) at cell offsets (fx, fy, fz).
    __device__ __forceinline__ float trilinear_weight(int corner, float fx, float fy, float fz)
    {
        const float wx = (corner & 1) ? fx : 1.0f - fx;
        const float wy = (corner & 2) ? fy : 1.0f - fy;
        const float wz = (corner & 4) ? fz : 1.0f - fz;
        return wx * wy * wz;
    }

    /// Block path of the lattice restriction: per slot the lattice cell {base point, fx, fy, fz}
    /// (base < 0: not restricted) and the LDS position of each of its 8 corner terms (16 bits
    /// each); per block the number of distinct targets nt and, at a fixed stride of LCC_TMAX,
    /// the block-local end of each target's term range (target t covers [tend[t-1], tend[t]),
    /// tend[-1] = 0; entries past nt repeat the block's term count), so that a thread's entry is
    /// addressed without a prior load.
    struct LatticeBlocks
    {
        const int4 *cell = nullptr;
        const uint4 *pos = nullptr;
        const int *block_nt = nullptr;
        const int *tend = nullptr; ///< [block][LCC_TMAX]
        double *partial = nullptr; ///< [block][LCC_TMAX][bd]
    };

} // namespace psmas
