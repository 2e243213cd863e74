// Device side of the learned correction's lattice block restriction, shared by the standalone
// restriction kernel and the residual update that produces r (no counterpart in the reference).
#pragma once

#include "device.hpp"
#include "lattice.hpp"

namespace psmas
{
    /// What a slot's thread needs of the block path, loaded by the caller before anything else
    /// so that the loads overlap its own (no address depends on a loaded value).
    struct LatticeSlot
    {
        int nt = 0;
        int4 c{-1, 0, 0, 0};
        uint4 P{0u, 0u, 0u, 0u};
        int tend = 0; ///< entry threadIdx.x % LCC_TMAX of the block's target ends
    };
    __device__ __forceinline__ LatticeSlot lattice_slot(const LatticeBlocks &lb, int b)
    {
        LatticeSlot s;
        const size_t slot = static_cast<size_t>(b) * 256 + threadIdx.x;
        s.nt = lb.block_nt[b];
        s.c = lb.cell[slot];
        s.P = lb.pos[slot];
        s.tend = lb.tend[static_cast<size_t>(b) * LCC_TMAX + threadIdx.x % LCC_TMAX];
        return s;
    }

    /// partial[b][t][d] = sum over block b's restricted slots of w_q * r[d] for the corners q
    /// whose block-local lattice point is t (t < nt).  Thread i of the 256 owns slot b*256 + i
    /// with values r and its LatticeSlot s; each writes its 8 weighted terms to their positions
    /// in `vals` (LDS, 256*8*BD), so target t's terms form a contiguous range.  Each 16-lane row
    /// of a wave then sums one target (wave w takes targets 4w .. 4w+3, then 16 further): lanes
    /// stride the range, accumulating in T, and a fixed-order row reduction finishes.  Called by
    /// all 256 threads (has a barrier).
    template <int BD, typename T>
    __device__ __forceinline__ void lattice_block_partials(const LatticeBlocks &lb, int b, const LatticeSlot &s,
                                                           const T (&r)[BD], T *vals)
    {
        const int nt = s.nt;
        if (nt == 0)
            return; // uniform over the block
        __shared__ int tp[LCC_TMAX + 1];
        if (threadIdx.x == 0)
            tp[0] = 0;
        if (threadIdx.x < LCC_TMAX)
            tp[threadIdx.x + 1] = s.tend;
        if (s.c.x >= 0)
        {
            const float fx = __int_as_float(s.c.y), fy = __int_as_float(s.c.z), fz = __int_as_float(s.c.w);
            const unsigned pk[4] = {s.P.x, s.P.y, s.P.z, s.P.w};
#pragma unroll
            for (int q = 0; q < 8; ++q)
            {
                const int at = static_cast<int>((pk[q >> 1] >> (16 * (q & 1))) & 0xffffu);
                const T w = static_cast<T>(trilinear_weight(q, fx, fy, fz));
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    vals[at * BD + d] = w * r[d];
            }
        }
        __syncthreads();
        const int lane = threadIdx.x % WAVE, rl = lane & 15;
        for (int t0 = static_cast<int>(threadIdx.x / WAVE) * 4; t0 < nt; t0 += 16) // wave-uniform
        {
            const int t = t0 + (lane >> 4);
            const bool tv = t < nt;
            const int e0 = tv ? tp[t] : 0, e1 = tv ? tp[t + 1] : 0;
            T a[BD];
#pragma unroll
            for (int d = 0; d < BD; ++d)
                a[d] = T(0);
            for (int e = e0 + rl; e < e1; e += 16)
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    a[d] += vals[e * BD + d];
#pragma unroll
            for (int d = 0; d < BD; ++d)
                a[d] = row_sum16(a[d]);
            if (tv && rl == 0)
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    lb.partial[(static_cast<size_t>(b) * LCC_TMAX + t) * BD + d] = static_cast<double>(a[d]);
        }
    }
} // namespace psmas
