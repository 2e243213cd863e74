// Device side of the learned correction's lattice block restriction, shared by the standalone
// restriction kernel and the residual update that produces r (no counterpart in the reference).
#pragma once

#include "device.hpp"
#include "lattice.hpp"

namespace psmas
{
    /// partial[b][t][d] = sum over block b's restricted slots of w_q * r[d] for the corners q
    /// whose block-local lattice point is t (t < nt = block_nt[b]).  Thread i of the 256 owns slot
    /// b*256 + i with values r, lattice cell c and corner positions P (loaded by the caller); each
    /// writes its 8 weighted terms to their positions in `vals` (LDS, 256*8*BD), so target t's
    /// terms form a contiguous range.  Wave w then sums targets w, w+4, ...: lanes stride the
    /// range, a fixed-order wave reduction finishes.  Called by all 256 threads (has a barrier).
    template <int BD, typename T>
    __device__ __forceinline__ void lattice_block_partials(const LatticeBlocks &lb, int b, int nt, int4 c, uint4 P,
                                                           const T (&r)[BD], T *vals)
    {
        if (nt == 0)
            return; // uniform over the block
        // The targets' term ranges (nt <= LCC_TMAX) go to LDS with one load per thread, issued
        // before the terms are written, so the waves' loops below make no global round trips.
        __shared__ int tp[LCC_TMAX + 1];
        const int base = lb.tbase[b];
        if (threadIdx.x <= nt)
            tp[threadIdx.x] = lb.tptr[base + threadIdx.x];
        if (c.x >= 0)
        {
            const float fx = __int_as_float(c.y), fy = __int_as_float(c.z), fz = __int_as_float(c.w);
            const unsigned pk[4] = {P.x, P.y, P.z, P.w};
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
        const int e_begin = tp[0];
        const int lane = threadIdx.x % WAVE;
        for (int t = threadIdx.x / WAVE; t < nt; t += 256 / WAVE)
        {
            const int e0 = tp[t] - e_begin, e1 = tp[t + 1] - e_begin;
            double a[BD];
#pragma unroll
            for (int d = 0; d < BD; ++d)
                a[d] = 0.0;
            for (int e = e0 + lane; e < e1; e += WAVE)
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    a[d] += static_cast<double>(vals[e * BD + d]);
#pragma unroll
            for (int d = 0; d < BD; ++d)
            {
                const double v = wave_sum(a[d]);
                if (lane == 0)
                    lb.partial[(static_cast<size_t>(b) * LCC_TMAX + t) * BD + d] = v;
            }
        }
    }
} // namespace psmas
