// Device-side building blocks shared by the MAS kernels: wave64 and workgroup reductions and a
// deterministic grid reduction (per-workgroup partials summed in a fixed order by the last
// workgroup to arrive), plus the solver's device-scalar slots.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace psmas
{
    constexpr int WAVE = 64;

    /// Device scalar slots (double) used by the PCG loops.  All reductions of one iteration
    /// land here so that alpha / beta never travel to the host.
    enum Scalar : int
    {
        S_RZ = 0,      ///< r . z (current)
        S_RZ_OLD = 1,  ///< r . z of the previous iteration
        S_PAP = 2,     ///< p . A p
        S_ALPHA = 3,
        S_BETA = 4,
        S_RR = 5,      ///< r . r of the recurrence residual
        S_RR_TRUE = 6, ///< r . r of an explicitly computed residual b - A x (FP64)
        S_RZ0 = 7,
        S_RR0 = 8,
        S_RZ_PREV_Z = 9, ///< r_new . z_old (flexible PCG)
        S_DELTA_R = 10,  ///< learned correction . restricted r
        S_DELTA_R_OLD = 11,
        S_GUARD = 12,    ///< division guard (see safe_div)
        S_COUNT = 16,
    };

    /// Reduction sites: each owns a partials array and an arrival counter.
    enum Site : int
    {
        R_PAP = 0,
        R_RR = 1,
        R_RZ = 2,
        R_RR_TRUE = 3,
        R_RZ_PREV = 4,
        R_COUNT = 5,
    };

    __device__ __forceinline__ double wave_sum(double v)
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off, WAVE);
        return v;
    }

    /// Sum over the workgroup; the result is valid in thread 0.  Fixed combination order.
    template <int BLOCK>
    __device__ __forceinline__ double block_sum(double v)
    {
        static_assert(BLOCK % WAVE == 0 && BLOCK <= 1024, "block size");
        constexpr int NW = BLOCK / WAVE;
        __shared__ double warp_part[NW];
        v = wave_sum(v);
        const int lane = threadIdx.x % WAVE;
        const int w = threadIdx.x / WAVE;
        if (NW == 1)
            return v;
        if (lane == 0)
            warp_part[w] = v;
        __syncthreads();
        double s = 0.0;
        if (threadIdx.x == 0)
        {
#pragma unroll
            for (int i = 0; i < NW; ++i)
                s += warp_part[i];
        }
        __syncthreads();
        return s;
    }

    /// Grid-wide deterministic sum.  Every workgroup calls this once with its threads' values.
    /// The last workgroup to arrive adds all partials in index order and calls epilogue(total)
    /// from thread 0, then rearms the counter.  Returns true in the workgroup that finished.
    template <int BLOCK, typename Epilogue>
    __device__ __forceinline__ void grid_sum(double v, double *partials, unsigned *counter, Epilogue epilogue)
    {
        __shared__ bool am_last;
        const double s = block_sum<BLOCK>(v);
        if (threadIdx.x == 0)
        {
            __hip_atomic_store(partials + blockIdx.x, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned prev =
                __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            am_last = (prev == gridDim.x - 1);
        }
        __syncthreads();
        if (!am_last)
            return;
        double t = 0.0;
        for (unsigned i = threadIdx.x; i < gridDim.x; i += BLOCK)
            t += __hip_atomic_load(partials + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = block_sum<BLOCK>(t);
        if (threadIdx.x == 0)
        {
            epilogue(t);
            __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    /// Per-system deterministic sum over the node blocks of system `sys` (the workgroup's own).
    /// The system's last workgroup to arrive sums the partials of its blocks in index order and
    /// calls epilogue(total) from thread 0.
    template <int BLOCK, typename Epilogue>
    __device__ __forceinline__ void system_sum(double v, int sys, const SysMap &sm, double *partials,
                                               unsigned *counters, Epilogue epilogue)
    {
        __shared__ bool am_last;
        const double s = block_sum<BLOCK>(v);
        if (threadIdx.x == 0)
        {
            __hip_atomic_store(partials + blockIdx.x, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned prev =
                __hip_atomic_fetch_add(counters + sys, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            am_last = (prev == static_cast<unsigned>(sm.blk_count[sys]) - 1u);
        }
        __syncthreads();
        if (!am_last)
            return;
        const int b0 = sm.blk_begin[sys], nb = sm.blk_count[sys];
        double t = 0.0;
        for (int i = threadIdx.x; i < nb; i += BLOCK)
            t += __hip_atomic_load(partials + b0 + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = block_sum<BLOCK>(t);
        if (threadIdx.x == 0)
        {
            epilogue(t);
            __hip_atomic_store(counters + sys, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    /// alpha / beta division.  The reference returns 0 when |den| < 1e-20 (MASSolver.cu:43-56)
    /// because convergence is only checked every few iterations and the step must vanish once
    /// the solution is reached.  guard > 0 reproduces that absolute rule exactly; guard == 0
    /// (default) returns 0 only when the quotient is not finite, which keeps the purpose without
    /// the absolute scale (with 1e-20 the iteration stalls once p^T A p < 1e-20, e.g. near
    /// ||r|| ~ 1e-10 for gravity-loaded elasticity with moduli ~1e5).
    __device__ __forceinline__ double safe_div(double num, double den, double guard)
    {
        if (guard > 0.0)
            return (fabs(den) < guard) ? 0.0 : num / den;
        const double q = num / den;
        return isfinite(q) ? q : 0.0;
    }

    template <typename T>
    __device__ __forceinline__ T ldg_stream(const T *p)
    {
        return __builtin_nontemporal_load(p);
    }

    __device__ __forceinline__ float to_float(double v) { return static_cast<float>(v); }
    __device__ __forceinline__ float to_float(float v) { return v; }
    __device__ __forceinline__ float to_float(__half v) { return __half2float(v); }
} // namespace psmas
