// Device-side building blocks shared by the MAS kernels: wave64 and workgroup reductions, the
// per-system partial-sum protocol of the PCG reductions, and the solver's device-scalar slots.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

namespace psmas
{
    constexpr int WAVE = 64;

    /// v from another lane of the same row, moved by DPP (CTRL: a DPP control word); a lane whose
    /// source is inactive reads 0.
    template <int CTRL>
    __device__ __forceinline__ double dpp_move(double v)
    {
        const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }

    __device__ __forceinline__ double read_lane(double v, int lane)
    {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                                __builtin_amdgcn_readlane(__double2loint(v), lane));
    }

    /// Sum over the 64 lanes of a wave, returned to every lane; all 64 lanes must be active.
    /// Within each 16-lane row, DPP moves add lane pairs (xor 1), quads (xor 2), then mirror the
    /// half rows and rows, so every lane holds its row's sum; the four row sums are then added
    /// as (r0 + r1) + (r2 + r3).  A fixed order, so every wave gets the same bits for the same
    /// values; no LDS traffic (a shuffle through ds_bpermute would be an LDS round trip per step).
    /// Sum over each 16-lane row of the wave, in every lane of the row (all lanes active).
    __device__ __forceinline__ double row_sum16(double v)
    {
        v += dpp_move<0xB1>(v);  // quad_perm [1, 0, 3, 2]
        v += dpp_move<0x4E>(v);  // quad_perm [2, 3, 0, 1]
        v += dpp_move<0x141>(v); // row_half_mirror
        v += dpp_move<0x140>(v); // row_mirror
        return v;
    }

    /// FP32 version of row_sum16 (same DPP row moves).
    __device__ __forceinline__ float row_sum16(float v)
    {
        auto mv = [](float x, auto ctrl) {
            return __int_as_float(
                __builtin_amdgcn_update_dpp(0, __float_as_int(x), decltype(ctrl)::value, 0xF, 0xF, false));
        };
        v += mv(v, std::integral_constant<int, 0xB1>());  // quad_perm [1, 0, 3, 2]
        v += mv(v, std::integral_constant<int, 0x4E>());  // quad_perm [2, 3, 0, 1]
        v += mv(v, std::integral_constant<int, 0x141>()); // row_half_mirror
        v += mv(v, std::integral_constant<int, 0x140>()); // row_mirror
        return v;
    }

    __device__ __forceinline__ double wave_sum(double v)
    {
        v = row_sum16(v);
        return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
    }

    /// FP32 version of wave_sum (same DPP row moves and row order).
    __device__ __forceinline__ float wave_sum(float v)
    {
        v = row_sum16(v);
        auto rl = [&](int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); };
        return (rl(0) + rl(16)) + (rl(32) + rl(48));
    }

    /// The two BF16 values packed in a 32-bit word, as FP32.
    __device__ __forceinline__ float bf16_lo(unsigned v) { return __uint_as_float(v << 16); }
    __device__ __forceinline__ float bf16_hi(unsigned v) { return __uint_as_float(v & 0xffff0000u); }

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

    /// The workgroup's sum of v -> partials[block] (block = the node block it processed).
    template <int BLOCK>
    __device__ __forceinline__ void store_partial(double v, double *partials, int block)
    {
        const double s = block_sum<BLOCK>(v);
        if (threadIdx.x == 0)
            partials[block] = s;
    }

    /// System of node block blk (SysMap; 0 for a single system, without a load).
    __device__ __forceinline__ int system_of_block(const SysMap &sm, int blk)
    {
        return sm.single_blocks > 0 ? 0 : sm.block_sys[blk];
    }

    /// Whether system sys has stopped iterating (never, while a single system's kernels run).
    __device__ __forceinline__ bool system_stopped(const SysMap &sm, int sys)
    {
        return sm.single_blocks == 0 && sm.active[sys] == 0;
    }

    /// NV workgroup sums at once -> parts[v][block] (parts[v] null: not stored); one barrier.
    /// Each value equals store_partial's for the same v bit for bit (same wave and wave order).
    template <int BLOCK, int NV>
    __device__ __forceinline__ void store_partials(const double (&v)[NV], double *const (&parts)[NV], int block)
    {
        static_assert(BLOCK % WAVE == 0 && BLOCK <= 1024, "block size");
        constexpr int NW = BLOCK / WAVE;
        __shared__ double part[NV][NW];
        const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
        double t[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k)
        {
            t[k] = wave_sum(v[k]);
            if (lane == 0)
                part[k][w] = t[k];
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
#pragma unroll
            for (int k = 0; k < NV; ++k)
            {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < NW; ++i)
                    s += part[k][i];
                if (parts[k] != nullptr)
                    parts[k][block] = NW == 1 ? t[k] : s;
            }
        }
    }

    /// Sum of the partials of system sys's node blocks, in block order, returned to every thread.
    /// The partials come from an earlier kernel; every workgroup computes the same value bit
    /// for bit.
    template <int BLOCK>
    __device__ __forceinline__ double system_total(const double *__restrict__ partials, int sys, const SysMap &sm)
    {
        __shared__ double total;
        const int b0 = sm.single_blocks > 0 ? 0 : sm.blk_begin[sys];
        const int nb = sm.single_blocks > 0 ? sm.single_blocks : sm.blk_count[sys];
        // The first PRE strides are loaded in one batch (one cache round trip instead of
        // one per stride; the partials come from another kernel, often from another XCD's
        // cache), then added in stride order.
        constexpr int PRE = 8;
        double pv[PRE];
#pragma unroll
        for (int k = 0; k < PRE; ++k)
            pv[k] = partials[b0 + min(static_cast<int>(threadIdx.x) + k * BLOCK, nb - 1)];
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < PRE; ++k)
            if (static_cast<int>(threadIdx.x) + k * BLOCK < nb)
                t += pv[k];
        for (int i = threadIdx.x + PRE * BLOCK; i < nb; i += BLOCK)
            t += partials[b0 + i];
        t = block_sum<BLOCK>(t);
        if (threadIdx.x == 0)
            total = t;
        __syncthreads();
        const double r = total;
        __syncthreads();
        return r;
    }

    /// NV system totals at once (one pass over the partials, one reduction): out[v] equals
    /// system_total(parts[v]) bit for bit (same per-thread order, wave reduction and wave order).
    template <int BLOCK, int NV>
    __device__ __forceinline__ void system_totals(const double *const (&parts)[NV], int sys, const SysMap &sm,
                                                  double (&out)[NV])
    {
        static_assert(BLOCK % WAVE == 0 && BLOCK <= 1024, "block size");
        constexpr int NW = BLOCK / WAVE;
        __shared__ double part[NV][NW];
        const int b0 = sm.single_blocks > 0 ? 0 : sm.blk_begin[sys];
        const int nb = sm.single_blocks > 0 ? sm.single_blocks : sm.blk_count[sys];
        // The first PRE strides of every array are loaded in one batch (one cache round trip
        // instead of one per stride: the partials come from another kernel, often from another
        // XCD's cache), then added in stride order; further strides follow one by one.
        constexpr int PRE = NV <= 2 ? 8 : 5;
        double pv[NV][PRE];
#pragma unroll
        for (int k = 0; k < PRE; ++k)
        {
            const int i = b0 + min(static_cast<int>(threadIdx.x) + k * BLOCK, nb - 1);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                pv[v][k] = parts[v][i];
        }
        double t[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
            t[v] = 0.0;
#pragma unroll
        for (int k = 0; k < PRE; ++k)
            if (static_cast<int>(threadIdx.x) + k * BLOCK < nb)
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    t[v] += pv[v][k];
        for (int i = threadIdx.x + PRE * BLOCK; i < nb; i += BLOCK)
#pragma unroll
            for (int v = 0; v < NV; ++v)
                t[v] += parts[v][b0 + i];
        const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
#pragma unroll
        for (int v = 0; v < NV; ++v)
        {
            t[v] = wave_sum(t[v]);
            if (lane == 0)
                part[v][w] = t[v];
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < NV; ++v)
        {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < NW; ++i)
                s += part[v][i];
            out[v] = NW == 1 ? t[v] : s;
        }
        __syncthreads();
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
