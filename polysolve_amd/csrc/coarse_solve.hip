// Exact FP64 solve on the coarsest MAS level.
//
// The reference's coarsest level is bank-block-diagonal like every other level
// (src/polysolve/linear/mas_utils/MASPreconditioner.cu:939-979): coarse nodes in different
// 32-node banks never couple.  Here the coarsest level's *active* nodes — those containing at
// least one fine node with an off-diagonal block (Dirichlet rows and other isolated nodes form
// singleton components that only ever see their own diagonal) — get their full Galerkin matrix
// R A R^T, factorised by a dense FP64 Cholesky, and its inverse is applied as one dense GEMV.
// Isolated coarsest nodes keep their bank treatment.  The operator stays symmetric positive
// definite, and it adds the global coarse solve that the additive bank hierarchy lacks.
#include "coarse_solve.hpp"
#include "common.hpp"
#include "device.hpp"

#include <cmath>
#include <stdexcept>

namespace psmas
{
    namespace
    {
        __device__ __forceinline__ int cmap_level(const int4 &c, int l)
        {
            return l == 0 ? c.x : (l == 1 ? c.y : (l == 2 ? c.z : c.w));
        }

        /// dense[(3a+bi)*nd + 3b+bj] += A_ij[bi][bj] for every stored block (i, j) whose coarsest
        /// ids map to active indices a, b.  Both triangles are accumulated (no symmetry shortcut).
        template <int BD>
        __global__ void k_dense_galerkin(int npad, int level, SellView<double> A, const int *__restrict__ row_len,
                                         const int *__restrict__ block_of_slot, const int4 *__restrict__ cmap,
                                         const int *__restrict__ active_index, int nd, double *__restrict__ dense)
        {
            const int i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= npad || block_of_slot[i] < 0)
                return;
            const int a = active_index[cmap_level(cmap[i], level)];
            if (a < 0)
                return;
            const int slice = i / WAVE, lane = i % WAVE;
            const int beg = A.slice_ptr[slice];
            const int len = row_len[i];
            // Consecutive blocks on the same coarsest column are merged before the atomics.
            int cur = -1;
            double acc[BD * BD];
#pragma unroll
            for (int e = 0; e < BD * BD; ++e)
                acc[e] = 0.0;
            auto flush = [&]() {
                if (cur < 0)
                    return;
#pragma unroll
                for (int bi = 0; bi < BD; ++bi)
#pragma unroll
                    for (int bj = 0; bj < BD; ++bj)
                    {
                        if (acc[bi * BD + bj] != 0.0)
                            atomicAdd(dense + static_cast<size_t>(BD * a + bi) * nd + BD * cur + bj, acc[bi * BD + bj]);
                        acc[bi * BD + bj] = 0.0;
                    }
            };
            for (int k = 0; k < len; ++k)
            {
                const size_t slot = static_cast<size_t>(beg + k);
                const int j = A.cols[slot * WAVE + lane];
                const int b = active_index[cmap_level(cmap[j], level)];
                if (b < 0)
                    continue;
                if (b != cur)
                {
                    flush();
                    cur = b;
                }
#pragma unroll
                for (int e = 0; e < BD * BD; ++e)
                    acc[e] += A.vals[(slot * BD * BD + e) * WAVE + lane];
            }
            flush();
        }

    } // namespace

    void launch_dense_galerkin(int bd, int npad, int level, const SellView<double> &A, const int *row_len,
                               const int *block_of_slot, const int4 *cmap, const int *active_index, int nd,
                               double *dense, hipStream_t s)
    {
        const dim3 grid(div_up(npad, 128)), block(128);
        if (bd == 1)
            hipLaunchKernelGGL(k_dense_galerkin<1>, grid, block, 0, s, npad, level, A, row_len, block_of_slot, cmap,
                               active_index, nd, dense);
        else if (bd == 2)
            hipLaunchKernelGGL(k_dense_galerkin<2>, grid, block, 0, s, npad, level, A, row_len, block_of_slot, cmap,
                               active_index, nd, dense);
        else
            hipLaunchKernelGGL(k_dense_galerkin<3>, grid, block, 0, s, npad, level, A, row_len, block_of_slot, cmap,
                               active_index, nd, dense);
        PSMAS_LAUNCH_CHECK();
    }

    bool cholesky_inverse(int n, std::vector<double> &A)
    {
        // A = L L^T (lower, in place), then Inv = L^-T L^-1, returned full and symmetric.
        std::vector<double> L(A);
        for (int j = 0; j < n; ++j)
        {
            double d = L[static_cast<size_t>(j) * n + j];
            for (int k = 0; k < j; ++k)
                d -= L[static_cast<size_t>(j) * n + k] * L[static_cast<size_t>(j) * n + k];
            if (!(d > 0.0) || !std::isfinite(d))
                return false;
            d = std::sqrt(d);
            L[static_cast<size_t>(j) * n + j] = d;
            for (int i = j + 1; i < n; ++i)
            {
                double v = L[static_cast<size_t>(i) * n + j];
                for (int k = 0; k < j; ++k)
                    v -= L[static_cast<size_t>(i) * n + k] * L[static_cast<size_t>(j) * n + k];
                L[static_cast<size_t>(i) * n + j] = v / d;
            }
        }
        // W = L^-1 (lower triangular), row by row: W_i = (e_i - sum_{k<i} L_ik W_k) / L_ii, where
        // row k of W is non-zero in columns 0..k only (contiguous row updates).
        std::vector<double> W(static_cast<size_t>(n) * n, 0.0);
        for (int i = 0; i < n; ++i)
        {
            double *wi = &W[static_cast<size_t>(i) * n];
            wi[i] = 1.0;
            for (int k = 0; k < i; ++k)
            {
                const double l = L[static_cast<size_t>(i) * n + k];
                const double *wk = &W[static_cast<size_t>(k) * n];
                for (int c = 0; c <= k; ++c)
                    wi[c] -= l * wk[c];
            }
            const double d = 1.0 / L[static_cast<size_t>(i) * n + i];
            for (int c = 0; c <= i; ++c)
                wi[c] *= d;
        }
        // Inv = W^T W = sum_k W_k^T W_k (outer products of the rows), row updates again.
        std::fill(A.begin(), A.end(), 0.0);
        for (int k = 0; k < n; ++k)
        {
            const double *wk = &W[static_cast<size_t>(k) * n];
            for (int i = 0; i <= k; ++i)
            {
                const double a = wk[i];
                double *ai = &A[static_cast<size_t>(i) * n];
                for (int c = 0; c <= k; ++c)
                    ai[c] += a * wk[c];
            }
        }
        return true;
    }

} // namespace psmas
