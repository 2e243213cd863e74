// Learned coarse correction: lattice coarse space, restriction / prolongation kernels and the
// BF16 GEMV inference of the correction network (see lcc.hpp).  No counterpart in the reference.
#include "lcc.hpp"
#include "coarse_solve.hpp"
#include "device.hpp"
#include "lattice_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace psmas
{
    namespace
    {
        constexpr int LCC_CHUNK = 1024; ///< restriction entries per workgroup
        constexpr int GEMV_SYS = 8;     ///< systems per wave in the inference GEMV


        /// partial[c][d] = sum over chunk c's entries (slot, w) of w * v[slot][d], in entry order.
        template <int BD, typename T>
        __global__ __launch_bounds__(256) void k_lcc_restrict_partial(const int *__restrict__ chunk_begin,
                                                                      const int2 *__restrict__ entries,
                                                                      const T *__restrict__ v,
                                                                      double *__restrict__ partial)
        {
            const int c = blockIdx.x;
            const int e0 = chunk_begin[c], e1 = chunk_begin[c + 1];
            double acc[BD];
#pragma unroll
            for (int d = 0; d < BD; ++d)
                acc[d] = 0.0;
            for (int e = e0 + threadIdx.x; e < e1; e += 256)
            {
                const int2 en = entries[e];
                const double w = static_cast<double>(__int_as_float(en.y));
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    acc[d] += w * static_cast<double>(v[static_cast<size_t>(en.x) * BD + d]);
            }
#pragma unroll
            for (int d = 0; d < BD; ++d)
            {
                const double s = block_sum<256>(acc[d]);
                if (threadIdx.x == 0)
                    partial[static_cast<size_t>(c) * BD + d] = s;
            }
        }

        /// One workgroup per 256-node block: the block's lattice partials from v (see
        /// lattice_block_partials).
        template <int BD, typename T>
        __global__ __launch_bounds__(256) void k_lcc_restrict_blocks(LatticeBlocks lb, const T *__restrict__ v)
        {
            __shared__ T vals[256 * 8 * BD];
            const int b = blockIdx.x;
            const LatticeSlot ls = lattice_slot(lb, b);
            T r[BD];
            const size_t p = static_cast<size_t>(b) * 256 + threadIdx.x;
#pragma unroll
            for (int d = 0; d < BD; ++d)
                r[d] = v[p * BD + d];
            lattice_block_partials<BD, T>(lb, b, ls, r, vals);
        }

        /// One wave per (global) lattice point J: u[J][d] = sum of its partials, lanes striding over
        /// the point's source list and a fixed-order wave reduction.
        template <int BD>
        __global__ __launch_bounds__(256) void k_lcc_gather(int npts, const int *__restrict__ point_ptr,
                                                            const int *__restrict__ point_src,
                                                            const double *__restrict__ partial, double *__restrict__ u)
        {
            const int J = blockIdx.x * 4 + threadIdx.x / WAVE;
            const int lane = threadIdx.x % WAVE;
            if (J >= npts)
                return;
            double a[BD];
#pragma unroll
            for (int d = 0; d < BD; ++d)
                a[d] = 0.0;
            for (int i = point_ptr[J] + lane; i < point_ptr[J + 1]; i += WAVE)
            {
                const size_t c = static_cast<size_t>(point_src[i]);
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    a[d] += partial[c * BD + d];
            }
#pragma unroll
            for (int d = 0; d < BD; ++d)
                a[d] = wave_sum(a[d]);
            if (lane == 0)
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    u[static_cast<size_t>(J) * BD + d] = a[d];
        }

        /// Layer 1 (ReLU) and the skip layer on the normalised lattice residual.  Workgroup (row
        /// chunk, system group): the threads compute |u_s| for the group's systems, stage
        /// x_s = u_s sqrt(m) / |u_s| (zero-padded to md) in LDS and, in the first chunk, write
        /// scale[s] = |u_s| a / sqrt(m); then each wave produces one row (rows < h: W1 -> h1,
        /// else the skip layer S -> skip).
        __global__ __launch_bounds__(256) void k_lcc_in(const double *__restrict__ u, int m, int md,
                                                        const unsigned short *__restrict__ W1,
                                                        const float *__restrict__ b1, int h,
                                                        const unsigned short *__restrict__ S, float *__restrict__ h1,
                                                        float *__restrict__ skip, float *__restrict__ scale,
                                                        float ascale, int nsys, int gsize,
                                                        const int *__restrict__ active)
        {
            extern __shared__ float xs[]; // [gsize][md]
            __shared__ double red[GEMV_SYS][4];
            __shared__ double fac[GEMV_SYS];
            const int ngroups = (nsys + gsize - 1) / gsize;
            const int g = blockIdx.x % ngroups, chunk = blockIdx.x / ngroups;
            const int s0 = g * gsize;
            const int ns = min(gsize, nsys - s0);
            const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
            for (int j = 0; j < ns; ++j)
            {
                const double *us = u + static_cast<size_t>(s0 + j) * m;
                double q = 0.0;
                for (int i = threadIdx.x; i < m; i += 256)
                    q += us[i] * us[i];
                q = wave_sum(q);
                if (lane == 0)
                    red[j][wave] = q;
            }
            __syncthreads();
            if (threadIdx.x < ns)
            {
                const int j = threadIdx.x;
                const double nu = sqrt(red[j][0] + red[j][1] + red[j][2] + red[j][3]);
                fac[j] = nu > 0.0 ? sqrt(static_cast<double>(m)) / nu : 0.0;
                if (chunk == 0)
                    scale[s0 + j] = static_cast<float>(nu * ascale / sqrt(static_cast<double>(m)));
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < ns * md; idx += 256)
            {
                const int j = idx / md, i = idx % md;
                xs[idx] = i < m ? static_cast<float>(u[static_cast<size_t>(s0 + j) * m + i] * fac[j]) : 0.0f;
            }
            __syncthreads();
            const int row_all = chunk * 4 + wave;
            if (row_all >= h + md)
                return;
            const bool first = row_all < h;
            const int row = first ? row_all : row_all - h;
            const uint4 *wrow = reinterpret_cast<const uint4 *>((first ? W1 : S) + static_cast<size_t>(row) * md);
            float acc[GEMV_SYS];
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
                acc[j] = 0.0f;
            for (int c = lane; c < md / 8; c += WAVE)
            {
                const uint4 wv = wrow[c];
                const float w[8] = {bf16_lo(wv.x), bf16_hi(wv.x), bf16_lo(wv.y), bf16_hi(wv.y),
                                    bf16_lo(wv.z), bf16_hi(wv.z), bf16_lo(wv.w), bf16_hi(wv.w)};
#pragma unroll
                for (int j = 0; j < GEMV_SYS; ++j)
                {
                    if (j < ns)
                    {
                        const float *x = xs + j * md + c * 8;
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            acc[j] += w[e] * x[e];
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
                if (j < ns) // ns is the same for the whole workgroup
                    acc[j] = wave_sum(acc[j]);
            if (lane != 0)
                return;
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
            {
                if (j >= ns)
                    break;
                const int s = s0 + j;
                if (active != nullptr && active[s] == 0)
                    continue;
                if (first)
                {
                    const float y = acc[j] + b1[row];
                    h1[static_cast<size_t>(s) * h + row] = y > 0.0f ? y : 0.0f;
                }
                else
                    skip[static_cast<size_t>(s) * md + row] = acc[j];
            }
        }

        constexpr int LCC_HMAX = 1024; ///< largest hidden width the fused output kernel holds in LDS

        /// k_lcc_out for hidden width 128 in one memory round trip: every load that depends on
        /// nothing (this workgroup's W2 rows, its output row of W3, h1, the biases, the skip
        /// outputs and the scales) is issued at the start.  A 16-lane unit u of the workgroup
        /// computes h2 rows u, u + 16, ... (lane l: columns 8l .. 8l + 7, then a row reduction),
        /// then output row 16 chunk + u the same way.  Workgroup (row chunk, system group).
        __global__ __launch_bounds__(256) void k_lcc_out128(const unsigned short *__restrict__ W2,
                                                            const float *__restrict__ b2,
                                                            const float *__restrict__ H1,
                                                            const unsigned short *__restrict__ W3,
                                                            const float *__restrict__ b3, int rows,
                                                            const float *__restrict__ skip, int ld_skip,
                                                            const float *__restrict__ scale, float *__restrict__ coef,
                                                            int ldc, int nsys, const int *__restrict__ active)
        {
            constexpr int H = 128, RPU = H / 16, HV = GEMV_SYS * H / 256;
            __shared__ __attribute__((aligned(16))) float h1s[GEMV_SYS][H];
            __shared__ __attribute__((aligned(16))) float h2s[GEMV_SYS][H];
            const int ngroups = (nsys + GEMV_SYS - 1) / GEMV_SYS;
            const int g = blockIdx.x % ngroups, chunk = blockIdx.x / ngroups;
            const int s0 = g * GEMV_SYS;
            const int ns = min(GEMV_SYS, nsys - s0);
            const int unit = threadIdx.x >> 4, rl = threadIdx.x & 15;
            const int orow = chunk * 16 + unit;
            const int orc = min(orow, rows - 1);
            // Loads first (addresses clamped where out of range; unused values are ignored).
            uint4 w2[RPU];
#pragma unroll
            for (int i = 0; i < RPU; ++i)
                w2[i] = reinterpret_cast<const uint4 *>(W2 + static_cast<size_t>(unit + 16 * i) * H)[rl];
            const uint4 w3 = reinterpret_cast<const uint4 *>(W3 + static_cast<size_t>(orc) * H)[rl];
            float h1v[HV];
#pragma unroll
            for (int i = 0; i < HV; ++i)
                h1v[i] = H1[static_cast<size_t>(s0) * H + min(static_cast<int>(threadIdx.x) + 256 * i, ns * H - 1)];
            float b2v[RPU];
#pragma unroll
            for (int i = 0; i < RPU; ++i)
                b2v[i] = b2[unit + 16 * i];
            const float b3v = b3[orc];
            float skv[GEMV_SYS], scv[GEMV_SYS];
            int actv[GEMV_SYS];
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
            {
                const int sj = s0 + min(j, ns - 1);
                skv[j] = skip[static_cast<size_t>(sj) * ld_skip + orc];
                scv[j] = scale[sj];
                actv[j] = active != nullptr ? active[sj] : 1;
            }
#pragma unroll
            for (int i = 0; i < HV; ++i)
            {
                const int idx = threadIdx.x + 256 * i;
                if (idx < ns * H)
                    h1s[idx / H][idx % H] = h1v[i];
            }
            __syncthreads();
            auto dot8 = [](const uint4 &wv, const float *x) {
                const float4 x0 = *reinterpret_cast<const float4 *>(x);
                const float4 x1 = *reinterpret_cast<const float4 *>(x + 4);
                return bf16_lo(wv.x) * x0.x + bf16_hi(wv.x) * x0.y + bf16_lo(wv.y) * x0.z + bf16_hi(wv.y) * x0.w +
                       bf16_lo(wv.z) * x1.x + bf16_hi(wv.z) * x1.y + bf16_lo(wv.w) * x1.z + bf16_hi(wv.w) * x1.w;
            };
#pragma unroll
            for (int i = 0; i < RPU; ++i)
            {
                const int r = unit + 16 * i;
#pragma unroll
                for (int j = 0; j < GEMV_SYS; ++j)
                {
                    if (j >= ns)
                        break; // uniform
                    const float a = row_sum16(dot8(w2[i], &h1s[j][8 * rl])) + b2v[i];
                    if (rl == 0)
                        h2s[j][r] = a > 0.0f ? a : 0.0f;
                }
            }
            __syncthreads();
            if (orow >= rows)
                return; // whole 16-lane units
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
            {
                if (j >= ns)
                    break;
                const float a = row_sum16(dot8(w3, &h2s[j][8 * rl]));
                if (rl == 0 && actv[j] != 0)
                    coef[static_cast<size_t>(s0 + j) * ldc + orow] = (a + b3v + skv[j]) * scv[j];
            }
        }

        /// Hidden layer 2 and the output layer in one launch.  Workgroup (row chunk, system group):
        /// the 256 threads first compute h2 = relu(W2 h1 + b2) for the group's systems into LDS (every
        /// workgroup of the group computes the same values), then each wave produces one output
        /// row: coef[s][row] = (W3[row] . h2[s] + b3[row] + skip[s][row]) * scale[s].
        __global__ __launch_bounds__(256) void k_lcc_out(const unsigned short *__restrict__ W2,
                                                         const float *__restrict__ b2, int h,
                                                         const float *__restrict__ H1,
                                                         const unsigned short *__restrict__ W3,
                                                         const float *__restrict__ b3, int rows,
                                                         const float *__restrict__ skip, int ld_skip,
                                                         const float *__restrict__ scale, float *__restrict__ coef,
                                                         int ldc, int nsys, const int *__restrict__ active)
        {
            __shared__ float h2s[GEMV_SYS][LCC_HMAX];
            const int ngroups = (nsys + GEMV_SYS - 1) / GEMV_SYS;
            const int g = blockIdx.x % ngroups, chunk = blockIdx.x / ngroups;
            const int s0 = g * GEMV_SYS;
            const int ns = min(GEMV_SYS, nsys - s0);
            for (int idx = threadIdx.x; idx < h * ns; idx += 256)
            {
                const int r = idx % h, j = idx / h;
                const uint4 *wrow = reinterpret_cast<const uint4 *>(W2 + static_cast<size_t>(r) * h);
                const float *x = H1 + static_cast<size_t>(s0 + j) * h;
                float a = 0.0f;
                for (int c = 0; c < h / 8; ++c)
                {
                    const uint4 wv = wrow[c];
                    const float4 x0 = reinterpret_cast<const float4 *>(x)[2 * c];
                    const float4 x1 = reinterpret_cast<const float4 *>(x)[2 * c + 1];
                    a += bf16_lo(wv.x) * x0.x + bf16_hi(wv.x) * x0.y + bf16_lo(wv.y) * x0.z + bf16_hi(wv.y) * x0.w +
                         bf16_lo(wv.z) * x1.x + bf16_hi(wv.z) * x1.y + bf16_lo(wv.w) * x1.z + bf16_hi(wv.w) * x1.w;
                }
                a += b2[r];
                h2s[j][r] = a > 0.0f ? a : 0.0f;
            }
            __syncthreads();
            const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
            const int row = chunk * 4 + wave;
            if (row >= rows)
                return;
            float acc[GEMV_SYS];
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
                acc[j] = 0.0f;
            const uint4 *wrow = reinterpret_cast<const uint4 *>(W3 + static_cast<size_t>(row) * h);
            for (int c = lane; c < h / 8; c += WAVE)
            {
                const uint4 wv = wrow[c];
                const float w[8] = {bf16_lo(wv.x), bf16_hi(wv.x), bf16_lo(wv.y), bf16_hi(wv.y),
                                    bf16_lo(wv.z), bf16_hi(wv.z), bf16_lo(wv.w), bf16_hi(wv.w)};
#pragma unroll
                for (int j = 0; j < GEMV_SYS; ++j)
                {
                    if (j < ns)
                    {
                        const float *x = &h2s[j][c * 8];
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            acc[j] += w[e] * x[e];
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
                if (j < ns) // ns is the same for the whole workgroup
                    acc[j] = wave_sum(acc[j]);
            if (lane != 0)
                return;
#pragma unroll
            for (int j = 0; j < GEMV_SYS; ++j)
            {
                if (j >= ns)
                    break;
                const int s = s0 + j;
                if (active != nullptr && active[s] == 0)
                    continue;
                const float y = acc[j] + b3[row] + skip[static_cast<size_t>(s) * ld_skip + row];
                coef[static_cast<size_t>(s) * ldc + row] = y * scale[s];
            }
        }

        /// e[p][d] = sum_corners w * coef[corner][d] for the slots of system 0; zero elsewhere.
        template <int BD>
        __global__ void k_lcc_prolong_fine(int npad, int np, const int4 *__restrict__ cell, int k,
                                           const double *__restrict__ coef, double *__restrict__ e)
        {
            const int p = blockIdx.x * blockDim.x + threadIdx.x;
            if (p >= npad)
                return;
            const int4 c = cell[p];
            double z[BD];
#pragma unroll
            for (int d = 0; d < BD; ++d)
                z[d] = 0.0;
            if (c.x >= 0 && c.x < np)
            {
                const float fx = __int_as_float(c.y), fy = __int_as_float(c.z), fz = __int_as_float(c.w);
#pragma unroll
                for (int q = 0; q < 8; ++q)
                {
                    const int id = c.x + (q & 1) + ((q >> 1) & 1) * k + ((q >> 2) & 1) * k * k;
                    const double w = static_cast<double>(trilinear_weight(q, fx, fy, fz));
#pragma unroll
                    for (int d = 0; d < BD; ++d)
                        z[d] += w * coef[static_cast<size_t>(id) * BD + d];
                }
            }
#pragma unroll
            for (int d = 0; d < BD; ++d)
                e[static_cast<size_t>(p) * BD + d] = z[d];
        }
    } // namespace

    // --------------------------------------------------------------------------- lattice
    void LearnedCoarse::build_lattice(int k, int bd, const std::vector<int> &node_off, const std::vector<int> &npad,
                                      const std::vector<const std::vector<double> *> &coords,
                                      const std::vector<int> &block_of_slot, const std::vector<char> &coupled,
                                      hipStream_t s)
    {
        if (k < 2 || k > 16)
            throw InvalidArgumentError("[MAS] learned correction: lattice_points must be in [2, 16].");
        const int nsys = static_cast<int>(node_off.size());
        const int np = k * k * k;
        const int total_pad = node_off.back() + npad.back();
        std::vector<int4> cell(total_pad, int4{-1, 0, 0, 0});
        for (int sy = 0; sy < nsys; ++sy)
        {
            const std::vector<double> &X = *coords[sy];
            const int nb = static_cast<int>(X.size() / 3);
            if (nb == 0)
                throw InvalidArgumentError("[MAS] learned correction needs 3-D node coordinates (psmas_set_coordinates).");
            double lo[3], hi[3];
            for (int d = 0; d < 3; ++d)
            {
                lo[d] = hi[d] = X[d];
                for (int i = 0; i < nb; ++i)
                {
                    lo[d] = std::min(lo[d], X[static_cast<size_t>(i) * 3 + d]);
                    hi[d] = std::max(hi[d], X[static_cast<size_t>(i) * 3 + d]);
                }
            }
            for (int p = node_off[sy]; p < node_off[sy] + npad[sy]; ++p)
            {
                const int b = block_of_slot[p];
                if (b < 0 || b >= nb || !coupled[p])
                    continue;
                int ci[3];
                float f[3];
                for (int d = 0; d < 3; ++d)
                {
                    const double ext = hi[d] - lo[d];
                    const double t = ext > 0.0 ? (X[static_cast<size_t>(b) * 3 + d] - lo[d]) / ext * (k - 1) : 0.0;
                    int c = static_cast<int>(std::floor(t));
                    c = std::min(std::max(c, 0), k - 2);
                    ci[d] = c;
                    f[d] = static_cast<float>(std::min(std::max(t - c, 0.0), 1.0));
                }
                int4 v;
                v.x = sy * np + ci[0] + k * (ci[1] + k * ci[2]);
                std::memcpy(&v.y, &f[0], 4);
                std::memcpy(&v.z, &f[1], 4);
                std::memcpy(&v.w, &f[2], 4);
                cell[p] = v;
            }
        }
        // Restriction entries grouped by lattice point, in slot order; chunks of LCC_CHUNK.
        const int npts = nsys * np;
        std::vector<int> count(npts + 1, 0);
        auto corners = [&](const int4 &c, int q, int &id, float &w) {
            float fx, fy, fz;
            std::memcpy(&fx, &c.y, 4);
            std::memcpy(&fy, &c.z, 4);
            std::memcpy(&fz, &c.w, 4);
            id = c.x + (q & 1) + ((q >> 1) & 1) * k + ((q >> 2) & 1) * k * k;
            const float wx = (q & 1) ? fx : 1.0f - fx;
            const float wy = (q & 2) ? fy : 1.0f - fy;
            const float wz = (q & 4) ? fz : 1.0f - fz;
            w = wx * wy * wz;
        };
        for (int p = 0; p < total_pad; ++p)
            if (cell[p].x >= 0)
                for (int q = 0; q < 8; ++q)
                {
                    int id;
                    float w;
                    corners(cell[p], q, id, w);
                    ++count[id + 1];
                }
        for (int j = 0; j < npts; ++j)
            count[j + 1] += count[j];
        std::vector<int2> entries(count[npts]);
        std::vector<int> fillp(count.begin(), count.end() - 1);
        for (int p = 0; p < total_pad; ++p)
            if (cell[p].x >= 0)
                for (int q = 0; q < 8; ++q)
                {
                    int id;
                    float w;
                    corners(cell[p], q, id, w);
                    int wb;
                    std::memcpy(&wb, &w, 4);
                    entries[fillp[id]++] = int2{p, wb};
                }
        std::vector<int> chunk_begin{0}, point_chunk(npts + 1, 0);
        for (int j = 0; j < npts; ++j)
        {
            for (int e = count[j]; e < count[j + 1]; e += LCC_CHUNK)
                chunk_begin.push_back(std::min(e + LCC_CHUNK, count[j + 1]));
            point_chunk[j + 1] = static_cast<int>(chunk_begin.size()) - 1;
        }
        nchunks_ = static_cast<int>(chunk_begin.size()) - 1;
        entries_count_ = static_cast<long long>(entries.size());

        // Block path: each 256-node block's distinct lattice points (sorted) get local ids; a slot
        // stores its 8 corners' local ids as bytes.  Used when no block touches more than
        // LCC_TMAX points; otherwise the chunked entry list above is used.
        const int nblocks = total_pad / 256;
        std::vector<int> block_nt(nblocks, 0), tend(static_cast<size_t>(nblocks) * LCC_TMAX, 0);
        std::vector<uint4> pos(total_pad, uint4{0u, 0u, 0u, 0u});
        std::vector<std::vector<int>> point_lists(npts);
        bool blocks_ok = true;
        for (int b = 0; b < nblocks && blocks_ok; ++b)
        {
            std::vector<int> pts;
            for (int p = b * 256; p < (b + 1) * 256; ++p)
                if (cell[p].x >= 0)
                    for (int q = 0; q < 8; ++q)
                    {
                        int id;
                        float w;
                        corners(cell[p], q, id, w);
                        pts.push_back(id);
                    }
            std::sort(pts.begin(), pts.end());
            pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
            if (static_cast<int>(pts.size()) > LCC_TMAX)
            {
                blocks_ok = false;
                break;
            }
            block_nt[b] = static_cast<int>(pts.size());
            for (int t = 0; t < block_nt[b]; ++t)
                point_lists[pts[t]].push_back(b * LCC_TMAX + t);
            // Per target: the (slot, corner) pairs that feed it, in slot-then-corner order; every
            // pair's position in the block's target-by-target order goes to pos.
            std::vector<std::vector<int>> lists(block_nt[b]);
            for (int p = b * 256; p < (b + 1) * 256; ++p)
                if (cell[p].x >= 0)
                    for (int q = 0; q < 8; ++q)
                    {
                        int id;
                        float w;
                        corners(cell[p], q, id, w);
                        const int t = static_cast<int>(std::lower_bound(pts.begin(), pts.end(), id) - pts.begin());
                        lists[t].push_back(((p - b * 256) << 3) | q);
                    }
            int at_end = 0; // block-local end of the targets so far
            for (size_t t = 0; t < lists.size(); ++t)
            {
                const auto &l = lists[t];
                for (size_t k = 0; k < l.size(); ++k)
                {
                    const int pq = l[k];
                    const int at = at_end + static_cast<int>(k);
                    const int p = b * 256 + (pq >> 3), q = pq & 7;
                    unsigned *word = (q >> 1) == 0 ? &pos[p].x : (q >> 1) == 1 ? &pos[p].y : (q >> 1) == 2 ? &pos[p].z : &pos[p].w;
                    *word |= static_cast<unsigned>(at) << (16 * (q & 1));
                }
                at_end += static_cast<int>(l.size());
                tend[static_cast<size_t>(b) * LCC_TMAX + t] = at_end;
            }
            for (int t = block_nt[b]; t < LCC_TMAX; ++t)
                tend[static_cast<size_t>(b) * LCC_TMAX + t] = at_end;
        }
        std::vector<int> point_ptr(npts + 1, 0), point_src;
        if (blocks_ok)
        {
            for (int j = 0; j < npts; ++j)
            {
                point_src.insert(point_src.end(), point_lists[j].begin(), point_lists[j].end());
                point_ptr[j + 1] = static_cast<int>(point_src.size());
            }
        }
        else
        {
            point_ptr = point_chunk;
            point_src.resize(nchunks_);
            for (int c = 0; c < nchunks_; ++c)
                point_src[c] = c;
        }
        use_blocks_ = blocks_ok;
        nblocks_ = nblocks;

        h_cell0_.assign(cell.begin() + node_off[0], cell.begin() + node_off[0] + npad[0]);
        ptp_inv_.clear(); // built on first use (pair collection only)

        k_ = k;
        bd_ = bd;
        np_ = np;
        nsys_ = nsys;
        md_ = (np * bd + 7) / 8 * 8;
        d_cell_.allocate(cell.size(), "MAS lattice cells");
        d_entries_.allocate(std::max<size_t>(entries.size(), 1), "MAS lattice restriction");
        d_chunk_begin_.allocate(chunk_begin.size(), "MAS lattice chunks");
        d_point_ptr_.allocate(point_ptr.size(), "MAS lattice point lists");
        d_point_src_.allocate(std::max<size_t>(point_src.size(), 1), "MAS lattice point lists");
        PSMAS_HIP(hipMemcpy(d_cell_.data(), cell.data(), d_cell_.bytes(), hipMemcpyHostToDevice));
        if (!entries.empty())
            PSMAS_HIP(hipMemcpy(d_entries_.data(), entries.data(), sizeof(int2) * entries.size(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_chunk_begin_.data(), chunk_begin.data(), d_chunk_begin_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_point_ptr_.data(), point_ptr.data(), d_point_ptr_.bytes(), hipMemcpyHostToDevice));
        if (!point_src.empty())
            PSMAS_HIP(hipMemcpy(d_point_src_.data(), point_src.data(), sizeof(int) * point_src.size(),
                                hipMemcpyHostToDevice));
        if (use_blocks_)
        {
            d_block_nt_.allocate(block_nt.size(), "MAS lattice block targets");
            d_tend_.allocate(tend.size(), "MAS lattice block lists");
            d_pos_.allocate(pos.size(), "MAS lattice block positions");
            PSMAS_HIP(hipMemcpy(d_block_nt_.data(), block_nt.data(), d_block_nt_.bytes(), hipMemcpyHostToDevice));
            PSMAS_HIP(hipMemcpy(d_tend_.data(), tend.data(), d_tend_.bytes(), hipMemcpyHostToDevice));
            PSMAS_HIP(hipMemcpy(d_pos_.data(), pos.data(), d_pos_.bytes(), hipMemcpyHostToDevice));
            d_partial_.allocate(static_cast<size_t>(nblocks) * LCC_TMAX * bd, "MAS lattice partials");
        }
        else
        {
            d_block_nt_.reset();
            d_tend_.reset();
            d_pos_.reset();
            d_partial_.allocate(static_cast<size_t>(std::max(nchunks_, 1)) * bd, "MAS lattice partials");
        }
        d_u_.allocate(static_cast<size_t>(nsys) * m(), "MAS lattice residual");
        d_scale_.allocate(nsys, "MAS network scale");
        d_coef_.allocate(static_cast<size_t>(nsys) * m(), "MAS lattice correction");
        d_coef_.zero(s);
        d_h1_.reset();
        d_h2_.reset();
        d_skip_.reset();
    }

    template <int BD, typename T>
    void LearnedCoarse::restrict_to_lattice(const T *v, const int *active, hipStream_t s, float output_factor)
    {
        if (use_blocks_)
            restrict_blocks<BD, T>(v, s);
        else if (nchunks_ > 0)
        {
            hipLaunchKernelGGL((k_lcc_restrict_partial<BD, T>), dim3(nchunks_), dim3(256), 0, s, d_chunk_begin_.data(),
                               d_entries_.data(), v, d_partial_.data());
            PSMAS_LAUNCH_CHECK();
        }
        const int npts = nsys_ * np_;
        hipLaunchKernelGGL((k_lcc_gather<BD>), dim3(div_up(npts, 4)), dim3(256), 0, s, npts, d_point_ptr_.data(),
                           d_point_src_.data(), d_partial_.data(), d_u_.data());
        PSMAS_LAUNCH_CHECK();
        out_factor_ = output_factor;
        (void)active;
    }

    template <int BD, typename T>
    void LearnedCoarse::restrict_blocks(const T *v, hipStream_t s)
    {
        hipLaunchKernelGGL((k_lcc_restrict_blocks<BD, T>), dim3(nblocks_), dim3(256), 0, s, block_view(), v);
        PSMAS_LAUNCH_CHECK();
    }

    LatticeBlocks LearnedCoarse::block_view() const
    {
        LatticeBlocks lb;
        if (!use_blocks_)
            return lb;
        lb.cell = d_cell_.data();
        lb.pos = d_pos_.data();
        lb.block_nt = d_block_nt_.data();
        lb.tend = d_tend_.data();
        lb.partial = const_cast<double *>(d_partial_.data());
        return lb;
    }

    void LearnedCoarse::gather_into(ApplySide &sd) const
    {
        sd.lat_npts = nsys_ * np_;
        sd.lat_ptr = d_point_ptr_.data();
        sd.lat_src = d_point_src_.data();
        sd.lat_partial = d_partial_.data();
        sd.lat_u = const_cast<double *>(d_u_.data());
    }

    void LearnedCoarse::check_network()
    {
        if (!net_)
            throw std::runtime_error("[MAS] learned correction: the network is not trained.");
        if (net_->d_in() != md_ || net_->d_out() != md_)
            throw std::runtime_error("[MAS] learned correction: network width does not match the lattice.");
        const int h = net_->hidden();
        if (h > LCC_HMAX || h % 8 != 0)
            throw std::runtime_error("[MAS] learned correction: hidden width must be a multiple of 8 up to 1024.");
        if (!d_h1_)
        {
            d_h1_.allocate(static_cast<size_t>(nsys_) * h, "MAS network activations");
            d_h2_.allocate(static_cast<size_t>(nsys_) * h, "MAS network activations");
            d_skip_.allocate(static_cast<size_t>(nsys_) * md_, "MAS network skip output");
        }
    }

    bool LearnedCoarse::in_side(ApplySide &sd, int stage_floats, int block, const int *active)
    {
        check_network();
        if (nsys_ != 1 || md_ > stage_floats || md_ / 8 > 64 * LIN_CHUNKS || m() > LIN_U_PER_THREAD * block)
            return false;
        const unsigned short *wb = net_->bf16_weights();
        const float *wf = net_->fp32_weights();
        const int h = net_->hidden();
        sd.lin_rows = h + md_;
        sd.lin_m = m();
        sd.lin_md = md_;
        sd.lin_h = h;
        sd.lin_u = d_u_.data();
        sd.lin_w1 = wb + net_->offset(0);
        sd.lin_b1 = wf + net_->offset(1);
        sd.lin_s = wb + net_->offset(6);
        sd.lin_h1 = d_h1_.data();
        sd.lin_skip = d_skip_.data();
        sd.lin_scale = d_scale_.data();
        sd.lin_ascale = ascale_ * out_factor_;
        sd.lin_active = active;
        // Workgroups of their own after the others' (a row per wave).
        sd.lin_wg0 = sd.wgs;
        sd.wgs += std::min(LIN_WGS_MAX, div_up(sd.lin_rows, block / 64));
        return true;
    }

    void LearnedCoarse::infer_out(const int *active, hipStream_t s)
    {
        const unsigned short *wb = net_->bf16_weights();
        const float *wf = net_->fp32_weights();
        const int h = net_->hidden();
        const int ngroups = (nsys_ + GEMV_SYS - 1) / GEMV_SYS;
        // Layer 2 fused into the output layer (only the m real outputs), plus the skip output,
        // times the scale.
        if (h == 128)
            hipLaunchKernelGGL(k_lcc_out128, dim3(div_up(m(), 16) * ngroups), dim3(256), 0, s, wb + net_->offset(2),
                               wf + net_->offset(3), static_cast<const float *>(d_h1_.data()), wb + net_->offset(4),
                               wf + net_->offset(5), m(), static_cast<const float *>(d_skip_.data()), md_,
                               static_cast<const float *>(d_scale_.data()), d_coef_.data(), m(), nsys_, active);
        else
            hipLaunchKernelGGL(k_lcc_out, dim3(div_up(m(), 4) * ngroups), dim3(256), 0, s, wb + net_->offset(2),
                               wf + net_->offset(3), h, static_cast<const float *>(d_h1_.data()), wb + net_->offset(4),
                               wf + net_->offset(5), m(), static_cast<const float *>(d_skip_.data()), md_,
                               static_cast<const float *>(d_scale_.data()), d_coef_.data(), m(), nsys_, active);
        PSMAS_LAUNCH_CHECK();
    }

    void LearnedCoarse::infer(const int *active, hipStream_t s)
    {
        check_network();
        const int h = net_->hidden();
        const unsigned short *wb = net_->bf16_weights();
        const float *wf = net_->fp32_weights();
        // Layer 1 and the skip layer read the same normalised input, built in the kernel.
        const int gsize = std::max(1, std::min(GEMV_SYS, static_cast<int>(60000 / (sizeof(float) * md_))));
        const int ngroups_in = div_up(nsys_, gsize);
        const size_t lds = sizeof(float) * static_cast<size_t>(std::min(gsize, nsys_)) * md_;
        hipLaunchKernelGGL(k_lcc_in, dim3(div_up(h + md_, 4) * ngroups_in), dim3(256), lds, s,
                           static_cast<const double *>(d_u_.data()), m(), md_, wb + net_->offset(0), wf + net_->offset(1),
                           h, wb + net_->offset(6), d_h1_.data(), d_skip_.data(), d_scale_.data(), ascale_ * out_factor_,
                           nsys_, gsize, active);
        PSMAS_LAUNCH_CHECK();
        infer_out(active, s);
    }

    LccApply LearnedCoarse::apply_view() const
    {
        LccApply v;
        v.cell = d_cell_.data();
        v.coef = d_coef_.data();
        v.k = k_;
        return v;
    }

    template <int BD>
    void LearnedCoarse::prolong_host_coef(const std::vector<double> &w, double *e, long long nvec, hipStream_t s)
    {
        const int npad0 = static_cast<int>(h_cell0_.size());
        if (static_cast<long long>(npad0) * BD > nvec || static_cast<int>(w.size()) != m())
            throw std::runtime_error("[MAS] learned correction: size mismatch in prolongation.");
        DeviceBuffer<double> dw(w.size(), "MAS lattice coefficients");
        PSMAS_HIP(hipMemcpyAsync(dw.data(), w.data(), dw.bytes(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_lcc_prolong_fine<BD>), dim3(div_up(npad0, 256)), dim3(256), 0, s, npad0, np_,
                           d_cell_.data(), k_, dw.data(), e);
        PSMAS_LAUNCH_CHECK();
        PSMAS_HIP(hipStreamSynchronize(s));
    }

    std::vector<double> LearnedCoarse::download_u(hipStream_t s) const
    {
        std::vector<double> u(m());
        PSMAS_HIP(hipMemcpyAsync(u.data(), d_u_.data(), sizeof(double) * u.size(), hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        return u;
    }

    void LearnedCoarse::ensure_ptp()
    {
        if (!ptp_inv_.empty())
            return;
        // (P^T P)^-1 of system 0 (the same for every component), with a Tikhonov shift of 1e-8
        // times the mean diagonal so that lattice points with little or no support (a lattice
        // finer than the mesh) keep the projection well posed.
        const int np = np_, k = k_;
        std::vector<double> ptp(static_cast<size_t>(np) * np, 0.0);
        for (const int4 &c : h_cell0_)
        {
            if (c.x < 0)
                continue;
            float f[3];
            std::memcpy(&f[0], &c.y, 4);
            std::memcpy(&f[1], &c.z, 4);
            std::memcpy(&f[2], &c.w, 4);
            int id[8];
            double w[8];
            for (int q = 0; q < 8; ++q)
            {
                id[q] = c.x + (q & 1) + ((q >> 1) & 1) * k + ((q >> 2) & 1) * k * k;
                w[q] = static_cast<double>(((q & 1) ? f[0] : 1.0f - f[0]) * ((q & 2) ? f[1] : 1.0f - f[1]) *
                                           ((q & 4) ? f[2] : 1.0f - f[2]));
            }
            for (int a = 0; a < 8; ++a)
                for (int b = 0; b < 8; ++b)
                    ptp[static_cast<size_t>(id[a]) * np + id[b]] += w[a] * w[b];
        }
        double tr = 0.0;
        for (int j = 0; j < np; ++j)
            tr += ptp[static_cast<size_t>(j) * np + j];
        const double shift = 1e-8 * (tr > 0.0 ? tr / np : 1.0);
        for (int j = 0; j < np; ++j)
            ptp[static_cast<size_t>(j) * np + j] += shift;
        if (!cholesky_inverse(np, ptp))
            throw std::runtime_error("[MAS] learned correction: lattice mass matrix is singular.");
        ptp_inv_ = std::move(ptp);
    }

    std::vector<double> LearnedCoarse::ls_project(const std::vector<double> &v)
    {
        ensure_ptp();
        std::vector<double> w(v.size(), 0.0);
        for (int i = 0; i < np_; ++i)
            for (int j = 0; j < np_; ++j)
            {
                const double g = ptp_inv_[static_cast<size_t>(i) * np_ + j];
                for (int d = 0; d < bd_; ++d)
                    w[static_cast<size_t>(i) * bd_ + d] += g * v[static_cast<size_t>(j) * bd_ + d];
            }
        return w;
    }

    // --------------------------------------------------------------------------- network
    void LearnedCoarse::add_pair(const std::vector<double> &u, const std::vector<double> &w)
    {
        if (m_pairs_ != 0 && m_pairs_ != m())
            clear_pairs();
        m_pairs_ = m();
        for (double v : u)
            pairs_u_.push_back(static_cast<float>(v));
        for (double v : w)
            pairs_w_.push_back(static_cast<float>(v));
    }

    void LearnedCoarse::clear_pairs()
    {
        pairs_u_.clear();
        pairs_w_.clear();
        m_pairs_ = 0;
    }

    void LearnedCoarse::reset_network()
    {
        net_.reset();
        trained_ = false;
    }

    double LearnedCoarse::train(int hidden, int epochs, int batch, float lr, unsigned seed, bool linear)
    {
        const int mm = m_pairs_;
        if (mm == 0 || pairs_u_.empty())
            throw std::runtime_error("[MAS] learned correction: no training pairs collected.");
        const size_t n = pairs_u_.size() / mm;
        const int md = (mm + 7) / 8 * 8;
        // Normalise: x = u sqrt(m) / |u|, t = w sqrt(m) / |u| / a with a the RMS of w sqrt(m) / |u|.
        std::vector<float> X(n * md, 0.0f), T(n * md, 0.0f);
        double sumsq = 0.0;
        size_t used = 0;
        for (size_t i = 0; i < n; ++i)
        {
            double nu = 0.0;
            for (int j = 0; j < mm; ++j)
                nu += static_cast<double>(pairs_u_[i * mm + j]) * pairs_u_[i * mm + j];
            nu = std::sqrt(nu);
            if (!(nu > 0.0) || !std::isfinite(nu))
                continue;
            const double f = std::sqrt(static_cast<double>(mm)) / nu;
            for (int j = 0; j < mm; ++j)
            {
                X[used * md + j] = static_cast<float>(pairs_u_[i * mm + j] * f);
                const double t = pairs_w_[i * mm + j] * f;
                T[used * md + j] = static_cast<float>(t);
                sumsq += t * t;
            }
            ++used;
        }
        if (used == 0)
            throw std::runtime_error("[MAS] learned correction: all training pairs are zero.");
        const double a = std::sqrt(sumsq / (static_cast<double>(used) * mm));
        for (size_t i = 0; i < used * md; ++i)
            T[i] = static_cast<float>(T[i] / a);
        net_ = std::make_unique<Mlp>(md, md, hidden, seed, true);
        if (linear)
            net_->set_linear_only(true);
        DeviceBuffer<float> dX(used * md, "MAS training inputs"), dT(used * md, "MAS training targets");
        PSMAS_HIP(hipMemcpy(dX.data(), X.data(), dX.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(dT.data(), T.data(), dT.bytes(), hipMemcpyHostToDevice));
        last_loss_ = net_->train(static_cast<int>(used), dX.data(), dT.data(), epochs, batch, lr, seed, true);
        ascale_ = static_cast<float>(a);
        trained_ = true;
        kind_ = linear ? "linear" : "mlp";
        return last_loss_;
    }

    void LearnedCoarse::install_linear(const std::vector<double> &M, int hidden, unsigned seed, const char *kind)
    {
        const int mm = m(), md = width();
        if (mm == 0 || M.size() != static_cast<size_t>(mm) * mm)
            throw std::runtime_error("[MAS] learned correction: linear map does not match the lattice.");
        if (!net_ || net_->d_in() != md || net_->hidden() != hidden)
            net_ = std::make_unique<Mlp>(md, md, hidden, seed, true);
        // With output scale a = 1 the correction is c = |u| / sqrt(m) * N(u sqrt(m) / |u|) = S u
        // when N is the skip layer alone.
        std::vector<float> theta(net_->num_parameters());
        net_->get_parameters(theta.data());
        std::fill(theta.begin() + static_cast<long>(net_->offset(4)), theta.begin() + static_cast<long>(net_->offset(6)),
                  0.0f); // W3, b3
        float *S = theta.data() + net_->offset(6);
        std::fill(S, S + static_cast<size_t>(md) * md, 0.0f);
        for (int i = 0; i < mm; ++i)
            for (int j = 0; j < mm; ++j)
                S[static_cast<size_t>(i) * md + j] = static_cast<float>(M[static_cast<size_t>(i) * mm + j]);
        net_->set_parameters(theta.data());
        ascale_ = 1.0f;
        last_loss_ = 0.0;
        trained_ = true;
        kind_ = kind;
    }

#define PSMAS_INST_LCC(BD)                                                                                          \
    template void LearnedCoarse::restrict_blocks<BD, double>(const double *, hipStream_t);                        \
    template void LearnedCoarse::restrict_blocks<BD, float>(const float *, hipStream_t);                          \
    template void LearnedCoarse::restrict_to_lattice<BD, double>(const double *, const int *, hipStream_t, float); \
    template void LearnedCoarse::restrict_to_lattice<BD, float>(const float *, const int *, hipStream_t, float);   \
    template void LearnedCoarse::prolong_host_coef<BD>(const std::vector<double> &, double *, long long, hipStream_t);
    PSMAS_INST_LCC(1)
    PSMAS_INST_LCC(2)
    PSMAS_INST_LCC(3)
#undef PSMAS_INST_LCC
} // namespace psmas
