// MAS preconditioner kernels for gfx950: build (bank-local connected components over <= 4
// coarse levels, Galerkin bank matrices, batched dense inverses, Morton keys) and apply
// (coarse restriction, batched symmetric bank GEMV staged in LDS, prolongation fused with r.z).
//
// Semantics follow the reference preconditioner (src/polysolve/linear/mas_utils/
// MASPreconditioner.cu): coarse space :117-481, Galerkin fill :484-591, zero-diagonal padding
// :779-796, inverse :799-936, apply :593-777,1099-1164.  The kernels are this build's own:
//  * the component search runs two 32-node banks per wave64 and reads each bank's half of a
//    64-bit ballot;
//  * restriction sums each component from its leader's member mask, in a fixed order, so no
//    atomics and no zero-fill of the coarse vectors are needed per apply;
//  * the inverse is an in-LDS Gauss-Jordan on the full matrix with 256 threads; a second kernel
//    emits the FP16 form S = D^-1 Inv D^-1, D = sqrt(diag(Inv)), used by the mixed-precision path.
#include "device.hpp"
#include "kernels.hpp"
#include "common.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <type_traits>
#include <utility>

namespace psmas
{
    namespace
    {
        // ------------------------------------------------------------------------------ apply
        template <int BD>
        constexpr int apply_block() { return BD == 1 ? 64 : (BD == 2 ? 128 : 192); }

        /// Input component d of multi-level node `node`: mr, or for a node whose restriction was
        /// cut into pieces the FP64 sum of its pieces in order (rounded like a stored value).  The
        /// two loads are independent, so only split nodes pay a further round trip.
        /// Component d of split node si's input: the FP64 sum of its pieces in order (pieces loaded
        /// eight at a time), rounded like a stored value.
        template <int BD, typename T>
        __device__ __forceinline__ T split_value(const ApplySide &sd, int si, int d)
        {
            const int4 c = sd.rcomb[si];
            double acc = 0.0;
            for (int k0 = 0; k0 < c.z; k0 += 8)
            {
                double pv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    pv[j] = k0 + j < c.z ? sd.rpart[static_cast<size_t>(c.y + k0 + j) * BD + d] : 0.0;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (k0 + j < c.z)
                        acc += pv[j];
            }
            return static_cast<T>(acc);
        }

        template <int BD, typename T>
        __device__ __forceinline__ T coarse_input(const ApplySide &sd, const T *__restrict__ mr, int node, int d)
        {
            T v = mr[static_cast<size_t>(node) * BD + d];
            if (sd.split_of != nullptr && node >= sd.node0)
            {
                const int si = sd.split_of[node - sd.node0];
                if (si >= 0)
                    v = split_value<BD, T>(sd, si, d);
            }
            return v;
        }

        /// The loads of a bank row's input, issued one matrix ahead and used only after the next
        /// barrier: a use in the load phase would make the in-order memory counter wait for the
        /// matrix tiles in flight as well.  Every load is unconditional (indices clamped, absent
        /// arrays replaced by sd.dummy), so the compiler can count the loads in flight and wait
        /// for exactly the earlier matrix's.  split: node >= node0 of a launch with split nodes
        /// (its split index then decides the input).
        template <typename T>
        struct RawInput
        {
            T x = T(0);
            int si = -1;
            bool split = false;
        };

        template <int BD, typename T>
        __device__ __forceinline__ RawInput<T> raw_input(const ApplySide &sd, const T *__restrict__ mr, int node, int d)
        {
            RawInput<T> in;
            in.x = mr[static_cast<size_t>(node) * BD + d];
            const int rel = node - sd.node0;
            const bool has = sd.split_of != nullptr;
            in.split = has && rel >= 0;
            const int last = has ? sd.split_count - 1 : 0;
            in.si = (has ? sd.split_of : sd.dummy)[min(max(rel, 0), last)];
            return in;
        }

        template <int BD, typename T>
        __device__ __forceinline__ T input_value(const ApplySide &sd, const RawInput<T> &in, int d)
        {
            return (in.split && in.si >= 0) ? split_value<BD, T>(sd, in.si, d) : in.x;
        }

        /// Whether the dense coarsest solve owns node `node` (its bank output is then not stored),
        /// in two halves: dense_probe issues the (unconditional, clamped) load, the caller tests
        /// `in && dix >= 0` after the next barrier.
        __device__ __forceinline__ void dense_probe(const ApplySide &sd, int node, bool &in, int &dix)
        {
            const bool has = sd.dense_index != nullptr;
            int rel = node - sd.dense_seg;
            if (sd.dense_copies > 1 && rel >= 0 && rel < sd.dense_copies * sd.dense_stride)
                rel %= sd.dense_stride;
            in = has && rel >= 0 && rel < sd.dense_count;
            const int last = has ? sd.dense_count - 1 : 0;
            dix = (has ? sd.dense_index : sd.dummy)[min(max(rel, 0), last)];
        }

        /// Stored matrix of bank m (CopyLayout).
        __device__ __forceinline__ int stored_of_bank(const CopyLayout &cl, int m)
        {
            int l = 0;
#pragma unroll
            for (int i = 1; i < CopyLayout::MAXL; ++i)
                if (i < cl.levels && m >= cl.moff[i])
                    l = i;
            return cl.soff[l] + (m - cl.moff[l]) % cl.cnt[l];
        }

        /// Copy 0's bank of stored matrix s and the bank stride between copies (CopyLayout).
        __device__ __forceinline__ void bank_of_stored(const CopyLayout &cl, int s, int &bank, int &stride)
        {
            int l = 0;
#pragma unroll
            for (int i = 1; i < CopyLayout::MAXL; ++i)
                if (i < cl.levels && s >= cl.soff[i])
                    l = i;
            bank = cl.moff[l] + (s - cl.soff[l]);
            stride = cl.cnt[l];
        }

        /// Layer 1 and the skip layer (ApplySide lin_*) by side workgroup w of nw.  Everything that
        /// depends on nothing is loaded first: the weights of this wave's first row (rows go to
        /// waves round-robin), u (once: |u| and the normalised input come from the same
        /// registers) and the activity flag; each later row's weights are issued before the
        /// previous row is computed.
        __device__ void lin_side_work(const ApplySide &sd, int w, int nw, float *xs)
        {
            __shared__ double lin_red[4];
            const int lane = threadIdx.x & 63;
            const int m = sd.lin_m, md = sd.lin_md, h = sd.lin_h;
            const int nch = md / 8; // 16-byte chunks per row (<= 64 LIN_CHUNKS)
            const int nwb = static_cast<int>(blockDim.x) / 64;
            const int wave_id = w * nwb + static_cast<int>(threadIdx.x) / 64, waves = nw * nwb;
            auto row_ptr = [&](int row) {
                const bool first = row < h;
                return reinterpret_cast<const uint4 *>((first ? sd.lin_w1 : sd.lin_s) +
                                                       static_cast<size_t>(first ? row : row - h) * md);
            };
            // A row's weights and, for rows of layer 1, its bias.
            auto load_row = [&](int row, uint4(&wv)[LIN_CHUNKS], float &bias) {
                const int r = min(row, sd.lin_rows - 1);
                const uint4 *wr = row_ptr(r);
#pragma unroll
                for (int k = 0; k < LIN_CHUNKS; ++k)
                    wv[k] = wr[min(lane + 64 * k, nch - 1)];
                bias = sd.lin_b1[min(r, h - 1)];
            };
            uint4 wv[LIN_CHUNKS];
            float bv;
            int row = wave_id;
            load_row(row, wv, bv);
            double uv[LIN_U_PER_THREAD];
#pragma unroll
            for (int k = 0; k < LIN_U_PER_THREAD; ++k)
                uv[k] = sd.lin_u[min(static_cast<int>(threadIdx.x + k * blockDim.x), m - 1)];
            const int act = sd.lin_active != nullptr ? sd.lin_active[0] : 1;
            if (act == 0)
                return; // uniform
            double q = 0.0;
#pragma unroll
            for (int k = 0; k < LIN_U_PER_THREAD; ++k)
                if (static_cast<int>(threadIdx.x + k * blockDim.x) < m)
                    q += uv[k] * uv[k];
            q = wave_sum(q);
            if (lane == 0)
                lin_red[threadIdx.x / 64] = q;
            __syncthreads();
            double nu2 = 0.0;
            for (int i = 0; i < nwb; ++i)
                nu2 += lin_red[i];
            const double nu = sqrt(nu2);
            const double fac = nu > 0.0 ? sqrt(static_cast<double>(m)) / nu : 0.0;
            if (w == 0 && threadIdx.x == 0)
                sd.lin_scale[0] = static_cast<float>(nu * sd.lin_ascale / sqrt(static_cast<double>(m)));
#pragma unroll
            for (int k = 0; k < LIN_U_PER_THREAD; ++k)
            {
                const int i = threadIdx.x + k * blockDim.x;
                if (i < m)
                    xs[i] = static_cast<float>(uv[k] * fac);
            }
            for (int i = m + threadIdx.x; i < md; i += blockDim.x)
                xs[i] = 0.0f;
            __syncthreads();
            for (; row < sd.lin_rows; row += waves) // wave-uniform
            {
                uint4 cur[LIN_CHUNKS];
#pragma unroll
                for (int k = 0; k < LIN_CHUNKS; ++k)
                    cur[k] = wv[k];
                const float bcur = bv;
                if (row + waves < sd.lin_rows)
                    load_row(row + waves, wv, bv);
                float acc = 0.0f;
#pragma unroll
                for (int k = 0; k < LIN_CHUNKS; ++k)
                {
                    const int c = lane + 64 * k;
                    if (c < nch)
                    {
                        const uint4 v = cur[k];
                        const float *x = xs + c * 8;
                        acc += bf16_lo(v.x) * x[0] + bf16_hi(v.x) * x[1] + bf16_lo(v.y) * x[2] + bf16_hi(v.y) * x[3] +
                               bf16_lo(v.z) * x[4] + bf16_hi(v.z) * x[5] + bf16_lo(v.w) * x[6] + bf16_hi(v.w) * x[7];
                    }
                }
                acc = wave_sum(acc);
                if (lane == 0)
                {
                    if (row < h)
                    {
                        const float y = acc + bcur;
                        sd.lin_h1[row] = y > 0.0f ? y : 0.0f;
                    }
                    else
                        sd.lin_skip[row - h] = acc;
                }
            }
        }

        /// Side work of workgroup `wg` < sd.wgs: restriction tasks (thread per small node, wave per
        /// large node), lattice gather, then dense coarsest rows (the inputs staged once per
        /// workgroup in `stage`, LDS of `stage_cap` elements, when they fit; wave per row, lanes
        /// over columns).  PART 1 compiles only the work of a first apply launch (restriction,
        /// lattice gather), PART 2 only that of a second one (dense rows, network layer 1),
        /// PART 0 all of it.
        template <int BD, typename T, int PART = 0>
        __device__ void apply_side_work(const ApplySide &sd, const T *__restrict__ mr, T *__restrict__ mz, int wg,
                                        T *stage, int stage_cap)
        {
            if constexpr (std::is_same<T, float>::value && PART != 1)
            {
                if (sd.lin_rows > 0 && wg >= sd.lin_wg0)
                {
                    lin_side_work(sd, wg - sd.lin_wg0, sd.wgs - sd.lin_wg0, stage);
                    return;
                }
            }
            const int nside = sd.lin_rows > 0 ? sd.lin_wg0 : sd.wgs;
            const int nt = nside * static_cast<int>(blockDim.x);
            const int t = wg * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
            const int lane = threadIdx.x & 63;
            if (PART != 2 && sd.nsmall + sd.nbig > 0)
            {
                const T *seg1 = static_cast<const T *>(sd.seg1);
                T *out = static_cast<T *>(sd.out);
                // A small node's list (at most RESTRICT_SMALL_MAX entries) is loaded in one batch,
                // then its values in one batch: two memory round trips per node.
                for (int i = t; i < sd.nsmall; i += nt)
                {
                    const int4 task = sd.rtask[i];
                    int q[RESTRICT_SMALL_MAX];
#pragma unroll
                    for (int k = 0; k < RESTRICT_SMALL_MAX; ++k)
                        q[k] = task.y + k < task.z ? sd.rlist[task.y + k] : -1;
                    T v[RESTRICT_SMALL_MAX][BD];
#pragma unroll
                    for (int k = 0; k < RESTRICT_SMALL_MAX; ++k)
#pragma unroll
                        for (int d = 0; d < BD; ++d)
                            v[k][d] = q[k] >= 0 ? seg1[static_cast<size_t>(q[k]) * BD + d] : T(0);
#pragma unroll
                    for (int d = 0; d < BD; ++d)
                    {
                        double acc = 0.0;
#pragma unroll
                        for (int k = 0; k < RESTRICT_SMALL_MAX; ++k)
                            if (q[k] >= 0)
                                acc += static_cast<double>(v[k][d]);
                        out[static_cast<size_t>(task.x) * BD + d] = static_cast<T>(acc);
                    }
                }
                // A piece (at most RESTRICT_PIECE entries) per wave: each lane loads its entries
                // in one batch, then their values in one batch.
                constexpr int PL = RESTRICT_PIECE / 64;
                for (int i = t >> 6; i < sd.nbig; i += nt >> 6)
                {
                    const int4 task = sd.rtask[sd.nsmall + i];
                    int q[PL];
#pragma unroll
                    for (int k = 0; k < PL; ++k)
                    {
                        const int e = task.y + lane + 64 * k;
                        q[k] = e < task.z ? sd.rlist[e] : -1;
                    }
                    T v[PL][BD];
#pragma unroll
                    for (int k = 0; k < PL; ++k)
#pragma unroll
                        for (int d = 0; d < BD; ++d)
                            v[k][d] = q[k] >= 0 ? seg1[static_cast<size_t>(q[k]) * BD + d] : T(0);
#pragma unroll
                    for (int d = 0; d < BD; ++d)
                    {
                        double acc = 0.0;
#pragma unroll
                        for (int k = 0; k < PL; ++k)
                            if (q[k] >= 0)
                                acc += static_cast<double>(v[k][d]);
                        acc = wave_sum(acc);
                        if (lane == 0)
                        {
                            if (task.w < 0)
                                out[static_cast<size_t>(task.x) * BD + d] = static_cast<T>(acc);
                            else
                                sd.rpart[static_cast<size_t>(task.w) * BD + d] = acc;
                        }
                    }
                }
            }
            if (PART != 2 && sd.lat_npts > 0)
            {
                // Lattice gather: a wave per point, counting waves from the far end so that the
                // waves with restriction pieces get points last.
                const int nw = nt >> 6;
                for (int J = nw - 1 - (t >> 6); J < sd.lat_npts; J += nw)
                {
                    double a[BD];
#pragma unroll
                    for (int d = 0; d < BD; ++d)
                        a[d] = 0.0;
                    for (int i = sd.lat_ptr[J] + lane; i < sd.lat_ptr[J + 1]; i += 64)
                    {
                        const size_t c = static_cast<size_t>(sd.lat_src[i]);
#pragma unroll
                        for (int d = 0; d < BD; ++d)
                            a[d] += sd.lat_partial[c * BD + d];
                    }
#pragma unroll
                    for (int d = 0; d < BD; ++d)
                    {
                        const double v = wave_sum(a[d]);
                        if (lane == 0)
                            sd.lat_u[static_cast<size_t>(J) * BD + d] = v;
                    }
                }
            }
            if (PART == 1)
                return;
            if (sd.dense_nd > 0 && sd.dense_copies <= 1)
            {
                const int nd = sd.dense_nd;
                const bool staged = nd <= stage_cap;
                if (staged)
                {
                    for (int c = threadIdx.x; c < nd; c += blockDim.x)
                        stage[c] = coarse_input<BD, T>(sd, mr, sd.dense_seg + sd.dense_ids[c / BD], c % BD);
                    __syncthreads();
                }
                for (int row = t >> 6; row < nd; row += nt >> 6)
                {
                    const double *a = sd.dense_inv + static_cast<size_t>(row) * nd;
                    double acc = 0.0;
                    for (int c = lane; c < nd; c += 64)
                    {
                        const T xc = staged ? stage[c]
                                            : coarse_input<BD, T>(sd, mr, sd.dense_seg + sd.dense_ids[c / BD], c % BD);
                        acc += a[c] * static_cast<double>(xc);
                    }
                    acc = wave_sum(acc);
                    if (lane == 0)
                    {
                        const int node = sd.dense_seg + sd.dense_ids[row / BD];
                        mz[static_cast<size_t>(node) * BD + row % BD] = static_cast<T>(acc);
                    }
                }
            }
            else if (sd.dense_nd > 0)
            {
                // Every copy's dense rows from one read of each inverse row (copy c's inputs at
                // stage[c nd ..] when they all fit); per copy the same arithmetic as one system.
                constexpr int MC = 8;
                const int nd = sd.dense_nd, nc = sd.dense_copies;
                const bool staged = nd * nc <= stage_cap;
                auto input = [&](int c, int col) {
                    return coarse_input<BD, T>(sd, mr, sd.dense_seg + c * sd.dense_stride + sd.dense_ids[col / BD],
                                               col % BD);
                };
                if (staged)
                {
                    for (int i = threadIdx.x; i < nd * nc; i += blockDim.x)
                        stage[i] = input(i / nd, i % nd);
                    __syncthreads();
                }
                for (int row = t >> 6; row < nd; row += nt >> 6)
                {
                    const double *a = sd.dense_inv + static_cast<size_t>(row) * nd;
                    double acc[MC];
#pragma unroll
                    for (int c = 0; c < MC; ++c)
                        acc[c] = 0.0;
                    for (int col = lane; col < nd; col += 64)
                    {
                        const double av = a[col];
#pragma unroll
                        for (int c = 0; c < MC; ++c)
                            if (c < nc)
                                acc[c] += av * static_cast<double>(staged ? stage[c * nd + col] : input(c, col));
                    }
#pragma unroll
                    for (int c = 0; c < MC; ++c)
                    {
                        if (c >= nc)
                            break;
                        const double v = wave_sum(acc[c]);
                        if (lane == 0)
                        {
                            const int node = sd.dense_seg + c * sd.dense_stride + sd.dense_ids[row / BD];
                            mz[static_cast<size_t>(node) * BD + row % BD] = static_cast<T>(v);
                        }
                    }
                }
            }
        }

        // FP16 tiled storage of a scaled bank matrix S (N = 32 BD, T = N / 8 tiles of 8 x 8 per
        // side, unit diagonal implicit), N*N/2 values in blocks of 64:
        //   block q < Q = T(T-1)/2: off-diagonal tile (I, J), I < J, row-major; tiles ordered by
        //   diagonal J - I = 1, 2, ... and by I within a diagonal (neighbouring blocks then start
        //   in different LDS banks when unpacked);
        //   block Q + p, p < T/2: diagonal tiles 2p and 2p+1 folded into one 8 x 8 block X with
        //   X[r][c] = S(16p + r, 16p + c) for c > r and X[r][c] = S(16p + 8 + c, 16p + 8 + r) for
        //   c < r (X[r][r] unused).
        template <int BD>
        constexpr int half_matrix_len() { return (32 * BD) * (32 * BD) / 2; }

        /// (I, J) of off-diagonal tile block q, I < J, for T tiles per side.
        template <int T>
        __device__ __forceinline__ void off_tile_of_block(int q, int &I, int &J)
        {
            int d = 1;
            while (q >= T - d) // diagonal d holds T - d tiles
            {
                q -= T - d;
                ++d;
            }
            I = q;
            J = q + d;
        }

        /// Matrix element (i, j), i <= j, held at position e of the tiled FP16 layout (i == j: unused).
        template <int BD>
        __device__ __forceinline__ void half_tile_element(int e, int &i, int &j)
        {
            constexpr int T = 32 * BD / 8;
            constexpr int Q = T * (T - 1) / 2;
            const int blk = e >> 6, r = (e >> 3) & 7, c = e & 7;
            if (blk < Q)
            {
                int I, J;
                off_tile_of_block<T>(blk, I, J);
                i = 8 * I + r;
                j = 8 * J + c;
            }
            else
            {
                const int base = 16 * (blk - Q);
                if (c >= r)
                {
                    i = base + r;
                    j = base + c;
                }
                else
                {
                    i = base + 8 + c;
                    j = base + 8 + r;
                }
            }
        }

        /// y_m = D_m S_m D_m x_m for FP16 tiled matrices [mat_begin, mat_begin + count), grid-stride.
        /// Each matrix is unpacked into a full padded row-major copy in LDS.  An off-diagonal tile
        /// row is one 16-byte store; its transpose is written by lane pairs (rows r, r^1 of a tile):
        /// after one lane exchange each lane packs the two rows' values of every other column into
        /// one 4-byte store.  Folded diagonal blocks are unpacked element by element.  The LDS
        /// addresses of every thread's chunks are decoded once per launch.  Thread (half, row) =
        /// (tid / N, tid % N) then sums half of row `row` with 16-byte LDS reads.  Two matrices are
        /// in flight per workgroup: the loads of the one after next are issued before the current
        /// matrix is computed.
        ///
        /// KC > 1 (cl.copies <= KC right-hand sides of one factorisation, CopyLayout): the items
        /// [mat_begin, mat_begin + count) are stored matrices; each tile is unpacked once and one
        /// pass over each LDS row multiplies it with every copy's input (KC accumulators), copy
        /// c's bank being copy 0's plus c times its level's bank count.  Per copy the arithmetic
        /// is that of a single system.
        template <int BD, int KC>
        __global__ __launch_bounds__(apply_block<BD>()) void k_mas_apply_h16(int mat_begin, int count,
                                                                              const __half *__restrict__ mats,
                                                                              const float *__restrict__ dscale,
                                                                              const float *__restrict__ mr,
                                                                              float *__restrict__ mz,
                                                                              const int *__restrict__ bank_sys,
                                                                              const int *__restrict__ active,
                                                                              ApplySide sd, CopyLayout cl)
        {
            constexpr int N = 32 * BD;
            constexpr int T = N / 8;
            constexpr int Q = T * (T - 1) / 2;
            constexpr int LEN = half_matrix_len<BD>();
            constexpr int BLOCK = apply_block<BD>();
            constexpr int CH = LEN / 8; // 16-byte chunks
            constexpr int PER = CH / BLOCK;
            constexpr int LDW = N + 8; // padded row, 16-byte multiple
            constexpr int HALF = N / 2;
            static_assert(BLOCK == 2 * N && CH % BLOCK == 0 && BLOCK % 8 == 0, "thread layout");
            using u32x4 = unsigned int __attribute__((ext_vector_type(4)));

            __shared__ __attribute__((aligned(16))) __half sF[N * LDW];
            // Copy rows 4 floats apart in bank order (N is a multiple of 32): the MFMA path reads
            // and writes one row per lane group, which would otherwise hit the same banks.
            constexpr int SXW = KC > 1 ? N + 4 : N;
            __shared__ __attribute__((aligned(16))) float sx[KC][SXW];
            __shared__ float part[KC][SXW];

            const int nc = KC > 1 ? cl.copies : 1;
            auto skip = [&](int m) {
                if (KC > 1 || bank_sys == nullptr)
                    return false; // (copies converge at different times: every copy is applied)
                const int sy = bank_sys[m]; // -1: a coarse matrix (its nodes may span systems)
                return sy >= 0 && active[sy] == 0;
            };
            if (static_cast<int>(blockIdx.x) < sd.wgs)
            {
                apply_side_work<BD, float>(sd, mr, mz, blockIdx.x, reinterpret_cast<float *>(sF),
                                           static_cast<int>(sizeof(sF) / (sizeof(float))));
                return;
            }
            const int nwg = static_cast<int>(gridDim.x) - sd.wgs; // apply workgroups
            const int end = mat_begin + count;
            auto step = [&](int m) {
                m += nwg;
                while (m < end && skip(m))
                    m += nwg;
                return m;
            };
            int cur = mat_begin + static_cast<int>(blockIdx.x) - sd.wgs;
            while (cur < end && skip(cur))
                cur += nwg;
            if (cur >= end)
                return;

            const int tid = threadIdx.x;
            // LDS targets of this thread's chunks (in halves).  Off-diagonal tile row r: ca = start
            // of the row copy, cb = first 4-byte transposed target (row 8J + (r & 1), columns
            // 8I + (r & ~1) .. +1), sel = byte selector pairing this lane's and its partner's
            // values.  Folded block: ca = base + r LDW, cb = base + r (base = top-left corner).
            int ca[PER], cb[PER], cr[PER];
            unsigned sel[PER];
            bool folded[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k)
            {
                const int c = tid + k * BLOCK;
                const int blk = c >> 3, r = c & 7;
                cr[k] = r;
                // even row: own low half, partner low half; odd row: partner high, own high
                sel[k] = (r & 1) ? 0x03020706u : 0x05040100u;
                if (blk < Q)
                {
                    int I, J;
                    off_tile_of_block<T>(blk, I, J);
                    ca[k] = (8 * I + r) * LDW + 8 * J;
                    cb[k] = (8 * J + (r & 1)) * LDW + 8 * I + (r & ~1);
                    folded[k] = false;
                }
                else
                {
                    const int base = 16 * (blk - Q) * (LDW + 1);
                    ca[k] = base + r * LDW;
                    cb[k] = base + r;
                    folded[k] = true;
                }
            }
            for (int i = tid; i < N; i += BLOCK)
                sF[i * LDW + i] = __float2half_rn(1.0f); // never overwritten below

            // A matrix in flight: its tile in registers, and for thread tid < N (the one that also
            // stores row tid) the raw loads of the row's scale, each copy's input and the dense
            // ownership of its node (RawInput), combined after the next barrier, so that nothing
            // global is read or waited for between the barriers.
            struct Row
            {
                RawInput<float> in[KC];
                float d = 0.0f;
                bool dense = false;
                int dix = -1;
                int bank = 0, stride = 0; ///< copy 0's bank of the item, bank stride between copies
                __device__ bool store() const { return !(dense && dix >= 0); }
            };
            // The row loads go first and the tile loads last, so that waiting for this matrix's
            // row values later leaves the next matrix's tile loads in flight.
            auto load = [&](u32x4(&reg)[PER], Row &rw, int mm) {
                const int r = tid % N; // (threads past N repeat row tid - N's loads, unused)
                if constexpr (KC > 1)
                    bank_of_stored(cl, mm, rw.bank, rw.stride);
                else
                    rw.bank = mm;
                rw.d = dscale[static_cast<size_t>(mm) * N + r];
#pragma unroll
                for (int c = 0; c < KC; ++c)
                    rw.in[c] = raw_input<BD, float>(sd, mr, (rw.bank + (c < nc ? c : nc - 1) * rw.stride) * 32 + r / BD,
                                                    r % BD);
                dense_probe(sd, rw.bank * 32 + r / BD, rw.dense, rw.dix);
                const u32x4 *src = reinterpret_cast<const u32x4 *>(mats + static_cast<size_t>(mm) * LEN);
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    reg[k] = __builtin_nontemporal_load(src + tid + k * BLOCK);
            };
            // What the compute phase needs of a matrix's row loads, taken out in unpack (which
            // waits for them anyway), so that the next load() may reuse the Row's registers
            // without waiting for anything.
            struct Out
            {
                float d;
                bool store;
                int bank, stride;
            };
            auto unpack = [&](const u32x4(&reg)[PER], const Row &rw) -> Out {
#pragma unroll
                for (int k = 0; k < PER; ++k)
                {
                    const u32x4 v = reg[k];
                    // Folded blocks hold the last chunks: slots with (k + 1) BLOCK <= 8 Q have none.
                    const bool fold = (k + 1) * BLOCK > 8 * Q && folded[k];
                    if (!fold)
                    {
                        *reinterpret_cast<u32x4 *>(sF + ca[k]) = v;
#pragma unroll
                        for (int w = 0; w < 4; ++w)
                        {
                            // quad_perm [1, 0, 3, 2]: the partner lane (row r ^ 1 of the same tile)
                            const unsigned pw = static_cast<unsigned>(__builtin_amdgcn_update_dpp(
                                0, static_cast<int>(v[w]), 0xB1, 0xF, 0xF, false));
                            *reinterpret_cast<unsigned *>(sF + cb[k] + 2 * w * LDW) =
                                __builtin_amdgcn_perm(pw, v[w], sel[k]);
                        }
                    }
                    else
                    {
                        __half h[8];
                        __builtin_memcpy(h, &v, 16);
                        const int r = cr[k];
#pragma unroll
                        for (int c = 0; c < 8; ++c)
                        {
                            const bool up = c > r; // element of the first tile, else of the second
                            const int p1 = up ? ca[k] + c : cb[k] + (8 + c) * LDW + 8;
                            const int p2 = up ? cb[k] + c * LDW : ca[k] + 8 * LDW + 8 + c;
                            if (c != r)
                            {
                                sF[p1] = h[c];
                                sF[p2] = h[c];
                            }
                        }
                    }
                }
                if (tid < N)
                {
#pragma unroll
                    for (int c = 0; c < KC; ++c)
                        sx[c][tid] = c < nc ? input_value<BD, float>(sd, rw.in[c], tid % BD) * rw.d : 0.0f;
                }
                const Out o{rw.d, rw.store(), rw.bank, rw.stride};
                __syncthreads();
                return o;
            };
            const int row = tid % N, hf = tid / N;
            auto compute_valu = [&](const Out &rw) {
                const __half *arow = sF + row * LDW + hf * HALF;
                float acc[KC];
#pragma unroll
                for (int c = 0; c < KC; ++c)
                    acc[c] = 0.0f;
#pragma unroll
                for (int t = 0; t < HALF; t += 8)
                {
                    const u32x4 av = *reinterpret_cast<const u32x4 *>(arow + t);
                    __half h[8];
                    __builtin_memcpy(h, &av, 16);
                    // Every accumulator runs (copies past nc have zero inputs): a loop without an
                    // exit unrolls fully and its LDS reads issue ahead of the FMAs.
#pragma unroll
                    for (int c = 0; c < KC; ++c)
                    {
                        const float *xs = sx[c] + hf * HALF;
                        const float4 x0 = *reinterpret_cast<const float4 *>(xs + t);
                        const float4 x1 = *reinterpret_cast<const float4 *>(xs + t + 4);
                        acc[c] = fmaf(__half2float(h[0]), x0.x, acc[c]);
                        acc[c] = fmaf(__half2float(h[1]), x0.y, acc[c]);
                        acc[c] = fmaf(__half2float(h[2]), x0.z, acc[c]);
                        acc[c] = fmaf(__half2float(h[3]), x0.w, acc[c]);
                        acc[c] = fmaf(__half2float(h[4]), x1.x, acc[c]);
                        acc[c] = fmaf(__half2float(h[5]), x1.y, acc[c]);
                        acc[c] = fmaf(__half2float(h[6]), x1.z, acc[c]);
                        acc[c] = fmaf(__half2float(h[7]), x1.w, acc[c]);
                    }
                }
                if (hf == 1)
                {
#pragma unroll
                    for (int c = 0; c < KC; ++c)
                        part[c][row] = acc[c];
                }
                __syncthreads(); // also: sF and sx may be rewritten after this point
                if (hf == 0 && rw.store) // (hf == 0: row == tid)
                {
#pragma unroll
                    for (int c = 0; c < KC; ++c)
                    {
                        if (c >= nc)
                            break;
                        mz[static_cast<size_t>(rw.bank + c * rw.stride) * N + row] = (acc[c] + part[c][row]) * rw.d;
                    }
                }
                // part[] is next written after the following unpack's barrier
            };
            // KC = 8: Y = S X on MFMA (v_mfma_f32_16x16x16f16, FP32 accumulation).  Wave w owns
            // rows 32w .. 32w + 31 (two 16-row tiles) over all N / 16 k-steps; the B operand is
            // X (N x 16, columns past the copies zero) in FP16 as hi + lo parts of x 2^-e_j, e_j
            // the exponent of copy j's largest |x| (exact power-of-two scaling), so the two
            // products carry about 22 bits of x.  Lane l: A[i = l % 16][k = 4 (l / 16) + v],
            // B[k = 4 (l / 16) + v][j = l % 16], D[i = 4 (l / 16) + v][j = l % 16].
            auto compute_mfma = [&](const Out &rw) {
                using h4 = _Float16 __attribute__((ext_vector_type(4)));
                using f4 = float __attribute__((ext_vector_type(4)));
                const int lane = tid & 63, w = tid >> 6, q = lane >> 4, j = lane & 15;
                // Largest |x| of copy lane % 8 over the rows (lanes l, l + 8, ... share a copy).
                float mx = 0.0f;
                for (int i = lane >> 3; i < N; i += 8)
                    mx = fmaxf(mx, fabsf(sx[lane & 7][i]));
                mx = fmaxf(mx, __shfl_xor(mx, 8));
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                mx = fmaxf(mx, __shfl_xor(mx, 32));
                int ex = 0;
                (void)frexpf(mx, &ex);
                const int e = __shfl(ex, j & 7); // copy j's exponent (lanes with j >= 8: unused)
                const bool jv = j < nc;
                f4 acc[2];
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < N / 16; ++ks)
                {
                    const int k = 16 * ks + 4 * q;
                    h4 bh = h4{0, 0, 0, 0}, bl = h4{0, 0, 0, 0};
                    if (jv)
                    {
                        const float4 xv = *reinterpret_cast<const float4 *>(sx[j] + k);
                        const float xs[4] = {ldexpf(xv.x, -e), ldexpf(xv.y, -e), ldexpf(xv.z, -e), ldexpf(xv.w, -e)};
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                        {
                            const _Float16 hi = static_cast<_Float16>(xs[v]);
                            bh[v] = hi;
                            bl[v] = static_cast<_Float16>(xs[v] - static_cast<float>(hi));
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                    {
                        const int r = 32 * w + 16 * t + j;
                        const h4 a = *reinterpret_cast<const h4 *>(sF + r * LDW + k);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(a, bh, acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(a, bl, acc[t], 0, 0, 0);
                    }
                }
                if (jv)
                {
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            part[j][32 * w + 16 * t + 4 * q + v] = ldexpf(acc[t][v], e);
                }
                __syncthreads(); // also: sF and sx may be rewritten after this point
                if (tid < N && rw.store)
                {
#pragma unroll
                    for (int c = 0; c < KC; ++c)
                        if (c < nc)
                            mz[static_cast<size_t>(rw.bank + c * rw.stride) * N + tid] = part[c][tid] * rw.d;
                }
                // part[] is next written after the following unpack's barrier
            };
            auto compute = [&](const Out &rw) {
                if constexpr (KC == 8)
                    compute_mfma(rw);
                else
                    compute_valu(rw);
            };

            // Invariant: regA holds matrix cur, regB matrix nxt (when nxt < end).
            u32x4 regA[PER], regB[PER];
            Row ra, rb;
            int nxt = step(cur);
            load(regA, ra, cur);
            if (nxt < end)
                load(regB, rb, nxt);
            while (true)
            {
                int nn = nxt < end ? step(nxt) : end;
                Out now = unpack(regA, ra);
                if (nn < end)
                    load(regA, ra, nn);
                compute(now);
                cur = nxt;
                nxt = nn;
                if (cur >= end)
                    break;
                nn = nxt < end ? step(nxt) : end;
                now = unpack(regB, rb);
                if (nn < end)
                    load(regB, rb, nn);
                compute(now);
                cur = nxt;
                nxt = nn;
                if (cur >= end)
                    break;
            }
        }
        /// Sum over the 8 lanes of an aligned 8-lane group of v[c]: lane r (= lane % 8) receives
        /// sum over the group's lanes r' of v_r'[r].  Three halving exchanges (half-row mirror,
        /// quad_perm [2, 3, 0, 1], quad_perm [1, 0, 3, 2]); at each a lane keeps the half of its
        /// values whose index bit matches its own lane bit and adds the partner's copy of that
        /// half.  Fixed order: the result is reproducible.
        __device__ __forceinline__ float group8_transpose_sum(const float (&v)[8], int lane)
        {
            auto mv = [](float x, auto ctrl) {
                return __int_as_float(
                    __builtin_amdgcn_update_dpp(0, __float_as_int(x), decltype(ctrl)::value, 0xF, 0xF, false));
            };
            const bool b2 = (lane & 4) != 0, b1 = (lane & 2) != 0, b0 = (lane & 1) != 0;
            float k4[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const float keep = b2 ? v[j + 4] : v[j];
                const float send = b2 ? v[j] : v[j + 4];
                k4[j] = keep + mv(send, std::integral_constant<int, 0x141>()); // lane i <-> 7 - i
            }
            float k2[2];
#pragma unroll
            for (int j = 0; j < 2; ++j)
            {
                const float keep = b1 ? k4[j + 2] : k4[j];
                const float send = b1 ? k4[j] : k4[j + 2];
                k2[j] = keep + mv(send, std::integral_constant<int, 0x4E>()); // lane i <-> i ^ 2
            }
            const float keep = b0 ? k2[1] : k2[0];
            const float send = b0 ? k2[0] : k2[1];
            return keep + mv(send, std::integral_constant<int, 0xB1>()); // lane i <-> i ^ 1
        }

        /// Tiles of the blocks g + 8 k (g = 0..7) of the tiled FP16 layout with NT tiles per side:
        /// byte g holds I | J << 4 of off-diagonal tile (I, J), or p | 15 << 4 for folded diagonal
        /// block p.  Evaluated at compile time.
        template <int NT>
        constexpr unsigned long long w16_tiles(int k)
        {
            constexpr int Q = NT * (NT - 1) / 2;
            unsigned long long v = 0;
            for (int g = 0; g < 8; ++g)
            {
                int q = g + 8 * k, I = 0, J = 15;
                if (q < Q)
                {
                    int d = 1;
                    while (q >= NT - d)
                    {
                        q -= NT - d;
                        ++d;
                    }
                    I = q;
                    J = q + d;
                }
                else
                    I = q - Q;
                v |= static_cast<unsigned long long>(I | (J << 4)) << (8 * g);
            }
            return v;
        }

        /// Calls f(std::integral_constant<int, k>()) for k = 0 .. K-1 (compile-time loop index).
        template <int K, typename F, int... Ks>
        __device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Ks...>)
        {
            (f(std::integral_constant<int, Ks>()), ...);
        }
        template <int K, typename F>
        __device__ __forceinline__ void static_for(F &&f)
        {
            static_for_impl<K>(f, std::make_integer_sequence<int, K>());
        }

        /// Shape of the wave-per-matrix FP16 apply: N = 32 BD, NT = N / 8 tiles per side, NB = NT^2 / 2
        /// blocks of 64 halves (Q off-diagonal tiles, then NT / 2 folded diagonal blocks), KB = NB / 8
        /// blocks per 8-lane group; per wave the scaled input (N floats) and the partial sums
        /// (N rows of NT slots, odd row stride PS) in LDS.
        template <int BD>
        struct W16Shape
        {
            static constexpr int N = 32 * BD;
            static constexpr int NT = N / 8;
            static constexpr int Q = NT * (NT - 1) / 2;
            static constexpr int NB = NT * NT / 2;
            static constexpr int KB = NB / 8;
            static constexpr int PS = NT + 1;
            static constexpr int WAVE_FLOATS = N + N * PS;
            static constexpr int LDS_FLOATS = APPLY_W16_WAVES * WAVE_FLOATS;
            static_assert(NB % 8 == 0 && NB == Q + NT / 2 && (N * PS) % 4 == 0, "tile layout");
        };

        /// y_m = D_m S_m D_m x_m for FP16 tiled matrices [mat_begin, mat_begin + count): a wave per
        /// matrix, four waves per workgroup, each wave on its own (no workgroup barrier).  Lane l
        /// loads rows l, l + 64, ... of the matrix's 16-byte tile rows in one batch (each load
        /// instruction of the wave reads 1 KB contiguous), so that the tile stays in registers and
        /// every resident wave keeps its whole matrix in flight; nothing is unpacked into LDS.
        /// Lane (group g = l / 8, row r = l % 8) holds row r of blocks g, g + 8, ...:
        ///  * off-diagonal tile (I, J), I < J: z[8I + r] += S_IJ[r][:] . x[8J ..] (its own row) and
        ///    z[8J + c] += S_IJ[r][c] x[8I + r] (summed over the group's 8 lanes by
        ///    group8_transpose_sum, lane c receiving column c);
        ///  * folded diagonal block p: the strict upper triangle of tile 2p (c > r) and the strict
        ///    lower one of tile 2p + 1 (c < r, stored transposed), each in both orientations.
        /// Every tile contribution to row i lands in its own slot of the row's partials
        /// (slot J for the row part of tile (I, J), slot I for the column part, slot I for the
        /// diagonal tile), and z_i = d_i (x~_i + sum of the slots in slot order) with x~ = D x and
        /// the unit diagonal of S: a fixed summation order.  Matrices past the first
        /// 4 (gridDim - wgs) are walked grid-stride (apply_workgroups caps the grid).
        template <int BD, bool COARSE>
        __global__ __launch_bounds__(APPLY_W16_BLOCK) void k_mas_apply_w16(int mat_begin, int count,
                                                                          const __half *__restrict__ mats,
                                                                          const float *__restrict__ dscale,
                                                                          const float *__restrict__ mr,
                                                                          float *__restrict__ mz,
                                                                          const int *__restrict__ bank_sys,
                                                                          const int *__restrict__ active,
                                                                          ApplySide sd)
        {
            using S = W16Shape<BD>;
            constexpr int N = S::N, NT = S::NT, Q = S::Q, KB = S::KB, PS = S::PS;
            constexpr int LEN = half_matrix_len<BD>();
            constexpr int RPL = (N + 63) / 64; // rows of the input and output per lane
            using u32x4 = unsigned int __attribute__((ext_vector_type(4)));
            __shared__ __attribute__((aligned(16))) float lds[S::LDS_FLOATS];
            if (static_cast<int>(blockIdx.x) < sd.wgs)
            {
                apply_side_work<BD, float, COARSE ? 2 : 1>(sd, mr, mz, blockIdx.x, lds, S::LDS_FLOATS);
                return;
            }
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            float *sx = lds + wave * S::WAVE_FLOATS; // x~ = D x
            float *sp = sx + N;                      // partials [N][PS]
            const int end = mat_begin + count;
            const int stride = (static_cast<int>(gridDim.x) - sd.wgs) * APPLY_W16_WAVES;
            for (int m = mat_begin + (static_cast<int>(blockIdx.x) - sd.wgs) * APPLY_W16_WAVES + wave; m < end;
                 m += stride) // wave-uniform
            {
                if (bank_sys != nullptr)
                {
                    const int sy = bank_sys[m]; // -1: a coarse matrix (its nodes may span systems)
                    if (sy >= 0 && active[sy] == 0)
                        continue;
                }
                // Row loads first (their wait then leaves the tile in flight), all unconditional.
                // Split inputs and the dense coarsest rows exist only in a second launch (COARSE).
                RawInput<float> in[RPL];
                float d[RPL];
                bool dn[RPL];
                int dx[RPL];
#pragma unroll
                for (int j = 0; j < RPL; ++j)
                {
                    const int i = min(lane + 64 * j, N - 1);
                    d[j] = dscale[static_cast<size_t>(m) * N + i];
                    if constexpr (COARSE)
                    {
                        in[j] = raw_input<BD, float>(sd, mr, m * 32 + i / BD, i % BD);
                        dense_probe(sd, m * 32 + i / BD, dn[j], dx[j]);
                    }
                    else
                    {
                        in[j].x = mr[static_cast<size_t>(m) * N + i];
                        dn[j] = false;
                        dx[j] = -1;
                    }
                }
                const u32x4 *src = reinterpret_cast<const u32x4 *>(mats + static_cast<size_t>(m) * LEN);
                u32x4 tile[KB];
#pragma unroll
                for (int k = 0; k < KB; ++k)
                    tile[k] = __builtin_nontemporal_load(src + lane + 64 * k);
                float xt[RPL];
#pragma unroll
                for (int j = 0; j < RPL; ++j)
                {
                    const int i = lane + 64 * j;
                    xt[j] = (COARSE ? input_value<BD, float>(sd, in[j], i % BD) : in[j].x) * d[j];
                    if (i < N)
                        sx[i] = xt[j];
                }
                __builtin_amdgcn_wave_barrier(); // (LDS operations of one wave complete in order)
                // The blocks' LDS addresses depend only on the lane: made opaque here so that they
                // are computed per block, not hoisted out of the matrix loop (nine blocks' worth
                // of addresses would stay live and cost occupancy).
                int g = lane >> 3, r = lane & 7;
                asm volatile("" : "+v"(g), "+v"(r));
                static_for<KB>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    // This group's tile, from a compile-time table; blocks 8k .. 8k + 7 all lie
                    // before the folded ones when 8k + 7 < Q.
                    constexpr unsigned long long tab = w16_tiles<NT>(k);
                    const int tij = static_cast<int>((tab >> (8 * g)) & 0xFFull);
                    const int I = tij & 15, J = tij >> 4;
                    float h[8];
                    {
                        __half hv[8];
                        __builtin_memcpy(hv, &tile[k], 16);
#pragma unroll
                        for (int c = 0; c < 8; ++c)
                            h[c] = __half2float(hv[c]);
                    }
                    if (8 * k + 7 < Q || J != 15)
                    {
                        const float4 x0 = *reinterpret_cast<const float4 *>(sx + 8 * J);
                        const float4 x1 = *reinterpret_cast<const float4 *>(sx + 8 * J + 4);
                        const float xi = sx[8 * I + r];
                        float row = h[0] * x0.x;
                        row = fmaf(h[1], x0.y, row);
                        row = fmaf(h[2], x0.z, row);
                        row = fmaf(h[3], x0.w, row);
                        row = fmaf(h[4], x1.x, row);
                        row = fmaf(h[5], x1.y, row);
                        row = fmaf(h[6], x1.z, row);
                        row = fmaf(h[7], x1.w, row);
                        float v[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c)
                            v[c] = h[c] * xi;
                        const float col = group8_transpose_sum(v, lane);
                        sp[(8 * I + r) * PS + J] = row;
                        sp[(8 * J + r) * PS + I] = col;
                    }
                    if (8 * k + 7 >= Q && J == 15)
                    {
                        const int p = I, b = 16 * p;
                        const float4 a0 = *reinterpret_cast<const float4 *>(sx + b);
                        const float4 a1 = *reinterpret_cast<const float4 *>(sx + b + 4);
                        const float4 c0 = *reinterpret_cast<const float4 *>(sx + b + 8);
                        const float4 c1 = *reinterpret_cast<const float4 *>(sx + b + 12);
                        const float xa[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
                        const float xb[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
                        const float xar = sx[b + r], xbr = sx[b + 8 + r];
                        float row1 = 0.0f, row2 = 0.0f, v1[8], v2[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c)
                        {
                            // c > r: S(b + r, b + c) of tile 2p; c < r: S(b + 8 + c, b + 8 + r) of tile 2p + 1
                            const bool up = c > r, lo = c < r;
                            row1 = fmaf(up ? h[c] : 0.0f, xa[c], row1);
                            row2 = fmaf(lo ? h[c] : 0.0f, xb[c], row2);
                            v1[c] = up ? h[c] * xar : 0.0f;
                            v2[c] = lo ? h[c] * xbr : 0.0f;
                        }
                        const float col1 = group8_transpose_sum(v1, lane);
                        const float col2 = group8_transpose_sum(v2, lane);
                        sp[(b + r) * PS + 2 * p] = row1 + col1;
                        sp[(b + 8 + r) * PS + 2 * p + 1] = row2 + col2;
                    }
                    // One block at a time: interleaving the blocks' LDS reads and products costs
                    // registers (occupancy) and gains nothing once the tile has arrived.
                    __builtin_amdgcn_sched_barrier(0);
                });
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int j = 0; j < RPL; ++j)
                {
                    const int i = lane + 64 * j;
                    if (i < N)
                    {
                        float acc = xt[j];
#pragma unroll
                        for (int t = 0; t < NT; ++t)
                            acc += sp[i * PS + t];
                        if (!(dn[j] && dx[j] >= 0))
                            mz[static_cast<size_t>(m) * N + i] = acc * d[j];
                    }
                }
                __builtin_amdgcn_wave_barrier(); // sx and sp are rewritten for the next matrix
            }
        }

        /// y_m = Inv_m x_m for FP64 packed matrices [mat_begin, mat_begin + count), grid-stride:
        /// each workgroup takes every nwg-th matrix.  The packed upper matrix moves global ->
        /// registers -> LDS; the next matrix's loads are issued before the current one is
        /// computed, so every workgroup keeps a matrix in flight.  Threads (row, half) each sum half
        /// of row `row`, reading element (t, row) for t < row (packed row t, contiguous across
        /// threads) and (row, t) otherwise.  (Bandwidth-bound in FP64; the FP16 copy uses the tiled
        /// kernel above.)
        template <int BD, typename TM, typename T>
        __global__ __launch_bounds__(apply_block<BD>()) void k_mas_apply(int mat_begin, int count,
                                                                          const TM *__restrict__ mats,
                                                                          const float *__restrict__ dscale,
                                                                          const T *__restrict__ mr, T *__restrict__ mz,
                                                                          const int *__restrict__ bank_sys,
                                                                          const int *__restrict__ active,
                                                                          ApplySide sd, CopyLayout cl)
        {
            constexpr int N = 32 * BD;
            constexpr int L = N * (N + 1) / 2;
            constexpr int BLOCK = apply_block<BD>();
            static_assert(std::is_same<TM, double>::value && std::is_same<T, double>::value,
                          "the packed apply is the FP64 path's");
            using TA = double;
            using u32x4 = unsigned int __attribute__((ext_vector_type(4)));
            static_assert((L * sizeof(TM)) % 16 == 0, "packed matrix must be a whole number of 16 B chunks");
            constexpr int CH = static_cast<int>(L * sizeof(TM) / 16);
            constexpr int PER = (CH + BLOCK - 1) / BLOCK; // 16-byte chunks per thread

            __shared__ __attribute__((aligned(16))) TM sA[L];
            __shared__ TA sx[N];
            __shared__ TA part[N];

            auto skip = [&](int m) {
                if (bank_sys == nullptr)
                    return false;
                const int sy = bank_sys[m]; // -1: a coarse matrix (its nodes may span systems)
                return sy >= 0 && active[sy] == 0;
            };
            if (static_cast<int>(blockIdx.x) < sd.wgs)
            {
                apply_side_work<BD, T>(sd, mr, mz, blockIdx.x, reinterpret_cast<T *>(sA),
                                       static_cast<int>(sizeof(sA) / (sizeof(T))));
                return;
            }
            const int nwg = static_cast<int>(gridDim.x) - sd.wgs; // apply workgroups
            const int end = mat_begin + count;
            int m = mat_begin + static_cast<int>(blockIdx.x) - sd.wgs;
            while (m < end && skip(m))
                m += nwg;
            if (m >= end)
                return;
            u32x4 reg[PER];
            // Row tid's raw input and dense probe (used after the next barrier, RawInput); the
            // row loads go before the tile loads.
            RawInput<T> rin;
            bool dn = false;
            int dx = -1;
            auto load = [&](int mm) {
                const int r = threadIdx.x % N; // (threads past N repeat row tid - N's loads, unused)
                rin = raw_input<BD, T>(sd, mr, mm * 32 + r / BD, r % BD);
                dense_probe(sd, mm * 32 + r / BD, dn, dx);
                // Several right-hand sides of one factorisation: bank mm's matrix is stored once.
                const int sm = cl.copies > 1 ? stored_of_bank(cl, mm) : mm;
                const u32x4 *src = reinterpret_cast<const u32x4 *>(mats + static_cast<size_t>(sm) * L);
#pragma unroll
                for (int k = 0; k < PER; ++k)
                {
                    const int i = threadIdx.x + k * BLOCK;
                    if (i < CH)
                        reg[k] = __builtin_nontemporal_load(src + i);
                }
            };
            load(m);
            const int row = threadIdx.x % N;
            const int half = threadIdx.x / N; // 0 or 1
            const int t0 = half * (N / 2), t1 = t0 + N / 2;
            const int rs_row = row * N - (row * (row + 1)) / 2;
            while (m < end)
            {
                u32x4 *dst = reinterpret_cast<u32x4 *>(sA);
#pragma unroll
                for (int k = 0; k < PER; ++k)
                {
                    const int i = threadIdx.x + k * BLOCK;
                    if (i < CH)
                        dst[i] = reg[k];
                }
                if (threadIdx.x < N)
                    sx[threadIdx.x] = input_value<BD, T>(sd, rin, threadIdx.x % BD);
                const bool store_row = !(dn && dx >= 0); // (row tid's, for the threads that store)
                __syncthreads();
                int next = m + nwg;
                while (next < end && skip(next))
                    next += nwg;
                if (next < end)
                    load(next); // in flight while this matrix is computed
                int rs_t = t0 * N - (t0 * (t0 + 1)) / 2;
                TA acc = TA(0);
                if (threadIdx.x < 2 * N)
                {
#pragma unroll 8
                    for (int t = t0; t < t1; ++t)
                    {
                        const int idx = (t < row) ? (rs_t + row) : (rs_row + t);
                        acc = fma(sA[idx], sx[t], acc);
                        rs_t += N - t - 1;
                    }
                }
                if (half == 1 && threadIdx.x < 2 * N)
                    part[row] = acc;
                __syncthreads();
                if (half == 0 && store_row) // (half == 0: row == threadIdx.x)
                    mz[static_cast<size_t>(m) * N + row] = acc + part[row];
                __syncthreads(); // sA, sx and part are rewritten for the next matrix
                m = next;
            }
        }

        // ----------------------------------------------------------------------- prolongation
        template <int BD, typename T>
        __global__ __launch_bounds__(NODE_BLOCK) void k_prolong_dot(int level_num, const int4 *__restrict__ cmap,
                                                                    const int *__restrict__ lev_off,
                                                                    const T *__restrict__ r, T *__restrict__ mz,
                                                                    Reductions red, bool init, SysMap sm,
                                                                    LccApply lcc, FlexDot<T> flex)
        {
            const int sys = system_of_block(sm, blockIdx.x);
            if (system_stopped(sm, sys))
                return;
            if (!init && threadIdx.x == 0)
                red.rz_old[blockIdx.x] = red.rz[blockIdx.x];
            const int p = blockIdx.x * NODE_BLOCK + threadIdx.x;
            const int4 c = cmap[p];
            const int cc[4] = {c.x, c.y, c.z, c.w};
            // r (and the previous z) do not depend on the coarse gathers: load them with the first
            // round trip.
            T rv[BD], zp[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                rv[i] = r[static_cast<size_t>(p) * BD + i];
                zp[i] = flex.zprev != nullptr ? flex.zprev[static_cast<size_t>(p) * BD + i] : T(0);
            }
            double z[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                z[i] = static_cast<double>(mz[static_cast<size_t>(p) * BD + i]);
            if (cc[0] >= 0)
            {
                for (int l = 0; l < level_num; ++l)
                {
                    const size_t o = static_cast<size_t>(lev_off[l + 1] + cc[l]) * BD;
#pragma unroll
                    for (int i = 0; i < BD; ++i)
                        z[i] += static_cast<double>(mz[o + i]);
                }
            }
            if (lcc.cell != nullptr)
            {
                // Learned lattice correction: trilinear interpolation of the coarse coefficients.
                const int4 lc = lcc.cell[p];
                if (lc.x >= 0)
                {
                    const float fx = __int_as_float(lc.y), fy = __int_as_float(lc.z), fz = __int_as_float(lc.w);
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                    {
                        const float w = ((q & 1) ? fx : 1.0f - fx) * ((q & 2) ? fy : 1.0f - fy) *
                                        ((q & 4) ? fz : 1.0f - fz);
                        const int id = lc.x + (q & 1) + ((q >> 1) & 1) * lcc.k + ((q >> 2) & 1) * lcc.k * lcc.k;
#pragma unroll
                        for (int i = 0; i < BD; ++i)
                            z[i] += static_cast<double>(w) * static_cast<double>(lcc.coef[static_cast<size_t>(id) * BD + i]);
                    }
                }
            }
            double dot = 0.0, dot_old = 0.0;
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                const T zi = static_cast<T>(z[i]);
                mz[static_cast<size_t>(p) * BD + i] = zi;
                const double ri = static_cast<double>(rv[i]);
                dot += ri * static_cast<double>(zi);
                if (flex.zprev != nullptr)
                    dot_old += ri * static_cast<double>(zp[i]);
            }
            store_partial<NODE_BLOCK>(dot, red.rz, blockIdx.x);
            if (flex.zprev != nullptr)
                store_partial<NODE_BLOCK>(dot_old, flex.rzo, blockIdx.x);
        }

        // ----------------------------------------------------------------------- build: levels
        /// Closure of a node's neighbour mask inside its bank (bit-BFS).
        __device__ __forceinline__ uint32_t closure(uint32_t conn, int lane32, const uint32_t *bank_masks)
        {
            uint32_t visited = 1u << lane32;
            uint32_t to_visit = conn & ~visited;
            while (to_visit)
            {
                const int v = __builtin_ctz(to_visit);
                conn |= bank_masks[v];
                visited |= 1u << v;
                to_visit = conn & ~visited;
            }
            return conn;
        }

        __device__ __forceinline__ void leaders_and_ids(bool valid, uint32_t conn, int lane32, uint32_t &lead_mask,
                                                        int &local_id, int &bank_count)
        {
            const bool is_lead = valid && (__builtin_ctz(conn) == lane32);
            const unsigned long long ball = __ballot(is_lead);
            const uint32_t half = static_cast<uint32_t>(ball >> (threadIdx.x & 32));
            const int lead_lane = valid ? __builtin_ctz(conn) : 0;
            const uint32_t below = static_cast<uint32_t>((1ull << lead_lane) - 1ull);
            local_id = valid ? __builtin_popcount(half & below) : -1;
            lead_mask = is_lead ? conn : 0u;
            bank_count = __builtin_popcount(half);
        }

        __global__ __launch_bounds__(256) void k_cc_level0(int npad, SellView<double> A, const int *__restrict__ row_len,
                                                           const int *__restrict__ block_of_slot, uint32_t *lead,
                                                           int *local_id, int *bank_count)
        {
            __shared__ uint32_t masks[256];
            const int p = blockIdx.x * 256 + threadIdx.x;
            const int lane32 = threadIdx.x & 31;
            const int bank = p / 32;
            const bool valid = p < npad && block_of_slot[p] >= 0;
            uint32_t m = valid ? (1u << lane32) : 0u;
            if (valid)
            {
                const int slice = p / WAVE, lane = p % WAVE;
                const int beg = A.slice_ptr[slice];
                const int len = row_len[p];
                for (int k = 0; k < len; ++k)
                {
                    const int q = A.cols[static_cast<size_t>(beg + k) * WAVE + lane];
                    if (q / 32 == bank)
                        m |= 1u << (q & 31);
                }
            }
            masks[threadIdx.x] = m;
            __syncthreads();
            const uint32_t conn = valid ? closure(m, lane32, masks + (threadIdx.x & ~31)) : 0u;
            uint32_t lm;
            int lid, cnt;
            leaders_and_ids(valid, conn, lane32, lm, lid, cnt);
            if (p < npad)
            {
                lead[p] = lm;
                local_id[p] = lid;
                if (lane32 == 0)
                    bank_count[bank] = cnt;
            }
        }

        __device__ __forceinline__ int cmap_get(const int4 &c, int l)
        {
            return l == 0 ? c.x : (l == 1 ? c.y : (l == 2 ? c.z : c.w));
        }
        __device__ __forceinline__ void cmap_set(int4 &c, int l, int v)
        {
            if (l == 0)
                c.x = v;
            else if (l == 1)
                c.y = v;
            else if (l == 2)
                c.z = v;
            else
                c.w = v;
        }

        __global__ void k_level_masks(int npad, int lv, SellView<double> A, const int *__restrict__ row_len,
                                      const int *__restrict__ block_of_slot, const int4 *__restrict__ cmap,
                                      uint32_t *nbr)
        {
            const int p = blockIdx.x * blockDim.x + threadIdx.x;
            if (p >= npad || block_of_slot[p] < 0)
                return;
            const int c = cmap_get(cmap[p], lv - 1);
            uint32_t m = 1u << (c & 31);
            const int slice = p / WAVE, lane = p % WAVE;
            const int beg = A.slice_ptr[slice];
            const int len = row_len[p];
            for (int k = 0; k < len; ++k)
            {
                const int q = A.cols[static_cast<size_t>(beg + k) * WAVE + lane];
                const int cj = cmap_get(cmap[q], lv - 1);
                if ((cj >> 5) == (c >> 5))
                    m |= 1u << (cj & 31);
            }
            atomicOr(nbr + c, m);
        }

        __global__ __launch_bounds__(256) void k_cc_coarse(int count, const uint32_t *__restrict__ nbr, uint32_t *lead,
                                                           int *local_id, int *bank_count)
        {
            const int q = blockIdx.x * 256 + threadIdx.x;
            const int lane32 = threadIdx.x & 31;
            const bool valid = q < count;
            const uint32_t m = valid ? nbr[q] : 0u;
            const uint32_t conn = valid ? closure(m, lane32, nbr + (q & ~31)) : 0u;
            uint32_t lm;
            int lid, cnt;
            leaders_and_ids(valid, conn, lane32, lm, lid, cnt);
            if (valid)
            {
                lead[q] = lm;
                local_id[q] = lid;
                if (lane32 == 0)
                    bank_count[q / 32] = cnt;
            }
        }

        __global__ __launch_bounds__(1024) void k_scan(int n, const int *__restrict__ in, int *__restrict__ out)
        {
            __shared__ int sums[1024];
            const int per = div_up(n, 1024);
            const int b = threadIdx.x * per, e = min(n, b + per);
            int s = 0;
            for (int i = b; i < e; ++i)
                s += in[i];
            sums[threadIdx.x] = s;
            __syncthreads();
            // Hillis-Steele inclusive scan of the 1024 chunk sums.
            for (int off = 1; off < 1024; off <<= 1)
            {
                const int v = threadIdx.x >= off ? sums[threadIdx.x - off] : 0;
                __syncthreads();
                sums[threadIdx.x] += v;
                __syncthreads();
            }
            int run = threadIdx.x > 0 ? sums[threadIdx.x - 1] : 0;
            for (int i = b; i < e; ++i)
            {
                out[i] = run;
                run += in[i];
            }
            if (threadIdx.x == 1023)
                out[n] = sums[1023];
        }

        __global__ void k_make_gid(int count, const int *__restrict__ local_id, const int *__restrict__ prefix,
                                   int *__restrict__ gid)
        {
            const int q = blockIdx.x * blockDim.x + threadIdx.x;
            if (q < count)
                gid[q] = local_id[q] + prefix[q / 32];
        }

        __global__ void k_fill_cmap(int npad, int lv, const int *__restrict__ block_of_slot, const int *__restrict__ gid,
                                    int4 *cmap)
        {
            const int p = blockIdx.x * blockDim.x + threadIdx.x;
            if (p >= npad)
                return;
            int4 c = cmap[p];
            if (block_of_slot[p] < 0)
                cmap_set(c, lv, -1);
            else
                cmap_set(c, lv, lv == 0 ? gid[p] : gid[cmap_get(c, lv - 1)]);
            cmap[p] = c;
        }

        // --------------------------------------------------------------------- build: Galerkin
        __device__ __forceinline__ int upper_index(int dim, int i, int j)
        {
            if (i > j)
            {
                const int t = i;
                i = j;
                j = t;
            }
            return i * dim - (i * (i + 1)) / 2 + j;
        }

        /// One thread per fine row i; blocks (i, j) with i <= j only (A symmetric), as the reference
        /// fill (MASPreconditioner.cu:494-591).  Fine-level entries have exactly one writer.  Coarse
        /// entries are summed with FP64 atomics, but each row first merges consecutive blocks that
        /// land on the same coarse column of a level (Morton-ordered rows give long runs), so an
        /// atomic is issued per run and entry instead of per block and entry.
        constexpr int GAL_LEVELS = 4; ///< coarse levels of the MAS hierarchy (MASPreconditioner.hpp:70)

        template <int BD>
        __global__ void k_galerkin(int npad, int level_num, const int *__restrict__ mat_off, SellView<double> A,
                                   const int *__restrict__ row_len, const int *__restrict__ block_of_slot,
                                   const int4 *__restrict__ cmap, double *__restrict__ mats)
        {
            constexpr int BS = BD * BD;
            const int i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= npad || block_of_slot[i] < 0)
                return;
            constexpr int dim = 32 * BD;
            constexpr size_t st = static_cast<size_t>(dim) * (dim + 1) / 2;
            const int slice = i / WAVE, lane = i % WAVE;
            const int beg = A.slice_ptr[slice];
            const int len = row_len[i];
            const int4 ci4 = cmap[i];
            int cur[GAL_LEVELS];
            double acc[GAL_LEVELS][BS];
#pragma unroll
            for (int lv = 0; lv < GAL_LEVELS; ++lv)
            {
                cur[lv] = -1;
#pragma unroll
                for (int e = 0; e < BS; ++e)
                    acc[lv][e] = 0.0;
            }
            auto flush = [&](int lv) {
                const int ci = cmap_get(ci4, lv), cj = cur[lv];
                if (cj < 0)
                    return;
                double *mat = mats + static_cast<size_t>(mat_off[lv + 1] + (ci >> 5)) * st;
                const int ri = (ci & 31) * BD, rj = (cj & 31) * BD;
#pragma unroll
                for (int bi = 0; bi < BD; ++bi)
#pragma unroll
                    for (int bj = 0; bj < BD; ++bj)
                    {
                        const int row = ri + bi, col = rj + bj;
                        const double a = acc[lv][bi * BD + bj];
                        if (row <= col && a != 0.0)
                            atomicAdd(mat + upper_index(dim, row, col), a);
                        acc[lv][bi * BD + bj] = 0.0;
                    }
            };
            for (int k = 0; k < len; ++k)
            {
                const size_t slot = static_cast<size_t>(beg + k);
                const int j = A.cols[slot * WAVE + lane];
                if (i > j)
                    continue;
                double v[BS];
#pragma unroll
                for (int e = 0; e < BS; ++e)
                    v[e] = A.vals[(slot * BS + e) * WAVE + lane];
                if ((i >> 5) == (j >> 5))
                {
                    double *mat = mats + static_cast<size_t>(i >> 5) * st;
                    const int ri = (i & 31) * BD, rj = (j & 31) * BD;
#pragma unroll
                    for (int bi = 0; bi < BD; ++bi)
#pragma unroll
                        for (int bj = 0; bj < BD; ++bj)
                        {
                            const int row = ri + bi, col = rj + bj;
                            if (row > col || (i == j && bi > bj))
                                continue;
                            mat[upper_index(dim, row, col)] = v[bi * BD + bj];
                        }
                }
                const int4 cj4 = cmap[j];
#pragma unroll
                for (int lv = 0; lv < GAL_LEVELS; ++lv)
                {
                    if (lv >= level_num)
                        break;
                    const int ci = cmap_get(ci4, lv), cj = cmap_get(cj4, lv);
                    if ((ci >> 5) != (cj >> 5))
                        continue;
                    if (cj != cur[lv])
                    {
                        flush(lv);
                        cur[lv] = cj;
                    }
#pragma unroll
                    for (int bi = 0; bi < BD; ++bi)
#pragma unroll
                        for (int bj = 0; bj < BD; ++bj)
                        {
                            double c = v[bi * BD + bj];
                            if (i != j && ci == cj)
                                c += v[bj * BD + bi];
                            if (i == j && bi > bj)
                                c = 0.0;
                            acc[lv][bi * BD + bj] += c;
                        }
                }
            }
#pragma unroll
            for (int lv = 0; lv < GAL_LEVELS; ++lv)
                if (lv < level_num)
                    flush(lv);
        }

        // ---------------------------------------------------------------------- build: inverse
        using f64x4 = double __attribute__((ext_vector_type(4)));

        /// In-place inverse of one packed SPD matrix per workgroup by block Gauss-Jordan with 4x4
        /// pivot blocks, the rank-4 updates on FP64 MFMA (v_mfma_f64_16x16x4f64).  A zero diagonal
        /// entry is first replaced by 1 (virtual slots become identity rows,
        /// MASPreconditioner.cu:779-796).
        ///
        /// The N x N matrix (N = 32 BD) stays in registers as (N/16)^2 tiles of 16 x 16, each wave
        /// of the four owning a BD x BD quadrant of tiles in the MFMA result layout (lane l,
        /// register r: row 4r + l/16, column l%16).  Step K (N/4 steps) sweeps the rows and columns
        /// s = 4K..4K+3: with the strip rows R = A[s, :] as they are (a panel in LDS, the strip rows
        /// being register K%4 of the owning tiles) and P = A[s, s],
        ///   A[i, j] -= A[i, s] P^-1 R[:, j]      (all i, j; one MFMA per tile),
        ///   A[s, j] = P^-1 R[:, j],  A[i, s] = -A[i, s] P^-1,  A[s, s] = P^-1,
        /// where A[i, s] = R[:, i]^T for rows not yet swept and -R[:, i]^T for swept rows (the
        /// Gauss-Jordan sweep keeps that relation), so the panel alone feeds both MFMA operands:
        /// A operand -(+-)R[k][i] (natural layout: lane i + 16k) and B operand X = P^-1 R (lane
        /// j + 16k).  P^-1 comes from a 4x4 Cholesky computed by the wave that owns the pivot tile,
        /// which gathers P with lane shuffles.  Panel and P^-1 are double-buffered, so a step
        /// needs one barrier.  Flags a non-positive or non-finite pivot, a non-finite entry or a
        /// non-positive diagonal of the inverse.
        template <int BD>
        __global__ __launch_bounds__(256) void k_invert(double *__restrict__ mats, int *fail)
        {
            constexpr int N = 32 * BD;
            constexpr int L = N * (N + 1) / 2;
            constexpr int H = BD;     // tiles per wave and dimension
            constexpr int PW = N + 4; // panel row stride (doubles)
            __shared__ double panel[2][4 * PW];
            __shared__ double pinv_s[2][16];
            __shared__ int bad;
            double *g = mats + static_cast<size_t>(blockIdx.x) * L;
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            const int lr = lane >> 4, lc = lane & 15; // result layout: row 4r + lr, column lc
            const int ti0 = H * (wave >> 1), tj0 = H * (wave & 1);
            if (threadIdx.x == 0)
                bad = 0;
            auto packed = [&](int i, int j) {
                if (i > j)
                {
                    const int t = i;
                    i = j;
                    j = t;
                }
                return i * N - (i * (i + 1)) / 2 + j;
            };
            f64x4 C[H][H];
#pragma unroll
            for (int a = 0; a < H; ++a)
#pragma unroll
                for (int b = 0; b < H; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                    {
                        const int row = 16 * (ti0 + a) + 4 * r + lr, col = 16 * (tj0 + b) + lc;
                        double v = g[packed(row, col)];
                        if (row == col && v == 0.0)
                            v = 1.0;
                        C[a][b][r] = v;
                    }
            int flag = 0;
            // One sweep step; q = K % 4 is a compile-time constant and tiles are selected by
            // predicates, so every register array index is static (no scratch indexing).
            auto step = [&](const int K, auto qc) {
                constexpr int q = decltype(qc)::value;
                const int buf = K & 1, tK = K >> 2, s0 = 4 * K;
                double *pn = panel[buf];
                // Strip rows -> panel (register q of the tiles in tile row tK).
#pragma unroll
                for (int a = 0; a < H; ++a)
                    if (ti0 + a == tK)
                    {
#pragma unroll
                        for (int b = 0; b < H; ++b)
                            pn[lr * PW + 16 * (tj0 + b) + lc] = C[a][b][q];
                    }
                // P^-1 by the wave owning the pivot tile: P[m][c] is register q of lane 16m + 4q + c.
#pragma unroll
                for (int a = 0; a < H; ++a)
                    if (ti0 + a == tK && tj0 + a == tK)
                    {
                        double P[4][4];
#pragma unroll
                        for (int m = 0; m < 4; ++m)
#pragma unroll
                            for (int c = 0; c < 4; ++c)
                                P[m][c] = __shfl(C[a][a][q], 16 * m + 4 * q + c);
                        // Cholesky P = L L^T (lower triangle read), W = L^-1, P^-1 = W^T W.
                        double l[4][4] = {};
                        int ok = 1;
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                        {
                            double d = P[j][j];
#pragma unroll
                            for (int k = 0; k < j; ++k)
                                d -= l[j][k] * l[j][k];
                            ok &= (d > 0.0 && isfinite(d)) ? 1 : 0;
                            d = sqrt(d > 0.0 ? d : 1.0);
                            l[j][j] = d;
                            const double inv = 1.0 / d;
#pragma unroll
                            for (int i = j + 1; i < 4; ++i)
                            {
                                double v = P[i][j];
#pragma unroll
                                for (int k = 0; k < j; ++k)
                                    v -= l[i][k] * l[j][k];
                                l[i][j] = v * inv;
                            }
                        }
                        double w[4][4] = {};
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                        {
                            w[i][i] = 1.0 / l[i][i];
#pragma unroll
                            for (int j = 0; j < i; ++j)
                            {
                                double v = 0.0;
#pragma unroll
                                for (int k = j; k < i; ++k)
                                    v -= l[i][k] * w[k][j];
                                w[i][j] = v * w[i][i];
                            }
                        }
                        if (lane == 0) // static indices only (no scratch)
                        {
#pragma unroll
                            for (int m = 0; m < 4; ++m)
#pragma unroll
                                for (int c = 0; c < 4; ++c)
                                {
                                    double v = 0.0;
#pragma unroll
                                    for (int k = 0; k < 4; ++k)
                                        v += w[k][m] * w[k][c];
                                    pinv_s[buf][4 * m + c] = v;
                                }
                        }
                        flag |= ok ? 0 : 1;
                    }
                __syncthreads();
                // Row lr of P^-1, X = P^-1 R for this wave's tile columns (B operands), the A
                // operands for its tile rows, then the rank-4 update.
                double pr[4];
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    pr[m] = pinv_s[buf][4 * lr + m];
                double X[H], Aop[H];
#pragma unroll
                for (int b = 0; b < H; ++b)
                {
                    const int col = 16 * (tj0 + b) + lc;
                    double v = 0.0;
#pragma unroll
                    for (int m = 0; m < 4; ++m)
                        v = fma(pr[m], pn[m * PW + col], v);
                    X[b] = v;
                }
#pragma unroll
                for (int a = 0; a < H; ++a)
                {
                    const int row = 16 * (ti0 + a) + lc; // A operand: row lc, k = lr
                    const double v = pn[lr * PW + row];
                    Aop[a] = row < s0 ? v : -v;
                }
#pragma unroll
                for (int a = 0; a < H; ++a)
#pragma unroll
                    for (int b = 0; b < H; ++b)
                        C[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(Aop[a], X[b], C[a][b], 0, 0, 0);
                // Strip rows become X (this lane's X[b] is exactly its register-q entry), the
                // strip columns -(+-)X^T, the pivot block P^-1.
#pragma unroll
                for (int a = 0; a < H; ++a)
                    if (ti0 + a == tK)
                    {
#pragma unroll
                        for (int b = 0; b < H; ++b)
                        {
                            const int col = 16 * (tj0 + b) + lc;
                            C[a][b][q] = (col >= s0 && col < s0 + 4) ? pinv_s[buf][4 * lr + (col - s0)] : X[b];
                        }
                    }
                if (lc >= 4 * q && lc < 4 * q + 4)
                {
                    const int c = lc - 4 * q;
#pragma unroll
                    for (int b = 0; b < H; ++b)
                        if (tj0 + b == tK)
                        {
                            double pc[4];
#pragma unroll
                            for (int m = 0; m < 4; ++m)
                                pc[m] = pinv_s[buf][4 * c + m];
#pragma unroll
                            for (int a = 0; a < H; ++a)
#pragma unroll
                                for (int r = 0; r < 4; ++r)
                                {
                                    const int row = 16 * (ti0 + a) + 4 * r + lr;
                                    if (row >= s0 && row < s0 + 4)
                                        continue;
                                    double v = 0.0;
#pragma unroll
                                    for (int m = 0; m < 4; ++m)
                                        v = fma(pc[m], pn[m * PW + row], v);
                                    C[a][b][r] = row < s0 ? v : -v;
                                }
                        }
                }
            };
            for (int K4 = 0; K4 < N / 16; ++K4)
            {
                step(4 * K4 + 0, std::integral_constant<int, 0>());
                step(4 * K4 + 1, std::integral_constant<int, 1>());
                step(4 * K4 + 2, std::integral_constant<int, 2>());
                step(4 * K4 + 3, std::integral_constant<int, 3>());
            }
            // Check and write back the upper triangle.
#pragma unroll
            for (int a = 0; a < H; ++a)
#pragma unroll
                for (int b = 0; b < H; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                    {
                        const int row = 16 * (ti0 + a) + 4 * r + lr, col = 16 * (tj0 + b) + lc;
                        const double v = C[a][b][r];
                        if (!isfinite(v) || (row == col && !(v > 0.0)))
                            flag = 1;
                        if (row <= col)
                            g[packed(row, col)] = v;
                    }
            if (flag)
                bad = 1; // benign race: every writer stores 1
            __syncthreads();
            if (threadIdx.x == 0 && bad)
                atomicExch(fail, 1);
        }

        /// FP16 form of packed inverses: Inv = D S D with D = diag(sqrt(Inv_ii)), so S has a unit
        /// diagonal and |S_ij| <= 1; D is stored in FP32 and S in FP16, tiled for k_mas_apply_h16
        /// (half_tile_index gives the layout; N*N/2 values, the unit diagonal is implicit).
        template <int BD>
        __global__ __launch_bounds__(256) void k_make_half(const double *__restrict__ mats, __half *__restrict__ out,
                                                           float *__restrict__ dscale)
        {
            constexpr int N = 32 * BD;
            constexpr int L = N * (N + 1) / 2;
            constexpr int LEN = half_matrix_len<BD>();
            __shared__ double d[N];
            const double *g = mats + static_cast<size_t>(blockIdx.x) * L;
            for (int i = threadIdx.x; i < N; i += 256)
            {
                const double v = sqrt(fmax(g[i * N - (i * (i + 1)) / 2 + i], 0.0));
                d[i] = v;
                dscale[static_cast<size_t>(blockIdx.x) * N + i] = static_cast<float>(v);
            }
            __syncthreads();
            __half *o = out + static_cast<size_t>(blockIdx.x) * LEN;
            for (int e = threadIdx.x; e < LEN; e += 256)
            {
                int i, j;
                half_tile_element<BD>(e, i, j);
                if (i == j)
                {
                    o[e] = __float2half_rn(1.0f); // unused slot
                    continue;
                }
                const double den = d[i] * d[j];
                const double v = g[i * N - (i * (i + 1)) / 2 + j];
                o[e] = __float2half_rn(static_cast<float>(den > 0.0 ? v / den : 0.0));
            }
        }

        // ---------------------------------------------------------------------- build: Morton
        __device__ __forceinline__ uint64_t spread3(uint64_t v)
        {
            v &= 0x1fffffull;
            v = (v | (v << 32)) & 0x1f00000000ffffull;
            v = (v | (v << 16)) & 0x1f0000ff0000ffull;
            v = (v | (v << 8)) & 0x100f00f00f00f00full;
            v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
            v = (v | (v << 2)) & 0x1249249249249249ull;
            return v;
        }

        __global__ void k_morton_keys(int nb, int dim, const double *__restrict__ xyz, const double *__restrict__ lo,
                                      const double *__restrict__ inv_extent, uint64_t *keys, int *ids)
        {
            const int i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= nb)
                return;
            uint64_t key = 0;
            for (int d = 0; d < dim; ++d)
            {
                double u = (xyz[static_cast<size_t>(i) * dim + d] - lo[d]) * inv_extent[d];
                u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
                const uint64_t q = static_cast<uint64_t>(u * 2097151.0);
                key |= spread3(q) << d;
            }
            keys[i] = key;
            ids[i] = i;
        }
    } // namespace

    template <int BD, typename TM, typename T>
    void launch_mas_apply(int mat_begin, int count, const TM *mats, const float *dscale, const T *mr, T *mz,
                          const int *bank_sys, const int *active, hipStream_t s, const ApplySide &side,
                          const CopyLayout &cl)
    {
        if (count <= 0 && side.wgs <= 0)
            return;
        require_device_ptrs("mas_apply", {mr, mz});
        if (count > 0)
            require_device_ptrs("mas_apply (matrices)",
                                {mats, std::is_same<TM, __half>::value ? static_cast<const void *>(dscale)
                                                                       : static_cast<const void *>(mats)});
        if (count > 0)
            require_device_ptrs("mas_apply (index placeholder)", {side.dummy});
        // Apply workgroups: as many as are resident at once beside the side workgroups (registers
        // and LDS decide); each walks its matrices grid-stride.  Side workgroups come first so the
        // dispatcher starts them at once.
        auto resident_of = [](const void *kernel) {
            int per_cu = 0, cus = 0, dev = 0;
            PSMAS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, apply_block<BD>(), 0));
            PSMAS_HIP(hipGetDevice(&dev));
            PSMAS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            return (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
        };
        auto grid_for = [&](int resident) {
            int avail = resident - side.wgs > 1 ? resident - side.wgs : 1;
            if (side.apply_cap > 0 && side.apply_cap < avail)
                avail = side.apply_cap;
            const int apply_wgs = count <= 0 ? 0 : (count < avail ? count : avail);
            return side.wgs + apply_wgs;
        };
        if constexpr (std::is_same<TM, __half>::value)
        {
#define PSMAS_H16_LAUNCH(KC)                                                                                           \
    {                                                                                                                  \
        static const int resident = resident_of(reinterpret_cast<const void *>(&k_mas_apply_h16<BD, KC>));             \
        hipLaunchKernelGGL((k_mas_apply_h16<BD, KC>), dim3(grid_for(resident)), dim3(apply_block<BD>()), 0, s,         \
                           mat_begin, count, mats, dscale, mr, mz, bank_sys, active, side, cl);                        \
    }
            if (cl.copies > 8)
                throw std::logic_error("[MAS] internal error: at most 8 right-hand sides per factorisation.");
            if (side.wave_apply && cl.copies > 1)
                throw std::logic_error("[MAS] internal error: the wave apply takes one right-hand side.");
            if (side.wave_apply)
            {
                // A wave per matrix, every matrix's workgroup launched at once (the hardware keeps
                // as many resident as fit); apply_cap makes waves walk matrices grid-stride.
                int apply_wgs = div_up(count > 0 ? count : 0, APPLY_W16_WAVES);
                if (side.apply_cap > 0 && side.apply_cap < apply_wgs)
                    apply_wgs = side.apply_cap;
                // A second launch (split inputs, dense rows, network layer 1) compiles that work in.
                if (side.split_of != nullptr || side.dense_nd > 0 || side.lin_rows > 0)
                    hipLaunchKernelGGL((k_mas_apply_w16<BD, true>), dim3(side.wgs + apply_wgs), dim3(APPLY_W16_BLOCK),
                                       0, s, mat_begin, count, mats, dscale, mr, mz, bank_sys, active, side);
                else
                    hipLaunchKernelGGL((k_mas_apply_w16<BD, false>), dim3(side.wgs + apply_wgs), dim3(APPLY_W16_BLOCK),
                                       0, s, mat_begin, count, mats, dscale, mr, mz, bank_sys, active, side);
            }
            else if (cl.copies <= 1)
                PSMAS_H16_LAUNCH(1)
            else if (cl.copies <= 2)
                PSMAS_H16_LAUNCH(2)
            else if (cl.copies <= 4)
                PSMAS_H16_LAUNCH(4)
            else
                PSMAS_H16_LAUNCH(8)
#undef PSMAS_H16_LAUNCH
        }
        else
        {
            static const int resident = resident_of(reinterpret_cast<const void *>(&k_mas_apply<BD, TM, T>));
            hipLaunchKernelGGL((k_mas_apply<BD, TM, T>), dim3(grid_for(resident)), dim3(apply_block<BD>()), 0, s,
                               mat_begin, count, mats, dscale, mr, mz, bank_sys, active, side, cl);
        }
        PSMAS_LAUNCH_CHECK();
    }

    int apply_h16_stage_floats(int bd, bool wave)
    {
        if (wave)
            return bd == 1 ? W16Shape<1>::LDS_FLOATS : (bd == 2 ? W16Shape<2>::LDS_FLOATS : W16Shape<3>::LDS_FLOATS);
        const int n = 32 * bd;
        return n * (n + 8) / 2; // the __half LDS copy of one matrix, as floats
    }

    int apply_threads(int bd, bool wave) { return wave ? APPLY_W16_BLOCK : 64 * bd; }

    template <int BD, typename T>
    void launch_prolong_dot(int nblocks, int level_num, const int4 *cmap, const int *lev_off, const T *r, T *mz,
                            const Reductions &red, bool init, const SysMap &sm, const LccApply &lcc,
                            const FlexDot<T> &flex, hipStream_t s)
    {
        require_device_ptrs("prolong_dot", {cmap, lev_off, r, mz});
        hipLaunchKernelGGL((k_prolong_dot<BD, T>), dim3(nblocks), dim3(NODE_BLOCK), 0, s, level_num, cmap, lev_off, r,
                           mz, red, init, sm, lcc, flex);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_cc_level0(int npad, const SellView<double> &A, const int *row_len, const int *block_of_slot,
                          uint32_t *lead, int *local_id, int *bank_count, hipStream_t s)
    {
        hipLaunchKernelGGL(k_cc_level0, dim3(div_up(npad, 256)), dim3(256), 0, s, npad, A, row_len, block_of_slot,
                           lead, local_id, bank_count);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_level_masks(int npad, int lv, const SellView<double> &A, const int *row_len, const int *block_of_slot,
                            const int4 *cmap, uint32_t *nbr, hipStream_t s)
    {
        hipLaunchKernelGGL(k_level_masks, dim3(div_up(npad, 256)), dim3(256), 0, s, npad, lv, A, row_len,
                           block_of_slot, cmap, nbr);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_cc_coarse(int count, const uint32_t *nbr, uint32_t *lead, int *local_id, int *bank_count,
                          hipStream_t s)
    {
        hipLaunchKernelGGL(k_cc_coarse, dim3(div_up(count, 256)), dim3(256), 0, s, count, nbr, lead, local_id,
                           bank_count);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_scan(int n, const int *in, int *out, hipStream_t s)
    {
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, n, in, out);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_make_gid(int count, const int *local_id, const int *prefix, int *gid, hipStream_t s)
    {
        if (count <= 0)
            return;
        hipLaunchKernelGGL(k_make_gid, dim3(div_up(count, 256)), dim3(256), 0, s, count, local_id, prefix, gid);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_fill_cmap(int npad, int lv, const int *block_of_slot, const int *gid, int4 *cmap, hipStream_t s)
    {
        hipLaunchKernelGGL(k_fill_cmap, dim3(div_up(npad, 256)), dim3(256), 0, s, npad, lv, block_of_slot, gid, cmap);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_galerkin(int npad, int bd, int level_num, const int *mat_off, const SellView<double> &A,
                         const int *row_len, const int *block_of_slot, const int4 *cmap, double *mats, hipStream_t s)
    {
        const dim3 grid(div_up(npad, 128)), block(128);
        if (bd == 1)
            hipLaunchKernelGGL(k_galerkin<1>, grid, block, 0, s, npad, level_num, mat_off, A, row_len, block_of_slot,
                               cmap, mats);
        else if (bd == 2)
            hipLaunchKernelGGL(k_galerkin<2>, grid, block, 0, s, npad, level_num, mat_off, A, row_len, block_of_slot,
                               cmap, mats);
        else
            hipLaunchKernelGGL(k_galerkin<3>, grid, block, 0, s, npad, level_num, mat_off, A, row_len, block_of_slot,
                               cmap, mats);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_invert(int bd, int count, double *mats, int *fail, hipStream_t s)
    {
        if (count <= 0)
            return;
        if (bd == 1)
            hipLaunchKernelGGL(k_invert<1>, dim3(count), dim3(256), 0, s, mats, fail);
        else if (bd == 2)
            hipLaunchKernelGGL(k_invert<2>, dim3(count), dim3(256), 0, s, mats, fail);
        else
            hipLaunchKernelGGL(k_invert<3>, dim3(count), dim3(256), 0, s, mats, fail);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_make_half(int bd, int count, const double *mats, __half *out, float *dscale, hipStream_t s)
    {
        if (count <= 0)
            return;
        if (bd == 1)
            hipLaunchKernelGGL(k_make_half<1>, dim3(count), dim3(256), 0, s, mats, out, dscale);
        else if (bd == 2)
            hipLaunchKernelGGL(k_make_half<2>, dim3(count), dim3(256), 0, s, mats, out, dscale);
        else
            hipLaunchKernelGGL(k_make_half<3>, dim3(count), dim3(256), 0, s, mats, out, dscale);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_morton_keys(int nb, int dim, const double *xyz, const double *lo, const double *inv_extent,
                            uint64_t *keys, int *ids, hipStream_t s)
    {
        hipLaunchKernelGGL(k_morton_keys, dim3(div_up(nb, 256)), dim3(256), 0, s, nb, dim, xyz, lo, inv_extent, keys,
                           ids);
        PSMAS_LAUNCH_CHECK();
    }

    size_t morton_sort_temp_bytes(int nb)
    {
        size_t bytes = 0;
        PSMAS_HIP(rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const uint64_t *>(nullptr),
                                            static_cast<uint64_t *>(nullptr), static_cast<const int *>(nullptr),
                                            static_cast<int *>(nullptr), static_cast<unsigned>(nb), 0u, 63u));
        return bytes;
    }

    void morton_sort(int nb, void *temp, size_t temp_bytes, const uint64_t *keys_in, uint64_t *keys_out,
                     const int *ids_in, int *ids_out, hipStream_t s)
    {
        PSMAS_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, ids_in, ids_out,
                                            static_cast<unsigned>(nb), 0u, 63u, s));
    }

#define PSMAS_INST_APPLY(BD)                                                                                     \
    template void launch_mas_apply<BD, double, double>(int, int, const double *, const float *, const double *,   \
                                                       double *, const int *, const int *, hipStream_t,          \
                                                       const ApplySide &, const CopyLayout &);                   \
    template void launch_mas_apply<BD, __half, float>(int, int, const __half *, const float *, const float *,     \
                                                      float *, const int *, const int *, hipStream_t,            \
                                                      const ApplySide &, const CopyLayout &);                    \
    template void launch_prolong_dot<BD, double>(int, int, const int4 *, const int *, const double *, double *,   \
                                                 const Reductions &, bool, const SysMap &, const LccApply &,     \
                                                 const FlexDot<double> &, hipStream_t);                          \
    template void launch_prolong_dot<BD, float>(int, int, const int4 *, const int *, const float *, float *,      \
                                                const Reductions &, bool, const SysMap &, const LccApply &,      \
                                                const FlexDot<float> &, hipStream_t);

    PSMAS_INST_APPLY(1)
    PSMAS_INST_APPLY(2)
    PSMAS_INST_APPLY(3)
#undef PSMAS_INST_APPLY
} // namespace psmas
