// SELL-64 block SpMV and the fused vector kernels of the PCG loop (gfx950, wave64).
//
// Replaces the reference's cuSPARSE BSR SpMV (src/polysolve/linear/MASSolver.cu:245-290), its
// inner_product (mas_utils/InnerProduct.cu:16-45), axpby / scalar_division
// (MASSolver.cu:46-81) and the restriction half of gather_multi_level_r
// (mas_utils/MASPreconditioner.cu:593-659).  Layout and fusion are this build's own:
//  * one lane per block row; a wave owns a slice of 64 rows = two 32-node banks, so every
//    matrix load of a wave is one contiguous 512 B (FP64) / 256 B (FP32) run;
//  * a workgroup owns 256 consecutive rows (a node block); a batch of systems is laid out so
//    that each system covers whole node blocks.  Every reduction writes one partial per
//    workgroup; the kernels that need alpha or beta re-sum their system's partials in block
//    order at their start, so the scalars never leave the device and need no atomics;
//  * the residual-producing kernels also restrict the residual to coarse level 0 through the
//    bank-local component masks, so the preconditioner never re-reads r for that step;
//  * systems that have converged (active == 0) skip their workgroups.
#include "device.hpp"
#include "lattice_device.hpp"
#include "kernels.hpp"
#include "common.hpp"

namespace psmas
{
    /// Dynamic LDS of the SpMV kernels: the workgroup's halo of x.
    extern __shared__ __align__(16) unsigned char spmv_halo_lds[];

    namespace
    {
        constexpr int SPMV_BLOCK = NODE_BLOCK; // 4 waves = 4 slices of 64 rows
        constexpr int VEC_BLOCK = NODE_BLOCK;

        template <typename T>
        struct Acc
        {
            using type = T;
        };

        /// Row `lane` of SELL slice `slice` as a stream of batches of U slots: acc += A_row x.
        /// open() issues the first batch's column and value loads and returns without waiting for
        /// them, so that they overlap what the caller does next (the halo stage); run() then
        /// consumes one batch per step.  Per step the x reads use column indices loaded one step
        /// earlier, the next columns are issued ahead of the FMAs and the next values after them,
        /// so a wave waits for one memory round trip per batch.  Every load is issued
        /// unconditionally (slot indices past the slice end are clamped to its last slot, whose
        /// lines are in cache, and their values are zeroed), which lets the compiler count the
        /// loads in flight.
        ///
        /// HALO: x comes from the workgroup's LDS copy xs of the entries its rows read, indexed by
        /// 16-bit local columns (halo_fill writes it); otherwise x is gathered from global memory
        /// through the 32-bit columns.
        template <int BD, int U, bool HALO, typename TV>
        struct SellRow
        {
            int beg = 0, end = 0, last = 0;
            const int *__restrict__ cols = nullptr;
            const unsigned short *__restrict__ lcols = nullptr;
            const TV *__restrict__ vals = nullptr;
            int c[U];
            TV v[U][BD * BD];

            __device__ __forceinline__ int column(int slot) const
            {
                if constexpr (HALO)
                    return static_cast<int>(ldg_stream(lcols + static_cast<size_t>(slot) * WAVE));
                else
                    return ldg_stream(cols + static_cast<size_t>(slot) * WAVE);
            }
            __device__ __forceinline__ void load_values(int k)
            {
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const size_t slot = static_cast<size_t>(min(k + u, last));
#pragma unroll
                    for (int e = 0; e < BD * BD; ++e)
                        v[u][e] = ldg_stream(vals + (slot * (BD * BD) + e) * WAVE);
                }
            }
            __device__ __forceinline__ void open(const SellView<TV> &A, int slice, int lane)
            {
                beg = __builtin_amdgcn_readfirstlane(A.slice_ptr[slice]);
                end = __builtin_amdgcn_readfirstlane(A.slice_ptr[slice + 1]);
                last = end - 1;
                cols = A.cols + lane;
                lcols = A.lcols + lane;
                vals = A.vals + lane;
                if (end > beg) // wave-uniform
                {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        c[u] = column(min(beg + u, last));
                    load_values(beg);
                }
            }
            template <typename TX, typename TA>
            __device__ __forceinline__ void run(const TX *__restrict__ x, const TX *xs, TA acc[BD])
            {
                for (int k = beg; k < end; k += U)
                {
                    TA xv[U][BD];
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int j = 0; j < BD; ++j)
                        {
                            if constexpr (HALO)
                                xv[u][j] = static_cast<TA>(xs[c[u] * BD + j]);
                            else
                                xv[u][j] = static_cast<TA>(x[static_cast<size_t>(c[u]) * BD + j]);
                        }
                    // The next batch's column loads go behind this batch's loads and ahead of the
                    // FMAs, so that they stay in flight while the FMAs wait for this batch.
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        c[u] = column(min(k + U + u, last));
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < U; ++u)
                    {
                        const bool ok = k + u < end; // wave-uniform
#pragma unroll
                        for (int i = 0; i < BD; ++i)
#pragma unroll
                            for (int j = 0; j < BD; ++j)
                                acc[i] = fma(ok ? static_cast<TA>(v[u][i * BD + j]) : TA(0), xv[u][j], acc[i]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    if (k + U < end) // wave-uniform
                        load_values(k + U);
                }
            }
        };

        /// Node block of this workgroup for the SpMV kernels.  Workgroups are dealt to the 8 XCDs
        /// round-robin; this gives each XCD one contiguous range of node blocks, so the x gathers
        /// of neighbouring rows hit the same L2.  A bijection on [0, gridDim.x).
        __device__ __forceinline__ int xcd_block()
        {
            constexpr int NXCD = 8;
            const int G = gridDim.x, b = blockIdx.x;
            const int q = G / NXCD, r = G % NXCD, xcd = b % NXCD;
            return xcd * q + (xcd < r ? xcd : r) + b / NXCD;
        }

        __device__ __forceinline__ int block_system(const SysMap &sm) { return system_of_block(sm, blockIdx.x); }

        /// Halo stage of node block blk in two parts: halo_indices issues the loads of the first
        /// HALO_R rounds of halo node indices (clamped, unconditional), halo_fill gathers x at
        /// them (one round trip for all rounds), stores the values in the dynamic LDS array,
        /// handles longer halos in a loop, and waits for the workgroup.  Every wave of the
        /// workgroup must call both (uniform control flow).
        constexpr int HALO_R = 4;
        struct HaloStage
        {
            int hb = 0, hn = 0;
            int g[HALO_R];
        };
        template <typename TV>
        __device__ __forceinline__ HaloStage halo_indices(const SellView<TV> &A, int blk)
        {
            HaloStage h;
            h.hb = A.halo_ptr[blk];
            h.hn = A.halo_ptr[blk + 1] - h.hb;
            const int lastl = max(h.hn - 1, 0);
#pragma unroll
            for (int r = 0; r < HALO_R; ++r)
                h.g[r] = h.hn > 0 ? A.halo[h.hb + min(static_cast<int>(threadIdx.x) + r * SPMV_BLOCK, lastl)] : 0;
            return h;
        }
        template <int BD, typename TV, typename TX>
        __device__ __forceinline__ const TX *halo_fill(const HaloStage &h, const SellView<TV> &A,
                                                       const TX *__restrict__ x)
        {
            TX *xs = reinterpret_cast<TX *>(spmv_halo_lds);
            if (h.hn > 0)
            {
                TX v[HALO_R][BD];
#pragma unroll
                for (int r = 0; r < HALO_R; ++r)
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                        v[r][j] = x[static_cast<size_t>(h.g[r]) * BD + j];
#pragma unroll
                for (int r = 0; r < HALO_R; ++r)
                {
                    const int i = threadIdx.x + r * SPMV_BLOCK;
                    if (i < h.hn)
#pragma unroll
                        for (int j = 0; j < BD; ++j)
                            xs[i * BD + j] = v[r][j];
                }
                for (int i = threadIdx.x + HALO_R * SPMV_BLOCK; i < h.hn; i += SPMV_BLOCK)
                {
                    const size_t g = static_cast<size_t>(A.halo[h.hb + i]) * BD;
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                        xs[i * BD + j] = x[g + j];
                }
            }
            __syncthreads();
            return xs;
        }

        template <int BD, int U, bool HALO, typename TV, typename TX>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_ap(SellView<TV> A, const TX *__restrict__ x,
                                                                TX *__restrict__ y, Reductions red, SysMap sm)
        {
            using TA = typename Acc<TV>::type;
            const int blk = xcd_block();
            const int sys = system_of_block(sm, blk);
            if (system_stopped(sm, sys))
                return;
            const int lane = threadIdx.x % WAVE;
            const int slice = blk * (SPMV_BLOCK / WAVE) + threadIdx.x / WAVE;
            HaloStage hs;
            if constexpr (HALO)
                hs = halo_indices(A, blk);
            const TX *xs = nullptr;
            if constexpr (HALO)
                xs = halo_fill<BD>(hs, A, x);
            SellRow<BD, U, HALO, TV> sr;
            sr.open(A, slice, lane);
            TA acc[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                acc[i] = TA(0);
            sr.run(x, xs, acc);
            const size_t row = static_cast<size_t>(slice) * WAVE + lane;
            double dot = 0.0;
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                const TX yi = static_cast<TX>(acc[i]);
                y[row * BD + i] = yi;
                dot += static_cast<double>(x[row * BD + i]) * static_cast<double>(yi);
            }
            store_partial<SPMV_BLOCK>(dot, red.pap, blk);
        }

        /// Byte offset of the incoming arrays of k_spmv_sym in the dynamic LDS (after the halo).
        template <int BD, typename TX>
        __host__ __device__ constexpr size_t sym_incoming_offset(int halo_max)
        {
            return (static_cast<size_t>(halo_max) * BD * sizeof(TX) + 15) / 16 * 16;
        }

        /// The rows of node block blk over the symmetric form S: halo stage, then every row works
        /// through its stored slots U at a time as SellRow does (x from the workgroup's halo copy,
        /// the next batch's metadata issued ahead of the FMAs, its values after them).  A slot of
        /// an owned block that carries an incoming slot also multiplies the block's transpose with
        /// the row's own x (slot 0 is the diagonal block, so its column is the row's own node) and
        /// stores the product in the incoming arrays (component i at inc[i * incoming_max +
        /// slot]); a foreign slot stores its product A_rc x_c there instead of adding it to its own
        /// row.  After the workgroup barrier each row adds its incoming products in slot order.
        /// Fixed order throughout: reproducible.  Returns the row's local index; acc = (A x)_row,
        /// xself = x_row.  Every thread of the workgroup must call it (it holds barriers).
        template <int BD, int U, typename TV, typename TX, typename TA>
        __device__ __forceinline__ int sym_rows(const SellView<TV> &A, const SymView<TV> &S, const TX *__restrict__ x,
                                                const int blk, TA (&acc)[BD], TA (&xself)[BD])
        {
            constexpr uint32_t NONE = 0x7FFFu, FOREIGN = 0x8000u; // (fields of meta >> 16)
            const int lane = threadIdx.x % WAVE;
            const int slice = blk * (SPMV_BLOCK / WAVE) + threadIdx.x / WAVE;
            const HaloStage hs = halo_indices(A, blk);
            const int beg = __builtin_amdgcn_readfirstlane(S.slice_ptr[slice]);
            const int end = __builtin_amdgcn_readfirstlane(S.slice_ptr[slice + 1]);
            const int last = end - 1;
            const uint32_t *__restrict__ meta = S.meta + lane;
            const TV *__restrict__ vals = S.vals + lane;
            const uint32_t info = S.row_info[static_cast<size_t>(slice) * WAVE + lane];
            uint32_t m[U];
            TV v[U][BD * BD];
            auto load_meta = [&](int k) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                    m[u] = ldg_stream(meta + static_cast<size_t>(min(k + u, last)) * WAVE);
            };
            auto load_values = [&](int k) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const size_t slot = static_cast<size_t>(min(k + u, last));
#pragma unroll
                    for (int e = 0; e < BD * BD; ++e)
                        v[u][e] = ldg_stream(vals + (slot * (BD * BD) + e) * WAVE);
                }
            };
            load_meta(beg); // every slice has at least one slot
            load_values(beg);
            const TX *xs = halo_fill<BD>(hs, A, x);
            TA *inc = reinterpret_cast<TA *>(spmv_halo_lds + sym_incoming_offset<BD, TX>(A.halo_max));
            const int im = S.incoming_max;
            {
                const int ls = static_cast<int>(m[0] & 0xFFFFu); // slot 0: the diagonal block
#pragma unroll
                for (int j = 0; j < BD; ++j)
                    xself[j] = static_cast<TA>(xs[ls * BD + j]);
            }
#pragma unroll
            for (int i = 0; i < BD; ++i)
                acc[i] = TA(0);
            // One batch of U slots; only the slice's last batch checks for slots past its end.
            auto batch = [&](const int k, auto check) {
                TA xv[U][BD];
                uint32_t ts[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const int c = static_cast<int>(m[u] & 0xFFFFu);
                    ts[u] = m[u] >> 16;
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                        xv[u][j] = static_cast<TA>(xs[c * BD + j]);
                }
                __builtin_amdgcn_sched_barrier(0);
                load_meta(k + U);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const bool ok = !decltype(check)::value || k + u < end; // wave-uniform
                    const bool fo = (ts[u] & FOREIGN) != 0;
                    const uint32_t t = ts[u] & NONE;
                    TA r[BD];
#pragma unroll
                    for (int i = 0; i < BD; ++i)
                    {
                        r[i] = fo ? TA(0) : acc[i];
#pragma unroll
                        for (int j = 0; j < BD; ++j)
                            r[i] = fma(ok ? static_cast<TA>(v[u][i * BD + j]) : TA(0), xv[u][j], r[i]);
                    }
                    if (fo)
                    {
                        if (ok)
#pragma unroll
                            for (int i = 0; i < BD; ++i)
                                inc[i * im + static_cast<int>(t)] = r[i];
                    }
                    else
                    {
#pragma unroll
                        for (int i = 0; i < BD; ++i)
                            acc[i] = r[i];
                    }
                    if (ok && !fo && t != NONE)
                    {
#pragma unroll
                        for (int j = 0; j < BD; ++j)
                        {
                            TA tr = TA(0);
#pragma unroll
                            for (int i = 0; i < BD; ++i)
                                tr = fma(static_cast<TA>(v[u][i * BD + j]), xself[i], tr);
                            inc[j * im + static_cast<int>(t)] = tr;
                        }
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                if (k + U < end) // wave-uniform
                    load_values(k + U);
            };
            int k = beg;
            for (; k + U <= end; k += U)
                batch(k, std::false_type());
            if (k < end)
                batch(k, std::true_type());
            __syncthreads();
            const int lr = static_cast<int>(info & 0xFFu), cnt = static_cast<int>((info >> 8) & 0xFFu);
            const int ib = static_cast<int>(info >> 16);
            for (int q = 0; q < cnt; ++q)
#pragma unroll
                for (int i = 0; i < BD; ++i)
                    acc[i] += inc[i * im + ib + q];
            return lr;
        }

        /// k_spmv_ap over the symmetric form S.
        template <int BD, int U, typename TV, typename TX>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_sym(SellView<TV> A, SymView<TV> S,
                                                                 const TX *__restrict__ x, TX *__restrict__ y,
                                                                 Reductions red, SysMap sm)
        {
            using TA = typename Acc<TV>::type;
            const int blk = xcd_block();
            const int sys = system_of_block(sm, blk);
            if (system_stopped(sm, sys))
                return;
            TA acc[BD], xself[BD];
            const int lr = sym_rows<BD, U>(A, S, x, blk, acc, xself);
            const size_t row = static_cast<size_t>(blk) * NODE_BLOCK + lr;
            double dot = 0.0;
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                const TX yi = static_cast<TX>(acc[i]);
                y[row * BD + i] = yi;
                dot += static_cast<double>(xself[i]) * static_cast<double>(yi);
            }
            store_partial<SPMV_BLOCK>(dot, red.pap, blk);
        }

        /// Values of the symmetric form: one thread per sym entry (slot * 64 + lane).
        template <typename TV>
        __global__ __launch_bounds__(256) void k_sym_gather(int bs, long long entries, const int *__restrict__ src,
                                                            const double *__restrict__ v64, TV *__restrict__ out)
        {
            const long long e = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
            if (e >= entries)
                return;
            const long long slot = e / WAVE, lane = e % WAVE;
            const int s = src[e];
            const long long ss = s / WAVE, sl = s % WAVE;
            for (int el = 0; el < bs; ++el)
                out[(slot * bs + el) * WAVE + lane] = s < 0 ? TV(0) : static_cast<TV>(v64[(ss * bs + el) * WAVE + sl]);
        }

        /// Symmetry of the blocks the symmetric form stores once: block src[e] against the
        /// transpose of block mirror[e], relative to the pair's largest entry.
        __global__ __launch_bounds__(256) void k_sym_check(int bd, long long entries, const int *__restrict__ src,
                                                           const int *__restrict__ mirror,
                                                           const double *__restrict__ v64, double rel_tol,
                                                           int *__restrict__ flag)
        {
            const long long e = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
            if (e >= entries)
                return;
            const int s = src[e], t = mirror[e];
            if (s < 0 || t < 0)
                return;
            const int bs = bd * bd;
            const long long ss = s / WAVE, sl = s % WAVE, ts = t / WAVE, tl = t % WAVE;
            double big = 0.0, diff = 0.0;
            for (int i = 0; i < bd; ++i)
                for (int j = 0; j < bd; ++j)
                {
                    const double a = v64[(ss * bs + i * bd + j) * WAVE + sl];
                    const double b = v64[(ts * bs + j * bd + i) * WAVE + tl];
                    big = fmax(big, fmax(fabs(a), fabs(b)));
                    diff = fmax(diff, fabs(a - b));
                }
            if (!(diff <= rel_tol * big))
                flag[0] = 1;
        }

        /// Streaming-read probe: every 16-byte chunk once (nontemporal); the sum is stored only
        /// under an impossible condition, which keeps the loads.
        typedef float probe_f4 __attribute__((ext_vector_type(4)));
        __global__ __launch_bounds__(256) void k_read_probe(const probe_f4 *__restrict__ a, long long n4,
                                                            float *__restrict__ sink)
        {
            float acc = 0.0f;
            const long long stride = static_cast<long long>(gridDim.x) * 256;
            for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < n4; i += stride)
            {
                const probe_f4 v = __builtin_nontemporal_load(a + i);
                acc += v.x + v.y + v.z + v.w;
            }
            if (acc == -1.2345e-37f)
                sink[0] = acc;
        }

        /// beta of the search-direction update for system sys (the same arithmetic as k_update_p).
        template <typename T>
        __device__ __forceinline__ T direction_beta(const Reductions &red, int sys, const SysMap &sm,
                                                    const double *rzo)
        {
            double rz, rz_old;
            if (rzo != nullptr) // flexible (Polak-Ribiere): r_new . (z_new - z_old)
            {
                const double *const parts[3] = {red.rz, red.rz_old, rzo};
                double t[3];
                system_totals<VEC_BLOCK, 3>(parts, sys, sm, t);
                rz = t[0] - t[2];
                rz_old = t[1];
            }
            else
            {
                const double *const parts[2] = {red.rz, red.rz_old};
                double t[2];
                system_totals<VEC_BLOCK, 2>(parts, sys, sm, t);
                rz = t[0];
                rz_old = t[1];
            }
            return static_cast<T>(safe_div(rz, rz_old, red.sc[static_cast<size_t>(sys) * S_COUNT + S_GUARD]));
        }

        /// k_spmv_ap with the direction update folded in (FusedDirection): the halo receives
        /// z + beta p_old, own rows write p_new (and z_prev).
        template <int BD, int U, typename TV, typename TX>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_ap_fused(SellView<TV> A, FusedDirection<TX> fd,
                                                                      TX *__restrict__ y, Reductions red, SysMap sm)
        {
            using TA = typename Acc<TV>::type;
            const int blk = xcd_block();
            const int sys = system_of_block(sm, blk);
            if (system_stopped(sm, sys))
                return;
            TX *xs = reinterpret_cast<TX *>(spmv_halo_lds);
            const int hb = A.halo_ptr[blk], hn = A.halo_ptr[blk + 1] - hb;
            // The first PRE rounds of halo values are loaded before beta's reduction so that both
            // latencies overlap.
            constexpr int PRE = 3;
            TX zr[PRE][BD], pr[PRE][BD];
#pragma unroll
            for (int r = 0; r < PRE; ++r)
            {
                const int i = threadIdx.x + r * SPMV_BLOCK;
                if (i < hn)
                {
                    const size_t g = static_cast<size_t>(A.halo[hb + i]) * BD;
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                    {
                        zr[r][j] = fd.z[g + j];
                        pr[r][j] = fd.init ? TX(0) : fd.p_old[g + j];
                    }
                }
            }
            const TX beta = fd.init ? TX(0) : direction_beta<TX>(red, sys, sm, fd.rzo);
            auto dir = [&](size_t o) { return fd.init ? fd.z[o] : fma(beta, fd.p_old[o], fd.z[o]); };
#pragma unroll
            for (int r = 0; r < PRE; ++r)
            {
                const int i = threadIdx.x + r * SPMV_BLOCK;
                if (i < hn)
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                        xs[i * BD + j] = fd.init ? zr[r][j] : fma(beta, pr[r][j], zr[r][j]);
            }
            for (int i = threadIdx.x + PRE * SPMV_BLOCK; i < hn; i += SPMV_BLOCK)
            {
                const size_t g = static_cast<size_t>(A.halo[hb + i]) * BD;
#pragma unroll
                for (int j = 0; j < BD; ++j)
                    xs[i * BD + j] = dir(g + j);
            }
            const int lane = threadIdx.x % WAVE;
            const int slice = blk * (SPMV_BLOCK / WAVE) + threadIdx.x / WAVE;
            __syncthreads();
            SellRow<BD, U, true, TV> sr;
            sr.open(A, slice, lane);
            TA acc[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                acc[i] = TA(0);
            sr.run(static_cast<const TX *>(nullptr), xs, acc);
            const size_t row = static_cast<size_t>(slice) * WAVE + lane;
            double dot = 0.0;
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                const size_t o = row * BD + i;
                const TX pn = dir(o);
                fd.p_new[o] = pn;
                if (fd.zprev != nullptr)
                    fd.zprev[o] = fd.z[o];
                const TX yi = static_cast<TX>(acc[i]);
                y[o] = yi;
                dot += static_cast<double>(pn) * static_cast<double>(yi);
            }
            store_partial<SPMV_BLOCK>(dot, red.pap, blk);
        }

        /// u_g of node g from its preloaded coarse ids c, fine z zf and lattice cell lc: FP64 sum in
        /// level order (then the lattice terms), as k_prolong_dot.  The coarse and lattice loads
        /// are unconditional (clamped indices, zero weights where a node has no coarse ids or no
        /// lattice cell), so that the compiler keeps every node's loads of a batch in flight.
        template <int BD, bool LAT>
        __device__ __forceinline__ void cg_prolong(const CgProlong &pr, const int4 &c, const float (&zf)[BD],
                                                   const int4 &lc, float (&u)[BD])
        {
            double z[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                z[i] = static_cast<double>(zf[i]);
            const bool has = c.x >= 0;
            const int cc[4] = {c.x, c.y, c.z, c.w};
            float cv[4][BD];
#pragma unroll
            for (int l = 0; l < 4; ++l)
            {
                const bool on = has && l < pr.level_num;
                const size_t o = on ? static_cast<size_t>(pr.lev_off[l + 1] + cc[l]) * BD : 0;
#pragma unroll
                for (int i = 0; i < BD; ++i)
                    cv[l][i] = l < pr.level_num ? pr.mz[o + i] : 0.0f;
            }
#pragma unroll
            for (int l = 0; l < 4; ++l)
                if (has && l < pr.level_num)
#pragma unroll
                    for (int i = 0; i < BD; ++i)
                        z[i] += static_cast<double>(cv[l][i]);
            if constexpr (LAT)
            {
                const bool in = lc.x >= 0;
                const float fx = __int_as_float(lc.y), fy = __int_as_float(lc.z), fz = __int_as_float(lc.w);
                const int k = pr.lcc.k;
                float cf[8][BD];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                {
                    const int id = in ? lc.x + (q & 1) + ((q >> 1) & 1) * k + ((q >> 2) & 1) * k * k : 0;
#pragma unroll
                    for (int i = 0; i < BD; ++i)
                        cf[q][i] = pr.lcc.coef[static_cast<size_t>(id) * BD + i];
                }
                if (in)
                {
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                    {
                        const float w = ((q & 1) ? fx : 1.0f - fx) * ((q & 2) ? fy : 1.0f - fy) *
                                        ((q & 4) ? fz : 1.0f - fz);
#pragma unroll
                        for (int i = 0; i < BD; ++i)
                            z[i] += static_cast<double>(w) * static_cast<double>(cf[q][i]);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < BD; ++i)
                u[i] = static_cast<float>(z[i]);
        }

        /// Rounds of halo nodes whose inputs the single-reduction SpMV loads in one batch.
        constexpr int HALO_CG_R = 3;

        /// Single-reduction PCG SpMV (launch_spmv_cg): every workgroup stages u = prolongation of
        /// z at its halo in LDS (its own rows' u among them, at own[blk] + row), multiplies, and
        /// writes u and w = A u for its own rows with the partials of r . u, w . u, u . s_prev and
        /// r . u_prev.  The block's previous rz and ps partials move to rz_old and ps_old first.
        template <int BD, int U, bool LAT>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_cg(SellView<float> A, CgProlong pr, float *__restrict__ u,
                                                                float *__restrict__ w, const float *__restrict__ r,
                                                                const float *__restrict__ s_prev,
                                                                const float *__restrict__ u_prev, CgReductions red,
                                                                SysMap sm, int init)
        {
            const int blk = xcd_block();
            const int sys = system_of_block(sm, blk);
            if (system_stopped(sm, sys))
                return;
            float *xs = reinterpret_cast<float *>(spmv_halo_lds);
            const size_t row = static_cast<size_t>(blk) * SPMV_BLOCK + threadIdx.x;
            // The own rows' inputs of the partials do not depend on the halo: loaded first.
            float rv[BD], sv[BD], uv[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                rv[i] = r[row * BD + i];
                sv[i] = init ? 0.0f : s_prev[row * BD + i];
                uv[i] = (init || u_prev == nullptr) ? 0.0f : u_prev[row * BD + i];
            }
            if (threadIdx.x == 0)
            {
                red.rz_old[blk] = red.rz[blk];
                red.ps_old[blk] = red.ps[blk];
            }
            const int hb = A.halo_ptr[blk], hn = A.halo_ptr[blk + 1] - hb;
            if (hn > 0)
            {
                const int last = hn - 1;
                int g[HALO_CG_R];
                int4 c[HALO_CG_R], lc[HALO_CG_R];
                float zf[HALO_CG_R][BD];
#pragma unroll
                for (int k = 0; k < HALO_CG_R; ++k)
                    g[k] = A.halo[hb + min(static_cast<int>(threadIdx.x) + k * SPMV_BLOCK, last)];
#pragma unroll
                for (int k = 0; k < HALO_CG_R; ++k)
                {
                    c[k] = pr.cmap[g[k]];
#pragma unroll
                    for (int i = 0; i < BD; ++i)
                        zf[k][i] = pr.mz[static_cast<size_t>(g[k]) * BD + i];
                    lc[k] = LAT ? pr.lcc.cell[g[k]] : int4{-1, 0, 0, 0};
                }
#pragma unroll
                for (int k = 0; k < HALO_CG_R; ++k)
                {
                    const int i = threadIdx.x + k * SPMV_BLOCK;
                    float uk[BD];
                    cg_prolong<BD, LAT>(pr, c[k], zf[k], lc[k], uk);
                    if (i < hn)
#pragma unroll
                        for (int j = 0; j < BD; ++j)
                            xs[i * BD + j] = uk[j];
                }
                for (int i = threadIdx.x + HALO_CG_R * SPMV_BLOCK; i < hn; i += SPMV_BLOCK)
                {
                    const int gi = A.halo[hb + i];
                    float zi[BD], ui[BD];
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                        zi[j] = pr.mz[static_cast<size_t>(gi) * BD + j];
                    cg_prolong<BD, LAT>(pr, pr.cmap[gi], zi, LAT ? pr.lcc.cell[gi] : int4{-1, 0, 0, 0}, ui);
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                        xs[i * BD + j] = ui[j];
                }
            }
            __syncthreads();
            const int lane = threadIdx.x % WAVE;
            const int slice = blk * (SPMV_BLOCK / WAVE) + threadIdx.x / WAVE;
            SellRow<BD, U, true, float> sr;
            sr.open(A, slice, lane);
            float acc[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                acc[i] = 0.0f;
            sr.run(static_cast<const float *>(nullptr), xs, acc);
            const int own = pr.own[blk] + static_cast<int>(threadIdx.x);
            double dots[4] = {0.0, 0.0, 0.0, 0.0}; // gamma, delta, eps, zeta
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                const float ui = xs[own * BD + i];
                u[row * BD + i] = ui;
                w[row * BD + i] = acc[i];
                const double ud = static_cast<double>(ui);
                dots[0] += static_cast<double>(rv[i]) * ud;
                dots[1] += static_cast<double>(acc[i]) * ud;
                dots[2] += ud * static_cast<double>(sv[i]);
                dots[3] += static_cast<double>(rv[i]) * static_cast<double>(uv[i]);
            }
            double *const dst[4] = {red.rz, red.dlt, init ? nullptr : red.eps,
                                    (init || u_prev == nullptr) ? nullptr : red.zeta};
            store_partials<SPMV_BLOCK, 4>(dots, dst, blk);
        }

        /// y_c = A x_c for the `copies` systems of a batch that repeats one matrix: system c's node
        /// block b is block c * nblocks1 + b and its slots are copy 0's with c * npad1 added to the
        /// columns.  Workgroup b reads each block of copy 0's slots once and applies it to every
        /// copy's x (KC >= copies accumulators per row); copy c's pAp partial goes to its block.
        /// Per copy the FMAs run in the same order as in k_spmv_ap.
        ///
        /// HALO: every copy's x at block b's halo (copy 0's list plus c * npad1) is first staged in
        /// the dynamic LDS array (copy after copy, hn entries each) and read through the 16-bit
        /// local columns.
        template <int BD, int KC, bool HALO, typename TV, typename TX>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_ap_copies(SellView<TV> A, int c0, int copies, int npad1,
                                                                       int nblocks1, const TX *__restrict__ x,
                                                                       TX *__restrict__ y, Reductions red, SysMap sm)
        {
            // This launch's copies are c0 .. c0 + copies - 1 of the batch.
            x += static_cast<size_t>(c0) * npad1 * BD;
            y += static_cast<size_t>(c0) * npad1 * BD;
            using TA = typename Acc<TV>::type;
            constexpr int U = 2;
            const int b = xcd_block();
            const int lane = threadIdx.x % WAVE;
            const int slice = b * (SPMV_BLOCK / WAVE) + threadIdx.x / WAVE;
            TX *xs = reinterpret_cast<TX *>(spmv_halo_lds);
            int hn = 0;
            if constexpr (HALO)
            {
                const int hb = A.halo_ptr[b];
                hn = A.halo_ptr[b + 1] - hb;
                for (int i = threadIdx.x; i < copies * hn; i += SPMV_BLOCK)
                {
                    const int c = i / hn, t = i - c * hn;
                    const size_t g = (static_cast<size_t>(c) * npad1 + A.halo[hb + t]) * BD;
#pragma unroll
                    for (int j = 0; j < BD; ++j)
                        xs[i * BD + j] = x[g + j];
                }
                __syncthreads();
            }
            bool act[KC];
#pragma unroll
            for (int c = 0; c < KC; ++c)
                act[c] = c < copies && sm.active[sm.block_sys[(c0 + c) * nblocks1 + b]] != 0;
            TA acc[KC][BD];
#pragma unroll
            for (int c = 0; c < KC; ++c)
#pragma unroll
                for (int i = 0; i < BD; ++i)
                    acc[c][i] = TA(0);
            const int beg = __builtin_amdgcn_readfirstlane(A.slice_ptr[slice]);
            const int end = __builtin_amdgcn_readfirstlane(A.slice_ptr[slice + 1]);
            const int last = end - 1;
            const int *__restrict__ cols = A.cols + lane;
            const unsigned short *__restrict__ lcols = A.lcols + lane;
            const TV *__restrict__ vals = A.vals + lane;
            for (int k = beg; k < end; k += U)
            {
                int col[U];
                TV v[U][BD * BD];
                TA xv[U][KC][BD];
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const size_t slot = static_cast<size_t>(min(k + u, last));
                    if constexpr (HALO)
                        col[u] = static_cast<int>(ldg_stream(lcols + slot * WAVE));
                    else
                        col[u] = ldg_stream(cols + slot * WAVE);
#pragma unroll
                    for (int e = 0; e < BD * BD; ++e)
                        v[u][e] = ldg_stream(vals + (slot * (BD * BD) + e) * WAVE);
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int c = 0; c < KC; ++c)
#pragma unroll
                        for (int j = 0; j < BD; ++j)
                        {
                            if (c >= copies)
                                xv[u][c][j] = TA(0);
                            else if constexpr (HALO)
                                xv[u][c][j] = static_cast<TA>(xs[(c * hn + col[u]) * BD + j]);
                            else
                                xv[u][c][j] = static_cast<TA>(x[(static_cast<size_t>(c) * npad1 + col[u]) * BD + j]);
                        }
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const bool ok = k + u < end; // wave-uniform
#pragma unroll
                    for (int c = 0; c < KC; ++c)
#pragma unroll
                        for (int i = 0; i < BD; ++i)
#pragma unroll
                            for (int j = 0; j < BD; ++j)
                                acc[c][i] = fma(ok ? static_cast<TA>(v[u][i * BD + j]) : TA(0), xv[u][c][j], acc[c][i]);
                }
            }
            const size_t row = static_cast<size_t>(slice) * WAVE + lane;
#pragma unroll
            for (int c = 0; c < KC; ++c)
            {
                if (c >= copies)
                    break;
                double dot = 0.0;
                const size_t g = static_cast<size_t>(c) * npad1 + row;
                if (act[c])
                {
#pragma unroll
                    for (int i = 0; i < BD; ++i)
                    {
                        const TX yi = static_cast<TX>(acc[c][i]);
                        y[g * BD + i] = yi;
                        dot += static_cast<double>(x[g * BD + i]) * static_cast<double>(yi);
                    }
                }
                if (act[c]) // uniform in the workgroup
                    store_partial<SPMV_BLOCK>(dot, red.pap, (c0 + c) * nblocks1 + b);
            }
        }

        /// Leader lanes of a bank sum their component's members from LDS into the coarse vector.
        template <int BD, typename T>
        __device__ __forceinline__ void restrict_bank(const T *sv, int bank_base_in_block, uint32_t mask, int gid,
                                                      T *__restrict__ yc0)
        {
            if (mask == 0u)
                return;
            double s[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                s[i] = 0.0;
            while (mask)
            {
                const int b = __builtin_ctz(mask);
                mask &= mask - 1u;
#pragma unroll
                for (int i = 0; i < BD; ++i)
                    s[i] += static_cast<double>(sv[(bank_base_in_block + b) * BD + i]);
            }
#pragma unroll
            for (int i = 0; i < BD; ++i)
                yc0[static_cast<size_t>(gid) * BD + i] = static_cast<T>(s[i]);
        }

        template <int BD, int U, bool HALO, typename TV, typename TX, typename TY>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_res(SellView<TV> A, const TX *__restrict__ x,
                                                                 const TX *__restrict__ b, TY *__restrict__ y,
                                                                 FineRestrict fr, TY *__restrict__ yc0,
                                                                 double *__restrict__ rr_partials, SysMap sm,
                                                                 int selected_only)
        {
            using TA = typename Acc<TV>::type;
            __shared__ TY sv[SPMV_BLOCK * BD];
            const int blk = xcd_block();
            const int sys = system_of_block(sm, blk);
            if (system_stopped(sm, sys) || (selected_only && sm.select[sys] == 0))
                return;
            const int lane = threadIdx.x % WAVE;
            const int slice = blk * (SPMV_BLOCK / WAVE) + threadIdx.x / WAVE;
            const size_t row = static_cast<size_t>(slice) * WAVE + lane;
            HaloStage hs;
            if constexpr (HALO)
                hs = halo_indices(A, blk);
            const TX *xs = nullptr;
            if constexpr (HALO)
                xs = halo_fill<BD>(hs, A, x);
            SellRow<BD, U, HALO, TV> sr;
            sr.open(A, slice, lane);
            TA acc[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                acc[i] = TA(0);
            sr.run(x, xs, acc);
            double dot = 0.0;
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                const double ri = static_cast<double>(b[row * BD + i]) - static_cast<double>(acc[i]);
                const TY yi = static_cast<TY>(ri);
                y[row * BD + i] = yi;
                sv[threadIdx.x * BD + i] = yi;
                dot += ri * ri;
            }
            const uint32_t mask = fr.lead[row];
            const int gid = fr.gid[row];
            __syncthreads();
            restrict_bank<BD, TY>(sv, (threadIdx.x / 32) * 32, mask, gid, yc0);
            store_partial<SPMV_BLOCK>(dot, rr_partials, blk);
        }

        /// k_spmv_res over the symmetric form S: the rows come in the form's order, so each
        /// residual is staged in LDS at its local row for the bank restriction.  The stage reuses
        /// the halo's LDS (free once every row has passed the barrier in sym_rows; the halo holds
        /// at least the block's own 256 nodes), which keeps the kernel's LDS that of k_spmv_sym.
        template <int BD, int U, typename TV, typename TX, typename TY>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_res_sym(SellView<TV> A, SymView<TV> S,
                                                                     const TX *__restrict__ x, const TX *__restrict__ b,
                                                                     TY *__restrict__ y, FineRestrict fr,
                                                                     TY *__restrict__ yc0, double *__restrict__ rr_partials,
                                                                     SysMap sm, int selected_only)
        {
            static_assert(sizeof(TY) <= sizeof(TX), "the residual stage fits the halo's LDS");
            using TA = typename Acc<TV>::type;
            TY *const sv = reinterpret_cast<TY *>(spmv_halo_lds);
            const int blk = xcd_block();
            const int sys = system_of_block(sm, blk);
            if (system_stopped(sm, sys) || (selected_only && sm.select[sys] == 0))
                return;
            TA acc[BD], xself[BD];
            const int lr = sym_rows<BD, U>(A, S, x, blk, acc, xself);
            const size_t row = static_cast<size_t>(blk) * NODE_BLOCK + lr;
            double dot = 0.0;
#pragma unroll
            for (int i = 0; i < BD; ++i)
            {
                const double ri = static_cast<double>(b[row * BD + i]) - static_cast<double>(acc[i]);
                const TY yi = static_cast<TY>(ri);
                y[row * BD + i] = yi;
                sv[lr * BD + i] = yi;
                dot += ri * ri;
            }
            const size_t own = static_cast<size_t>(blk) * NODE_BLOCK + threadIdx.x; // (bank lists by row)
            const uint32_t mask = fr.lead[own];
            const int gid = fr.gid[own];
            __syncthreads();
            restrict_bank<BD, TY>(sv, (threadIdx.x / 32) * 32, mask, gid, yc0);
            store_partial<SPMV_BLOCK>(dot, rr_partials, blk);
        }

        template <int BD, int U, typename TV, typename TX>
        __global__ __launch_bounds__(SPMV_BLOCK) void k_spmv_plain(SellView<TV> A, const TX *__restrict__ x,
                                                                   TX *__restrict__ y)
        {
            using TA = typename Acc<TV>::type;
            const int lane = threadIdx.x % WAVE;
            const int slice = blockIdx.x * (SPMV_BLOCK / WAVE) + threadIdx.x / WAVE;
            if (slice >= A.nslices)
                return;
            TA acc[BD];
#pragma unroll
            for (int i = 0; i < BD; ++i)
                acc[i] = TA(0);
            SellRow<BD, U, false, TV> sr;
            sr.open(A, slice, lane);
            sr.run(x, static_cast<const TX *>(nullptr), acc);
            const size_t row = static_cast<size_t>(slice) * WAVE + lane;
#pragma unroll
            for (int i = 0; i < BD; ++i)
                y[row * BD + i] = static_cast<TX>(acc[i]);
        }

        /// alpha = rz / pap of the block's system (reference scalar_division, MASSolver.cu:46-56).
        __device__ __forceinline__ double system_alpha(const Reductions &red, int sys, const SysMap &sm)
        {
            const double *const parts[2] = {red.rz, red.pap};
            double t[2];
            system_totals<VEC_BLOCK, 2>(parts, sys, sm, t);
            return safe_div(t[0], t[1], red.sc[static_cast<size_t>(sys) * S_COUNT + S_GUARD]);
        }

        /// x += alpha p; r -= alpha Ap on the block's 256*BD contiguous scalars (coalesced), then
        /// r . r and the restriction of r to coarse level 0 from LDS.
        template <int BD, typename T, bool LAT>
        __global__ __launch_bounds__(VEC_BLOCK) void k_update_xr(const T *__restrict__ p, const T *__restrict__ Ap,
                                                                 T *__restrict__ x, T *__restrict__ r, FineRestrict fr,
                                                                 T *__restrict__ rc0, Reductions red, SysMap sm,
                                                                 LatticeBlocks lb)
        {
            static_assert(VEC_BLOCK == 256, "lattice blocks are 256-node blocks");
            __shared__ T sv[VEC_BLOCK * BD];
            const int sys = block_system(sm);
            if (system_stopped(sm, sys))
                return;
            // Lattice inputs first: their loads overlap the vector update.
            LatticeSlot ls;
            if constexpr (LAT)
                ls = lattice_slot(lb, blockIdx.x);
            // The block's vectors and restriction map are loaded before alpha's reduction (whose
            // barriers the loads would not cross) so that all these latencies overlap.
            const size_t node = static_cast<size_t>(blockIdx.x) * VEC_BLOCK + threadIdx.x;
            const uint32_t mask = fr.lead[node];
            const int gid = fr.gid[node];
            const size_t base = static_cast<size_t>(blockIdx.x) * VEC_BLOCK * BD;
            T pv[BD], apv[BD], xv[BD], rv0[BD];
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const size_t o = base + threadIdx.x + k * VEC_BLOCK;
                pv[k] = p[o];
                apv[k] = Ap[o];
                xv[k] = x[o];
                rv0[k] = r[o];
            }
            const T alpha = static_cast<T>(system_alpha(red, sys, sm));
            double dot = 0.0;
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const int j = threadIdx.x + k * VEC_BLOCK;
                const size_t o = base + j;
                x[o] = fma(alpha, pv[k], xv[k]);
                const T rv = fma(-alpha, apv[k], rv0[k]);
                r[o] = rv;
                sv[j] = rv;
                dot += static_cast<double>(rv) * static_cast<double>(rv);
            }
            __syncthreads();
            restrict_bank<BD, T>(sv, (threadIdx.x / 32) * 32, mask, gid, rc0);
            store_partial<VEC_BLOCK>(dot, red.rr, blockIdx.x);
            if constexpr (LAT)
            {
                __shared__ T vals[256 * 8 * BD];
                T rn[BD];
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    rn[d] = sv[threadIdx.x * BD + d];
                lattice_block_partials<BD, T>(lb, blockIdx.x, ls, rn, vals);
            }
        }

        /// Single-reduction PCG update (launch_update_cg).
        template <int BD, bool LAT>
        __global__ __launch_bounds__(VEC_BLOCK) void k_update_cg(const float *__restrict__ u, const float *__restrict__ w,
                                                                 float *__restrict__ p, float *__restrict__ s,
                                                                 float *__restrict__ x, float *__restrict__ r,
                                                                 FineRestrict fr, float *__restrict__ rc0,
                                                                 float *__restrict__ u_prev, CgReductions red,
                                                                 SysMap sm, LatticeBlocks lb, int init)
        {
            static_assert(VEC_BLOCK == 256, "lattice blocks are 256-node blocks");
            __shared__ float sv[VEC_BLOCK * BD];
            const int sys = block_system(sm);
            if (system_stopped(sm, sys))
                return;
            LatticeSlot ls;
            if constexpr (LAT)
                ls = lattice_slot(lb, blockIdx.x);
            // The block's vectors and restriction map are loaded before the scalars' reduction
            // (whose barriers the loads would not cross) so that the latencies overlap.
            const size_t node = static_cast<size_t>(blockIdx.x) * VEC_BLOCK + threadIdx.x;
            const uint32_t mask = fr.lead[node];
            const int gid = fr.gid[node];
            const size_t base = static_cast<size_t>(blockIdx.x) * VEC_BLOCK * BD;
            float uv[BD], wv[BD], pv[BD], svv[BD], xv[BD], rv0[BD];
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const size_t o = base + threadIdx.x + k * VEC_BLOCK;
                uv[k] = u[o];
                wv[k] = w[o];
                pv[k] = init ? 0.0f : p[o];
                svv[k] = init ? 0.0f : s[o];
                xv[k] = x[o];
                rv0[k] = r[o];
            }
            // Every scalar comes from partials (no state kept between launches): gamma = r . u,
            // delta = u . A u, eps = u . s_prev, gamma_old, (p . s)_old and zeta = r . u_prev.
            const double guard = red.sc[static_cast<size_t>(sys) * S_COUNT + S_GUARD];
            const bool flex = u_prev != nullptr;
            double alpha, beta = 0.0;
            if (init)
            {
                const double *const parts[2] = {red.rz, red.dlt};
                double t[2];
                system_totals<VEC_BLOCK, 2>(parts, sys, sm, t);
                alpha = safe_div(t[0], t[1], guard);
            }
            else
            {
                const double *const parts[6] = {red.rz, red.dlt, red.eps, red.rz_old, red.ps_old,
                                                 flex ? red.zeta : red.rz};
                double t[6];
                system_totals<VEC_BLOCK, 6>(parts, sys, sm, t);
                beta = safe_div(flex ? t[0] - t[5] : t[0], t[3], guard);
                // p . A p of the new direction from the recurrences: (u + beta p)(w + beta s).
                const double pap = t[1] + 2.0 * beta * t[2] + beta * beta * t[4];
                alpha = safe_div(t[0], pap, guard);
            }
            const float al = static_cast<float>(alpha), be = static_cast<float>(beta);
            double rr = 0.0, ps = 0.0;
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const int j = threadIdx.x + k * VEC_BLOCK;
                const size_t o = base + j;
                const float pn = fmaf(be, pv[k], uv[k]);
                const float sn = fmaf(be, svv[k], wv[k]);
                p[o] = pn;
                s[o] = sn;
                x[o] = fmaf(al, pn, xv[k]);
                const float rn = fmaf(-al, sn, rv0[k]);
                r[o] = rn;
                sv[j] = rn;
                if (flex)
                    u_prev[o] = uv[k];
                rr += static_cast<double>(rn) * static_cast<double>(rn);
                ps += static_cast<double>(pn) * static_cast<double>(sn);
            }
            __syncthreads();
            restrict_bank<BD, float>(sv, (threadIdx.x / 32) * 32, mask, gid, rc0);
            const double v2[2] = {rr, ps};
            double *const dst[2] = {red.rr, red.ps};
            store_partials<VEC_BLOCK, 2>(v2, dst, blockIdx.x);
            if constexpr (LAT)
            {
                __shared__ float vals[256 * 8 * BD];
                float rn[BD];
#pragma unroll
                for (int d = 0; d < BD; ++d)
                    rn[d] = sv[threadIdx.x * BD + d];
                lattice_block_partials<BD, float>(lb, blockIdx.x, ls, rn, vals);
            }
        }

        template <int BD, typename T>
        __global__ __launch_bounds__(VEC_BLOCK) void k_axpy_alpha(const T *__restrict__ p, T *__restrict__ x,
                                                                  Reductions red, SysMap sm)
        {
            const int sys = block_system(sm);
            if (system_stopped(sm, sys))
                return;
            const T alpha = static_cast<T>(system_alpha(red, sys, sm));
            const size_t base = static_cast<size_t>(blockIdx.x) * VEC_BLOCK * BD;
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const size_t o = base + threadIdx.x + k * VEC_BLOCK;
                x[o] = fma(alpha, p[o], x[o]);
            }
        }

        template <int BD, typename T>
        __global__ __launch_bounds__(VEC_BLOCK) void k_update_p(const T *__restrict__ z, T *__restrict__ p,
                                                                Reductions red, SysMap sm,
                                                                const double *__restrict__ rzo, T *__restrict__ zprev)
        {
            const int sys = block_system(sm);
            if (system_stopped(sm, sys))
                return;
            // The block's vectors are loaded before beta's reduction so both latencies overlap.
            const size_t base = static_cast<size_t>(blockIdx.x) * VEC_BLOCK * BD;
            T zv[BD], pv[BD];
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const size_t o = base + threadIdx.x + k * VEC_BLOCK;
                zv[k] = z[o];
                pv[k] = p[o];
            }
            double rz, rz_old;
            if (rzo != nullptr) // flexible (Polak-Ribiere): r_new . (z_new - z_old)
            {
                const double *const parts[3] = {red.rz, red.rz_old, rzo};
                double t[3];
                system_totals<VEC_BLOCK, 3>(parts, sys, sm, t);
                rz = t[0] - t[2];
                rz_old = t[1];
            }
            else
            {
                const double *const parts[2] = {red.rz, red.rz_old};
                double t[2];
                system_totals<VEC_BLOCK, 2>(parts, sys, sm, t);
                rz = t[0];
                rz_old = t[1];
            }
            const T beta = static_cast<T>(safe_div(rz, rz_old, red.sc[static_cast<size_t>(sys) * S_COUNT + S_GUARD]));
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const size_t o = base + threadIdx.x + k * VEC_BLOCK;
                p[o] = fma(beta, pv[k], zv[k]);
                if (zprev != nullptr)
                    zprev[o] = zv[k];
            }
        }

        template <int BD>
        __global__ __launch_bounds__(VEC_BLOCK) void k_accumulate_correction(float *__restrict__ x32,
                                                                             double *__restrict__ x64, SysMap sm)
        {
            const int sys = block_system(sm);
            if (system_stopped(sm, sys) || sm.select[sys] == 0)
                return;
            const size_t base = static_cast<size_t>(blockIdx.x) * VEC_BLOCK * BD;
#pragma unroll
            for (int k = 0; k < BD; ++k)
            {
                const size_t o = base + threadIdx.x + k * VEC_BLOCK;
                x64[o] += static_cast<double>(x32[o]);
                x32[o] = 0.0f;
            }
        }

        template <typename TS, typename TD>
        __global__ void k_copy_convert(long long n, const TS *__restrict__ src, TD *__restrict__ dst)
        {
            for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
                dst[i] = static_cast<TD>(src[i]);
        }

        template <typename T>
        __global__ void k_fill(long long n, T *__restrict__ dst, T value)
        {
            for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
                dst[i] = value;
        }

        template <typename T>
        __global__ void k_gather_in(int npad, int bd, int n, const int *__restrict__ block_of_slot,
                                    const T *__restrict__ src, T *__restrict__ dst)
        {
            const long long total = static_cast<long long>(npad) * bd;
            for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
            {
                const int slot = static_cast<int>(i / bd), c = static_cast<int>(i % bd);
                const int blk = block_of_slot[slot];
                const long long s = static_cast<long long>(blk) * bd + c;
                dst[i] = (blk >= 0 && s < n) ? src[s] : T(0);
            }
        }

        template <typename T>
        __global__ void k_scatter_out(int nb, int bd, int n, const int *__restrict__ slot_of_block,
                                      const T *__restrict__ src, T *__restrict__ dst)
        {
            for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
            {
                const int blk = static_cast<int>(i / bd), c = static_cast<int>(i % bd);
                dst[i] = src[static_cast<long long>(slot_of_block[blk]) * bd + c];
            }
        }

        __global__ void k_set_scalar(double *sc, int stride, int count, int slot, double value)
        {
            const int s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s < count)
                sc[static_cast<size_t>(s) * stride + slot] = value;
        }

        __global__ void k_snapshot(long long n, const double *__restrict__ x64, const float *__restrict__ x32,
                                   double *__restrict__ out)
        {
            for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
                out[i] = x64[i] + static_cast<double>(x32[i]);
        }

        __global__ void k_difference(long long n, const double *__restrict__ a, const double *__restrict__ b,
                                     double *__restrict__ out)
        {
            for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
                 i += static_cast<long long>(gridDim.x) * blockDim.x)
                out[i] = a[i] - b[i];
        }

        /// One workgroup per system: totals of every reduction site -> the system's scalars.
        __global__ __launch_bounds__(VEC_BLOCK) void k_finalize(Reductions red, SysMap sm)
        {
            const int sys = blockIdx.x;
            const double *const parts[5] = {red.pap, red.rr, red.rz, red.rz_old, red.rr_true};
            double t[5];
            system_totals<VEC_BLOCK, 5>(parts, sys, sm, t);
            const double pap = t[0], rr = t[1], rz = t[2], rz_old = t[3], rr_true = t[4];
            if (threadIdx.x == 0)
            {
                double *scs = red.sc + static_cast<size_t>(sys) * S_COUNT;
                scs[S_PAP] = pap;
                scs[S_RR] = rr;
                scs[S_RZ] = rz;
                scs[S_RZ_OLD] = rz_old;
                scs[S_RR_TRUE] = rr_true;
            }
        }

        int stream_grid(long long n)
        {
            long long g = div_up_ll(n, 256);
            if (g > 8192)
                g = 8192;
            return static_cast<int>(g < 1 ? 1 : g);
        }
    } // namespace

    // The row loop's batch depth is a template parameter; the launchers pick the instance from
    // SellView::unroll (2, 4 or 8).
#define PSMAS_UNROLL_DISPATCH(U_, CALL)                                                                             \
    switch (U_)                                                                                                    \
    {                                                                                                              \
    case 2:                                                                                                        \
    {                                                                                                              \
        constexpr int U = 2;                                                                                       \
        CALL;                                                                                                      \
        break;                                                                                                     \
    }                                                                                                              \
    case 8:                                                                                                        \
    {                                                                                                              \
        constexpr int U = 8;                                                                                       \
        CALL;                                                                                                      \
        break;                                                                                                     \
    }                                                                                                              \
    default:                                                                                                       \
    {                                                                                                              \
        constexpr int U = 4;                                                                                       \
        CALL;                                                                                                      \
        break;                                                                                                     \
    }                                                                                                              \
    }

    template <int BD, typename TV, typename TX>
    void launch_spmv_ap(const SellView<TV> &A, const TX *x, TX *y, const Reductions &red, const SysMap &sm,
                        hipStream_t s)
    {
        require_device_ptrs("spmv_ap", {A.slice_ptr, A.cols, A.vals, x, y});
        const int grid = A.nslices / (SPMV_BLOCK / WAVE);
        if (A.halo_max > 0)
        {
            const size_t lds = static_cast<size_t>(A.halo_max) * BD * sizeof(TX);
            PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_ap<BD, U, true, TV, TX>), dim3(grid),
                                                               dim3(SPMV_BLOCK), lds, s, A, x, y, red, sm));
        }
        else
        {
            PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_ap<BD, U, false, TV, TX>), dim3(grid),
                                                               dim3(SPMV_BLOCK), 0, s, A, x, y, red, sm));
        }
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename TV, typename TX>
    void launch_spmv_ap_fused(const SellView<TV> &A, const FusedDirection<TX> &fd, TX *y, const Reductions &red,
                              const SysMap &sm, hipStream_t s)
    {
        require_device_ptrs("spmv_ap_fused", {A.slice_ptr, A.lcols, A.halo_ptr, A.halo, A.vals, fd.z, fd.p_new, y});
        if (!fd.init)
            require_device_ptrs("spmv_ap_fused (previous direction)", {fd.p_old});
        if (A.halo_max <= 0)
            throw std::logic_error("[MAS] internal error: the fused direction update needs the SpMV halo lists.");
        if (fd.p_old == fd.p_new)
            throw std::logic_error("[MAS] internal error: the fused direction update needs two direction buffers.");
        const int grid = A.nslices / (SPMV_BLOCK / WAVE);
        const size_t lds = static_cast<size_t>(A.halo_max) * BD * sizeof(TX);
        PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_ap_fused<BD, U, TV, TX>), dim3(grid),
                                                           dim3(SPMV_BLOCK), lds, s, A, fd, y, red, sm));
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename TV, typename TX>
    size_t spmv_sym_lds(int halo_max, int incoming_max)
    {
        return sym_incoming_offset<BD, TX>(halo_max) +
               static_cast<size_t>(incoming_max) * BD * sizeof(typename Acc<TV>::type);
    }

    template <int BD, typename TV, typename TX>
    void launch_spmv_sym(const SellView<TV> &A, const SymView<TV> &S, const TX *x, TX *y, const Reductions &red,
                         const SysMap &sm, hipStream_t s)
    {
        require_device_ptrs("spmv_sym", {A.halo_ptr, A.halo, S.slice_ptr, S.meta, S.vals, S.row_info, x, y});
        if (A.halo_max <= 0 || S.nslices != A.nslices || S.incoming_max < 0)
            throw std::logic_error("[MAS] internal error: the symmetric SpMV needs the halo lists of its matrix.");
        const size_t lds = spmv_sym_lds<BD, TV, TX>(A.halo_max, S.incoming_max);
        if (lds > SPMV_SYM_LDS_CAP)
            throw std::logic_error("[MAS] internal error: the symmetric SpMV's LDS exceeds its cap.");
        const int grid = A.nslices / (SPMV_BLOCK / WAVE);
        // Above 64 KB of dynamic LDS a kernel must be allowed it explicitly (once per instance).
        auto allow = [](const void *fn) {
            PSMAS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, SPMV_SYM_LDS_CAP));
        };
#define PSMAS_SYM_LAUNCH                                                                                               \
    {                                                                                                                  \
        if (lds > 65536)                                                                                               \
        {                                                                                                              \
            static const bool allowed = (allow(reinterpret_cast<const void *>(&k_spmv_sym<BD, U, TV, TX>)), true);     \
            (void)allowed;                                                                                             \
        }                                                                                                              \
        hipLaunchKernelGGL((k_spmv_sym<BD, U, TV, TX>), dim3(grid), dim3(SPMV_BLOCK), lds, s, A, S, x, y, red, sm);    \
    }
        PSMAS_UNROLL_DISPATCH(S.unroll, PSMAS_SYM_LAUNCH);
#undef PSMAS_SYM_LAUNCH
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename TV, typename TX, typename TY>
    void launch_spmv_res_sym(const SellView<TV> &A, const SymView<TV> &S, const TX *x, const TX *b, TY *y,
                             FineRestrict fr, TY *yc0, double *rr_partials, const SysMap &sm, bool selected_only,
                             hipStream_t s)
    {
        require_device_ptrs("spmv_res_sym", {A.halo_ptr, A.halo, S.slice_ptr, S.meta, S.vals, S.row_info, x, b, y,
                                             fr.lead, fr.gid, yc0, rr_partials});
        if (A.halo_max < NODE_BLOCK || S.nslices != A.nslices || S.incoming_max < 0)
            throw std::logic_error("[MAS] internal error: the symmetric SpMV needs the halo lists of its matrix.");
        const size_t lds = spmv_sym_lds<BD, TV, TX>(A.halo_max, S.incoming_max);
        if (lds > SPMV_SYM_LDS_CAP)
            throw std::logic_error("[MAS] internal error: the symmetric SpMV's LDS exceeds its cap.");
        const int grid = A.nslices / (SPMV_BLOCK / WAVE);
        const int sel = selected_only ? 1 : 0;
        auto allow = [](const void *fn) {
            PSMAS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, SPMV_SYM_LDS_CAP));
        };
#define PSMAS_SYM_RES_LAUNCH                                                                                           \
    {                                                                                                                  \
        if (lds > 65536)                                                                                               \
        {                                                                                                              \
            static const bool allowed =                                                                                \
                (allow(reinterpret_cast<const void *>(&k_spmv_res_sym<BD, U, TV, TX, TY>)), true);                     \
            (void)allowed;                                                                                             \
        }                                                                                                              \
        hipLaunchKernelGGL((k_spmv_res_sym<BD, U, TV, TX, TY>), dim3(grid), dim3(SPMV_BLOCK), lds, s, A, S, x, b, y,  \
                           fr, yc0, rr_partials, sm, sel);                                                             \
    }
        PSMAS_UNROLL_DISPATCH(S.unroll, PSMAS_SYM_RES_LAUNCH);
#undef PSMAS_SYM_RES_LAUNCH
        PSMAS_LAUNCH_CHECK();
    }

    template <typename TV>
    void launch_sym_gather(int bd, long long entries, const int *src, const double *vals64, TV *out, hipStream_t s)
    {
        if (entries <= 0)
            return;
        require_device_ptrs("sym_gather", {src, vals64, out});
        hipLaunchKernelGGL((k_sym_gather<TV>), dim3(static_cast<unsigned>((entries + 255) / 256)), dim3(256), 0, s,
                           bd * bd, entries, src, vals64, out);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_sym_check(int bd, long long entries, const int *src, const int *mirror, const double *vals64,
                          double rel_tol, int *flag, hipStream_t s)
    {
        if (entries <= 0)
            return;
        require_device_ptrs("sym_check", {src, mirror, vals64, flag});
        hipLaunchKernelGGL(k_sym_check, dim3(static_cast<unsigned>((entries + 255) / 256)), dim3(256), 0, s, bd,
                           entries, src, mirror, vals64, rel_tol, flag);
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD>
    void launch_spmv_cg(const SellView<float> &A, const CgProlong &pr, float *u, float *w, const float *r,
                        const float *s_prev, const float *u_prev, const CgReductions &red, const SysMap &sm,
                        bool init, hipStream_t s)
    {
        require_device_ptrs("spmv_cg", {A.slice_ptr, A.lcols, A.halo_ptr, A.halo, A.vals, pr.mz, pr.cmap, pr.lev_off,
                                        pr.own, u, w, r, red.rz, red.rz_old, red.dlt, red.ps, red.ps_old});
        if (!init)
            require_device_ptrs("spmv_cg (previous direction)", {s_prev, red.eps});
        if (!init && u_prev != nullptr)
            require_device_ptrs("spmv_cg (flexible)", {red.zeta});
        if (A.halo_max <= 0)
            throw std::logic_error("[MAS] internal error: the single-reduction SpMV needs the SpMV halo lists.");
        if (pr.level_num < 0 || pr.level_num > 4)
            throw std::logic_error("[MAS] internal error: at most 4 coarse levels.");
        const int grid = A.nslices / (SPMV_BLOCK / WAVE);
        const size_t lds = static_cast<size_t>(A.halo_max) * BD * sizeof(float);
        const int in = init ? 1 : 0;
        if (pr.lcc.cell != nullptr)
        {
            require_device_ptrs("spmv_cg (lattice)", {pr.lcc.coef});
            PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_cg<BD, U, true>), dim3(grid), dim3(SPMV_BLOCK),
                                                               lds, s, A, pr, u, w, r, s_prev, u_prev, red, sm, in));
        }
        else
        {
            PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_cg<BD, U, false>), dim3(grid), dim3(SPMV_BLOCK),
                                                               lds, s, A, pr, u, w, r, s_prev, u_prev, red, sm, in));
        }
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD>
    void launch_update_cg(int nblocks, const float *u, const float *w, float *p, float *s_vec, float *x, float *r,
                          FineRestrict fr, float *rc0, float *u_prev, const CgReductions &red, const SysMap &sm,
                          bool init, hipStream_t s, const LatticeBlocks *lb)
    {
        require_device_ptrs("update_cg", {u, w, p, s_vec, x, r, fr.lead, fr.gid, rc0, red.rz, red.dlt, red.rr, red.ps,
                                          red.sc});
        if (!init)
            require_device_ptrs("update_cg (previous iteration)", {red.eps, red.rz_old, red.ps_old});
        if (!init && u_prev != nullptr)
            require_device_ptrs("update_cg (flexible)", {red.zeta});
        const int in = init ? 1 : 0;
        if (lb != nullptr && lb->cell != nullptr)
            hipLaunchKernelGGL((k_update_cg<BD, true>), dim3(nblocks), dim3(VEC_BLOCK), 0, s, u, w, p, s_vec, x, r, fr,
                               rc0, u_prev, red, sm, *lb, in);
        else
            hipLaunchKernelGGL((k_update_cg<BD, false>), dim3(nblocks), dim3(VEC_BLOCK), 0, s, u, w, p, s_vec, x, r,
                               fr, rc0, u_prev, red, sm, LatticeBlocks{}, in);
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename TV, typename TX, typename TY>
    void launch_spmv_res(const SellView<TV> &A, const TX *x, const TX *b, TY *y, FineRestrict fr, TY *yc0,
                         double *rr_partials, const SysMap &sm, bool selected_only, hipStream_t s)
    {
        require_device_ptrs("spmv_res", {A.slice_ptr, A.cols, A.vals, x, b, y});
        const int grid = A.nslices / (SPMV_BLOCK / WAVE);
        const int sel = selected_only ? 1 : 0;
        if (A.halo_max > 0)
        {
            const size_t lds = static_cast<size_t>(A.halo_max) * BD * sizeof(TX);
            PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_res<BD, U, true, TV, TX, TY>), dim3(grid),
                                                               dim3(SPMV_BLOCK), lds, s, A, x, b, y, fr, yc0,
                                                               rr_partials, sm, sel));
        }
        else
        {
            PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_res<BD, U, false, TV, TX, TY>), dim3(grid),
                                                               dim3(SPMV_BLOCK), 0, s, A, x, b, y, fr, yc0,
                                                               rr_partials, sm, sel));
        }
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename TV, typename TX>
    void launch_spmv_plain(const SellView<TV> &A, const TX *x, TX *y, hipStream_t s)
    {
        require_device_ptrs("spmv_plain", {A.slice_ptr, A.cols, A.vals, x, y});
        const int grid = div_up(A.nslices, SPMV_BLOCK / WAVE);
        PSMAS_UNROLL_DISPATCH(A.unroll, hipLaunchKernelGGL((k_spmv_plain<BD, U, TV, TX>), dim3(grid), dim3(SPMV_BLOCK),
                                                           0, s, A, x, y));
        PSMAS_LAUNCH_CHECK();
    }
    template <int BD, typename TV, typename TX>
    void launch_spmv_ap_copies(const SellView<TV> &A, int copies, int npad1, const TX *x, TX *y, const Reductions &red,
                               const SysMap &sm, hipStream_t s)
    {
        require_device_ptrs("spmv_ap_copies", {A.slice_ptr, A.cols, A.vals, x, y});
        // Groups of at most 4 copies per launch: their halos (copy 0's lists plus c * npad1) stay
        // within 64 KB of LDS, and each group reads the matrix once.
        const int nblocks1 = npad1 / SPMV_BLOCK;
        for (int c0 = 0; c0 < copies; c0 += 4)
        {
            const int cn = std::min(4, copies - c0);
            const size_t lds = static_cast<size_t>(cn) * A.halo_max * BD * sizeof(TX);
            const bool halo = A.halo_max > 0 && lds <= 65536;
#define PSMAS_COPIES_LAUNCH(KC)                                                                                        \
    if (halo)                                                                                                          \
        hipLaunchKernelGGL((k_spmv_ap_copies<BD, KC, true, TV, TX>), dim3(nblocks1), dim3(SPMV_BLOCK), lds, s, A, c0, \
                           cn, npad1, nblocks1, x, y, red, sm);                                                        \
    else                                                                                                               \
        hipLaunchKernelGGL((k_spmv_ap_copies<BD, KC, false, TV, TX>), dim3(nblocks1), dim3(SPMV_BLOCK), 0, s, A, c0, \
                           cn, npad1, nblocks1, x, y, red, sm);
            if (cn <= 2)
            {
                PSMAS_COPIES_LAUNCH(2)
            }
            else
            {
                PSMAS_COPIES_LAUNCH(4)
            }
#undef PSMAS_COPIES_LAUNCH
            PSMAS_LAUNCH_CHECK();
        }
    }
#undef PSMAS_UNROLL_DISPATCH

    namespace
    {
        __global__ __launch_bounds__(256) void k_scatter_values(long long n, const int *__restrict__ map,
                                                                const double *__restrict__ src, double *__restrict__ dst)
        {
            const long long k = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
            if (k < n)
            {
                const int v = map[k];
                if (v >= 0)
                    dst[v] = src[k];
            }
        }

        __global__ __launch_bounds__(256) void k_gather_values(int n, const int *__restrict__ idx,
                                                               const double *__restrict__ src, double *__restrict__ dst)
        {
            const int k = blockIdx.x * 256 + threadIdx.x;
            if (k < n)
                dst[k] = src[idx[k]];
        }

        // g_i = f_i - sum_k a_k f_{col_k} over row i's couplings, summed in list order with
        // separate roundings (no fused multiply-add), as a sequential sparse product does.
        __global__ __launch_bounds__(256) void k_dirichlet_lift(int nrows, const int *__restrict__ crow,
                                                                const int *__restrict__ cptr,
                                                                const int *__restrict__ ccol,
                                                                const double *__restrict__ cval, const double *f,
                                                                double *g)
        {
#pragma clang fp contract(off)
            const int t = blockIdx.x * 256 + threadIdx.x;
            if (t >= nrows)
                return;
            const int i = crow[t];
            double s = 0.0;
            for (int k = cptr[t]; k < cptr[t + 1]; ++k)
            {
                const double prod = cval[k] * f[ccol[k]];
                s = s + prod;
            }
            g[i] = f[i] - s;
        }
    } // namespace

    void launch_gather_values(int n, const int *idx, const double *src, double *dst, hipStream_t s)
    {
        if (n <= 0)
            return;
        require_device_ptrs("gather_values", {idx, src, dst});
        hipLaunchKernelGGL(k_gather_values, dim3(static_cast<unsigned>(div_up(n, 256))), dim3(256), 0, s, n, idx, src,
                           dst);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_dirichlet_lift(int nrows, const int *crow, const int *cptr, const int *ccol, const double *cval,
                               const double *f, double *g, hipStream_t s)
    {
        if (nrows <= 0)
            return;
        hipLaunchKernelGGL(k_dirichlet_lift, dim3(static_cast<unsigned>(div_up(nrows, 256))), dim3(256), 0, s, nrows,
                           crow, cptr, ccol, cval, f, g);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_scatter_values(long long n, const int *map, const double *src, double *dst, hipStream_t s)
    {
        if (n <= 0)
            return;
        require_device_ptrs("scatter_values", {map, src, dst});
        hipLaunchKernelGGL(k_scatter_values, dim3(static_cast<unsigned>(div_up_ll(n, 256))), dim3(256), 0, s, n, map,
                           src, dst);
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename T>
    void launch_update_xr(int nblocks, const T *p, const T *Ap, T *x, T *r, FineRestrict fr, T *rc0,
                          const Reductions &red, const SysMap &sm, hipStream_t s, const LatticeBlocks *lb)
    {
        require_device_ptrs("update_xr", {p, Ap, x, r});
        if (lb != nullptr && lb->cell != nullptr)
            hipLaunchKernelGGL((k_update_xr<BD, T, true>), dim3(nblocks), dim3(VEC_BLOCK), 0, s, p, Ap, x, r, fr, rc0,
                               red, sm, *lb);
        else
            hipLaunchKernelGGL((k_update_xr<BD, T, false>), dim3(nblocks), dim3(VEC_BLOCK), 0, s, p, Ap, x, r, fr,
                               rc0, red, sm, LatticeBlocks{});
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename T>
    void launch_axpy_alpha(int nblocks, const T *p, T *x, const Reductions &red, const SysMap &sm, hipStream_t s)
    {
        require_device_ptrs("axpy_alpha", {p, x});
        hipLaunchKernelGGL((k_axpy_alpha<BD, T>), dim3(nblocks), dim3(VEC_BLOCK), 0, s, p, x, red, sm);
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD, typename T>
    void launch_update_p(int nblocks, const T *z, T *p, const Reductions &red, const SysMap &sm, hipStream_t s,
                         const double *rzo, T *zprev)
    {
        require_device_ptrs("update_p", {z, p});
        hipLaunchKernelGGL((k_update_p<BD, T>), dim3(nblocks), dim3(VEC_BLOCK), 0, s, z, p, red, sm, rzo, zprev);
        PSMAS_LAUNCH_CHECK();
    }

    template <int BD>
    void launch_accumulate_correction(int nblocks, float *x32, double *x64, const SysMap &sm, hipStream_t s)
    {
        require_device_ptrs("accumulate_correction", {x32, x64});
        hipLaunchKernelGGL((k_accumulate_correction<BD>), dim3(nblocks), dim3(VEC_BLOCK), 0, s, x32, x64, sm);
        PSMAS_LAUNCH_CHECK();
    }

    template <typename TS, typename TD>
    void launch_copy_convert(long long n, const TS *src, TD *dst, hipStream_t s)
    {
        if (n <= 0)
            return;
        require_device_ptrs("copy_convert", {src, dst});
        hipLaunchKernelGGL((k_copy_convert<TS, TD>), dim3(stream_grid(n)), dim3(256), 0, s, n, src, dst);
        PSMAS_LAUNCH_CHECK();
    }

    template <typename T>
    void launch_fill(long long n, T *dst, T value, hipStream_t s)
    {
        if (n <= 0)
            return;
        require_device_ptrs("fill", {dst});
        hipLaunchKernelGGL((k_fill<T>), dim3(stream_grid(n)), dim3(256), 0, s, n, dst, value);
        PSMAS_LAUNCH_CHECK();
    }

    template <typename T>
    void launch_gather_in(int npad, int bd, int n, const int *block_of_slot, const T *src, T *dst, hipStream_t s)
    {
        require_device_ptrs("gather_in", {block_of_slot, src, dst});
        hipLaunchKernelGGL((k_gather_in<T>), dim3(stream_grid(static_cast<long long>(npad) * bd)), dim3(256), 0, s,
                           npad, bd, n, block_of_slot, src, dst);
        PSMAS_LAUNCH_CHECK();
    }

    template <typename T>
    void launch_scatter_out(int nb, int bd, int n, const int *slot_of_block, const T *src, T *dst, hipStream_t s)
    {
        require_device_ptrs("scatter_out", {slot_of_block, src, dst});
        hipLaunchKernelGGL((k_scatter_out<T>), dim3(stream_grid(n)), dim3(256), 0, s, nb, bd, n, slot_of_block, src,
                           dst);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_set_scalar(double *sc, int count, int slot, double value, hipStream_t s)
    {
        require_device_ptrs("set_scalar", {sc});
        hipLaunchKernelGGL(k_set_scalar, dim3(div_up(count, 256)), dim3(256), 0, s, sc, static_cast<int>(S_COUNT),
                           count, slot, value);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_snapshot(long long n, const double *x64, const float *x32, double *out, hipStream_t s)
    {
        require_device_ptrs("snapshot", {x64, x32, out});
        hipLaunchKernelGGL(k_snapshot, dim3(stream_grid(n)), dim3(256), 0, s, n, x64, x32, out);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_difference(long long n, const double *a, const double *b, double *out, hipStream_t s)
    {
        require_device_ptrs("difference", {a, b, out});
        hipLaunchKernelGGL(k_difference, dim3(stream_grid(n)), dim3(256), 0, s, n, a, b, out);
        PSMAS_LAUNCH_CHECK();
    }

    void launch_finalize(int nsys, const Reductions &red, const SysMap &sm, hipStream_t s)
    {
        hipLaunchKernelGGL(k_finalize, dim3(nsys), dim3(VEC_BLOCK), 0, s, red, sm);
        PSMAS_LAUNCH_CHECK();
    }

#define PSMAS_INST_BD(BD)                                                                                          \
    template void launch_spmv_sym<BD, double, double>(const SellView<double> &, const SymView<double> &,          \
                                                      const double *, double *, const Reductions &, const SysMap &, \
                                                      hipStream_t);                                               \
    template void launch_spmv_sym<BD, float, float>(const SellView<float> &, const SymView<float> &, const float *, \
                                                    float *, const Reductions &, const SysMap &, hipStream_t);    \
    template size_t spmv_sym_lds<BD, double, double>(int, int);                                                  \
    template void launch_spmv_res_sym<BD, double, double, double>(const SellView<double> &, const SymView<double> &, \
                                                                  const double *, const double *, double *,       \
                                                                  FineRestrict, double *, double *, const SysMap &, \
                                                                  bool, hipStream_t);                             \
    template void launch_spmv_res_sym<BD, double, double, float>(const SellView<double> &, const SymView<double> &, \
                                                                 const double *, const double *, float *,         \
                                                                 FineRestrict, float *, double *, const SysMap &, \
                                                                 bool, hipStream_t);                              \
    template size_t spmv_sym_lds<BD, float, float>(int, int);                                                    \
    template void launch_spmv_ap<BD, double, double>(const SellView<double> &, const double *, double *,          \
                                                     const Reductions &, const SysMap &, hipStream_t);            \
    template void launch_spmv_ap<BD, float, float>(const SellView<float> &, const float *, float *,               \
                                                   const Reductions &, const SysMap &, hipStream_t);              \
    template void launch_spmv_ap_fused<BD, double, double>(const SellView<double> &,                             \
                                                           const FusedDirection<double> &, double *,              \
                                                           const Reductions &, const SysMap &, hipStream_t);      \
    template void launch_spmv_ap_fused<BD, float, float>(const SellView<float> &, const FusedDirection<float> &,  \
                                                         float *, const Reductions &, const SysMap &, hipStream_t); \
    template void launch_spmv_res<BD, double, double, double>(const SellView<double> &, const double *,           \
                                                              const double *, double *, FineRestrict, double *,   \
                                                              double *, const SysMap &, bool, hipStream_t);       \
    template void launch_spmv_res<BD, double, double, float>(const SellView<double> &, const double *,            \
                                                             const double *, float *, FineRestrict, float *,      \
                                                             double *, const SysMap &, bool, hipStream_t);        \
    template void launch_spmv_ap_copies<BD, double, double>(const SellView<double> &, int, int, const double *,   \
                                                            double *, const Reductions &, const SysMap &,         \
                                                            hipStream_t);                                         \
    template void launch_spmv_ap_copies<BD, float, float>(const SellView<float> &, int, int, const float *, float *, \
                                                          const Reductions &, const SysMap &, hipStream_t);       \
    template void launch_spmv_plain<BD, double, double>(const SellView<double> &, const double *, double *,       \
                                                        hipStream_t);                                             \
    template void launch_spmv_plain<BD, float, float>(const SellView<float> &, const float *, float *,            \
                                                      hipStream_t);                                               \
    template void launch_update_xr<BD, double>(int, const double *, const double *, double *, double *,           \
                                               FineRestrict, double *, const Reductions &, const SysMap &,        \
                                               hipStream_t, const LatticeBlocks *);                               \
    template void launch_update_xr<BD, float>(int, const float *, const float *, float *, float *, FineRestrict,  \
                                              float *, const Reductions &, const SysMap &, hipStream_t,           \
                                              const LatticeBlocks *);                                             \
    template void launch_axpy_alpha<BD, double>(int, const double *, double *, const Reductions &, const SysMap &, \
                                                hipStream_t);                                                     \
    template void launch_update_p<BD, double>(int, const double *, double *, const Reductions &, const SysMap &,   \
                                              hipStream_t, const double *, double *);                             \
    template void launch_update_p<BD, float>(int, const float *, float *, const Reductions &, const SysMap &,      \
                                             hipStream_t, const double *, float *);                               \
    template void launch_accumulate_correction<BD>(int, float *, double *, const SysMap &, hipStream_t);            \
    template void launch_spmv_cg<BD>(const SellView<float> &, const CgProlong &, float *, float *, const float *,   \
                                     const float *, const float *, const CgReductions &, const SysMap &, bool,     \
                                     hipStream_t);                                                                 \
    template void launch_update_cg<BD>(int, const float *, const float *, float *, float *, float *, float *,       \
                                       FineRestrict, float *, float *, const CgReductions &, const SysMap &, bool, \
                                       hipStream_t, const LatticeBlocks *);

    PSMAS_INST_BD(1)
    PSMAS_INST_BD(2)
    PSMAS_INST_BD(3)
#undef PSMAS_INST_BD

    template void launch_sym_gather<double>(int, long long, const int *, const double *, double *, hipStream_t);
    template void launch_sym_gather<float>(int, long long, const int *, const double *, float *, hipStream_t);
    template void launch_copy_convert<double, float>(long long, const double *, float *, hipStream_t);
    template void launch_copy_convert<float, double>(long long, const float *, double *, hipStream_t);
    template void launch_copy_convert<double, double>(long long, const double *, double *, hipStream_t);
    template void launch_copy_convert<float, float>(long long, const float *, float *, hipStream_t);
    template void launch_fill<double>(long long, double *, double, hipStream_t);
    template void launch_fill<float>(long long, float *, float, hipStream_t);
    template void launch_fill<int>(long long, int *, int, hipStream_t);
    template void launch_fill<unsigned>(long long, unsigned *, unsigned, hipStream_t);
    template void launch_gather_in<double>(int, int, int, const int *, const double *, double *, hipStream_t);
    template void launch_scatter_out<double>(int, int, int, const int *, const double *, double *, hipStream_t);

    double read_bandwidth_probe(long long bytes, int reps)
    {
        if (bytes < (1LL << 20) || reps < 1)
            throw InvalidArgumentError("[MAS] read probe: at least 1 MiB and one repetition.");
        const long long n4 = bytes / 16;
        DeviceBuffer<float4> buf(static_cast<size_t>(n4), "MAS read probe");
        DeviceBuffer<float> sink(1, "MAS read probe");
        PSMAS_HIP(hipMemset(buf.data(), 0, static_cast<size_t>(n4) * 16));
        struct Events // destroyed on every exit, exceptions included
        {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Events()
            {
                for (hipEvent_t x : e)
                    if (x != nullptr)
                        (void)hipEventDestroy(x);
            }
        } ev;
        PSMAS_HIP(hipEventCreate(&ev.e[0]));
        PSMAS_HIP(hipEventCreate(&ev.e[1]));
        const probe_f4 *a = reinterpret_cast<const probe_f4 *>(buf.data());
        float best = 1e30f;
        for (int r = 0; r <= reps; ++r) // (the first pass warms up)
        {
            PSMAS_HIP(hipEventRecord(ev.e[0], nullptr));
            hipLaunchKernelGGL(k_read_probe, dim3(2048), dim3(256), 0, nullptr, a, n4, sink.data());
            PSMAS_LAUNCH_CHECK();
            PSMAS_HIP(hipEventRecord(ev.e[1], nullptr));
            PSMAS_HIP(hipEventSynchronize(ev.e[1]));
            float ms = 0.0f;
            PSMAS_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
            if (r > 0 && ms < best)
                best = ms;
        }
        return static_cast<double>(n4) * 16.0 / (static_cast<double>(best) * 1e-3) / 1e9;
    }
} // namespace psmas
