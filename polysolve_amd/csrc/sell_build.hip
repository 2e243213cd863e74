// Bank-padded SELL-64 matrix built on the GPU from a column-compressed matrix in device memory.
//
// The reference converts its CSC input to BSR with sorts on the GPU after a host round trip of
// the pattern (BSRAdjacency.cu:180-197, BSRMatrix.cu:214-231).  Here one radix sort of block
// keys (block-row slot, block-column slot) over all non-zeros gives the permuted block pattern in
// row-major order directly; a scan over key changes numbers the blocks, row lengths and slice
// lengths follow, and the values are scattered into the SELL-64 layout (one lane per row, slot
// after slot) that the host path (host_setup.cpp build_padded_bsr, Solver::upload_matrix) also
// produces.
#include "sell_build.hpp"

#include <climits>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace psmas
{
    namespace
    {
        constexpr int TB = 256;

        __host__ __device__ inline unsigned grid_for(long long count)
        {
            return static_cast<unsigned>(count <= 0 ? 1 : (count + TB - 1) / TB);
        }

        /// Entry k of column c: key (slot of r's block) << cbits | (slot of c's block), its index k
        /// and its element (r % bd, c % bd).  A wave per column, lanes over its entries, so that
        /// every read and write of a wave is contiguous (a thread per column wrote each lane's
        /// entries a column apart).  The trailing partial block's padded diagonal scalars follow
        /// as entries nnz .. nnz+pad-1 (value 1).
        __global__ __launch_bounds__(TB) void k_entry_keys(int n, int nb, int bd, int cbits, long long nnz,
                                                           const int *__restrict__ colptr,
                                                           const int *__restrict__ rowidx,
                                                           const int *__restrict__ sob, uint64_t *__restrict__ keys,
                                                           int *__restrict__ idx, unsigned char *__restrict__ ecode)
        {
            const int c = blockIdx.x * (TB / 64) + static_cast<int>(threadIdx.x) / 64; // wave-uniform
            const int lane = threadIdx.x % 64;
            if (c < n)
            {
                const uint64_t C = static_cast<uint64_t>(sob[c / bd]);
                const int cm = c % bd;
                const int k1 = colptr[c + 1];
                for (int k = colptr[c] + lane; k < k1; k += 64)
                {
                    const int r = rowidx[k];
                    keys[k] = (static_cast<uint64_t>(sob[r / bd]) << cbits) | C;
                    idx[k] = k;
                    ecode[k] = static_cast<unsigned char>((r % bd) * bd + cm);
                }
            }
            const int pad = nb * bd - n;
            if (blockIdx.x == 0 && static_cast<int>(threadIdx.x) < pad)
            {
                const int t = threadIdx.x;
                const long long k = nnz + t;
                const uint64_t T = static_cast<uint64_t>(sob[nb - 1]);
                const int local = (bd - pad) + t;
                keys[k] = (T << cbits) | T;
                idx[k] = static_cast<int>(k);
                ecode[k] = static_cast<unsigned char>(local * bd + local);
            }
        }

        __global__ __launch_bounds__(TB) void k_block_flags(long long N, const uint64_t *__restrict__ keys,
                                                            int *__restrict__ flag)
        {
            const long long k = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (k < N)
                flag[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
        }

        /// Per block (first entry of a key run; blk = inclusive scan of the run starts, so the
        /// block of entry k is blk[k] - 1): its row and column slots; row lengths (atomic counts);
        /// rows with a block off the diagonal.
        __global__ __launch_bounds__(TB) void k_blocks(long long N, int cbits, const uint64_t *__restrict__ keys,
                                                       const int *__restrict__ flag, const int *__restrict__ blk,
                                                       int *__restrict__ brow, int *__restrict__ bcol,
                                                       int *__restrict__ row_len, unsigned char *__restrict__ nonisol)
        {
            const long long k = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (k >= N || !flag[k])
                return;
            const int b = blk[k] - 1;
            const int R = static_cast<int>(keys[k] >> cbits);
            const int C = static_cast<int>(keys[k] & ((uint64_t(1) << cbits) - 1));
            brow[b] = R;
            bcol[b] = C;
            atomicAdd(&row_len[R], 1);
            if (R != C)
                nonisol[R] = 1;
        }

        /// Slots of slice s: the longest row among its 64, at least 1.
        __global__ __launch_bounds__(TB) void k_slice_len(int nslices, const int *__restrict__ row_len,
                                                          int *__restrict__ slen)
        {
            const int s = blockIdx.x * TB + threadIdx.x;
            if (s >= nslices)
                return;
            int m = 1;
            for (int l = 0; l < 64; ++l)
                m = max(m, row_len[s * 64 + l]);
            slen[s] = m;
        }

        /// Padding: every slot of row (s, lane) first holds the row's own column.
        __global__ __launch_bounds__(TB) void k_sell_pad(int nslices, const int *__restrict__ slice_ptr,
                                                         int *__restrict__ cols)
        {
            const long long t = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (t >= static_cast<long long>(nslices) * 64)
                return;
            const int s = static_cast<int>(t / 64), lane = static_cast<int>(t % 64);
            for (int slot = slice_ptr[s]; slot < slice_ptr[s + 1]; ++slot)
                cols[static_cast<size_t>(slot) * 64 + lane] = s * 64 + lane;
        }

        __device__ __forceinline__ size_t slot_of(int R, int b, const int *row_ptr, const int *slice_ptr)
        {
            return static_cast<size_t>(slice_ptr[R / 64]) + static_cast<size_t>(b - row_ptr[R]);
        }

        __global__ __launch_bounds__(TB) void k_sell_cols(int nnzb, const int *__restrict__ brow,
                                                          const int *__restrict__ bcol, const int *__restrict__ row_ptr,
                                                          const int *__restrict__ slice_ptr, int *__restrict__ cols)
        {
            const int b = blockIdx.x * TB + threadIdx.x;
            if (b >= nnzb)
                return;
            const int R = brow[b];
            cols[slot_of(R, b, row_ptr, slice_ptr) * 64 + R % 64] = bcol[b];
        }

        /// Scatter (sum) every entry into its SELL value; record its index; flag duplicate
        /// entries (an earlier entry of the same block with the same element, or more than bs
        /// entries in one block).
        __global__ __launch_bounds__(TB) void k_sell_vals(long long N, long long nnz, int bs, int cbits,
                                                          const uint64_t *__restrict__ keys,
                                                          const int *__restrict__ idx, const int *__restrict__ blk,
                                                          const int *__restrict__ row_ptr,
                                                          const int *__restrict__ slice_ptr,
                                                          const unsigned char *__restrict__ ecode,
                                                          const double *__restrict__ vals, double *sell_vals,
                                                          int *__restrict__ map, int *__restrict__ dup)
        {
            const long long k = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (k >= N)
                return;
            const uint64_t key = keys[k];
            const int R = static_cast<int>(key >> cbits);
            const int i = idx[k];
            const int e = ecode[i];
            const size_t vi = (slot_of(R, blk[k] - 1, row_ptr, slice_ptr) * bs + e) * 64 + R % 64;
            atomicAdd(&sell_vals[vi], i < nnz ? vals[i] : 1.0);
            if (map != nullptr && i < nnz)
                map[i] = static_cast<int>(vi);
            bool d = (k >= bs && keys[k - bs] == key);
            for (long long j = k - 1; !d && j >= 0 && j >= k - bs && keys[j] == key; --j)
                d = (ecode[idx[j]] == e);
            if (d)
                *dup = 1;
        }

        __global__ __launch_bounds__(TB) void k_compose_map(long long n, const int *__restrict__ keep,
                                                            const int *__restrict__ map, int *__restrict__ out)
        {
            const long long k = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (k < n)
                out[k] = keep[k] >= 0 ? map[keep[k]] : -1;
        }

        __global__ __launch_bounds__(TB) void k_shift_cols(long long len, int off, const int *__restrict__ cols,
                                                           int *__restrict__ out)
        {
            const long long i = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (i < len)
                out[i] = cols[i] + off;
        }

        /// Flags: 1 colptr[0] != 0 or colptr[n] != nnz; 2 a decreasing column pointer; 4 a row
        /// index out of [0, n).
        __global__ __launch_bounds__(TB) void k_validate(int n, long long nnz, const int *__restrict__ colptr,
                                                         const int *__restrict__ rowidx, int *__restrict__ bad)
        {
            const long long t = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (t == 0 && (colptr[0] != 0 || colptr[n] != nnz))
                atomicOr(bad, 1);
            if (t < n && colptr[t + 1] < colptr[t])
                atomicOr(bad, 2);
            if (t < nnz && (rowidx[t] < 0 || rowidx[t] >= n))
                atomicOr(bad, 4);
        }
    } // namespace

    void validate_csc_device(int n, long long nnz, const int *d_colptr, const int *d_rowidx, hipStream_t s)
    {
        if (n <= 0 || nnz < 0 || d_colptr == nullptr || (nnz > 0 && d_rowidx == nullptr))
            throw InvalidArgumentError("[MAS] invalid A: empty matrix or null arrays.");
        if (nnz > INT_MAX)
            throw InvalidArgumentError("[MAS] A is too large. Non-zero number exceeding int32 max.");
        DeviceBuffer<int> bad(1, "MAS input check");
        bad.zero(s);
        hipLaunchKernelGGL(k_validate, dim3(grid_for(std::max<long long>(nnz, n))), dim3(TB), 0, s, n, nnz, d_colptr,
                           d_rowidx, bad.data());
        PSMAS_LAUNCH_CHECK();
        int h = 0;
        PSMAS_HIP(hipMemcpyAsync(&h, bad.data(), sizeof(int), hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        if (h & 1)
            throw InvalidArgumentError("[MAS] invalid A: column pointers do not span the non-zeros.");
        if (h & 2)
            throw InvalidArgumentError("[MAS] invalid A: column pointers are not monotone.");
        if (h & 4)
            throw InvalidArgumentError("[MAS] invalid A: row index out of range.");
    }

    void build_sell_device(int n, long long nnz, int bd, const int *d_colptr, const int *d_rowidx,
                           const double *d_vals, const int *d_slot_of_block, int npad, bool want_map,
                           hipStream_t s, SellDevice &out)
    {
        const int nb = div_up(n, bd), bs = bd * bd, pad = nb * bd - n;
        const long long N = nnz + pad;
        if (N > INT_MAX || npad <= 0 || npad % 64 != 0)
            throw InvalidArgumentError("[MAS] A is too large for the device matrix build.");
        const int nslices = npad / 64;
        int cbits = 1;
        while ((1LL << cbits) < npad)
            ++cbits;
        const unsigned end_bit = static_cast<unsigned>(2 * cbits);

        DeviceBuffer<uint64_t> keys(N, "MAS matrix build keys"), keys_sorted(N, "MAS matrix build keys");
        DeviceBuffer<int> idx(N, "MAS matrix build order"), idx_sorted(N, "MAS matrix build order");
        DeviceBuffer<unsigned char> ecode(N, "MAS matrix build elements");
        hipLaunchKernelGGL(k_entry_keys, dim3(grid_for(static_cast<long long>(n) * 64)), dim3(TB), 0, s, n, nb, bd,
                           cbits, nnz, d_colptr, d_rowidx, d_slot_of_block, keys.data(), idx.data(), ecode.data());
        PSMAS_LAUNCH_CHECK();

        DeviceBuffer<int> flag(N, "MAS matrix build flags"), blk(N, "MAS matrix build blocks");
        DeviceBuffer<int> row_len(static_cast<size_t>(npad) + 1, "MAS SELL row lengths");
        DeviceBuffer<int> row_ptr(static_cast<size_t>(npad) + 1, "MAS matrix build rows");
        DeviceBuffer<int> slen(static_cast<size_t>(nslices) + 1, "MAS matrix build slices");
        out.d_slice_ptr.allocate(static_cast<size_t>(nslices) + 1, "MAS SELL slice pointers");
        size_t b_sort = 0, b_scan_n = 0, b_scan_r = 0, b_scan_s = 0;
        PSMAS_HIP(rocprim::radix_sort_pairs(nullptr, b_sort, keys.data(), keys_sorted.data(), idx.data(),
                                            idx_sorted.data(), static_cast<unsigned>(N), 0u, end_bit, s));
        PSMAS_HIP(rocprim::inclusive_scan(nullptr, b_scan_n, flag.data(), blk.data(), static_cast<size_t>(N),
                                          rocprim::plus<int>(), s));
        PSMAS_HIP(rocprim::exclusive_scan(nullptr, b_scan_r, row_len.data(), row_ptr.data(), 0,
                                          static_cast<size_t>(npad) + 1, rocprim::plus<int>(), s));
        PSMAS_HIP(rocprim::exclusive_scan(nullptr, b_scan_s, slen.data(), out.d_slice_ptr.data(), 0,
                                          static_cast<size_t>(nslices) + 1, rocprim::plus<int>(), s));
        const size_t temp_bytes = std::max(std::max(b_sort, b_scan_n), std::max(b_scan_r, b_scan_s));
        DeviceBuffer<unsigned char> temp(std::max<size_t>(temp_bytes, 1), "MAS matrix build scratch");

        PSMAS_HIP(rocprim::radix_sort_pairs(temp.data(), b_sort, keys.data(), keys_sorted.data(), idx.data(),
                                            idx_sorted.data(), static_cast<unsigned>(N), 0u, end_bit, s));
        keys.reset();
        idx.reset();
        hipLaunchKernelGGL(k_block_flags, dim3(grid_for(N)), dim3(TB), 0, s, N, keys_sorted.data(), flag.data());
        PSMAS_LAUNCH_CHECK();
        PSMAS_HIP(rocprim::inclusive_scan(temp.data(), b_scan_n, flag.data(), blk.data(), static_cast<size_t>(N),
                                          rocprim::plus<int>(), s));
        int nnzb = 0;
        PSMAS_HIP(hipMemcpyAsync(&nnzb, blk.data() + (N - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        out.nnzb = nnzb;

        DeviceBuffer<int> brow(nnzb, "MAS matrix build blocks"), bcol(nnzb, "MAS matrix build blocks");
        DeviceBuffer<unsigned char> nonisol(npad, "MAS matrix build rows");
        row_len.zero(s);
        nonisol.zero(s);
        hipLaunchKernelGGL(k_blocks, dim3(grid_for(N)), dim3(TB), 0, s, N, cbits, keys_sorted.data(), flag.data(),
                           blk.data(), brow.data(), bcol.data(), row_len.data(), nonisol.data());
        PSMAS_LAUNCH_CHECK();
        flag.reset();
        PSMAS_HIP(rocprim::exclusive_scan(temp.data(), b_scan_r, row_len.data(), row_ptr.data(), 0,
                                          static_cast<size_t>(npad) + 1, rocprim::plus<int>(), s));
        slen.zero(s);
        hipLaunchKernelGGL(k_slice_len, dim3(grid_for(nslices)), dim3(TB), 0, s, nslices, row_len.data(), slen.data());
        PSMAS_LAUNCH_CHECK();
        PSMAS_HIP(rocprim::exclusive_scan(temp.data(), b_scan_s, slen.data(), out.d_slice_ptr.data(), 0,
                                          static_cast<size_t>(nslices) + 1, rocprim::plus<int>(), s));
        out.slice_ptr.resize(static_cast<size_t>(nslices) + 1);
        PSMAS_HIP(hipMemcpyAsync(out.slice_ptr.data(), out.d_slice_ptr.data(), sizeof(int) * out.slice_ptr.size(),
                                 hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        out.nslots = out.slice_ptr[nslices];

        const size_t nvals = static_cast<size_t>(out.nslots) * 64 * bs;
        out.d_cols.allocate(static_cast<size_t>(out.nslots) * 64, "MAS SELL columns");
        out.d_vals.allocate(nvals, "MAS SELL values (FP64)");
        out.d_vals.zero(s);
        hipLaunchKernelGGL(k_sell_pad, dim3(grid_for(static_cast<long long>(nslices) * 64)), dim3(TB), 0, s, nslices,
                           out.d_slice_ptr.data(), out.d_cols.data());
        PSMAS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sell_cols, dim3(grid_for(nnzb)), dim3(TB), 0, s, nnzb, brow.data(), bcol.data(),
                           row_ptr.data(), out.d_slice_ptr.data(), out.d_cols.data());
        PSMAS_LAUNCH_CHECK();
        const bool map = want_map && nvals <= static_cast<size_t>(INT_MAX) && nnz > 0;
        out.d_map.reset();
        if (map)
            out.d_map.allocate(static_cast<size_t>(nnz), "MAS value map");
        DeviceBuffer<int> dup(1, "MAS matrix build flags");
        dup.zero(s);
        hipLaunchKernelGGL(k_sell_vals, dim3(grid_for(N)), dim3(TB), 0, s, N, nnz, bs, cbits, keys_sorted.data(),
                           idx_sorted.data(), blk.data(), row_ptr.data(), out.d_slice_ptr.data(), ecode.data(), d_vals,
                           out.d_vals.data(), map ? out.d_map.data() : nullptr, dup.data());
        PSMAS_LAUNCH_CHECK();
        int h_dup = 0;
        std::vector<unsigned char> h_nonisol(npad);
        PSMAS_HIP(hipMemcpyAsync(&h_dup, dup.data(), sizeof(int), hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipMemcpyAsync(h_nonisol.data(), nonisol.data(), npad, hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        out.nonisolated.assign(h_nonisol.begin(), h_nonisol.end());
        out.map_ok = map && h_dup == 0;
        if (!out.map_ok)
            out.d_map.reset();
        out.d_row_len = std::move(row_len);
    }

    DeviceBuffer<int> compose_value_map(const DeviceBuffer<int> &keep, size_t n, const DeviceBuffer<int> &map,
                                        hipStream_t s)
    {
        DeviceBuffer<int> out(std::max<size_t>(n, 1), "MAS value map");
        if (n > 0)
        {
            hipLaunchKernelGGL(k_compose_map, dim3(grid_for(static_cast<long long>(n))), dim3(TB), 0, s,
                               static_cast<long long>(n), keep.data(), map.data(), out.data());
            PSMAS_LAUNCH_CHECK();
        }
        PSMAS_HIP(hipStreamSynchronize(s));
        return out;
    }

    DeviceBuffer<int> compose_value_map(const std::vector<int> &keep, const DeviceBuffer<int> &map, hipStream_t s)
    {
        const size_t n = keep.size();
        DeviceBuffer<int> d_keep(std::max<size_t>(n, 1), "MAS Dirichlet value map");
        DeviceBuffer<int> out(std::max<size_t>(n, 1), "MAS value map");
        if (n > 0)
        {
            PSMAS_HIP(hipMemcpyAsync(d_keep.data(), keep.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_compose_map, dim3(grid_for(static_cast<long long>(n))), dim3(TB), 0, s,
                               static_cast<long long>(n), d_keep.data(), map.data(), out.data());
            PSMAS_LAUNCH_CHECK();
        }
        PSMAS_HIP(hipStreamSynchronize(s));
        return out;
    }

    void concat_sell(std::vector<SellDevice> &parts, const std::vector<int> &npads, hipStream_t s, SellDevice &out)
    {
        long long slots = 0, blocks = 0, rows = 0;
        size_t nvals = 0;
        for (size_t p = 0; p < parts.size(); ++p)
        {
            slots += parts[p].nslots;
            blocks += parts[p].nnzb;
            rows += npads[p];
            nvals += parts[p].d_vals.size();
        }
        if (slots > INT_MAX || rows > INT_MAX)
            throw InvalidArgumentError("[MAS] the batch is too large for 32-bit slot indices.");
        out = SellDevice();
        out.nnzb = blocks;
        out.nslots = slots;
        out.slice_ptr.assign(1, 0);
        out.d_slice_ptr.allocate(static_cast<size_t>(rows / 64 + 1), "MAS SELL slice pointers");
        out.d_cols.allocate(static_cast<size_t>(slots) * 64, "MAS SELL columns");
        out.d_vals.allocate(std::max<size_t>(nvals, 1), "MAS SELL values (FP64)");
        out.d_row_len.allocate(static_cast<size_t>(rows), "MAS SELL row lengths");
        long long slot_off = 0;
        size_t val_off = 0;
        int row_off = 0;
        for (size_t p = 0; p < parts.size(); ++p)
        {
            SellDevice &m = parts[p];
            const int nslices = npads[p] / 64;
            for (int k = 1; k <= nslices; ++k)
                out.slice_ptr.push_back(m.slice_ptr[k] + static_cast<int>(slot_off));
            out.nonisolated.insert(out.nonisolated.end(), m.nonisolated.begin(), m.nonisolated.end());
            const long long ncols = m.nslots * 64;
            hipLaunchKernelGGL(k_shift_cols, dim3(grid_for(ncols)), dim3(TB), 0, s, ncols, row_off, m.d_cols.data(),
                               out.d_cols.data() + slot_off * 64);
            PSMAS_LAUNCH_CHECK();
            PSMAS_HIP(hipMemcpyAsync(out.d_vals.data() + val_off, m.d_vals.data(), m.d_vals.bytes(),
                                     hipMemcpyDeviceToDevice, s));
            PSMAS_HIP(hipMemcpyAsync(out.d_row_len.data() + row_off, m.d_row_len.data(), sizeof(int) * npads[p],
                                     hipMemcpyDeviceToDevice, s));
            slot_off += m.nslots;
            val_off += m.d_vals.size();
            row_off += npads[p];
        }
        PSMAS_HIP(hipMemcpyAsync(out.d_slice_ptr.data(), out.slice_ptr.data(), out.d_slice_ptr.bytes(),
                                 hipMemcpyHostToDevice, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        parts.clear();
    }
} // namespace psmas
