// Graph aggregation into <=32-node domains on the GPU (see aggregate.hpp).
//
// Block graph: every scalar entry off the block diagonal gives the directed block edges
// (rb, cb) and (cb, rb) with the quantised weight 1 + |a| / max|a| * 2^20; a radix sort of the
// 64-bit keys (row << bits | column) and integer atomic sums per key run give the graph in CSR
// form, row after row.  Matching: each unmatched domain proposes to its heaviest unmatched
// neighbour (ties: the smaller id) whose size fits beside its own; mutual proposals merge.  A
// round runs four proposal passes, then numbers the merged domains (the smaller id of a pair
// leads), composes the node map and contracts the graph the same way as it was built.  Integer
// weights keep every sum exact, so the result does not depend on summation order.
#include "aggregate.hpp"
#include "common.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace psmas
{
    namespace
    {
        constexpr int TB = 256;
        constexpr int PASSES = 4;     // proposal passes per round
        constexpr int MAX_ROUNDS = 12;
        using u64 = unsigned long long;

        inline unsigned grid_of(long long count)
        {
            return static_cast<unsigned>(count <= 0 ? 1 : (count + TB - 1) / TB);
        }

        inline int bits_for(long long m) // smallest b with 2^b > m
        {
            int b = 1;
            while ((1LL << b) <= m)
                ++b;
            return b;
        }

        __global__ __launch_bounds__(TB) void k_absmax(long long nnz, const double *__restrict__ vals,
                                                       u64 *__restrict__ mx)
        {
            double m = 0.0;
            for (long long k = static_cast<long long>(blockIdx.x) * TB + threadIdx.x; k < nnz;
                 k += static_cast<long long>(gridDim.x) * TB)
                m = fmax(m, fabs(vals[k]));
            // non-negative doubles order like their bit patterns
            atomicMax(mx, static_cast<u64>(__double_as_longlong(m)));
        }

        __global__ __launch_bounds__(TB) void k_graph_entries(int n, int bd, int bits, const int *__restrict__ colptr,
                                                              const int *__restrict__ rowidx,
                                                              const double *__restrict__ vals, double scale,
                                                              u64 sentinel, u64 *__restrict__ keys, u64 *__restrict__ w)
        {
            // A wave per column, lanes over its entries: every read and write of a wave is
            // contiguous (a thread per column wrote each lane's entries a column apart).
            const int c = blockIdx.x * (TB / 64) + static_cast<int>(threadIdx.x) / 64; // wave-uniform
            const int lane = threadIdx.x % 64;
            if (c >= n)
                return;
            const u64 cb = static_cast<u64>(c / bd);
            const int k1 = colptr[c + 1];
            for (int k = colptr[c] + lane; k < k1; k += 64)
            {
                const u64 rb = static_cast<u64>(rowidx[k] / bd);
                const size_t e = 2 * static_cast<size_t>(k);
                if (rb == cb)
                {
                    keys[e] = keys[e + 1] = sentinel;
                    w[e] = w[e + 1] = 0;
                }
                else
                {
                    const u64 q = 1 + static_cast<u64>(llrint(fabs(vals[k]) * scale));
                    keys[e] = (rb << bits) | cb;
                    keys[e + 1] = (cb << bits) | rb;
                    w[e] = w[e + 1] = q;
                }
            }
        }

        __global__ __launch_bounds__(TB) void k_run_flags(long long N, const u64 *__restrict__ keys,
                                                          int *__restrict__ flag)
        {
            const long long k = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (k < N)
                flag[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
        }

        /// Run of entry k = seg[k] - 1: its key (first entry) and the integer sum of its weights.
        __global__ __launch_bounds__(TB) void k_run_sums(long long N, const u64 *__restrict__ keys,
                                                         const u64 *__restrict__ w, const int *__restrict__ flag,
                                                         const int *__restrict__ seg, u64 *__restrict__ ukey,
                                                         u64 *__restrict__ usum)
        {
            const long long k = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (k >= N)
                return;
            const int r = seg[k] - 1;
            if (flag[k])
                ukey[r] = keys[k];
            atomicAdd(&usum[r], w[k]);
        }

        __global__ __launch_bounds__(TB) void k_row_counts(int runs, int bits, const u64 *__restrict__ ukey,
                                                           int *__restrict__ cnt, int *__restrict__ col)
        {
            const int r = blockIdx.x * TB + threadIdx.x;
            if (r >= runs)
                return;
            atomicAdd(&cnt[static_cast<int>(ukey[r] >> bits)], 1);
            col[r] = static_cast<int>(ukey[r] & ((u64(1) << bits) - 1));
        }

        __global__ __launch_bounds__(TB) void k_propose(int m, const int *__restrict__ ptr, const int *__restrict__ col,
                                                        const u64 *__restrict__ w, const int *__restrict__ size,
                                                        const int *__restrict__ match, int *__restrict__ prop,
                                                        int max_size)
        {
            const int u = blockIdx.x * TB + threadIdx.x;
            if (u >= m)
                return;
            int best = -1;
            if (match[u] < 0)
            {
                u64 bw = 0;
                for (int e = ptr[u]; e < ptr[u + 1]; ++e)
                {
                    const int v = col[e];
                    if (match[v] >= 0 || size[u] + size[v] > max_size)
                        continue;
                    if (w[e] > bw || (w[e] == bw && v < best))
                    {
                        bw = w[e];
                        best = v;
                    }
                }
            }
            prop[u] = best;
        }

        __global__ __launch_bounds__(TB) void k_accept(int m, const int *__restrict__ prop, int *__restrict__ match,
                                                       int *__restrict__ merged)
        {
            const int u = blockIdx.x * TB + threadIdx.x;
            if (u >= m || match[u] >= 0)
                return;
            const int v = prop[u];
            if (v >= 0 && prop[v] == u)
            {
                match[u] = v;
                *merged = 1;
            }
        }

        __global__ __launch_bounds__(TB) void k_leaders(int m, const int *__restrict__ match, int *__restrict__ lead)
        {
            const int u = blockIdx.x * TB + threadIdx.x;
            if (u < m)
                lead[u] = (match[u] < 0 || u < match[u]) ? 1 : 0;
        }

        __global__ __launch_bounds__(TB) void k_new_ids(int m, const int *__restrict__ match,
                                                        const int *__restrict__ lead, const int *__restrict__ nid,
                                                        const int *__restrict__ size, int *__restrict__ id,
                                                        int *__restrict__ size2)
        {
            const int u = blockIdx.x * TB + threadIdx.x;
            if (u >= m)
                return;
            const int v = match[u];
            id[u] = lead[u] ? nid[u] : nid[v];
            if (lead[u])
                size2[nid[u]] = size[u] + (v >= 0 ? size[v] : 0);
        }

        __global__ __launch_bounds__(TB) void k_compose(int nb, const int *__restrict__ id, int *__restrict__ agg)
        {
            const int b = blockIdx.x * TB + threadIdx.x;
            if (b < nb)
                agg[b] = id[agg[b]];
        }

        __global__ __launch_bounds__(TB) void k_coarse_entries(int m, const int *__restrict__ ptr,
                                                               const int *__restrict__ col, const u64 *__restrict__ w,
                                                               const int *__restrict__ id, int bits, u64 sentinel,
                                                               u64 *__restrict__ keys, u64 *__restrict__ wout)
        {
            const int u = blockIdx.x * TB + threadIdx.x;
            if (u >= m)
                return;
            const u64 a = static_cast<u64>(id[u]);
            for (int e = ptr[u]; e < ptr[u + 1]; ++e)
            {
                const u64 b = static_cast<u64>(id[col[e]]);
                keys[e] = a != b ? ((a << bits) | b) : sentinel;
                wout[e] = w[e];
            }
        }

        struct Graph
        {
            int m = 0;
            DeviceBuffer<int> ptr, col;
            DeviceBuffer<u64> w;
        };

        /// Graph on m nodes from N (key, weight) entries (row << bits | column; the sentinel
        /// (m << bits | m) marks entries to drop): sort, sum key runs, CSR.
        void build_graph(long long N, DeviceBuffer<u64> &keys, DeviceBuffer<u64> &w, int m, int bits, hipStream_t s,
                         Graph &g)
        {
            g.m = m;
            g.ptr.allocate(static_cast<size_t>(m) + 1, "MAS aggregation graph");
            if (N == 0)
            {
                g.ptr.zero(s);
                g.col.reset();
                g.w.reset();
                return;
            }
            const unsigned end_bit = static_cast<unsigned>(2 * bits);
            DeviceBuffer<u64> ks(N, "MAS aggregation keys"), ws(N, "MAS aggregation weights");
            DeviceBuffer<int> flag(N, "MAS aggregation runs"), seg(N, "MAS aggregation runs");
            DeviceBuffer<int> cnt(static_cast<size_t>(m) + 1, "MAS aggregation graph");
            size_t b_sort = 0, b_scan = 0, b_rows = 0;
            PSMAS_HIP(rocprim::radix_sort_pairs(nullptr, b_sort, keys.data(), ks.data(), w.data(), ws.data(),
                                                static_cast<unsigned>(N), 0u, end_bit, s));
            PSMAS_HIP(rocprim::inclusive_scan(nullptr, b_scan, flag.data(), seg.data(), static_cast<size_t>(N),
                                              rocprim::plus<int>(), s));
            PSMAS_HIP(rocprim::exclusive_scan(nullptr, b_rows, cnt.data(), g.ptr.data(), 0,
                                              static_cast<size_t>(m) + 1, rocprim::plus<int>(), s));
            DeviceBuffer<unsigned char> temp(std::max<size_t>(std::max(b_sort, std::max(b_scan, b_rows)), 1),
                                             "MAS aggregation scratch");
            PSMAS_HIP(rocprim::radix_sort_pairs(temp.data(), b_sort, keys.data(), ks.data(), w.data(), ws.data(),
                                                static_cast<unsigned>(N), 0u, end_bit, s));
            keys.reset();
            w.reset();
            hipLaunchKernelGGL(k_run_flags, dim3(grid_of(N)), dim3(TB), 0, s, N, ks.data(), flag.data());
            PSMAS_LAUNCH_CHECK();
            PSMAS_HIP(rocprim::inclusive_scan(temp.data(), b_scan, flag.data(), seg.data(), static_cast<size_t>(N),
                                              rocprim::plus<int>(), s));
            int runs = 0;
            u64 last_key = 0;
            PSMAS_HIP(hipMemcpyAsync(&runs, seg.data() + (N - 1), sizeof(int), hipMemcpyDeviceToHost, s));
            PSMAS_HIP(hipMemcpyAsync(&last_key, ks.data() + (N - 1), sizeof(u64), hipMemcpyDeviceToHost, s));
            PSMAS_HIP(hipStreamSynchronize(s));
            DeviceBuffer<u64> ukey(runs, "MAS aggregation graph"), usum(runs, "MAS aggregation graph");
            usum.zero(s);
            hipLaunchKernelGGL(k_run_sums, dim3(grid_of(N)), dim3(TB), 0, s, N, ks.data(), ws.data(), flag.data(),
                               seg.data(), ukey.data(), usum.data());
            PSMAS_LAUNCH_CHECK();
            const u64 sentinel = (static_cast<u64>(m) << bits) | static_cast<u64>(m);
            const int edges = runs - (last_key == sentinel ? 1 : 0); // the sentinel run sorts last
            cnt.zero(s);
            g.col.allocate(std::max(edges, 1), "MAS aggregation graph");
            hipLaunchKernelGGL(k_row_counts, dim3(grid_of(edges)), dim3(TB), 0, s, edges, bits, ukey.data(), cnt.data(),
                               g.col.data());
            PSMAS_LAUNCH_CHECK();
            PSMAS_HIP(rocprim::exclusive_scan(temp.data(), b_rows, cnt.data(), g.ptr.data(), 0,
                                              static_cast<size_t>(m) + 1, rocprim::plus<int>(), s));
            g.w = std::move(usum); // the first `edges` runs are the real ones, in row-major order
            PSMAS_HIP(hipStreamSynchronize(s));
        }
    } // namespace

    Partition gpu_aggregate(int n, long long nnz, int bd, const int *d_colptr, const int *d_rowidx,
                            const double *d_vals, int max_size, hipStream_t s)
    {
        const int nb = div_up(n, bd);
        if (2 * nnz > static_cast<long long>(UINT_MAX))
            throw InvalidArgumentError("[MAS] A is too large for the device aggregation.");
        // Block graph.
        DeviceBuffer<u64> mx(1, "MAS aggregation scale");
        mx.zero(s);
        hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(TB), 0, s, nnz, d_vals, mx.data());
        PSMAS_LAUNCH_CHECK();
        u64 mbits = 0;
        PSMAS_HIP(hipMemcpyAsync(&mbits, mx.data(), sizeof(u64), hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        double amax = 0.0;
        std::memcpy(&amax, &mbits, sizeof(double));
        const double scale = amax > 0.0 ? 1048576.0 / amax : 0.0;
        int bits = bits_for(nb);
        Graph g;
        {
            const long long N = 2 * nnz;
            DeviceBuffer<u64> keys(std::max<long long>(N, 1), "MAS aggregation keys"),
                w(std::max<long long>(N, 1), "MAS aggregation weights");
            const u64 sentinel = (static_cast<u64>(nb) << bits) | static_cast<u64>(nb);
            hipLaunchKernelGGL(k_graph_entries, dim3(grid_of(static_cast<long long>(n) * 64)), dim3(TB), 0, s, n, bd,
                               bits, d_colptr, d_rowidx, d_vals, scale, sentinel, keys.data(), w.data());
            PSMAS_LAUNCH_CHECK();
            build_graph(N, keys, w, nb, bits, s, g);
        }

        // Matching rounds.
        DeviceBuffer<int> agg(nb, "MAS aggregation map"), size(nb, "MAS aggregation sizes");
        {
            std::vector<int> ident(nb), ones(nb, 1);
            for (int b = 0; b < nb; ++b)
                ident[b] = b;
            PSMAS_HIP(hipMemcpyAsync(agg.data(), ident.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
            PSMAS_HIP(hipMemcpyAsync(size.data(), ones.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
            PSMAS_HIP(hipStreamSynchronize(s));
        }
        DeviceBuffer<int> merged(1, "MAS aggregation flags");
        for (int round = 0; round < MAX_ROUNDS && g.m > 1; ++round)
        {
            const int m = g.m;
            DeviceBuffer<int> match(m, "MAS aggregation match"), prop(m, "MAS aggregation match");
            PSMAS_HIP(hipMemsetAsync(match.data(), 0xff, sizeof(int) * m, s)); // -1
            merged.zero(s);
            for (int pass = 0; pass < PASSES; ++pass)
            {
                hipLaunchKernelGGL(k_propose, dim3(grid_of(m)), dim3(TB), 0, s, m, g.ptr.data(), g.col.data(), g.w.data(),
                                   size.data(), match.data(), prop.data(), max_size);
                PSMAS_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_accept, dim3(grid_of(m)), dim3(TB), 0, s, m, prop.data(), match.data(),
                                   merged.data());
                PSMAS_LAUNCH_CHECK();
            }
            int any = 0;
            PSMAS_HIP(hipMemcpyAsync(&any, merged.data(), sizeof(int), hipMemcpyDeviceToHost, s));
            PSMAS_HIP(hipStreamSynchronize(s));
            if (!any)
                break;
            DeviceBuffer<int> lead(static_cast<size_t>(m) + 1, "MAS aggregation ids"),
                nid(static_cast<size_t>(m) + 1, "MAS aggregation ids");
            lead.zero(s);
            hipLaunchKernelGGL(k_leaders, dim3(grid_of(m)), dim3(TB), 0, s, m, match.data(), lead.data());
            PSMAS_LAUNCH_CHECK();
            size_t b_scan = 0;
            PSMAS_HIP(rocprim::exclusive_scan(nullptr, b_scan, lead.data(), nid.data(), 0, static_cast<size_t>(m) + 1,
                                              rocprim::plus<int>(), s));
            DeviceBuffer<unsigned char> temp(std::max<size_t>(b_scan, 1), "MAS aggregation scratch");
            PSMAS_HIP(rocprim::exclusive_scan(temp.data(), b_scan, lead.data(), nid.data(), 0,
                                              static_cast<size_t>(m) + 1, rocprim::plus<int>(), s));
            int m2 = 0;
            PSMAS_HIP(hipMemcpyAsync(&m2, nid.data() + m, sizeof(int), hipMemcpyDeviceToHost, s));
            PSMAS_HIP(hipStreamSynchronize(s));
            DeviceBuffer<int> id(m, "MAS aggregation ids"), size2(m2, "MAS aggregation sizes");
            hipLaunchKernelGGL(k_new_ids, dim3(grid_of(m)), dim3(TB), 0, s, m, match.data(), lead.data(), nid.data(),
                               size.data(), id.data(), size2.data());
            PSMAS_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_compose, dim3(grid_of(nb)), dim3(TB), 0, s, nb, id.data(), agg.data());
            PSMAS_LAUNCH_CHECK();
            // Contract the graph.
            int E = 0;
            PSMAS_HIP(hipMemcpyAsync(&E, g.ptr.data() + m, sizeof(int), hipMemcpyDeviceToHost, s));
            PSMAS_HIP(hipStreamSynchronize(s));
            const int bits2 = bits_for(m2);
            DeviceBuffer<u64> keys(std::max(E, 1), "MAS aggregation keys"), w(std::max(E, 1), "MAS aggregation weights");
            const u64 sentinel2 = (static_cast<u64>(m2) << bits2) | static_cast<u64>(m2);
            hipLaunchKernelGGL(k_coarse_entries, dim3(grid_of(m)), dim3(TB), 0, s, m, g.ptr.data(), g.col.data(),
                               g.w.data(), id.data(), bits2, sentinel2, keys.data(), w.data());
            PSMAS_LAUNCH_CHECK();
            Graph g2;
            build_graph(E, keys, w, m2, bits2, s, g2);
            g = std::move(g2);
            size = std::move(size2);
        }

        // Domains: the final ids, small ones packed first-fit (in id order) into shared domains.
        const int m = g.m;
        std::vector<int> h_agg(nb), h_size(m);
        PSMAS_HIP(hipMemcpyAsync(h_agg.data(), agg.data(), sizeof(int) * nb, hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipMemcpyAsync(h_size.data(), size.data(), sizeof(int) * m, hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        std::vector<int> part_of(m);
        int nparts = 0, open = -1, open_size = 0;
        for (int a = 0; a < m; ++a)
        {
            if (h_size[a] > max_size / 2)
            {
                part_of[a] = nparts++;
                continue;
            }
            if (open < 0 || open_size + h_size[a] > max_size)
            {
                open = nparts++;
                open_size = 0;
            }
            part_of[a] = open;
            open_size += h_size[a];
        }
        Partition P;
        P.nparts = nparts;
        P.offsets.assign(static_cast<size_t>(nparts) + 1, 0);
        for (int b = 0; b < nb; ++b)
            P.offsets[part_of[h_agg[b]] + 1]++;
        for (int p = 0; p < nparts; ++p)
            P.offsets[p + 1] += P.offsets[p];
        std::vector<int> cur(P.offsets.begin(), P.offsets.end() - 1);
        P.perm.resize(nb);
        for (int b = 0; b < nb; ++b)
            P.perm[b] = cur[part_of[h_agg[b]]]++;
        return P;
    }
} // namespace psmas
