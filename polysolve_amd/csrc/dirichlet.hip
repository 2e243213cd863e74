// FEM Dirichlet elimination on the GPU.
//
// The reference eliminates on the host with Eigen (FEMSolver.cpp:123-160: A′ = (I-N) A (I-N) + N
// and the lifted right-hand side).  Here one thread per column counts and then writes that
// column of A′ in the order of the host restatement (host_setup.cpp eliminate_dirichlet), so both
// give the same arrays; the couplings (free row, fixed column) are collected per fixed column,
// sorted by (row, column) with one radix sort, and grouped by row with a scan of row changes.
#include "dirichlet.hpp"

#include <climits>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace psmas
{
    namespace
    {
        constexpr int TB = 256;

        inline unsigned grid_for(long long count)
        {
            return static_cast<unsigned>(count <= 0 ? 1 : (count + TB - 1) / TB);
        }

        __global__ __launch_bounds__(TB) void k_mark_fixed(int nfix, const int *__restrict__ nodes,
                                                           unsigned char *__restrict__ fixed)
        {
            const int t = blockIdx.x * TB + threadIdx.x;
            if (t < nfix)
                fixed[nodes[t]] = 1;
        }

        /// Column c: entries of A′ (cnt, at position c + 1 for the scan) and couplings (ccnt).
        __global__ __launch_bounds__(TB) void k_count(int n, const int *__restrict__ colptr,
                                                      const int *__restrict__ rowidx,
                                                      const unsigned char *__restrict__ fixed, int *__restrict__ cnt,
                                                      int *__restrict__ ccnt)
        {
            const int c = blockIdx.x * TB + threadIdx.x;
            if (c >= n)
                return;
            int kept = 0, coupled = 0;
            bool diag = false;
            const bool fc = fixed[c] != 0;
            for (int k = colptr[c]; k < colptr[c + 1]; ++k)
            {
                const int r = rowidx[k];
                if (!fixed[r])
                {
                    ++kept;
                    diag |= (r == c);
                }
            }
            if (fc)
            {
                coupled = kept; // free rows of a fixed column are its couplings
                kept = 1;
            }
            else if (!diag)
                ++kept;
            cnt[c + 1] = kept;
            ccnt[c + 1] = coupled;
        }

        __global__ __launch_bounds__(TB) void k_fill(int n, const int *__restrict__ colptr,
                                                     const int *__restrict__ rowidx, const double *__restrict__ vals,
                                                     const unsigned char *__restrict__ fixed,
                                                     const int *__restrict__ optr, const int *__restrict__ cofs,
                                                     int *__restrict__ orow, double *__restrict__ oval,
                                                     int *__restrict__ keep, uint64_t *__restrict__ ckey,
                                                     int *__restrict__ cidx)
        {
            const int c = blockIdx.x * TB + threadIdx.x;
            if (c >= n)
                return;
            int o = optr[c];
            if (fixed[c])
            {
                orow[o] = c;
                oval[o] = 1.0;
                int q = cofs[c];
                for (int k = colptr[c]; k < colptr[c + 1]; ++k)
                {
                    const int r = rowidx[k];
                    keep[k] = -1;
                    if (!fixed[r])
                    {
                        ckey[q] = (static_cast<uint64_t>(r) << 32) | static_cast<uint32_t>(c);
                        cidx[q++] = k;
                    }
                }
                return;
            }
            // A free column without a diagonal gets an explicit zero one in row order (k_count
            // gave the column one more entry than it keeps).
            int kept = 0;
            for (int k = colptr[c]; k < colptr[c + 1]; ++k)
                kept += fixed[rowidx[k]] ? 0 : 1;
            bool pending = (optr[c + 1] - o) > kept;
            for (int k = colptr[c]; k < colptr[c + 1]; ++k)
            {
                const int r = rowidx[k];
                if (fixed[r])
                {
                    keep[k] = -1;
                    continue;
                }
                if (pending && r > c)
                {
                    orow[o] = c;
                    oval[o++] = 0.0;
                    pending = false;
                }
                orow[o] = r;
                oval[o] = vals[k];
                keep[k] = o++;
            }
            if (pending)
            {
                orow[o] = c;
                oval[o] = 0.0;
            }
        }

        __global__ __launch_bounds__(TB) void k_row_flags(long long m, const uint64_t *__restrict__ key,
                                                          int *__restrict__ flag)
        {
            const long long e = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (e < m)
                flag[e] = (e == 0 || (key[e] >> 32) != (key[e - 1] >> 32)) ? 1 : 0;
        }

        /// Entry e (sorted): its column; the first entry of each row records the row and start.
        __global__ __launch_bounds__(TB) void k_rows(long long m, const uint64_t *__restrict__ key,
                                                     const int *__restrict__ rid, int *__restrict__ crow,
                                                     int *__restrict__ cptr, int *__restrict__ ccol)
        {
            const long long e = static_cast<long long>(blockIdx.x) * TB + threadIdx.x;
            if (e >= m)
                return;
            const uint64_t k = key[e];
            ccol[e] = static_cast<int>(k & 0xffffffffu);
            if (e == 0 || (k >> 32) != (key[e - 1] >> 32))
            {
                const int t = rid[e] - 1;
                crow[t] = static_cast<int>(k >> 32);
                cptr[t] = static_cast<int>(e);
            }
            if (e == m - 1)
                cptr[rid[e]] = static_cast<int>(m);
        }

        /// Inclusive scan of in[0..count) into out[0..count) (rocPRIM, temporary storage sized here).
        void scan_ints(const int *in, int *out, size_t count, hipStream_t s)
        {
            size_t tb = 0;
            PSMAS_HIP(rocprim::inclusive_scan(nullptr, tb, in, out, count, rocprim::plus<int>(), s));
            DeviceBuffer<unsigned char> tmp(std::max<size_t>(tb, 1), "MAS Dirichlet scan");
            PSMAS_HIP(rocprim::inclusive_scan(tmp.data(), tb, in, out, count, rocprim::plus<int>(), s));
        }
    } // namespace

    void eliminate_dirichlet_device(int n, long long nnz, const int *d_colptr, const int *d_rowidx,
                                    const double *d_vals, const int *d_nodes, int nfix, hipStream_t s,
                                    DirichletDevice &out)
    {
        DeviceBuffer<unsigned char> fixed(static_cast<size_t>(n), "MAS Dirichlet mask");
        fixed.zero(s);
        if (nfix > 0)
        {
            hipLaunchKernelGGL(k_mark_fixed, dim3(grid_for(nfix)), dim3(TB), 0, s, nfix, d_nodes, fixed.data());
            PSMAS_LAUNCH_CHECK();
        }
        DeviceBuffer<int> cnt(static_cast<size_t>(n) + 1, "MAS Dirichlet counts");
        DeviceBuffer<int> ccnt(static_cast<size_t>(n) + 1, "MAS Dirichlet counts");
        cnt.zero(s);
        ccnt.zero(s);
        hipLaunchKernelGGL(k_count, dim3(grid_for(n)), dim3(TB), 0, s, n, d_colptr, d_rowidx, fixed.data(), cnt.data(),
                           ccnt.data());
        PSMAS_LAUNCH_CHECK();
        out.colptr.allocate(static_cast<size_t>(n) + 1, "MAS Dirichlet matrix");
        DeviceBuffer<int> cofs(static_cast<size_t>(n) + 1, "MAS Dirichlet couplings");
        scan_ints(cnt.data(), out.colptr.data(), static_cast<size_t>(n) + 1, s);
        scan_ints(ccnt.data(), cofs.data(), static_cast<size_t>(n) + 1, s);
        int tot[2] = {0, 0};
        PSMAS_HIP(hipMemcpyAsync(&tot[0], out.colptr.data() + n, sizeof(int), hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipMemcpyAsync(&tot[1], cofs.data() + n, sizeof(int), hipMemcpyDeviceToHost, s));
        PSMAS_HIP(hipStreamSynchronize(s));
        out.nnz = tot[0];
        out.ncoupl = tot[1];
        const size_t m = static_cast<size_t>(out.ncoupl);
        out.rowidx.allocate(static_cast<size_t>(std::max<long long>(out.nnz, 1)), "MAS Dirichlet matrix");
        out.vals.allocate(static_cast<size_t>(std::max<long long>(out.nnz, 1)), "MAS Dirichlet matrix");
        out.keep.allocate(static_cast<size_t>(std::max<long long>(nnz, 1)), "MAS Dirichlet value map");
        DeviceBuffer<uint64_t> key(std::max<size_t>(m, 1), "MAS Dirichlet couplings");
        DeviceBuffer<int> idx(std::max<size_t>(m, 1), "MAS Dirichlet couplings");
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(TB), 0, s, n, d_colptr, d_rowidx, d_vals, fixed.data(),
                           out.colptr.data(), cofs.data(), out.rowidx.data(), out.vals.data(), out.keep.data(),
                           key.data(), idx.data());
        PSMAS_LAUNCH_CHECK();
        out.ccol.allocate(std::max<size_t>(m, 1), "MAS Dirichlet couplings");
        out.cidx.allocate(std::max<size_t>(m, 1), "MAS Dirichlet couplings");
        out.crow.allocate(std::max<size_t>(m, 1), "MAS Dirichlet couplings");
        out.cptr.allocate(m + 1, "MAS Dirichlet couplings");
        out.nrows = 0;
        if (m == 0)
        {
            out.cptr.zero(s);
            PSMAS_HIP(hipStreamSynchronize(s));
            return;
        }
        DeviceBuffer<uint64_t> key_sorted(m, "MAS Dirichlet couplings");
        size_t tb = 0;
        PSMAS_HIP(rocprim::radix_sort_pairs(nullptr, tb, key.data(), key_sorted.data(), idx.data(), out.cidx.data(), m,
                                            0, 64, s));
        {
            DeviceBuffer<unsigned char> tmp(std::max<size_t>(tb, 1), "MAS Dirichlet sort");
            PSMAS_HIP(rocprim::radix_sort_pairs(tmp.data(), tb, key.data(), key_sorted.data(), idx.data(),
                                                out.cidx.data(), m, 0, 64, s));
        }
        DeviceBuffer<int> flag(m, "MAS Dirichlet couplings"), rid(m, "MAS Dirichlet couplings");
        hipLaunchKernelGGL(k_row_flags, dim3(grid_for(static_cast<long long>(m))), dim3(TB), 0, s,
                           static_cast<long long>(m), key_sorted.data(), flag.data());
        PSMAS_LAUNCH_CHECK();
        scan_ints(flag.data(), rid.data(), m, s);
        PSMAS_HIP(hipMemcpyAsync(&out.nrows, rid.data() + (m - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_rows, dim3(grid_for(static_cast<long long>(m))), dim3(TB), 0, s,
                           static_cast<long long>(m), key_sorted.data(), rid.data(), out.crow.data(), out.cptr.data(),
                           out.ccol.data());
        PSMAS_LAUNCH_CHECK();
        PSMAS_HIP(hipStreamSynchronize(s));
    }
} // namespace psmas
