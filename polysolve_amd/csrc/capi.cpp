// extern "C" layer: maps the C ABI of include/psmas.h onto psmas::Solver and turns exceptions
// into status codes (reference exception classes: CudaUtils.cuh:33-80, MASSolver.cu).
#include "../../include/psmas.h"

#include "host_setup.hpp"
#include "kernels.hpp"
#include "solver.hpp"

#include <algorithm>
#include <cstring>
#include <new>
#include <sstream>
#include <string>

struct psmas_solver
{
    psmas::Solver solver;
    std::string last_error;
};

namespace
{
    thread_local std::string g_create_error;

    template <typename F>
    int guarded(std::string &err, F &&f)
    {
        try
        {
            f();
            err.clear();
            return PSMAS_OK;
        }
        catch (const psmas::OutOfMemoryError &e)
        {
            err = e.what();
            return PSMAS_ERROR_OUT_OF_MEMORY;
        }
        catch (const psmas::NoDeviceError &e)
        {
            err = e.what();
            return PSMAS_ERROR_NO_DEVICE;
        }
        catch (const psmas::InvalidArgumentError &e)
        {
            err = e.what();
            return PSMAS_ERROR_INVALID_ARGUMENT;
        }
        catch (const std::bad_alloc &)
        {
            err = "[MAS] host allocation failed.";
            return PSMAS_ERROR_OUT_OF_MEMORY;
        }
        catch (const std::exception &e)
        {
            err = e.what();
            return PSMAS_ERROR_RUNTIME;
        }
        catch (...)
        {
            err = "[MAS] unknown error.";
            return PSMAS_ERROR_RUNTIME;
        }
    }

    int write_string(const std::string &s, char *buf, size_t cap, size_t *needed)
    {
        if (needed)
            *needed = s.size() + 1;
        if (buf == nullptr || cap == 0)
            return PSMAS_ERROR_BUFFER_TOO_SMALL;
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
        return (s.size() + 1 > cap) ? PSMAS_ERROR_BUFFER_TOO_SMALL : PSMAS_OK;
    }

    int null_handle() { return PSMAS_ERROR_INVALID_ARGUMENT; }
} // namespace

extern "C"
{
    int psmas_create(psmas_handle *out)
    {
        if (out == nullptr)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        *out = nullptr;
        return guarded(g_create_error, [&] { *out = new psmas_solver(); });
    }

    int psmas_destroy(psmas_handle h)
    {
        delete h;
        return PSMAS_OK;
    }

    const char *psmas_name(void) { return "MAS"; }

    int psmas_set_parameters(psmas_handle h, const char *json)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            if (json == nullptr)
                throw psmas::InvalidArgumentError("[MAS] null parameter string.");
            h->solver.set_parameters_json(json);
        });
    }

    int psmas_get_parameters(psmas_handle h, char *buf, size_t cap, size_t *needed)
    {
        if (!h)
            return null_handle();
        return write_string(h->solver.parameters_json(), buf, cap, needed);
    }

    int psmas_get_info(psmas_handle h, char *buf, size_t cap, size_t *needed)
    {
        if (!h)
            return null_handle();
        return write_string(h->solver.get_info_json(), buf, cap, needed);
    }

    int psmas_analyze_pattern(psmas_handle h, int, const int *, const int *, int)
    {
        if (!h)
            return null_handle();
        return PSMAS_OK;
    }

    int psmas_factorize_csc(psmas_handle h, int n, long long nnz, const int *colptr, const int *rowidx,
                            const double *vals)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            psmas::CscView A;
            A.n = n;
            A.nnz = nnz;
            A.colptr = colptr;
            A.rowidx = rowidx;
            A.vals = vals;
            h->solver.factorize(A);
        });
    }

    int psmas_solve(psmas_handle h, int n, const double *b, double *x)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.solve_host(n, b, x); });
    }

    int psmas_solve_device(psmas_handle h, int n, const double *d_b, double *d_x)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.solve_device(n, d_b, d_x); });
    }

    int psmas_solve_device_noalloc(psmas_handle h, int n, const double *d_b, double *d_x)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.solve_device_noalloc(n, d_b, d_x); });
    }

    int psmas_dirichlet_factorize_csc(psmas_handle h, int n, long long nnz, const int *colptr, const int *rowidx,
                                      const double *vals, int ndir, const int *dirichlet_nodes)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            psmas::CscView A;
            A.n = n;
            A.nnz = nnz;
            A.colptr = colptr;
            A.rowidx = rowidx;
            A.vals = vals;
            h->solver.dirichlet_factorize(A, dirichlet_nodes, ndir);
        });
    }

    int psmas_dirichlet_factorize_csc_device(psmas_handle h, int n, long long nnz, const int *d_colptr,
                                             const int *d_rowidx, const double *d_vals, int ndir,
                                             const int *dirichlet_nodes)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            h->solver.dirichlet_factorize_device_input(n, nnz, d_colptr, d_rowidx, d_vals, dirichlet_nodes, ndir);
        });
    }

    int psmas_dirichlet_solve(psmas_handle h, int n, double *f, double *x)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.dirichlet_solve_host(n, f, x); });
    }

    int psmas_dirichlet_solve_device(psmas_handle h, int n, double *d_f, double *d_x)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.dirichlet_solve_device(n, d_f, d_x); });
    }

    int psmas_factorize_csc_device(psmas_handle h, int n, long long nnz, const int *d_colptr, const int *d_rowidx,
                                   const double *d_vals)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.factorize_device(n, nnz, d_colptr, d_rowidx, d_vals); });
    }

    int psmas_factorize_multi_csc(psmas_handle h, int nrhs, int n, long long nnz, const int *colptr, const int *rowidx,
                                  const double *vals)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            psmas::CscView A;
            A.n = n;
            A.nnz = nnz;
            A.colptr = colptr;
            A.rowidx = rowidx;
            A.vals = vals;
            h->solver.factorize_copies(A, nrhs);
        });
    }

    int psmas_factorize_multi_csc_device(psmas_handle h, int nrhs, int n, long long nnz, const int *d_colptr,
                                         const int *d_rowidx, const double *d_vals)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            const psmas::CscView A{n, nnz, nullptr, nullptr, nullptr};
            const psmas::DeviceCsc dv{n, nnz, d_colptr, d_rowidx, d_vals};
            h->solver.factorize_copies(A, nrhs, &dv);
        });
    }

    static int solve_multi(psmas_handle h, int n, int nrhs, const double *B, double *X, bool device)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            if (nrhs != h->solver.num_systems() || B == nullptr || X == nullptr)
                throw psmas::InvalidArgumentError(
                    "[MAS] solve_multi: nrhs must equal the right-hand sides of the last factorize_multi.");
            std::vector<const double *> b(nrhs);
            std::vector<double *> x(nrhs);
            for (int c = 0; c < nrhs; ++c)
            {
                b[c] = B + static_cast<size_t>(c) * n;
                x[c] = X + static_cast<size_t>(c) * n;
            }
            if (n != h->solver.dim())
                throw psmas::InvalidArgumentError("[MAS] Size mismatch. Did you forget to call factorize?");
            h->solver.solve_batch(b, x, device);
        });
    }

    int psmas_solve_multi(psmas_handle h, int n, int nrhs, const double *B, double *X)
    {
        return solve_multi(h, n, nrhs, B, X, false);
    }

    int psmas_solve_multi_device(psmas_handle h, int n, int nrhs, const double *d_B, double *d_X)
    {
        return solve_multi(h, n, nrhs, d_B, d_X, true);
    }

    int psmas_factorize_values_device(psmas_handle h, long long nnz, const double *d_vals)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.refresh_values_device(nnz, d_vals); });
    }

    const char *psmas_last_error(psmas_handle h)
    {
        return h ? h->last_error.c_str() : g_create_error.c_str();
    }

    int psmas_set_coordinates(psmas_handle h, int nb, int dim, const double *xyz)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.set_coordinates(0, nb, dim, xyz); });
    }

    int psmas_batch_set_coordinates(psmas_handle h, int sys, int nb, int dim, const double *xyz)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.set_coordinates(sys, nb, dim, xyz); });
    }

    int psmas_batch_factorize_csc(psmas_handle h, int nsys, const int *n, const long long *nnz,
                                  const int *const *colptr, const int *const *rowidx, const double *const *vals)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            if (nsys <= 0 || !n || !nnz || !colptr || !rowidx || !vals)
                throw psmas::InvalidArgumentError("[MAS] invalid batch arguments.");
            std::vector<psmas::CscView> As(nsys);
            for (int s = 0; s < nsys; ++s)
            {
                As[s].n = n[s];
                As[s].nnz = nnz[s];
                As[s].colptr = colptr[s];
                As[s].rowidx = rowidx[s];
                As[s].vals = vals[s];
            }
            h->solver.factorize_batch(As);
        });
    }

    int psmas_batch_solve(psmas_handle h, int nsys, const double *const *b, double *const *x, int device)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            if (nsys != h->solver.num_systems() || !b || !x)
                throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
            std::vector<const double *> bv(b, b + nsys);
            std::vector<double *> xv(x, x + nsys);
            h->solver.solve_batch(bv, xv, device != 0);
        });
    }

    int psmas_batch_get_info(psmas_handle h, int sys, char *buf, size_t cap, size_t *needed)
    {
        if (!h)
            return null_handle();
        std::string s;
        const int rc = guarded(h->last_error, [&] { s = h->solver.system_info_json(sys); });
        if (rc != PSMAS_OK)
            return rc;
        return write_string(s, buf, cap, needed);
    }

    int psmas_device_count(void)
    {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess)
        {
            (void)hipGetLastError();
            return 0;
        }
        return c;
    }

    int psmas_device_alloc(size_t bytes, void **ptr)
    {
        if (!ptr)
            return PSMAS_ERROR_INVALID_ARGUMENT;
        *ptr = nullptr;
        return guarded(g_create_error, [&] { *ptr = psmas::device_alloc(bytes, "caller buffer"); });
    }

    int psmas_device_free(void *ptr)
    {
        psmas::device_free(ptr);
        return PSMAS_OK;
    }

    int psmas_memcpy(void *dst, const void *src, size_t bytes, int kind)
    {
        return guarded(g_create_error, [&] {
            const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                              : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
            PSMAS_HIP(hipMemcpy(dst, src, bytes, k));
        });
    }

    int psmas_device_synchronize(void)
    {
        return guarded(g_create_error, [&] { PSMAS_HIP(hipDeviceSynchronize()); });
    }

    int psmas_read_bandwidth(long long bytes, int reps, double *gbps)
    {
        return guarded(g_create_error, [&] {
            if (gbps == nullptr)
                throw psmas::InvalidArgumentError("[MAS] null output.");
            *gbps = psmas::read_bandwidth_probe(bytes, reps);
        });
    }

    int psmas_parse_parameters(const char *json, char *buf, size_t cap, size_t *needed)
    {
        std::string err;
        psmas::Params p;
        const int rc = guarded(err, [&] {
            if (json == nullptr)
                throw psmas::InvalidArgumentError("[MAS] null parameter string.");
            psmas::apply_json_parameters(p, json);
        });
        if (rc != PSMAS_OK)
        {
            write_string(err, buf, cap, needed);
            return rc;
        }
        return write_string(psmas::params_to_json(p), buf, cap, needed);
    }

    int psmas_block_graph(int n, long long nnz, const int *colptr, const int *rowidx, const double *vals, int bd,
                          int *ptr, int ptr_cap, int *adj, double *w, long long edge_cap, long long *nedges)
    {
        bool fits = true;
        const int rc = guarded(g_create_error, [&] { // (handle-less: psmas_last_error(NULL) reports it)
            if (bd < 1 || bd > 3)
                throw psmas::InvalidArgumentError("[MAS] invalid block size (1, 2 or 3).");
            if (ptr == nullptr || nedges == nullptr)
                throw psmas::InvalidArgumentError("[MAS] null output array.");
            const psmas::CscView A{n, nnz, colptr, rowidx, vals};
            psmas::validate_csc(A);
            const psmas::BlockGraph g = psmas::build_block_graph(A, bd);
            if (ptr_cap < g.nb + 1)
                throw psmas::InvalidArgumentError("[MAS] block graph pointer buffer too small.");
            std::copy(g.ptr.begin(), g.ptr.end(), ptr);
            *nedges = static_cast<long long>(g.adj.size());
            fits = edge_cap >= *nedges && (*nedges == 0 || (adj != nullptr && w != nullptr));
            if (fits)
            {
                std::copy(g.adj.begin(), g.adj.end(), adj);
                std::copy(g.w.begin(), g.w.end(), w);
            }
        });
        return rc != PSMAS_OK ? rc : (fits ? PSMAS_OK : PSMAS_ERROR_BUFFER_TOO_SMALL);
    }

    int psmas_get_level_info(psmas_handle h, int *info, int cap)
    {
        if (!h || !info)
            return null_handle();
        return guarded(h->last_error, [&] {
            const std::vector<int> v = h->solver.level_info();
            if (cap < static_cast<int>(v.size()))
                throw psmas::InvalidArgumentError("[MAS] level info buffer too small.");
            std::memcpy(info, v.data(), sizeof(int) * v.size());
        });
    }

    int psmas_get_partition(psmas_handle h, int *nparts, int *offsets, int offsets_cap, int *perm, int perm_cap)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            std::vector<int> off, pm;
            h->solver.get_partition(off, pm);
            if (nparts)
                *nparts = static_cast<int>(off.size()) - 1;
            if (offsets)
            {
                if (offsets_cap < static_cast<int>(off.size()))
                    throw psmas::InvalidArgumentError("[MAS] partition offsets buffer too small.");
                std::memcpy(offsets, off.data(), sizeof(int) * off.size());
            }
            if (perm)
            {
                if (perm_cap < static_cast<int>(pm.size()))
                    throw psmas::InvalidArgumentError("[MAS] permutation buffer too small.");
                std::memcpy(perm, pm.data(), sizeof(int) * pm.size());
            }
        });
    }

    int psmas_get_coarse_map(psmas_handle h, int *map, long long cap)
    {
        if (!h || !map)
            return null_handle();
        return guarded(h->last_error, [&] {
            const std::vector<int> v = h->solver.coarse_map();
            if (cap < static_cast<long long>(v.size()))
                throw psmas::InvalidArgumentError("[MAS] coarse map buffer too small.");
            std::memcpy(map, v.data(), sizeof(int) * v.size());
        });
    }

    int psmas_get_bank_inverses(psmas_handle h, double *out, long long cap)
    {
        if (!h || !out)
            return null_handle();
        return guarded(h->last_error, [&] {
            const std::vector<double> v = h->solver.bank_inverses();
            if (cap < static_cast<long long>(v.size()))
                throw psmas::InvalidArgumentError("[MAS] bank matrix buffer too small.");
            std::memcpy(out, v.data(), sizeof(double) * v.size());
        });
    }

    int psmas_apply_preconditioner(psmas_handle h, int n, const double *r, double *z, int mixed)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            if (n != h->solver.dim())
                throw std::runtime_error("[MAS] Invalid vector size for MAS preconditioner.");
            if (!r || !z)
                throw psmas::InvalidArgumentError("[MAS] null vector.");
            h->solver.apply_preconditioner(r, z, mixed != 0);
        });
    }

    int psmas_spmv(psmas_handle h, int n, const double *x, double *y, int fp32)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            if (n != h->solver.dim())
                throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
            if (!x || !y)
                throw psmas::InvalidArgumentError("[MAS] null vector.");
            h->solver.spmv(x, y, fp32 != 0);
        });
    }

    int psmas_autotune(psmas_handle h, int reps, char *buf, size_t cap, size_t *needed)
    {
        if (!h)
            return null_handle();
        std::string report;
        const int rc = guarded(h->last_error, [&] { report = h->solver.autotune(reps); });
        if (rc != PSMAS_OK)
            return rc;
        return write_string(report, buf, cap, needed);
    }

    int psmas_lattice_size(psmas_handle h, int *m)
    {
        if (!h || !m)
            return h ? PSMAS_ERROR_INVALID_ARGUMENT : null_handle();
        return guarded(h->last_error, [&] { *m = h->solver.lattice_size(); });
    }

    int psmas_lattice_restrict(psmas_handle h, int n, const double *v, double *u)
    {
        if (!h || !v || !u)
            return h ? PSMAS_ERROR_INVALID_ARGUMENT : null_handle();
        return guarded(h->last_error, [&] {
            if (n != h->solver.dim())
                throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
            h->solver.lattice_restrict(v, u);
        });
    }

    int psmas_lattice_prolong(psmas_handle h, int n, const double *w, double *e)
    {
        if (!h || !w || !e)
            return h ? PSMAS_ERROR_INVALID_ARGUMENT : null_handle();
        return guarded(h->last_error, [&] {
            if (n != h->solver.dim())
                throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
            h->solver.lattice_prolong(w, e);
        });
    }

    int psmas_learned_train(psmas_handle h, int epochs, int batch, float lr, unsigned seed, double *loss)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            const double l = h->solver.learned_train(epochs, batch, lr, seed);
            if (loss)
                *loss = l;
        });
    }

    int psmas_learned_reset(psmas_handle h, int drop_pairs)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] { h->solver.learned_reset(drop_pairs != 0); });
    }

    int psmas_learned_exact(psmas_handle h, double *ms)
    {
        if (!h)
            return null_handle();
        return guarded(h->last_error, [&] {
            const double t = h->solver.learned_exact();
            if (ms)
                *ms = t;
        });
    }

    int psmas_set_kernel_timing(psmas_handle h, int on)
    {
        if (!h)
            return null_handle();
        h->solver.set_kernel_timing(on != 0);
        if (on)
            h->solver.reset_kernel_timings();
        return PSMAS_OK;
    }

    int psmas_get_kernel_timing(psmas_handle h, char *buf, size_t cap, size_t *needed)
    {
        if (!h)
            return null_handle();
        std::ostringstream o;
        o.precision(17);
        o << "{";
        bool first = true;
        for (const auto &t : h->solver.kernel_timings())
        {
            if (!first)
                o << ", ";
            first = false;
            o << "\"" << t.name << "\": {\"launches\": " << t.launches << ", \"total_ms\": " << psmas::json_number(t.total_ms)
              << ", \"bytes\": " << psmas::json_number(t.bytes_per_launch) << "}";
        }
        o << "}";
        return write_string(o.str(), buf, cap, needed);
    }
}
