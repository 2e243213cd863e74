/*
 * CPU baseline (test infrastructure only): Jacobi-preconditioned CG with the semantics of
 * Eigen::ConjugateGradient<StiffnessMatrix, Lower|Upper, DiagonalPreconditioner<double>>, the
 * reference's iterative CPU solver (wired in src/polysolve/linear/Solver.cpp:274-279,436-439;
 * DiagonalPreconditioner is the default preconditioner, Solver.cpp:609-612; solveWithGuess from
 * EigenSolver.tpp:196-243), restated with OpenMP over `nthreads` threads.
 *
 * Eigen 3.4's conjugate_gradient (ConjugateGradient.h), which this follows step by step:
 *   r = b - A x;  threshold = max(tol^2 |b|^2, DBL_MIN);  stop at once if |r|^2 < threshold;
 *   p = D^-1 r;  absNew = r.p;  loop i < maxIters:
 *     t = A p;  alpha = absNew / p.t;  x += alpha p;  r -= alpha t;
 *     if |r|^2 < threshold: stop;  z = D^-1 r;  absOld = absNew;  absNew = r.z;
 *     p = z + (absNew / absOld) p;  i++
 *   tol_error = sqrt(|r|^2 / |b|^2), iterations = i.
 * D^-1 is 1/a_ii, or 1 where a_ii == 0 (Eigen's DiagonalPreconditioner).
 *
 * A is symmetric and stored with both triangles in CSC, so column c holds row c: the product is
 * computed row by row (each thread owns whole rows, no atomics).  Eigen's product on a
 * column-major matrix runs on one thread; the restatement parallelises it, which only makes the
 * CPU baseline faster.  Sums run in a fixed order for a given thread count.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static double pdot(const double *a, const double *b, int n)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i)
        s += a[i] * b[i];
    return s;
}

/* t = A p (rows of the symmetric matrix = its columns); returns p.t */
static double spmv_dot(int n, const int *colptr, const int *rowidx, const double *vals, const double *p, double *t)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int r = 0; r < n; ++r)
    {
        double acc = 0.0;
        for (int k = colptr[r]; k < colptr[r + 1]; ++k)
            acc += vals[k] * p[rowidx[k]];
        t[r] = acc;
        s += p[r] * acc;
    }
    return s;
}

int orc_jacobi_cg_omp(int n, const int *colptr, const int *rowidx, const double *vals, const double *b, double *x,
                      double tol, int max_iter, int nthreads, double *err_out)
{
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    double *inv_diag = (double *)malloc(sizeof(double) * (size_t)n);
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    int it = 0;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c)
    {
        double d = 0.0;
        for (int k = colptr[c]; k < colptr[c + 1]; ++k)
            if (rowidx[k] == c)
                d += vals[k];
        inv_diag[c] = (d != 0.0) ? 1.0 / d : 1.0;
    }
    (void)spmv_dot(n, colptr, rowidx, vals, x, t);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        r[i] = b[i] - t[i];
    const double rhs2 = pdot(b, b, n);
    double res2 = pdot(r, r, n);
    if (rhs2 == 0.0)
    {
        memset(x, 0, sizeof(double) * (size_t)n);
        *err_out = 0.0;
        goto out;
    }
    {
        const double tiny = 2.2250738585072014e-308; /* DBL_MIN, Eigen's considerAsZero */
        double threshold = tol * tol * rhs2;
        if (threshold < tiny)
            threshold = tiny;
        if (res2 < threshold)
        {
            *err_out = sqrt(res2 / rhs2);
            goto out;
        }
        double abs_new = 0.0;
#pragma omp parallel for reduction(+ : abs_new) schedule(static)
        for (int i = 0; i < n; ++i)
        {
            p[i] = inv_diag[i] * r[i];
            abs_new += r[i] * p[i];
        }
        while (it < max_iter)
        {
            const double alpha = abs_new / spmv_dot(n, colptr, rowidx, vals, p, t);
            res2 = 0.0;
#pragma omp parallel for reduction(+ : res2) schedule(static)
            for (int i = 0; i < n; ++i)
            {
                x[i] += alpha * p[i];
                r[i] -= alpha * t[i];
                res2 += r[i] * r[i];
            }
            if (res2 < threshold)
                break;
            const double abs_old = abs_new;
            abs_new = 0.0;
#pragma omp parallel for reduction(+ : abs_new) schedule(static)
            for (int i = 0; i < n; ++i)
                abs_new += r[i] * (inv_diag[i] * r[i]);
            const double beta = abs_new / abs_old;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i)
                p[i] = inv_diag[i] * r[i] + beta * p[i];
            ++it;
        }
        *err_out = sqrt(res2 / rhs2);
    }
out:
    free(inv_diag);
    free(r);
    free(p);
    free(t);
    return it;
}

int orc_omp_max_threads(void) { return omp_get_max_threads(); }

/*
 * Extend-add of the sparse Cholesky stand-in (oracle/sparse_cholesky.py): a child's update matrix
 * U (m x m, column-major, leading dimension m, lower triangle valid) is added into its parent's
 * front, whose rows/columns li[0..m) are ascending positions in [0, k + mb): positions below k
 * are the parent's pivots, the others its boundary.  The front is held as F11 (k x k), F21
 * (mb x k) and F22 (mb x mb), column-major.  Lower triangles only (li ascending keeps lower on
 * lower).
 */
void orc_extend_add(int m, const double *U, const int *li, int k, int mb, double *F11, double *F21, double *F22)
{
    /* One thread: the BLAS threads of the dense kernels around this call still spin on the
     * cores after their calls, and a second thread team competing with them is slower than a
     * single pass over U, which is memory-bound. */
    for (int j = 0; j < m; ++j)
    {
        const int c = li[j];
        const double *u = U + (size_t)j * (size_t)m;
        for (int i = j; i < m; ++i)
        {
            const int r = li[i];
            if (c < k)
            {
                if (r < k)
                    F11[(size_t)c * (size_t)k + (size_t)r] += u[i];
                else
                    F21[(size_t)c * (size_t)mb + (size_t)(r - k)] += u[i];
            }
            else
                F22[(size_t)(c - k) * (size_t)mb + (size_t)(r - k)] += u[i];
        }
    }
}
