/*
 * A plain C program that uses libpsmas.so only through include/psmas.h, the way a host
 * application or a polysolve adapter does (create, set_parameters, factorize, solve, get_info,
 * destroy), built by tests/test_capi_consumer.py with gcc -std=c99 -pedantic -Werror.
 *
 * get_info is read with the size query the header documents: a first call with no buffer
 * returns PSMAS_ERROR_BUFFER_TOO_SMALL and the size in *needed, the second call fills a buffer of
 * that size.  polysolve's Newton loop reads get_info after every solve
 * (src/polysolve/nonlinear/descent_strategies/Newton.cpp:209-211), so the info text must be
 * readable whatever its length: with the coarsest-level Cholesky and a Dirichlet factorisation it
 * exceeds 1 KB.
 *
 * Exit codes: 0 success, 77 no GPU visible, 1 any failed check (the message goes to stderr).
 */
#include "psmas.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int fail(psmas_handle h, const char *what, int rc)
{
    fprintf(stderr, "consumer: %s failed (status %d): %s\n", what, rc, psmas_last_error(h));
    return 1;
}

/* get_info through the size query; returns a malloc'd string or NULL. */
static char *read_info(psmas_handle h, size_t *len)
{
    size_t needed = 0;
    int rc = psmas_get_info(h, NULL, 0, &needed);
    if (rc != PSMAS_ERROR_BUFFER_TOO_SMALL || needed == 0)
        return NULL;
    char *buf = (char *)malloc(needed);
    if (buf == NULL)
        return NULL;
    rc = psmas_get_info(h, buf, needed, &needed);
    if (rc != PSMAS_OK)
    {
        free(buf);
        return NULL;
    }
    *len = strlen(buf);
    return buf;
}

/* SPD test matrix with 3x3 blocks on an m^3 node grid (6-neighbour couplings), CSC, both
 * triangles: diagonal block (deg + 0.5) I + 0.2 (J - I), off-diagonal block -I. */
typedef struct
{
    int n;
    int *colptr, *rowidx;
    double *vals;
} csc;

static int node(int m, int i, int j, int k) { return i + m * (j + m * k); }

static csc grid_matrix(int m)
{
    csc A;
    const int nn = m * m * m;
    A.n = 3 * nn;
    A.colptr = (int *)malloc(sizeof(int) * (size_t)(A.n + 1));
    A.rowidx = (int *)malloc(sizeof(int) * (size_t)A.n * 21);
    A.vals = (double *)malloc(sizeof(double) * (size_t)A.n * 21);
    int nz = 0;
    for (int col = 0; col < A.n; ++col)
    {
        const int p = col / 3, c = col % 3;
        const int i = p % m, j = (p / m) % m, k = p / (m * m);
        const int nb[7][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        int deg = 0;
        for (int t = 0; t < 7; ++t)
        {
            const int a = i + nb[t][0], b = j + nb[t][1], d = k + nb[t][2];
            if ((nb[t][0] | nb[t][1] | nb[t][2]) != 0 && a >= 0 && a < m && b >= 0 && b < m && d >= 0 && d < m)
                ++deg;
        }
        A.colptr[col] = nz;
        for (int t = 0; t < 7; ++t) /* neighbours in ascending node order */
        {
            const int a = i + nb[t][0], b = j + nb[t][1], d = k + nb[t][2];
            if (a < 0 || a >= m || b < 0 || b >= m || d < 0 || d >= m)
                continue;
            const int q = node(m, a, b, d);
            if (q == p)
            {
                for (int r = 0; r < 3; ++r)
                {
                    A.rowidx[nz] = 3 * p + r;
                    A.vals[nz++] = (r == c) ? deg + 0.5 : 0.2;
                }
            }
            else
            {
                A.rowidx[nz] = 3 * q + c;
                A.vals[nz++] = -1.0;
            }
        }
    }
    A.colptr[A.n] = nz;
    return A;
}

static double rel_residual(const csc *A, const double *b, const double *x)
{
    double *y = (double *)calloc((size_t)A->n, sizeof(double));
    for (int c = 0; c < A->n; ++c)
        for (int t = A->colptr[c]; t < A->colptr[c + 1]; ++t)
            y[A->rowidx[t]] += A->vals[t] * x[c];
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < A->n; ++i)
    {
        rr += (b[i] - y[i]) * (b[i] - y[i]);
        bb += b[i] * b[i];
    }
    free(y);
    return sqrt(rr / bb);
}

int main(void)
{
    psmas_handle h = NULL;
    int rc = psmas_create(&h);
    if (rc == PSMAS_ERROR_NO_DEVICE)
    {
        printf("consumer: no GPU\n");
        return 77;
    }
    if (rc != PSMAS_OK)
        return fail(NULL, "psmas_create", rc);
    if (strcmp(psmas_name(), "MAS") != 0)
        return fail(h, "psmas_name", -1);

    const csc A = grid_matrix(12);
    const long long nnz = A.colptr[A.n];
    double *b = (double *)malloc(sizeof(double) * (size_t)A.n);
    double *x = (double *)calloc((size_t)A.n, sizeof(double));
    for (int i = 0; i < A.n; ++i)
        b[i] = sin(0.37 * i) + 0.1;

    rc = psmas_set_parameters(h, "{\"MAS\": {\"block_dim\": 3, \"relative_tolerance\": 1e-10, "
                                 "\"absolute_tolerance\": 0.0, \"max_iter\": 2000, \"coarse_solver\": \"cholesky\"}}");
    if (rc != PSMAS_OK)
        return fail(h, "psmas_set_parameters", rc);
    if ((rc = psmas_analyze_pattern(h, A.n, A.colptr, A.rowidx, 0)) != PSMAS_OK)
        return fail(h, "psmas_analyze_pattern", rc);
    if ((rc = psmas_factorize_csc(h, A.n, nnz, A.colptr, A.rowidx, A.vals)) != PSMAS_OK)
        return fail(h, "psmas_factorize_csc", rc);
    if ((rc = psmas_solve(h, A.n, b, x)) != PSMAS_OK)
        return fail(h, "psmas_solve", rc);
    const double res = rel_residual(&A, b, x);
    size_t len1 = 0;
    char *info = read_info(h, &len1);
    if (info == NULL)
        return fail(h, "psmas_get_info (size query)", -1);
    if (strstr(info, "\"solver_status\": \"Reach relative tolerance\"") == NULL ||
        strstr(info, "dense FP64 Cholesky") == NULL || res > 1e-9)
    {
        fprintf(stderr, "consumer: solve: residual %.3e, info %s\n", res, info);
        return 1;
    }
    free(info);

    /* Dirichlet elimination (FEMSolver's prefactorize / dirichlet_solve_prefactorized): the
     * bottom node layer is fixed at f. */
    const int ndir = 3 * 12 * 12;
    int *fixed = (int *)malloc(sizeof(int) * (size_t)ndir);
    for (int i = 0; i < ndir; ++i)
        fixed[i] = i;
    rc = psmas_dirichlet_factorize_csc(h, A.n, nnz, A.colptr, A.rowidx, A.vals, ndir, fixed);
    if (rc != PSMAS_OK)
        return fail(h, "psmas_dirichlet_factorize_csc", rc);
    double *f = (double *)malloc(sizeof(double) * (size_t)A.n);
    for (int i = 0; i < A.n; ++i)
        f[i] = (i < ndir) ? 0.01 * (i % 7) : b[i];
    double *f_in = (double *)malloc(sizeof(double) * (size_t)A.n);
    memcpy(f_in, f, sizeof(double) * (size_t)A.n);
    if ((rc = psmas_dirichlet_solve(h, A.n, f, x)) != PSMAS_OK)
        return fail(h, "psmas_dirichlet_solve", rc);
    double worst = 0.0;
    for (int i = 0; i < ndir; ++i)
        worst = fmax(worst, fabs(x[i] - f_in[i]));
    size_t len2 = 0;
    info = read_info(h, &len2);
    if (info == NULL)
        return fail(h, "psmas_get_info after the Dirichlet solve", -1);
    if (strstr(info, "\"active\": true") == NULL || worst > 1e-8)
    {
        fprintf(stderr, "consumer: Dirichlet: fixed-dof error %.3e, info %s\n", worst, info);
        return 1;
    }
    free(info);

    /* A too-small caller buffer is reported, not overrun, and *needed says what would fit. */
    char small[16];
    size_t needed = 0;
    rc = psmas_get_info(h, small, sizeof small, &needed);
    if (rc != PSMAS_ERROR_BUFFER_TOO_SMALL || needed != len2 + 1 || strlen(small) != sizeof small - 1)
        return fail(h, "psmas_get_info truncation contract", rc);

    if ((rc = psmas_destroy(h)) != PSMAS_OK)
        return fail(NULL, "psmas_destroy", rc);
    printf("consumer ok: n=%d residual=%.3e info_bytes=%zu dirichlet_info_bytes=%zu fixed_err=%.1e\n", A.n, res,
           len1, len2, worst);
    free(b);
    free(x);
    free(f);
    free(f_in);
    free(fixed);
    free(A.colptr);
    free(A.rowidx);
    free(A.vals);
    return 0;
}
