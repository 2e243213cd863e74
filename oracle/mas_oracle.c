/*
 * CPU oracle for the MAS-preconditioned PCG path.  TEST INFRASTRUCTURE ONLY: it is the checker
 * that tests/, the smoke check and the benchmark's CPU-baseline leg compare against; nothing in
 * the product library links, loads or calls it.
 *
 * A plain-C, double-precision, serial restatement of the reference algorithm, function by
 * function (paths relative to src/polysolve/linear/ of iiiian/polysolve):
 *
 *   orc_mas_create   BSR build under a block permutation   mas_utils/BSRMatrix.cu:91-193,433-474
 *                    padded topology (32-node banks)       mas_utils/MASPreconditioner.cu:58-110,1003-1052
 *                    coarse space (bank-local CCs, <=4)    mas_utils/MASPreconditioner.cu:117-481
 *                    Galerkin bank matrices (upper)        mas_utils/MASPreconditioner.cu:484-591,939-979
 *                    zero-diagonal padding                 mas_utils/MASPreconditioner.cu:779-796
 *                    packed symmetric sweep inverse        mas_utils/MASPreconditioner.cu:799-936
 *   orc_mas_apply    restrict / bank SYMV / prolong        mas_utils/MASPreconditioner.cu:593-777,1099-1164
 *   orc_mas_spmv     y = A x on the permuted BSR           MASSolver.cu:271-290 (cuSPARSE BSR SpMV)
 *   orc_mas_pcg      PCG with the reference schedule       MASSolver.cu:46-130,376-409,469-595
 *   orc_jacobi_cg    Eigen::ConjugateGradient<Lower|Upper, DiagonalPreconditioner> semantics
 *                    (Eigen 3.4 ConjugateGradient.h; wired in Solver.cpp:274-279,436-439)
 *
 * Differences that do not change the mathematics: the reference sums with device atomics in a
 * nondeterministic order, here sums run in index order; the reference compacts the adjacency
 * list while building coarse levels (an optimisation that drops only edges that are already
 * inside one component), here every edge is visited at every level.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BANK 32
#define ORC_MAX_COARSE 4

typedef struct
{
    int n;      /* scalar dimension of A */
    int bd;     /* block dim */
    int nb;     /* block rows = ceil(n / bd) */
    int nnzb;   /* stored blocks */
    int *rows;  /* nb + 1 */
    int *cols;  /* nnzb */
    double *vals; /* nnzb * bd * bd, row-major blocks */
    int *perm;  /* old block -> new block */
    int *iperm; /* new block -> old block */

    int nparts;
    int npad;           /* nparts * 32 */
    int *real_to_padded; /* nb */
    int *padded_to_real; /* npad, -1 = virtual */

    int level_num;                 /* coarse levels built (1..4) */
    int cco_nums[ORC_MAX_COARSE];
    int *map;                      /* nb * 4 */

    int mat_dim, mat_storage;
    int mat_counts[ORC_MAX_COARSE + 1];
    int mat_offsets[ORC_MAX_COARSE + 1];
    int total_mats;
    double *mats_fwd; /* bank matrices before inversion (after diagonal padding) */
    double *mats;     /* inverses, packed upper row-major */

    int level_offsets[ORC_MAX_COARSE + 1]; /* in nodes */
    int total_level_nodes;

    int status; /* 0 ok, 1 inverse failed, 2 invalid input */
} orc_mas;

static int div_up(int a, int b) { return (a + b - 1) / b; }

static int upper_index(int dim, int i, int j)
{
    if (i > j)
    {
        int t = i;
        i = j;
        j = t;
    }
    return i * dim - (i * (i + 1)) / 2 + j;
}

static int popcount32(uint32_t v) { return __builtin_popcount(v); }
static int ctz32(uint32_t v) { return __builtin_ctz(v); }

/* ---------------------------------------------------------------------------------------------
 * BSR from CSC under a block permutation (BSRMatrix.cu:91-193).  Every scalar (r, c, v) goes to
 * block (perm[r/bd], perm[c/bd]) at row-major offset (r%bd, c%bd); duplicates are summed; when
 * n is not a multiple of bd the trailing block's missing diagonal scalars are set to 1.0
 * (BSRMatrix.cu:123-134,189-191).  Blocks of a row are sorted by column.
 * ------------------------------------------------------------------------------------------- */
static int build_bsr(orc_mas *M, const int *colptr, const int *rowidx, const double *vals)
{
    const int n = M->n, bd = M->bd, nb = M->nb, bs = bd * bd;
    const int padded = bd * nb - n;
    /* Pass 1: count unique (R, C) pairs per row R using a stamp per row. */
    int *cnt = (int *)calloc((size_t)nb + 1, sizeof(int));
    int *stamp = (int *)malloc(sizeof(int) * (size_t)nb);
    if (!cnt || !stamp)
        return 2;
    /* Collect pairs column-block by column-block. */
    int cap = 1024, npairs = 0;
    int *pr = (int *)malloc(sizeof(int) * cap);
    int *pc = (int *)malloc(sizeof(int) * cap);
    for (int i = 0; i < nb; ++i)
        stamp[i] = -1;
    for (int cb = 0; cb < nb; ++cb)
    {
        const int C = M->perm[cb];
        for (int c = cb * bd; c < cb * bd + bd; ++c)
        {
            if (c >= n)
                continue;
            for (int k = colptr[c]; k < colptr[c + 1]; ++k)
            {
                const int R = M->perm[rowidx[k] / bd];
                if (stamp[R] == cb)
                    continue;
                stamp[R] = cb;
                if (npairs == cap)
                {
                    cap *= 2;
                    pr = (int *)realloc(pr, sizeof(int) * cap);
                    pc = (int *)realloc(pc, sizeof(int) * cap);
                }
                pr[npairs] = R;
                pc[npairs] = C;
                ++npairs;
            }
        }
        if (padded > 0 && cb == nb - 1)
        {
            const int R = M->perm[cb];
            if (stamp[R] != cb)
            {
                stamp[R] = cb;
                if (npairs == cap)
                {
                    cap *= 2;
                    pr = (int *)realloc(pr, sizeof(int) * cap);
                    pc = (int *)realloc(pc, sizeof(int) * cap);
                }
                pr[npairs] = R;
                pc[npairs] = C;
                ++npairs;
            }
        }
    }
    for (int p = 0; p < npairs; ++p)
        cnt[pr[p] + 1]++;
    for (int i = 0; i < nb; ++i)
        cnt[i + 1] += cnt[i];
    M->nnzb = npairs;
    M->rows = (int *)malloc(sizeof(int) * ((size_t)nb + 1));
    M->cols = (int *)malloc(sizeof(int) * (size_t)npairs);
    M->vals = (double *)calloc((size_t)npairs * bs, sizeof(double));
    memcpy(M->rows, cnt, sizeof(int) * ((size_t)nb + 1));
    for (int p = 0; p < npairs; ++p)
        M->cols[cnt[pr[p]]++] = pc[p];
    free(pr);
    free(pc);
    free(cnt);
    free(stamp);
    /* Sort each row by column (insertion sort, rows are short). */
    for (int r = 0; r < nb; ++r)
    {
        for (int a = M->rows[r] + 1; a < M->rows[r + 1]; ++a)
        {
            int v = M->cols[a], b = a - 1;
            while (b >= M->rows[r] && M->cols[b] > v)
            {
                M->cols[b + 1] = M->cols[b];
                --b;
            }
            M->cols[b + 1] = v;
        }
    }
    /* Scatter values: locate the block by binary search in its row. */
    for (int c = 0; c < n; ++c)
    {
        const int C = M->perm[c / bd];
        for (int k = colptr[c]; k < colptr[c + 1]; ++k)
        {
            const int r = rowidx[k];
            const int R = M->perm[r / bd];
            int lo = M->rows[R], hi = M->rows[R + 1] - 1;
            while (lo < hi)
            {
                int mid = (lo + hi) / 2;
                if (M->cols[mid] < C)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            M->vals[(size_t)lo * bs + (r % bd) * bd + (c % bd)] += vals[k];
        }
    }
    if (padded > 0)
    {
        const int T = M->perm[nb - 1];
        int lo = M->rows[T];
        while (M->cols[lo] != T)
            ++lo;
        for (int t = 0; t < padded; ++t)
        {
            int local = (bd - padded) + t;
            M->vals[(size_t)lo * bs + local * bd + local] = 1.0;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Coarse space (MASPreconditioner.cu:117-481).
 * Level 0: connected components of each 32-slot bank's induced subgraph; component id inside a
 * bank is the rank of its lowest lane among the bank's component leaders; global id = bank
 * prefix + local id.  Level l: the level-(l-1) ids are grouped in banks of 32 consecutive ids and
 * two ids in one bank are linked when a fine edge joins nodes carrying them; components again.
 * Stop when a level merges nothing.
 * ------------------------------------------------------------------------------------------- */
static void bank_components(const uint32_t *nbr, int count, int *local_id, int *cco_in_bank)
{
    /* nbr[0..count): neighbour masks (incl. self) of one bank; count <= 32. */
    uint32_t leaders = 0;
    uint32_t conn[ORC_BANK];
    for (int l = 0; l < count; ++l)
    {
        uint32_t c = nbr[l];
        uint32_t visited = 1u << l;
        uint32_t to_visit = c ^ visited;
        while (to_visit)
        {
            int v = ctz32(to_visit);
            c |= nbr[v];
            visited |= 1u << v;
            to_visit = c ^ visited;
        }
        conn[l] = c;
        if ((uint32_t)ctz32(c) == (uint32_t)l)
            leaders |= 1u << l;
    }
    for (int l = 0; l < count; ++l)
    {
        uint32_t lead = (uint32_t)ctz32(conn[l]);
        uint32_t below = (uint32_t)((1ull << lead) - 1ull);
        local_id[l] = popcount32(leaders & below);
    }
    *cco_in_bank = popcount32(leaders);
}

static int build_coarse_space(orc_mas *M)
{
    const int nb = M->nb, npad = M->npad;
    const int banks = npad / ORC_BANK;
    M->map = (int *)malloc(sizeof(int) * (size_t)nb * ORC_MAX_COARSE);
    for (size_t i = 0; i < (size_t)nb * ORC_MAX_COARSE; ++i)
        M->map[i] = -1; /* levels past the last one */
    int *local = (int *)malloc(sizeof(int) * (size_t)npad);
    int *count = (int *)calloc((size_t)banks + 1, sizeof(int));
    uint32_t *nbr = (uint32_t *)calloc((size_t)npad, sizeof(uint32_t));

    /* Level 0: neighbour masks inside each bank (self included). */
    for (int p = 0; p < npad; ++p)
    {
        const int rid = M->padded_to_real[p];
        if (rid < 0)
            continue;
        uint32_t m = 1u << (p % ORC_BANK);
        for (int k = M->rows[rid]; k < M->rows[rid + 1]; ++k)
        {
            const int q = M->real_to_padded[M->cols[k]];
            if (q / ORC_BANK == p / ORC_BANK)
                m |= 1u << (q % ORC_BANK);
        }
        nbr[p] = m;
    }
    for (int b = 0; b < banks; ++b)
    {
        /* Virtual lanes carry no bits; their component id is never read. */
        int ids[ORC_BANK];
        uint32_t masks[ORC_BANK];
        int leaders_only = 0;
        for (int l = 0; l < ORC_BANK; ++l)
            masks[l] = nbr[b * ORC_BANK + l];
        /* Components over real lanes only: treat virtual lanes as absent. */
        uint32_t leaders = 0;
        uint32_t conn[ORC_BANK];
        for (int l = 0; l < ORC_BANK; ++l)
        {
            if (M->padded_to_real[b * ORC_BANK + l] < 0)
            {
                conn[l] = 0;
                continue;
            }
            uint32_t c = masks[l], visited = 1u << l, to_visit = c ^ visited;
            while (to_visit)
            {
                int v = ctz32(to_visit);
                c |= masks[v];
                visited |= 1u << v;
                to_visit = c ^ visited;
            }
            conn[l] = c;
            if (ctz32(c) == l)
                leaders |= 1u << l;
        }
        for (int l = 0; l < ORC_BANK; ++l)
        {
            if (M->padded_to_real[b * ORC_BANK + l] < 0)
            {
                ids[l] = -1;
                continue;
            }
            uint32_t lead = (uint32_t)ctz32(conn[l]);
            ids[l] = popcount32(leaders & (uint32_t)((1ull << lead) - 1ull));
        }
        (void)leaders_only;
        for (int l = 0; l < ORC_BANK; ++l)
            local[b * ORC_BANK + l] = ids[l];
        count[b] = popcount32(leaders);
    }
    /* Inclusive prefix -> global ids. */
    int *prefix = (int *)malloc(sizeof(int) * ((size_t)banks + 1));
    int run = 0;
    for (int b = 0; b < banks; ++b)
    {
        run += count[b];
        prefix[b] = run;
    }
    for (int r = 0; r < nb; ++r)
    {
        const int p = M->real_to_padded[r];
        const int b = p / ORC_BANK;
        M->map[r * ORC_MAX_COARSE + 0] = local[p] + prefix[b] - count[b];
    }
    int cco_num = run;
    M->cco_nums[0] = cco_num;
    M->level_num = 1;
    free(prefix);

    for (int lv = 1; lv < ORC_MAX_COARSE; ++lv)
    {
        memset(nbr, 0, sizeof(uint32_t) * (size_t)npad);
        for (int r = 0; r < nb; ++r)
        {
            const int c = M->map[r * ORC_MAX_COARSE + lv - 1];
            uint32_t m = 1u << (c % ORC_BANK);
            for (int k = M->rows[r]; k < M->rows[r + 1]; ++k)
            {
                const int cj = M->map[M->cols[k] * ORC_MAX_COARSE + lv - 1];
                if (cj / ORC_BANK == c / ORC_BANK)
                    m |= 1u << (cj % ORC_BANK);
            }
            nbr[c] |= m;
        }
        const int nbank = div_up(cco_num, ORC_BANK);
        for (int b = 0; b < nbank; ++b)
        {
            int cnt_in = cco_num - b * ORC_BANK;
            if (cnt_in > ORC_BANK)
                cnt_in = ORC_BANK;
            int c_here;
            bank_components(nbr + b * ORC_BANK, cnt_in, local + b * ORC_BANK, &c_here);
            count[b] = c_here;
        }
        int *pre = (int *)malloc(sizeof(int) * ((size_t)nbank + 1));
        run = 0;
        for (int b = 0; b < nbank; ++b)
        {
            run += count[b];
            pre[b] = run;
        }
        for (int r = 0; r < nb; ++r)
        {
            const int prev = M->map[r * ORC_MAX_COARSE + lv - 1];
            const int b = prev / ORC_BANK;
            M->map[r * ORC_MAX_COARSE + lv] = local[prev] + pre[b] - count[b];
        }
        free(pre);
        if (run == cco_num)
            break;
        cco_num = run;
        M->cco_nums[lv] = cco_num;
        M->level_num = lv + 1;
    }
    free(local);
    free(count);
    free(nbr);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Galerkin bank matrices (MASPreconditioner.cu:494-591).  Only blocks (i, j) with i <= j are read;
 * entry (row, col) of the bank matrix is accumulated only when row <= col, and a diagonal block
 * contributes only its upper part.  When two different fine nodes fall on the same coarse node
 * the coarse diagonal block receives A_ij + A_ij^T.
 * ------------------------------------------------------------------------------------------- */
static void fill_matrices(orc_mas *M)
{
    const int bd = M->bd, bs = bd * bd, dim = M->mat_dim, st = M->mat_storage;
    for (int i = 0; i < M->nb; ++i)
    {
        const int pi = M->real_to_padded[i];
        for (int nz = M->rows[i]; nz < M->rows[i + 1]; ++nz)
        {
            const int j = M->cols[nz];
            if (i > j)
                continue;
            const int pj = M->real_to_padded[j];
            const double *v = M->vals + (size_t)nz * bs;
            if (pi / ORC_BANK == pj / ORC_BANK)
            {
                double *mat = M->mats_fwd + (size_t)(pi / ORC_BANK) * st;
                const int ri = (pi % ORC_BANK) * bd, rj = (pj % ORC_BANK) * bd;
                for (int bi = 0; bi < bd; ++bi)
                    for (int bj = 0; bj < bd; ++bj)
                    {
                        const int row = ri + bi, col = rj + bj;
                        if (row > col)
                            continue;
                        if (i == j && bi > bj)
                            continue;
                        mat[upper_index(dim, row, col)] += v[bi * bd + bj];
                    }
            }
            for (int lv = 0; lv < M->level_num; ++lv)
            {
                const int ci = M->map[i * ORC_MAX_COARSE + lv];
                const int cj = M->map[j * ORC_MAX_COARSE + lv];
                if (ci / ORC_BANK != cj / ORC_BANK)
                    continue;
                double *mat = M->mats_fwd + (size_t)(M->mat_offsets[lv + 1] + ci / ORC_BANK) * st;
                const int ri = (ci % ORC_BANK) * bd, rj = (cj % ORC_BANK) * bd;
                for (int bi = 0; bi < bd; ++bi)
                    for (int bj = 0; bj < bd; ++bj)
                    {
                        const int row = ri + bi, col = rj + bj;
                        if (row > col)
                            continue;
                        if (i == j && bi > bj)
                            continue;
                        double val = v[bi * bd + bj];
                        if (i != j && ci == cj)
                            val += v[bj * bd + bi];
                        mat[upper_index(dim, row, col)] += val;
                    }
            }
        }
    }
    /* pad_zero_diagonal (MASPreconditioner.cu:779-796) */
    for (int m = 0; m < M->total_mats; ++m)
        for (int r = 0; r < dim; ++r)
        {
            double *d = M->mats_fwd + (size_t)m * st + upper_index(dim, r, r);
            if (*d == 0.0)
                *d = 1.0;
        }
}

/* Symmetric sweep inverse of one packed upper matrix (MASPreconditioner.cu:799-900). */
static int sweep_inverse(double *A, int N)
{
    double col[96];
    int ok = 1;
    for (int p = 0; p < N; ++p)
    {
        const double piv = A[upper_index(N, p, p)];
        if (!isfinite(piv) || piv == 0.0)
            ok = 0;
        for (int t = 0; t < N; ++t)
            col[t] = (t == p) ? 0.0 : A[upper_index(N, t, p)];
        for (int t = 0; t < N; ++t)
        {
            if (t == p)
                continue;
            const double scale = col[t] / piv;
            for (int c = t; c < N; ++c)
            {
                if (c == p)
                    continue;
                const double u = fma(-scale, col[c], A[upper_index(N, t, c)]);
                if (!isfinite(u))
                    ok = 0;
                A[upper_index(N, t, c)] = u;
            }
        }
        for (int t = 0; t < N; ++t)
        {
            if (t == p)
            {
                const double u = -1.0 / piv;
                if (!isfinite(u))
                    ok = 0;
                else
                    A[upper_index(N, p, p)] = u;
            }
            else
            {
                const double u = col[t] / piv;
                if (!isfinite(u))
                    ok = 0;
                else
                    A[upper_index(N, t, p)] = u;
            }
        }
    }
    const int L = N * (N + 1) / 2;
    for (int i = 0; i < L; ++i)
    {
        A[i] = -A[i];
        if (!isfinite(A[i]))
            ok = 0;
    }
    return ok;
}

orc_mas *orc_mas_create(int n, const int *colptr, const int *rowidx, const double *vals, int bd,
                        int nparts, const int *part_offsets, const int *perm)
{
    orc_mas *M = (orc_mas *)calloc(1, sizeof(orc_mas));
    if (!M)
        return NULL;
    if (n <= 0 || bd < 1 || bd > 3 || nparts < 1)
    {
        M->status = 2;
        return M;
    }
    M->n = n;
    M->bd = bd;
    M->nb = div_up(n, bd);
    M->perm = (int *)malloc(sizeof(int) * (size_t)M->nb);
    M->iperm = (int *)malloc(sizeof(int) * (size_t)M->nb);
    memcpy(M->perm, perm, sizeof(int) * (size_t)M->nb);
    for (int i = 0; i < M->nb; ++i)
        M->iperm[M->perm[i]] = i;
    if (part_offsets[nparts] != M->nb)
    {
        M->status = 2;
        return M;
    }
    for (int p = 0; p < nparts; ++p)
        if (part_offsets[p + 1] - part_offsets[p] > ORC_BANK || part_offsets[p + 1] < part_offsets[p])
        {
            M->status = 2;
            return M;
        }
    if (build_bsr(M, colptr, rowidx, vals) != 0)
    {
        M->status = 2;
        return M;
    }

    /* Padded topology (MASPreconditioner.cu:58-88): part p owns slots [32p, 32p + size). */
    M->nparts = nparts;
    M->npad = nparts * ORC_BANK;
    M->real_to_padded = (int *)malloc(sizeof(int) * (size_t)M->nb);
    M->padded_to_real = (int *)malloc(sizeof(int) * (size_t)M->npad);
    for (int p = 0; p < M->npad; ++p)
    {
        const int part = p / ORC_BANK, local = p % ORC_BANK;
        const int b = part_offsets[part], e = part_offsets[part + 1];
        if (local < e - b)
        {
            M->real_to_padded[b + local] = p;
            M->padded_to_real[p] = b + local;
        }
        else
        {
            M->padded_to_real[p] = -1;
        }
    }

    build_coarse_space(M);

    M->mat_dim = ORC_BANK * bd;
    M->mat_storage = M->mat_dim * (M->mat_dim + 1) / 2;
    M->mat_offsets[0] = 0;
    M->mat_counts[0] = M->npad / ORC_BANK;
    M->total_mats = M->mat_counts[0];
    for (int i = 0; i < M->level_num; ++i)
    {
        M->mat_offsets[i + 1] = M->total_mats;
        M->mat_counts[i + 1] = div_up(M->cco_nums[i], ORC_BANK);
        M->total_mats += M->mat_counts[i + 1];
    }
    const size_t total = (size_t)M->total_mats * M->mat_storage;
    M->mats_fwd = (double *)calloc(total, sizeof(double));
    M->mats = (double *)malloc(sizeof(double) * total);
    fill_matrices(M);
    memcpy(M->mats, M->mats_fwd, sizeof(double) * total);
    int ok = 1;
    for (int m = 0; m < M->total_mats; ++m)
        ok &= sweep_inverse(M->mats + (size_t)m * M->mat_storage, M->mat_dim);
    M->status = ok ? 0 : 1;

    M->level_offsets[0] = 0;
    M->total_level_nodes = M->npad;
    for (int i = 0; i < M->level_num; ++i)
    {
        M->level_offsets[i + 1] = M->total_level_nodes;
        M->total_level_nodes += M->mat_counts[i + 1] * ORC_BANK;
    }
    return M;
}

void orc_mas_free(orc_mas *M)
{
    if (!M)
        return;
    free(M->rows);
    free(M->cols);
    free(M->vals);
    free(M->perm);
    free(M->iperm);
    free(M->real_to_padded);
    free(M->padded_to_real);
    free(M->map);
    free(M->mats_fwd);
    free(M->mats);
    free(M);
}

int orc_mas_status(const orc_mas *M) { return M->status; }

/* info[0..]: nb, nnzb, npad, level_num, cco_nums[4], total_mats, mat_dim, mat_storage,
 * mat_counts[5], total_level_nodes */
void orc_mas_info(const orc_mas *M, int *info)
{
    info[0] = M->nb;
    info[1] = M->nnzb;
    info[2] = M->npad;
    info[3] = M->level_num;
    for (int i = 0; i < 4; ++i)
        info[4 + i] = M->cco_nums[i];
    info[8] = M->total_mats;
    info[9] = M->mat_dim;
    info[10] = M->mat_storage;
    for (int i = 0; i < 5; ++i)
        info[11 + i] = M->mat_counts[i];
    info[16] = M->total_level_nodes;
}

void orc_mas_get_bsr(const orc_mas *M, int *rows, int *cols, double *vals)
{
    memcpy(rows, M->rows, sizeof(int) * ((size_t)M->nb + 1));
    memcpy(cols, M->cols, sizeof(int) * (size_t)M->nnzb);
    memcpy(vals, M->vals, sizeof(double) * (size_t)M->nnzb * M->bd * M->bd);
}

void orc_mas_get_map(const orc_mas *M, int *map, int *real_to_padded)
{
    memcpy(map, M->map, sizeof(int) * (size_t)M->nb * ORC_MAX_COARSE);
    memcpy(real_to_padded, M->real_to_padded, sizeof(int) * (size_t)M->nb);
}

void orc_mas_get_matrices(const orc_mas *M, double *fwd, double *inv)
{
    const size_t total = (size_t)M->total_mats * M->mat_storage;
    if (fwd)
        memcpy(fwd, M->mats_fwd, sizeof(double) * total);
    if (inv)
        memcpy(inv, M->mats, sizeof(double) * total);
}

/* ---------------------------------------------------------------------------------------------
 * Apply in permuted space (MASPreconditioner.cu:593-777): z = sum_l R_l^T Inv_l R_l r.
 * r and z have length nb * bd.
 * ------------------------------------------------------------------------------------------- */
static void apply_permuted(const orc_mas *M, const double *r, double *z)
{
    const int bd = M->bd, dim = M->mat_dim;
    const size_t tot = (size_t)M->total_level_nodes * bd;
    double *mr = (double *)calloc(tot, sizeof(double));
    double *mz = (double *)calloc(tot, sizeof(double));
    for (int p = 0; p < M->npad; ++p)
    {
        const int rid = M->padded_to_real[p];
        if (rid < 0)
            continue;
        for (int i = 0; i < bd; ++i)
            mr[(size_t)p * bd + i] = r[(size_t)rid * bd + i];
        for (int lv = 0; lv < M->level_num; ++lv)
        {
            const int c = M->map[rid * ORC_MAX_COARSE + lv];
            for (int i = 0; i < bd; ++i)
                mr[((size_t)M->level_offsets[lv + 1] + c) * bd + i] += r[(size_t)rid * bd + i];
        }
    }
    /* Matrix m of every level reads/writes nodes [32 m, 32 m + 32) of the multi-level vector. */
    for (int m = 0; m < M->total_mats; ++m)
    {
        const double *A = M->mats + (size_t)m * M->mat_storage;
        const double *x = mr + (size_t)m * dim;
        double *y = mz + (size_t)m * dim;
        for (int row = 0; row < dim; ++row)
        {
            double s = 0.0;
            for (int col = 0; col < dim; ++col)
                s += A[upper_index(dim, row, col)] * x[col];
            y[row] = s;
        }
    }
    for (int rid = 0; rid < M->nb; ++rid)
    {
        const int p = M->real_to_padded[rid];
        for (int i = 0; i < bd; ++i)
            z[(size_t)rid * bd + i] = mz[(size_t)p * bd + i];
        for (int lv = 0; lv < M->level_num; ++lv)
        {
            const int c = M->map[rid * ORC_MAX_COARSE + lv];
            for (int i = 0; i < bd; ++i)
                z[(size_t)rid * bd + i] += mz[((size_t)M->level_offsets[lv + 1] + c) * bd + i];
        }
    }
    free(mr);
    free(mz);
}

static void spmv_permuted(const orc_mas *M, const double *x, double *y)
{
    const int bd = M->bd, bs = bd * bd;
    for (int r = 0; r < M->nb; ++r)
    {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int k = M->rows[r]; k < M->rows[r + 1]; ++k)
        {
            const double *v = M->vals + (size_t)k * bs;
            const double *xc = x + (size_t)M->cols[k] * bd;
            for (int bi = 0; bi < bd; ++bi)
                for (int bj = 0; bj < bd; ++bj)
                    acc[bi] += v[bi * bd + bj] * xc[bj];
        }
        for (int bi = 0; bi < bd; ++bi)
            y[(size_t)r * bd + bi] = acc[bi];
    }
}

/* Original-order vector (length n) <-> permuted, padded vector (length nb * bd); the solve path
 * does the same with permute_vector (MASSolver.cu:116-130,387-393,402-407). */
static void to_permuted(const orc_mas *M, const double *in, double *out)
{
    const int bd = M->bd, np = M->nb * bd;
    for (int idx = 0; idx < np; ++idx)
    {
        const int old = M->iperm[idx / bd] * bd + idx % bd;
        out[idx] = (old < M->n) ? in[old] : 0.0;
    }
}

static void from_permuted(const orc_mas *M, const double *in, double *out)
{
    const int bd = M->bd;
    for (int old = 0; old < M->n; ++old)
        out[old] = in[(size_t)M->perm[old / bd] * bd + old % bd];
}

void orc_mas_apply(const orc_mas *M, const double *r, double *z)
{
    const size_t np = (size_t)M->nb * M->bd;
    double *rp = (double *)malloc(sizeof(double) * np);
    double *zp = (double *)malloc(sizeof(double) * np);
    to_permuted(M, r, rp);
    apply_permuted(M, rp, zp);
    from_permuted(M, zp, z);
    free(rp);
    free(zp);
}

void orc_mas_spmv(const orc_mas *M, const double *x, double *y)
{
    const size_t np = (size_t)M->nb * M->bd;
    double *xp = (double *)malloc(sizeof(double) * np);
    double *yp = (double *)malloc(sizeof(double) * np);
    to_permuted(M, x, xp);
    spmv_permuted(M, xp, yp);
    from_permuted(M, yp, y);
    free(xp);
    free(yp);
}

static double dot(const double *a, const double *b, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i)
        s += a[i] * b[i];
    return s;
}

/* scalar_division (MASSolver.cu:43-56): 0 when |den| < guard.  guard = 1e-20 is the reference;
 * guard = 0 selects the scale-free variant of the product (0 only for a non-finite quotient). */
static double safe_div(double num, double den, double guard)
{
    if (guard > 0.0)
        return (fabs(den) < guard) ? 0.0 : num / den;
    const double q = num / den;
    return isfinite(q) ? q : 0.0;
}

/* PCG with the reference schedule (MASSolver.cu:469-595): x0 = 0; the true residual b - A x
 * replaces the recurrence every 4th iteration; convergence is tested every 10th iteration on
 * ||r||^2 (or r^T z when use_precond_norm) against rel^2 * initial and abs^2; status 1 =
 * relative tolerance, 2 = absolute tolerance, 3 = max iterations; the run reaching max_iter is
 * reported as 3 even when it also met a tolerance.  Returns 0, or 4 for an invalid initial r^T z
 * (MASSolver.cu:483-486).  b and x are in the original order, length n.  guard: see safe_div. */
int orc_mas_pcg(const orc_mas *M, const double *b_in, double *x_out, int max_iter, double rel_tol,
                double abs_tol, int use_precond_norm, int true_residual_period, double guard, int *iters_out,
                double *resnorm_out, int *status_out)
{
    const size_t np = (size_t)M->nb * M->bd;
    double *b = (double *)malloc(sizeof(double) * np);
    double *x = (double *)calloc(np, sizeof(double));
    double *r = (double *)malloc(sizeof(double) * np);
    double *p = (double *)malloc(sizeof(double) * np);
    double *z = (double *)malloc(sizeof(double) * np);
    double *Ap = (double *)malloc(sizeof(double) * np);
    to_permuted(M, b_in, b);
    int status = 0, iters = 0, rc = 0;
    double resnorm = 0.0;

    spmv_permuted(M, x, r);
    for (size_t i = 0; i < np; ++i)
        r[i] = b[i] - r[i];
    apply_permuted(M, r, z);
    memcpy(p, z, sizeof(double) * np);
    double rz = dot(r, z, np);
    const double rz0 = rz;
    if (isnan(rz0) || !isfinite(rz0))
    {
        rc = 4;
        goto done;
    }
    double rr0 = 0.0;
    if (!use_precond_norm)
        rr0 = dot(r, r, np);

    for (int k = 1; k <= max_iter; ++k)
    {
        spmv_permuted(M, p, Ap);
        const double pAp = dot(p, Ap, np);
        const double alpha = safe_div(rz, pAp, guard);
        for (size_t i = 0; i < np; ++i)
            x[i] = alpha * p[i] + x[i];
        if (k % true_residual_period == 0)
        {
            spmv_permuted(M, x, r);
            for (size_t i = 0; i < np; ++i)
                r[i] = b[i] - r[i];
        }
        else
        {
            for (size_t i = 0; i < np; ++i)
                r[i] = -alpha * Ap[i] + r[i];
        }
        apply_permuted(M, r, z);
        const double rz_old = rz;
        rz = dot(r, z, np);
        iters = k;
        int converged = 0;
        if (k % 10 == 0)
        {
            if (use_precond_norm)
            {
                resnorm = sqrt(rz);
                if (rz <= rel_tol * rel_tol * rz0 || rz <= abs_tol * abs_tol)
                {
                    status = (rz <= abs_tol * abs_tol) ? 2 : 1;
                    converged = 1;
                }
            }
            else
            {
                const double rr = dot(r, r, np);
                resnorm = sqrt(rr);
                if (rr <= rel_tol * rel_tol * rr0 || rr <= abs_tol * abs_tol)
                {
                    status = (rr <= abs_tol * abs_tol) ? 2 : 1;
                    converged = 1;
                }
            }
        }
        if (converged)
            break;
        const double beta = safe_div(rz, rz_old, guard);
        for (size_t i = 0; i < np; ++i)
            p[i] = z[i] + beta * p[i];
    }
    if (iters == max_iter)
        status = 3;
done:
    from_permuted(M, x, x_out);
    *iters_out = iters;
    *resnorm_out = resnorm;
    *status_out = status;
    free(b);
    free(x);
    free(r);
    free(p);
    free(z);
    free(Ap);
    return rc;
}

/* ---------------------------------------------------------------------------------------------
 * Jacobi-preconditioned CG with Eigen::ConjugateGradient semantics on the full (Lower|Upper)
 * CSC matrix: stop when ||r||^2 < max(tol^2 ||b||^2, tiny); x is the initial guess.  The
 * diagonal preconditioner uses 1/a_ii, or 1 where a_ii == 0.  Returns iterations; *err_out is
 * ||r|| / ||b||.  max_wall_iters caps the run for bounded timing samples (-1 = no cap).
 * ------------------------------------------------------------------------------------------- */
static void csc_spmv(int n, const int *colptr, const int *rowidx, const double *vals,
                     const double *x, double *y)
{
    memset(y, 0, sizeof(double) * (size_t)n);
    for (int c = 0; c < n; ++c)
    {
        const double xc = x[c];
        for (int k = colptr[c]; k < colptr[c + 1]; ++k)
            y[rowidx[k]] += vals[k] * xc;
    }
}

int orc_jacobi_cg(int n, const int *colptr, const int *rowidx, const double *vals, const double *b,
                  double *x, double tol, int max_iter, double *err_out)
{
    double *inv_diag = (double *)malloc(sizeof(double) * (size_t)n);
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    double *z = (double *)malloc(sizeof(double) * (size_t)n);
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    for (int c = 0; c < n; ++c)
    {
        double d = 0.0;
        for (int k = colptr[c]; k < colptr[c + 1]; ++k)
            if (rowidx[k] == c)
                d += vals[k];
        inv_diag[c] = (d != 0.0) ? 1.0 / d : 1.0;
    }
    csc_spmv(n, colptr, rowidx, vals, x, t);
    for (int i = 0; i < n; ++i)
        r[i] = b[i] - t[i];
    const double rhs2 = dot(b, b, (size_t)n);
    int it = 0;
    double res2 = dot(r, r, (size_t)n);
    if (rhs2 == 0.0)
    {
        memset(x, 0, sizeof(double) * (size_t)n);
        *err_out = 0.0;
        goto out;
    }
    {
        const double tiny = 2.2250738585072014e-308; /* (std::numeric_limits<double>::min)() */
        double threshold = tol * tol * rhs2;
        if (threshold < tiny)
            threshold = tiny;
        if (res2 < threshold)
        {
            *err_out = sqrt(res2 / rhs2);
            goto out;
        }
        for (int i = 0; i < n; ++i)
            p[i] = inv_diag[i] * r[i];
        double abs_new = dot(r, p, (size_t)n);
        while (it < max_iter)
        {
            csc_spmv(n, colptr, rowidx, vals, p, t);
            const double alpha = abs_new / dot(p, t, (size_t)n);
            for (int i = 0; i < n; ++i)
            {
                x[i] += alpha * p[i];
                r[i] -= alpha * t[i];
            }
            res2 = dot(r, r, (size_t)n);
            if (res2 < threshold)
                break;
            for (int i = 0; i < n; ++i)
                z[i] = inv_diag[i] * r[i];
            const double abs_old = abs_new;
            abs_new = dot(r, z, (size_t)n);
            const double beta = abs_new / abs_old;
            for (int i = 0; i < n; ++i)
                p[i] = z[i] + beta * p[i];
            ++it;
        }
        *err_out = sqrt(res2 / rhs2);
    }
out:
    free(inv_diag);
    free(r);
    free(p);
    free(z);
    free(t);
    return it;
}
