// Q1-hex linear-elasticity system assembler for the MAS solver's benchmark and test problems.
//
// Unit cube, ne^3 trilinear hexahedra, (ne+1)^3 nodes with 3 displacement DOFs each, node id
// i + j*(ne+1) + k*(ne+1)^2 (x fastest), DOF id 3*node + component.  Element stiffness is
// 2x2x2 Gauss quadrature of B^T D B for isotropic linear elasticity with a per-element Young's
// modulus.  Optional SPD 6x6 pair blocks kappa*g*g^T (contact/barrier-like Hessian terms) are added
// on top.  The bottom face z = 0 is Dirichlet-fixed (unless asked not to) using the convention of polysolve's
// FEM helper (src/polysolve/linear/FEMSolver.cpp:97-160): every entry whose row or column is a
// Dirichlet DOF is dropped, and the Dirichlet diagonal is set to 1; with a zero Dirichlet value
// the right-hand side of those DOFs is 0.
//
// Output is a compressed column-major sparse matrix (int32 column pointers and row indices,
// double values, rows sorted inside each column), exactly the shape of polysolve's
// StiffnessMatrix (src/polysolve/Types.hpp:11-15).  The assembled matrix is exactly symmetric
// bit for bit: the element matrix is symmetrised before scaling and each block pair is summed in
// the same element order.
//
// This file has no counterpart in the reference; it only manufactures inputs.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

namespace
{
    struct Problem
    {
        int n = 0;
        size_t nnz = 0;
        std::vector<int> colptr;
        std::unique_ptr<int[]> rowidx;  // nnz, written in parallel (no serial zero fill)
        std::unique_ptr<double[]> vals; // nnz
        std::vector<double> rhs;
    };

    // 24x24 stiffness of the unit cube element with E = 1, scaled later by E_e * h.
    void unit_element_stiffness(double nu, double K[24][24])
    {
        const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        const double mu = 1.0 / (2.0 * (1.0 + nu));
        double D[6][6] = {};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                D[a][b] = lam;
        for (int a = 0; a < 3; ++a)
            D[a][a] += 2.0 * mu;
        for (int a = 3; a < 6; ++a)
            D[a][a] = mu;

        std::memset(K, 0, sizeof(double) * 24 * 24);
        const double g = 1.0 / std::sqrt(3.0);
        const double gp[2] = {-g, g};
        for (int qz = 0; qz < 2; ++qz)
            for (int qy = 0; qy < 2; ++qy)
                for (int qx = 0; qx < 2; ++qx)
                {
                    const double xi = gp[qx], eta = gp[qy], zeta = gp[qz];
                    // Physical element [0,1]^3: x = (xi+1)/2, dN/dx = 2 dN/dxi, detJ = 1/8.
                    double dN[8][3];
                    for (int a = 0; a < 8; ++a)
                    {
                        const double sa = (a & 1) ? 1.0 : -1.0;
                        const double ta = (a & 2) ? 1.0 : -1.0;
                        const double ua = (a & 4) ? 1.0 : -1.0;
                        dN[a][0] = 2.0 * sa * (1.0 + ta * eta) * (1.0 + ua * zeta) / 8.0;
                        dN[a][1] = 2.0 * ta * (1.0 + sa * xi) * (1.0 + ua * zeta) / 8.0;
                        dN[a][2] = 2.0 * ua * (1.0 + sa * xi) * (1.0 + ta * eta) / 8.0;
                    }
                    double B[6][24] = {};
                    for (int a = 0; a < 8; ++a)
                    {
                        const double nx = dN[a][0], ny = dN[a][1], nz = dN[a][2];
                        B[0][3 * a + 0] = nx;
                        B[1][3 * a + 1] = ny;
                        B[2][3 * a + 2] = nz;
                        B[3][3 * a + 0] = ny; // gamma_xy
                        B[3][3 * a + 1] = nx;
                        B[4][3 * a + 1] = nz; // gamma_yz
                        B[4][3 * a + 2] = ny;
                        B[5][3 * a + 0] = nz; // gamma_xz
                        B[5][3 * a + 2] = nx;
                    }
                    double DB[6][24];
                    for (int r = 0; r < 6; ++r)
                        for (int c = 0; c < 24; ++c)
                        {
                            double s = 0.0;
                            for (int k = 0; k < 6; ++k)
                                s += D[r][k] * B[k][c];
                            DB[r][c] = s;
                        }
                    const double w = 1.0 / 8.0; // detJ * unit Gauss weights
                    for (int p = 0; p < 24; ++p)
                        for (int q = 0; q < 24; ++q)
                        {
                            double s = 0.0;
                            for (int k = 0; k < 6; ++k)
                                s += B[k][p] * DB[k][q];
                            K[p][q] += w * s;
                        }
                }
        for (int p = 0; p < 24; ++p)
            for (int q = p + 1; q < 24; ++q)
            {
                const double s = 0.5 * (K[p][q] + K[q][p]);
                K[p][q] = s;
                K[q][p] = s;
            }
    }

    struct ExtraBlock
    {
        int row, col;
        double v[9];
    };

    Problem *assemble(int ne, const double *E_elem, double nu, double body_force_z,
                      int npairs, const int *pair_i, const int *pair_j,
                      const double *pair_g, const double *pair_kappa, bool fix_bottom)
    {
        const int N = ne + 1;
        const int nn = N * N * N;
        const double h = 1.0 / ne;
        double Ku[24][24];
        unit_element_stiffness(nu, Ku);

        // 27-point stencil blocks per node, slot s = (dk+1)*9 + (dj+1)*3 + (di+1); slot order
        // equals ascending neighbour id.  blocks[node][slot] holds A(node, neighbour).
        // Left uninitialised here: each node's thread zeroes its own blocks first (parallel first
        // touch of ~0.5 GB on 64^3).
        std::unique_ptr<double[]> blocks(new double[static_cast<size_t>(nn) * 27 * 9]);
        std::vector<double> rhs(static_cast<size_t>(nn) * 3, 0.0);
        const double fnode = body_force_z * h * h * h / 8.0;

        // Element contributions gathered per node (one thread per node, no write conflicts): node
        // a visits its adjacent elements in ascending element id, which is the order an element
        // loop would add them in, so every block and load entry gets the same sums bit for bit.
#pragma omp parallel for schedule(static)
        for (int a = 0; a < nn; ++a)
        {
            std::fill(&blocks[static_cast<size_t>(a) * 27 * 9], &blocks[static_cast<size_t>(a + 1) * 27 * 9], 0.0);
            const int ai = a % N, aj = (a / N) % N, ak = a / (N * N);
            for (int ek = ak - 1; ek <= ak; ++ek)
                for (int ej = aj - 1; ej <= aj; ++ej)
                    for (int ei = ai - 1; ei <= ai; ++ei)
                    {
                        if (ei < 0 || ei >= ne || ej < 0 || ej >= ne || ek < 0 || ek >= ne)
                            continue;
                        const int e = ei + ne * (ej + ne * ek);
                        const double sc = E_elem[e] * h;
                        // local index of node a in element e
                        const int la = (ai - ei) | ((aj - ej) << 1) | ((ak - ek) << 2);
                        rhs[3 * static_cast<size_t>(a) + 2] += fnode;
                        for (int b = 0; b < 8; ++b)
                        {
                            const int di = (b & 1) - (la & 1);
                            const int dj = ((b >> 1) & 1) - ((la >> 1) & 1);
                            const int dk = ((b >> 2) & 1) - ((la >> 2) & 1);
                            const int slot = (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
                            double *blk = &blocks[(static_cast<size_t>(a) * 27 + slot) * 9];
                            for (int r = 0; r < 3; ++r)
                                for (int c = 0; c < 3; ++c)
                                    blk[r * 3 + c] += sc * Ku[3 * la + r][3 * b + c];
                        }
                    }
        }

        // Pair blocks: kappa * g g^T with g = (g_i, g_j) in R^6.
        std::vector<ExtraBlock> extra;
        auto stencil_slot = [&](int a, int b) -> int {
            const int ai = a % N, aj = (a / N) % N, ak = a / (N * N);
            const int bi = b % N, bj = (b / N) % N, bk = b / (N * N);
            const int di = bi - ai, dj = bj - aj, dk = bk - ak;
            if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1)
                return -1;
            return (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
        };
        auto add_block = [&](int a, int b, const double *ga, const double *gb, double kappa) {
            double v[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    v[r * 3 + c] = kappa * ga[r] * gb[c];
            const int slot = stencil_slot(a, b);
            if (slot >= 0)
            {
                double *blk = &blocks[(static_cast<size_t>(a) * 27 + slot) * 9];
                for (int t = 0; t < 9; ++t)
                    blk[t] += v[t];
            }
            else
            {
                ExtraBlock eb;
                eb.row = a;
                eb.col = b;
                std::memcpy(eb.v, v, sizeof(v));
                extra.push_back(eb);
            }
        };
        for (int p = 0; p < npairs; ++p)
        {
            const int a = pair_i[p], b = pair_j[p];
            const double *g = pair_g + 6 * p;
            const double kap = pair_kappa[p];
            add_block(a, a, g, g, kap);
            add_block(a, b, g, g + 3, kap);
            add_block(b, a, g + 3, g, kap);
            add_block(b, b, g + 3, g + 3, kap);
        }
        // Merge duplicate off-stencil blocks (same row, col), keeping insertion order of sums.
        std::stable_sort(extra.begin(), extra.end(), [](const ExtraBlock &x, const ExtraBlock &y) {
            return x.row != y.row ? x.row < y.row : x.col < y.col;
        });
        std::vector<ExtraBlock> merged;
        for (const auto &eb : extra)
        {
            if (!merged.empty() && merged.back().row == eb.row && merged.back().col == eb.col)
            {
                for (int t = 0; t < 9; ++t)
                    merged.back().v[t] += eb.v[t];
            }
            else
            {
                merged.push_back(eb);
            }
        }
        std::vector<int> extra_ptr(nn + 1, 0);
        for (const auto &eb : merged)
            extra_ptr[eb.row + 1] += 1;
        for (int a = 0; a < nn; ++a)
            extra_ptr[a + 1] += extra_ptr[a];

        auto dirichlet = [&](int node) { return fix_bottom && node < N * N; }; // bottom face k = 0

        Problem *P = new Problem();
        P->n = 3 * nn;
        P->rhs.assign(rhs.begin(), rhs.end());
        for (int a = 0; a < nn; ++a)
            if (dirichlet(a))
                for (int c = 0; c < 3; ++c)
                    P->rhs[3 * a + c] = 0.0;

        // Emit CSC.  Column (a, comp) lists rows (b, r) for neighbours b of a in ascending order,
        // value A(b, a)[r][comp] = A(a, b)[comp][r] (the matrix is symmetric).  Two passes over
        // the nodes (entry counts, then the entries at their scanned offsets), each parallel.
        auto neighbours = [&](int a, std::vector<std::pair<int, const double *>> &all) {
            const int ai = a % N, aj = (a / N) % N, ak = a / (N * N);
            all.clear();
            for (int s = 0; s < 27; ++s)
            {
                const int di = s % 3 - 1, dj = (s / 3) % 3 - 1, dk = s / 9 - 1;
                const int bi = ai + di, bj = aj + dj, bk = ak + dk;
                if (bi < 0 || bi >= N || bj < 0 || bj >= N || bk < 0 || bk >= N)
                    continue;
                all.emplace_back(bi + N * (bj + N * bk), &blocks[(static_cast<size_t>(a) * 27 + s) * 9]);
            }
            if (extra_ptr[a + 1] > extra_ptr[a])
            {
                for (int t = extra_ptr[a]; t < extra_ptr[a + 1]; ++t) // off-stencil blocks
                    all.emplace_back(merged[t].col, merged[t].v);
                std::stable_sort(all.begin(), all.end(),
                                 [](const std::pair<int, const double *> &x, const std::pair<int, const double *> &y) {
                                     return x.first < y.first;
                                 });
            }
        };
        // entries of column (a, comp): 3 per free neighbour, 1 (the diagonal) for a Dirichlet node
        std::vector<long long> per_node(static_cast<size_t>(nn) + 1, 0);
#pragma omp parallel
        {
            std::vector<std::pair<int, const double *>> all;
#pragma omp for schedule(static)
            for (int a = 0; a < nn; ++a)
            {
                neighbours(a, all);
                long long cnt = 0;
                for (const auto &nb : all)
                    cnt += (dirichlet(a) || dirichlet(nb.first)) ? (a == nb.first ? 1 : 0) : 3;
                per_node[a + 1] = cnt; // per component
            }
        }
        for (int a = 0; a < nn; ++a)
            per_node[a + 1] = per_node[a] + 3 * per_node[a + 1];
        if (per_node[nn] > 2147483647LL)
            throw std::bad_alloc();
        P->colptr.resize(static_cast<size_t>(P->n) + 1);
        P->nnz = static_cast<size_t>(per_node[nn]);
        P->rowidx.reset(new int[P->nnz]);
        P->vals.reset(new double[P->nnz]);
        P->colptr[0] = 0;
#pragma omp parallel
        {
            std::vector<std::pair<int, const double *>> all;
#pragma omp for schedule(static)
            for (int a = 0; a < nn; ++a)
            {
                neighbours(a, all);
                size_t at = static_cast<size_t>(per_node[a]);
                for (int comp = 0; comp < 3; ++comp)
                {
                    for (const auto &nb : all)
                    {
                        const int b = nb.first;
                        if (dirichlet(a) || dirichlet(b))
                        {
                            if (a == b)
                            {
                                // Dirichlet diagonal: only the (comp, comp) entry, value 1.
                                P->rowidx[at] = 3 * b + comp;
                                P->vals[at++] = 1.0;
                            }
                            continue;
                        }
                        for (int r = 0; r < 3; ++r)
                        {
                            P->rowidx[at] = 3 * b + r;
                            P->vals[at++] = nb.second[comp * 3 + r];
                        }
                    }
                    P->colptr[3 * static_cast<size_t>(a) + comp + 1] = static_cast<int>(at);
                }
            }
        }
        return P;
    }
} // namespace

extern "C"
{
    /// Assemble a Q1-hex elasticity system.  E_elem holds ne^3 per-element Young's moduli in
    /// element order i + ne*(j + ne*k).  body_force_z is rho*g_z per unit volume (lumped to the
    /// nodes).  Pairs (pair_i[p], pair_j[p]) add kappa_p * g_p g_p^T with g_p = pair_g[6p:6p+6].
    /// fix_bottom != 0 applies the Dirichlet convention on the bottom face; 0 returns the bare
    /// (singular) stiffness matrix and load.
    /// Returns an opaque handle (NULL on failure); sizes through n_out and nnz_out.
    void *psgen_hex_assemble(int ne, const double *E_elem, double nu, double body_force_z,
                             int npairs, const int *pair_i, const int *pair_j,
                             const double *pair_g, const double *pair_kappa, int fix_bottom,
                             int *n_out, long long *nnz_out)
    {
        if (ne < 1 || E_elem == nullptr || npairs < 0)
            return nullptr;
        const long long N = ne + 1;
        if (3 * N * N * N * 27 * 3 > 2147483647LL)
            return nullptr;
        try
        {
            Problem *P = assemble(ne, E_elem, nu, body_force_z, npairs, pair_i, pair_j, pair_g, pair_kappa,
                                  fix_bottom != 0);
            *n_out = P->n;
            *nnz_out = static_cast<long long>(P->nnz);
            return P;
        }
        catch (const std::bad_alloc &)
        {
            return nullptr;
        }
    }

    /// Copy the assembled system out: colptr[n+1], rowidx[nnz], vals[nnz], rhs[n].
    int psgen_fetch(void *handle, int *colptr, int *rowidx, double *vals, double *rhs)
    {
        if (handle == nullptr)
            return 1;
        const Problem *P = static_cast<const Problem *>(handle);
        std::memcpy(colptr, P->colptr.data(), sizeof(int) * P->colptr.size());
        const long long nnz = static_cast<long long>(P->nnz);
        const long long chunk = 1 << 20;
#pragma omp parallel for schedule(static)
        for (long long c0 = 0; c0 < nnz; c0 += chunk)
        {
            const size_t len = static_cast<size_t>(std::min(chunk, nnz - c0));
            std::memcpy(rowidx + c0, P->rowidx.get() + c0, sizeof(int) * len);
            std::memcpy(vals + c0, P->vals.get() + c0, sizeof(double) * len);
        }
        std::memcpy(rhs, P->rhs.data(), sizeof(double) * P->rhs.size());
        return 0;
    }

    void psgen_free(void *handle)
    {
        delete static_cast<Problem *>(handle);
    }

    /// The 24x24 unit-cube element stiffness (E = 1), row-major, for tests.
    void psgen_unit_element_stiffness(double nu, double *out)
    {
        double K[24][24];
        unit_element_stiffness(nu, K);
        std::memcpy(out, K, sizeof(K));
    }
}
