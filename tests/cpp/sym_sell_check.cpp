// Host check of build_symmetric_sell (polysolve_amd/csrc/host_setup.cpp): on SELL-64 matrices of
// random symmetric 27-point grids (node order shuffled, so node blocks couple irregularly), the
// symmetric form, evaluated the way the GPU kernel does (stored slots, transposed and foreign
// products through the incoming slots, summed after all rows), must give the full matrix's A x;
// every incoming slot is written exactly once and every row appears once.  A structurally
// asymmetric pattern within a node block must be refused.
#include "host_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <random>
#include <vector>

using namespace psmas;

namespace
{
    struct Full
    {
        int npad = 0, bd = 0;
        std::vector<int> slice_ptr, row_len, cols;
        std::vector<double> vals;
        std::vector<unsigned short> lcols;
        std::vector<std::vector<int>> halo;
    };

    Full make(int nx, int ny, int nz, int bd, unsigned seed)
    {
        const int bs = bd * bd, nn = nx * ny * nz;
        Full F;
        F.bd = bd;
        F.npad = (nn + 255) / 256 * 256;
        std::mt19937 rng(seed);
        std::uniform_real_distribution<double> U(-1, 1);
        std::vector<int> perm(nn);
        for (int i = 0; i < nn; ++i)
            perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), rng);
        std::vector<std::map<int, std::vector<double>>> rows(F.npad);
        auto id = [&](int x, int y, int z) { return perm[x + nx * (y + ny * z)]; };
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x)
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx)
                            {
                                const int X = x + dx, Y = y + dy, Z = z + dz;
                                if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz)
                                    continue;
                                const int i = id(x, y, z), j = id(X, Y, Z);
                                if (j < i)
                                    continue;
                                std::vector<double> b(bs), t(bs);
                                for (auto &v : b)
                                    v = U(rng);
                                for (int a = 0; a < bd; ++a)
                                    for (int c = 0; c < bd; ++c)
                                        t[c * bd + a] = b[a * bd + c];
                                rows[i][j] = b;
                                rows[j][i] = t;
                            }
        const int ns = F.npad / 64;
        F.slice_ptr.assign(ns + 1, 0);
        F.row_len.resize(F.npad);
        for (int r = 0; r < F.npad; ++r)
            F.row_len[r] = static_cast<int>(rows[r].size());
        for (int s = 0; s < ns; ++s)
        {
            int m = 1;
            for (int l = 0; l < 64; ++l)
                m = std::max(m, F.row_len[s * 64 + l]);
            F.slice_ptr[s + 1] = F.slice_ptr[s] + m;
        }
        const size_t nslots = F.slice_ptr[ns];
        F.cols.assign(nslots * 64, 0);
        F.vals.assign(nslots * 64 * bs, 0.0);
        for (int r = 0; r < F.npad; ++r)
        {
            const int s = r / 64, l = r % 64;
            for (int q = F.slice_ptr[s]; q < F.slice_ptr[s + 1]; ++q)
                F.cols[static_cast<size_t>(q) * 64 + l] = r;
            int k = 0;
            for (const auto &e : rows[r])
            {
                const size_t slot = F.slice_ptr[s] + k++;
                F.cols[slot * 64 + l] = e.first;
                for (int t = 0; t < bs; ++t)
                    F.vals[(slot * bs + t) * 64 + l] = e.second[t];
            }
        }
        F.lcols.resize(F.cols.size());
        F.halo.resize(F.npad / 256);
        for (int b = 0; b < F.npad / 256; ++b)
        {
            const size_t i0 = static_cast<size_t>(F.slice_ptr[b * 4]) * 64, i1 = static_cast<size_t>(F.slice_ptr[b * 4 + 4]) * 64;
            std::vector<int> u(F.cols.begin() + i0, F.cols.begin() + i1);
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            for (size_t i = i0; i < i1; ++i)
                F.lcols[i] = static_cast<unsigned short>(std::lower_bound(u.begin(), u.end(), F.cols[i]) - u.begin());
            F.halo[b] = u;
        }
        return F;
    }

    int check(const Full &F, unsigned seed)
    {
        const int bd = F.bd, bs = bd * bd, npad = F.npad;
        SymSell S;
        if (!build_symmetric_sell(npad, F.slice_ptr, F.cols, F.lcols, F.row_len, S))
        {
            std::printf("refused: %s\n", S.reason.c_str());
            return 1;
        }
        std::mt19937 rng(seed);
        std::uniform_real_distribution<double> U(-1, 1);
        std::vector<double> x(static_cast<size_t>(npad) * bd), yf(x.size(), 0.0), ys(x.size(), 0.0);
        for (auto &v : x)
            v = U(rng);
        for (int r = 0; r < npad; ++r)
        {
            const int s = r / 64, l = r % 64;
            for (int q = F.slice_ptr[s]; q < F.slice_ptr[s + 1]; ++q)
            {
                const size_t at = static_cast<size_t>(q) * 64 + l;
                for (int i = 0; i < bd; ++i)
                    for (int j = 0; j < bd; ++j)
                        yf[r * bd + i] += F.vals[(static_cast<size_t>(q) * bs + i * bd + j) * 64 + l] * x[F.cols[at] * bd + j];
            }
        }
        std::vector<int> seen(npad, 0);
        for (int b = 0; b < npad / 256; ++b)
        {
            const std::vector<int> &h = F.halo[b];
            const int im = std::max(1, S.incoming_max);
            std::vector<double> inc(static_cast<size_t>(bd) * im, NAN), acc(256 * bd, 0.0);
            std::vector<int> lrow(256);
            for (int p = 0; p < 256; ++p)
            {
                const int s = b * 4 + p / 64, l = p % 64;
                const uint32_t info = S.row_info[static_cast<size_t>(b) * 256 + p];
                const int lr = info & 0xFF;
                lrow[p] = lr;
                seen[b * 256 + lr]++;
                const int self = h[S.meta[static_cast<size_t>(S.slice_ptr[s]) * 64 + l] & 0xFFFF];
                if (self != b * 256 + lr)
                {
                    std::printf("slot 0 of row %d is not its diagonal\n", b * 256 + lr);
                    return 1;
                }
                for (int q = S.slice_ptr[s]; q < S.slice_ptr[s + 1]; ++q)
                {
                    const size_t at = static_cast<size_t>(q) * 64 + l;
                    const uint32_t m = S.meta[at];
                    const int src = S.src[at], c = h[m & 0xFFFF], ts = (m >> 16) & 0x7FFF;
                    const bool fo = (m >> 31) != 0;
                    double blk[9] = {0};
                    if (src >= 0)
                    {
                        if (F.cols[src] != c)
                        {
                            std::printf("column of slot %zu differs from its source block\n", at);
                            return 1;
                        }
                        for (int t = 0; t < bs; ++t)
                            blk[t] = F.vals[(static_cast<size_t>(src / 64) * bs + t) * 64 + src % 64];
                    }
                    auto put = [&](int slot, int i, double v) {
                        double &d = inc[static_cast<size_t>(i) * im + slot];
                        if (!std::isnan(d))
                            return false;
                        d = v;
                        return true;
                    };
                    if (fo)
                    {
                        if (src < 0 || ts == 0x7FFF)
                            return std::printf("foreign slot without block or target\n"), 1;
                        for (int i = 0; i < bd; ++i)
                        {
                            double t = 0;
                            for (int j = 0; j < bd; ++j)
                                t += blk[i * bd + j] * x[c * bd + j];
                            if (!put(ts, i, t))
                                return std::printf("incoming slot written twice\n"), 1;
                        }
                        continue;
                    }
                    for (int i = 0; i < bd; ++i)
                        for (int j = 0; j < bd; ++j)
                            acc[p * bd + i] += blk[i * bd + j] * x[c * bd + j];
                    if (ts != 0x7FFF)
                    {
                        if (src < 0 || S.mirror[at] < 0)
                            return std::printf("transposed slot without block or mirror\n"), 1;
                        for (int j = 0; j < bd; ++j)
                        {
                            double t = 0;
                            for (int i = 0; i < bd; ++i)
                                t += blk[i * bd + j] * x[self * bd + i];
                            if (!put(ts, j, t))
                                return std::printf("incoming slot written twice\n"), 1;
                        }
                    }
                }
            }
            for (int p = 0; p < 256; ++p)
            {
                const uint32_t info = S.row_info[static_cast<size_t>(b) * 256 + p];
                const int cnt = (info >> 8) & 0xFF, ib = static_cast<int>(info >> 16);
                for (int q = 0; q < cnt; ++q)
                    for (int i = 0; i < bd; ++i)
                    {
                        const double v = inc[static_cast<size_t>(i) * im + ib + q];
                        if (std::isnan(v))
                            return std::printf("incoming slot never written\n"), 1;
                        acc[p * bd + i] += v;
                    }
                for (int i = 0; i < bd; ++i)
                    ys[(b * 256 + lrow[p]) * bd + i] = acc[p * bd + i];
            }
        }
        for (int r = 0; r < npad; ++r)
            if (seen[r] != 1)
                return std::printf("row %d placed %d times\n", r, seen[r]), 1;
        double err = 0, nrm = 0;
        for (size_t i = 0; i < yf.size(); ++i)
        {
            err = std::max(err, std::fabs(yf[i] - ys[i]));
            nrm = std::max(nrm, std::fabs(yf[i]));
        }
        std::printf("bd %d: %zu full slots, %lld symmetric, incoming max %d, max |y_sym - y| %.2e of %.2e\n", bd,
                    F.cols.size() / 64, S.slots, S.incoming_max, err, nrm);
        return err <= 1e-12 * nrm ? 0 : 1;
    }
} // namespace

int main()
{
    int fails = 0;
    fails += check(make(9, 11, 13, 3, 1), 2);
    fails += check(make(8, 8, 8, 2, 3), 4);
    fails += check(make(7, 6, 5, 1, 5), 6);
    // Structurally asymmetric within a node block: drop one block of a pair (the row keeps its mirror).
    Full F = make(9, 11, 13, 3, 7);
    for (int r = 0; r < F.npad; ++r)
    {
        const int s = r / 64, l = r % 64, L = F.row_len[r];
        int pos = -1;
        for (int k = 0; k < L && pos < 0; ++k)
        {
            const int c = F.cols[static_cast<size_t>(F.slice_ptr[s] + k) * 64 + l];
            if (c != r && c / 256 == r / 256)
                pos = k;
        }
        if (pos < 0)
            continue;
        for (int k = pos; k < L - 1; ++k)
            F.cols[static_cast<size_t>(F.slice_ptr[s] + k) * 64 + l] = F.cols[static_cast<size_t>(F.slice_ptr[s] + k + 1) * 64 + l];
        F.cols[static_cast<size_t>(F.slice_ptr[s] + L - 1) * 64 + l] = r;
        F.row_len[r] = L - 1;
        break;
    }
    SymSell S;
    if (build_symmetric_sell(F.npad, F.slice_ptr, F.cols, F.lcols, F.row_len, S))
    {
        std::printf("a structurally asymmetric block pattern was accepted\n");
        ++fails;
    }
    else
        std::printf("asymmetric pattern refused: %s\n", S.reason.c_str());
    std::printf(fails ? "FAILED\n" : "OK\n");
    return fails ? 1 : 0;
}
