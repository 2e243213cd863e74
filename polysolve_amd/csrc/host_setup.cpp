#include "host_setup.hpp"
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <climits>
#include <cmath>
#include <numeric>
#include <atomic>
#include <thread>

namespace psmas
{
    void validate_csc(const CscView &A)
    {
        if (A.n <= 0 || A.colptr == nullptr || (A.nnz > 0 && (A.rowidx == nullptr || A.vals == nullptr)))
            throw InvalidArgumentError("[MAS] invalid A: empty matrix or null arrays.");
        if (A.nnz > INT_MAX)
            throw InvalidArgumentError("[MAS] A is too large. Non-zero number exceeding int32 max.");
        if (A.colptr[0] != 0 || A.colptr[A.n] != A.nnz)
            throw InvalidArgumentError("[MAS] invalid A: column pointers do not span the non-zeros.");
        for (int c = 0; c < A.n; ++c)
            if (A.colptr[c + 1] < A.colptr[c])
                throw InvalidArgumentError("[MAS] invalid A: column pointers are not monotone.");
        for (long long k = 0; k < A.nnz; ++k)
            if (A.rowidx[k] < 0 || A.rowidx[k] >= A.n)
                throw InvalidArgumentError("[MAS] invalid A: row index out of range.");
    }

    BlockGraph build_block_graph(const CscView &A, int bd)
    {
        const int n = A.n;
        const int nb = div_up(n, bd);
        BlockGraph g;
        g.nb = nb;
        // Block column cb: unique block rows with accumulated |a|; edges stored in both
        // directions, weights summed over (u,v) and (v,u).
        std::vector<int> stamp(nb, -1);
        std::vector<double> acc(nb, 0.0);
        std::vector<int> touched;
        std::vector<std::vector<std::pair<int, double>>> lists(nb);
        for (int cb = 0; cb < nb; ++cb)
        {
            touched.clear();
            for (int c = cb * bd; c < std::min(n, cb * bd + bd); ++c)
                for (int k = A.colptr[c]; k < A.colptr[c + 1]; ++k)
                {
                    const int rb = A.rowidx[k] / bd;
                    if (rb == cb)
                        continue;
                    if (stamp[rb] != cb)
                    {
                        stamp[rb] = cb;
                        acc[rb] = 0.0;
                        touched.push_back(rb);
                    }
                    acc[rb] += std::fabs(A.vals[k]);
                }
            for (int rb : touched)
            {
                lists[cb].emplace_back(rb, acc[rb]);
                lists[rb].emplace_back(cb, acc[rb]);
            }
        }
        g.ptr.assign(nb + 1, 0);
        for (int u = 0; u < nb; ++u)
        {
            auto &L = lists[u];
            std::sort(L.begin(), L.end());
            int out = 0;
            for (size_t i = 0; i < L.size(); ++i)
            {
                if (out > 0 && L[out - 1].first == L[i].first)
                    L[out - 1].second += L[i].second;
                else
                    L[out++] = L[i];
            }
            L.resize(out);
            g.ptr[u + 1] = g.ptr[u] + out;
        }
        g.adj.resize(g.ptr[nb]);
        g.w.resize(g.ptr[nb]);
        for (int u = 0; u < nb; ++u)
        {
            int o = g.ptr[u];
            for (const auto &e : lists[u])
            {
                g.adj[o] = e.first;
                g.w[o] = e.second;
                ++o;
            }
        }
        return g;
    }

    namespace
    {
        /// BFS order of every component, each started from a pseudo-peripheral node.
        std::vector<int> bfs_order(const BlockGraph &g)
        {
            const int nb = g.nb;
            std::vector<int> order;
            order.reserve(nb);
            std::vector<int> level(nb, -1);
            std::vector<char> done(nb, 0);
            std::vector<int> queue;
            auto bfs = [&](int s, std::vector<int> &out, bool mark_done) {
                out.clear();
                out.push_back(s);
                level[s] = 0;
                for (size_t h = 0; h < out.size(); ++h)
                {
                    const int u = out[h];
                    for (int k = g.ptr[u]; k < g.ptr[u + 1]; ++k)
                    {
                        const int v = g.adj[k];
                        if (level[v] < 0)
                        {
                            level[v] = level[u] + 1;
                            out.push_back(v);
                        }
                    }
                }
                for (int v : out)
                {
                    level[v] = -1;
                    if (mark_done)
                        done[v] = 1;
                }
            };
            for (int s = 0; s < nb; ++s)
            {
                if (done[s])
                    continue;
                bfs(s, queue, false);
                const int far = queue.back();
                bfs(far, queue, true);
                order.insert(order.end(), queue.begin(), queue.end());
            }
            return order;
        }
    } // namespace

    Partition graph_aggregate(const BlockGraph &g, int max_size)
    {
        const int nb = g.nb;
        Partition P;
        if (nb <= max_size)
        {
            P.nparts = 1;
            P.offsets = {0, nb};
            P.perm.resize(nb);
            std::iota(P.perm.begin(), P.perm.end(), 0);
            return P;
        }
        const std::vector<int> order = bfs_order(g);
        std::vector<int> rank(nb);
        for (int i = 0; i < nb; ++i)
            rank[order[i]] = i;

        std::vector<int> agg(nb, -1);
        std::vector<std::vector<int>> members;
        std::vector<int> conn(nb, 0);
        std::vector<double> wsum(nb, 0.0);
        std::vector<int> cand_stamp(nb, -1);
        std::vector<int> cand;
        for (int s : order)
        {
            if (agg[s] != -1)
                continue;
            const int a = static_cast<int>(members.size());
            members.emplace_back();
            auto &mem = members.back();
            cand.clear();
            auto add = [&](int v) {
                agg[v] = a;
                mem.push_back(v);
                for (int k = g.ptr[v]; k < g.ptr[v + 1]; ++k)
                {
                    const int u = g.adj[k];
                    if (agg[u] != -1)
                        continue;
                    if (cand_stamp[u] != a)
                    {
                        cand_stamp[u] = a;
                        conn[u] = 0;
                        wsum[u] = 0.0;
                        cand.push_back(u);
                    }
                    conn[u] += 1;
                    wsum[u] += g.w[k];
                }
            };
            add(s);
            while (static_cast<int>(mem.size()) < max_size)
            {
                int best = -1;
                for (int u : cand)
                {
                    if (agg[u] != -1)
                        continue;
                    if (best < 0 || conn[u] > conn[best] ||
                        (conn[u] == conn[best] && (wsum[u] > wsum[best] ||
                                                   (wsum[u] == wsum[best] && rank[u] < rank[best]))))
                        best = u;
                }
                if (best < 0)
                    break;
                add(best);
            }
        }

        // Merge small aggregates into their most strongly connected neighbour with room.
        const int nagg = static_cast<int>(members.size());
        std::vector<int> target(nagg);
        std::iota(target.begin(), target.end(), 0);
        std::vector<int> size(nagg);
        for (int a = 0; a < nagg; ++a)
            size[a] = static_cast<int>(members[a].size());
        auto find = [&](int a) {
            while (target[a] != a)
                a = target[a];
            return a;
        };
        std::vector<double> link(nagg, 0.0);
        std::vector<int> link_stamp(nagg, -1);
        for (int a = 0; a < nagg; ++a)
        {
            if (size[a] >= max_size / 4 || find(a) != a)
                continue;
            std::vector<int> nbrs;
            for (int v : members[a])
                for (int k = g.ptr[v]; k < g.ptr[v + 1]; ++k)
                {
                    const int b = find(agg[g.adj[k]]);
                    if (b == a)
                        continue;
                    if (link_stamp[b] != a)
                    {
                        link_stamp[b] = a;
                        link[b] = 0.0;
                        nbrs.push_back(b);
                    }
                    link[b] += g.w[k];
                }
            int best = -1;
            for (int b : nbrs)
                if (size[b] + size[a] <= max_size && (best < 0 || link[b] > link[best]))
                    best = b;
            if (best >= 0)
            {
                target[a] = best;
                size[best] += size[a];
            }
        }
        std::vector<std::vector<int>> final_members;
        std::vector<int> final_id(nagg, -1);
        for (int a = 0; a < nagg; ++a)
        {
            const int r = find(a);
            if (final_id[r] < 0)
            {
                final_id[r] = static_cast<int>(final_members.size());
                final_members.emplace_back();
            }
        }
        for (int a = 0; a < nagg; ++a)
        {
            auto &dst = final_members[final_id[find(a)]];
            dst.insert(dst.end(), members[a].begin(), members[a].end());
        }
        P.nparts = static_cast<int>(final_members.size());
        P.offsets.assign(P.nparts + 1, 0);
        P.perm.assign(nb, -1);
        int pos = 0;
        for (int p = 0; p < P.nparts; ++p)
        {
            auto &m = final_members[p];
            std::sort(m.begin(), m.end());
            for (int v : m)
                P.perm[v] = pos++;
            P.offsets[p + 1] = pos;
        }
        return P;
    }

    Partition chunk_partition(const std::vector<int> &order, int max_size)
    {
        const int nb = static_cast<int>(order.size());
        Partition P;
        P.nparts = std::max(1, div_up(nb, max_size));
        P.offsets.resize(P.nparts + 1);
        for (int p = 0; p <= P.nparts; ++p)
            P.offsets[p] = std::min(nb, p * max_size);
        P.perm.resize(nb);
        for (int i = 0; i < nb; ++i)
            P.perm[order[i]] = i;
        return P;
    }

    namespace
    {
        template <typename F>
        void parallel_for_blocks(int count, F &&f)
        {
            unsigned hw = std::thread::hardware_concurrency();
            int nt = static_cast<int>(std::min<unsigned>(hw == 0 ? 1 : hw, 16));
            if (count < 4096 || nt <= 1)
            {
                f(0, count);
                return;
            }
            std::vector<std::thread> th;
            const int chunk = div_up(count, nt);
            for (int t = 0; t < nt; ++t)
            {
                const int b = t * chunk, e = std::min(count, b + chunk);
                if (b >= e)
                    break;
                th.emplace_back([&f, b, e] { f(b, e); });
            }
            for (auto &x : th)
                x.join();
        }
    } // namespace

    void assign_slots(int nb, int bd, const Partition &part, PaddedBsr &out)
    {
        out.bd = bd;
        out.nb = nb;
        out.nparts = div_up(std::max(part.nparts, 1), 8) * 8; // whole 256-node blocks
        out.npad = out.nparts * BANK;
        out.slot_of_block.assign(nb, -1);
        out.block_of_slot.assign(out.npad, -1);
        for (int p = 0; p < part.nparts; ++p)
        {
            const int size = part.offsets[p + 1] - part.offsets[p];
            if (size < 0 || size > BANK)
                throw InvalidArgumentError("[MAS] internal: domain larger than 32 nodes.");
        }
        // permuted id -> part
        std::vector<int> part_of(nb);
        for (int p = 0; p < part.nparts; ++p)
            for (int i = part.offsets[p]; i < part.offsets[p + 1]; ++i)
                part_of[i] = p;
        for (int b = 0; b < nb; ++b)
        {
            const int pid = part.perm[b];
            const int p = part_of[pid];
            const int slot = BANK * p + (pid - part.offsets[p]);
            out.slot_of_block[b] = slot;
            out.block_of_slot[slot] = b;
        }
    }

    void build_padded_bsr(const CscView &A, int bd, const Partition &part, PaddedBsr &out)
    {
        const int n = A.n;
        const int nb = div_up(n, bd);
        const int bs = bd * bd;
        const int padded_scalars = bd * nb - n;
        assign_slots(nb, bd, part, out);
        const std::vector<int> &slot = out.slot_of_block;

        // Pattern: per block column, unique block rows; transpose into rows.
        std::vector<int> stamp(nb, -1);
        std::vector<int> pr, pc;
        pr.reserve(static_cast<size_t>(A.nnz / std::max(1, bs)) + nb);
        pc.reserve(pr.capacity());
        for (int cb = 0; cb < nb; ++cb)
        {
            const int C = slot[cb];
            for (int c = cb * bd; c < std::min(n, cb * bd + bd); ++c)
                for (int k = A.colptr[c]; k < A.colptr[c + 1]; ++k)
                {
                    const int rb = A.rowidx[k] / bd;
                    if (stamp[rb] == cb)
                        continue;
                    stamp[rb] = cb;
                    pr.push_back(slot[rb]);
                    pc.push_back(C);
                }
            if (padded_scalars > 0 && cb == nb - 1 && stamp[cb] != cb)
            {
                stamp[cb] = cb;
                pr.push_back(C);
                pc.push_back(C);
            }
        }
        out.rows.assign(out.npad + 1, 0);
        for (int R : pr)
            out.rows[R + 1]++;
        for (int i = 0; i < out.npad; ++i)
            out.rows[i + 1] += out.rows[i];
        out.cols.resize(pr.size());
        {
            std::vector<int> cur(out.rows.begin(), out.rows.end() - 1);
            for (size_t t = 0; t < pr.size(); ++t)
                out.cols[cur[pr[t]]++] = pc[t];
        }
        std::vector<int>().swap(pr);
        std::vector<int>().swap(pc);
        parallel_for_blocks(out.npad, [&](int b, int e) {
            for (int r = b; r < e; ++r)
                std::sort(out.cols.begin() + out.rows[r], out.cols.begin() + out.rows[r + 1]);
        });
        out.vals.assign(out.cols.size() * bs, 0.0);
        if (!refresh_padded_bsr_values(A, out))
            throw InvalidArgumentError("[MAS] internal: BSR pattern mismatch.");
    }

    bool build_value_map(const CscView &A, int bd, const std::vector<int> &slot_of_block, const std::vector<int> &rows,
                         const std::vector<int> &cols, const std::vector<int> &slice_ptr, std::vector<int> &map)
    {
        const int n = A.n, nb = div_up(n, bd), bs = bd * bd;
        const size_t nvals = static_cast<size_t>(slice_ptr.back()) * bs * 64;
        if (nvals > static_cast<size_t>(INT_MAX))
            return false;
        map.assign(static_cast<size_t>(A.nnz), -1);
        std::vector<unsigned char> taken(nvals, 0);
        std::atomic<bool> ok{true};
        // Parallel over block columns: distinct block columns own distinct blocks, so the marks
        // of one value are only touched by the thread that owns its block column.
        parallel_for_blocks(nb, [&](int b0, int b1) {
            for (int cb = b0; cb < b1; ++cb)
            {
                const int C = slot_of_block[cb];
                for (int c = cb * bd; c < std::min(n, cb * bd + bd); ++c)
                    for (int k = A.colptr[c]; k < A.colptr[c + 1]; ++k)
                    {
                        const int r = A.rowidx[k];
                        const int R = slot_of_block[r / bd];
                        const int *beg = cols.data() + rows[R];
                        const int *end = cols.data() + rows[R + 1];
                        const int *it = std::lower_bound(beg, end, C);
                        if (it == end || *it != C)
                        {
                            ok = false;
                            continue;
                        }
                        const size_t slot = static_cast<size_t>(slice_ptr[R / 64]) + static_cast<size_t>(it - beg);
                        const size_t v = (slot * bs + static_cast<size_t>((r % bd) * bd + (c % bd))) * 64 + R % 64;
                        if (taken[v])
                            ok = false; // duplicate entry: values would have to be added
                        taken[v] = 1;
                        map[static_cast<size_t>(k)] = static_cast<int>(v);
                    }
            }
        });
        return ok;
    }

    void eliminate_dirichlet(const CscView &A, const std::vector<unsigned char> &fixed, DirichletSystem &out)
    {
        const int n = A.n;
        out.colptr.assign(static_cast<size_t>(n) + 1, 0);
        std::vector<unsigned char> add_diag(n, 0);
        // Kept entries per column; a free column without a diagonal gets an explicit zero one.
        parallel_for_blocks(n, [&](int c0, int c1) {
            for (int c = c0; c < c1; ++c)
            {
                if (fixed[c])
                {
                    out.colptr[c + 1] = 1;
                    continue;
                }
                int cnt = 0;
                bool diag = false;
                for (int k = A.colptr[c]; k < A.colptr[c + 1]; ++k)
                {
                    const int r = A.rowidx[k];
                    if (!fixed[r])
                    {
                        ++cnt;
                        diag |= (r == c);
                    }
                }
                add_diag[c] = !diag;
                out.colptr[c + 1] = cnt + (diag ? 0 : 1);
            }
        });
        for (int c = 0; c < n; ++c)
            out.colptr[c + 1] += out.colptr[c];
        const size_t nnz = static_cast<size_t>(out.colptr[n]);
        out.rowidx.resize(nnz);
        out.vals.resize(nnz);
        out.keep.assign(static_cast<size_t>(A.nnz), -1);
        parallel_for_blocks(n, [&](int c0, int c1) {
            for (int c = c0; c < c1; ++c)
            {
                int o = out.colptr[c];
                if (fixed[c])
                {
                    out.rowidx[o] = c;
                    out.vals[o] = 1.0;
                    continue;
                }
                bool pending = add_diag[c];
                for (int k = A.colptr[c]; k < A.colptr[c + 1]; ++k)
                {
                    const int r = A.rowidx[k];
                    if (fixed[r])
                        continue;
                    if (pending && r > c)
                    {
                        out.rowidx[o] = c;
                        out.vals[o++] = 0.0;
                        pending = false;
                    }
                    out.rowidx[o] = r;
                    out.vals[o] = A.vals[k];
                    out.keep[k] = o++;
                }
                if (pending)
                {
                    out.rowidx[o] = c;
                    out.vals[o] = 0.0;
                }
            }
        });
        // Couplings a_ij (i free, j fixed) live in the fixed columns; bucket them by row, columns
        // ascending within a row.
        std::vector<int> count(n, 0);
        for (int j = 0; j < n; ++j)
            if (fixed[j])
                for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k)
                    if (!fixed[A.rowidx[k]])
                        ++count[A.rowidx[k]];
        out.crow.clear();
        out.cptr.assign(1, 0);
        std::vector<int> pos(n, -1);
        for (int i = 0; i < n; ++i)
            if (count[i] > 0)
            {
                pos[i] = out.cptr.back();
                out.crow.push_back(i);
                out.cptr.push_back(out.cptr.back() + count[i]);
            }
        out.ccol.resize(static_cast<size_t>(out.cptr.back()));
        out.cidx.resize(static_cast<size_t>(out.cptr.back()));
        for (int j = 0; j < n; ++j)
            if (fixed[j])
                for (int k = A.colptr[j]; k < A.colptr[j + 1]; ++k)
                {
                    const int i = A.rowidx[k];
                    if (fixed[i])
                        continue;
                    out.ccol[pos[i]] = j;
                    out.cidx[pos[i]++] = k;
                }
    }

    bool same_ints(const int *a, const int *b, size_t n)
    {
        constexpr size_t CHUNK = 4096; // enough chunks for the parallel split
        const int chunks = static_cast<int>((n + CHUNK - 1) / CHUNK);
        std::atomic<bool> same{true};
        parallel_for_blocks(chunks, [&](int c0, int c1) {
            for (int c = c0; c < c1 && same; ++c)
            {
                const size_t lo = static_cast<size_t>(c) * CHUNK, hi = std::min(n, lo + CHUNK);
                if (std::memcmp(a + lo, b + lo, (hi - lo) * sizeof(int)) != 0)
                    same = false;
            }
        });
        return same;
    }

    bool refresh_padded_bsr_values(const CscView &A, PaddedBsr &bsr)
    {
        const int n = A.n, bd = bsr.bd, nb = bsr.nb, bs = bd * bd;
        if (div_up(n, bd) != nb)
            return false;
        std::fill(bsr.vals.begin(), bsr.vals.end(), 0.0);
        std::atomic<bool> ok{true};
        // Parallel over block columns: distinct block columns write distinct blocks.
        parallel_for_blocks(nb, [&](int b0, int b1) {
            for (int cb = b0; cb < b1; ++cb)
            {
                const int C = bsr.slot_of_block[cb];
                for (int c = cb * bd; c < std::min(n, cb * bd + bd); ++c)
                    for (int k = A.colptr[c]; k < A.colptr[c + 1]; ++k)
                    {
                        const int r = A.rowidx[k];
                        const int R = bsr.slot_of_block[r / bd];
                        const int *beg = bsr.cols.data() + bsr.rows[R];
                        const int *end = bsr.cols.data() + bsr.rows[R + 1];
                        const int *it = std::lower_bound(beg, end, C);
                        if (it == end || *it != C)
                        {
                            ok = false;
                            continue;
                        }
                        const size_t pos = static_cast<size_t>(it - bsr.cols.data());
                        bsr.vals[pos * bs + (r % bd) * bd + (c % bd)] += A.vals[k];
                    }
            }
        });
        const int padded_scalars = bd * nb - n;
        if (padded_scalars > 0)
        {
            const int T = bsr.slot_of_block[nb - 1];
            const int *beg = bsr.cols.data() + bsr.rows[T];
            const int *end = bsr.cols.data() + bsr.rows[T + 1];
            const int *it = std::lower_bound(beg, end, T);
            if (it == end || *it != T)
                return false;
            const size_t pos = static_cast<size_t>(it - bsr.cols.data());
            for (int t = 0; t < padded_scalars; ++t)
            {
                const int local = (bd - padded_scalars) + t;
                bsr.vals[pos * bs + local * bd + local] = 1.0;
            }
        }
        return ok;
    }
    bool build_symmetric_sell(int npad, const std::vector<int> &slice_ptr, const std::vector<int> &cols,
                              const std::vector<unsigned short> &lcols, const std::vector<int> &row_len, SymSell &out)
    {
        constexpr int NB = 256, W = 64, SL = NB / W;
        constexpr uint32_t NONE = 0x7FFFu, FOREIGN = 1u << 31;
        out = SymSell();
        if (npad % NB != 0 || static_cast<int>(slice_ptr.size()) != npad / W + 1 ||
            static_cast<int>(row_len.size()) != npad || lcols.size() != cols.size())
        {
            out.reason = "inconsistent SELL arrays";
            return false;
        }
        if (cols.size() > static_cast<size_t>(INT_MAX))
        {
            out.reason = "too many slots for 32-bit slot indices";
            return false;
        }
        const int nblk = npad / NB;
        struct Ent
        {
            int col; ///< global column (ordering)
            int src, mirror;
            uint32_t meta;
        };
        struct Block
        {
            int len[SL] = {0, 0, 0, 0};
            int incoming = 0;
            std::vector<Ent> ent;              ///< position p's entries: ent[off[p] .. off[p + 1])
            std::vector<int> off;
            std::vector<uint32_t> info;
        };
        std::vector<Block> blocks(static_cast<size_t>(nblk));
        std::atomic<bool> ok{true};
        std::string reason;
        std::atomic<int> failed_block{-1};
        auto work = [&](int b0, int b1) {
            // Per row of the block: diagonal slot, cross-block slots, intra-block (column, slot).
            std::vector<std::pair<int, size_t>> intra[NB];
            std::vector<size_t> cross[NB];
            size_t diag[NB];
            bool real_diag[NB];
            int len[NB], cnt[NB], beg[NB];
            struct Own
            {
                int other;
                size_t src, mirror;
                int rank;
            };
            std::vector<Own> owned[NB];
            struct Foreign
            {
                int row;
                size_t idx;
                int rank;
            };
            std::vector<Foreign> foreign[NB], pool;
            std::vector<Ent> rest;
            for (int b = b0; b < b1 && ok; ++b)
            {
                const int r0 = b * NB;
                bool fine = true;
                for (int lr = 0; lr < NB && fine; ++lr)
                {
                    intra[lr].clear();
                    cross[lr].clear();
                    owned[lr].clear();
                    const int r = r0 + lr, s = r / W, lane = r % W;
                    const int sp = slice_ptr[s], L = row_len[r], slots = slice_ptr[s + 1] - sp;
                    diag[lr] = SIZE_MAX;
                    real_diag[lr] = false;
                    for (int k = 0; k < L; ++k)
                    {
                        const size_t idx = static_cast<size_t>(sp + k) * W + lane;
                        const int c = cols[idx];
                        if (c == r)
                        {
                            diag[lr] = idx;
                            real_diag[lr] = true;
                        }
                        else if (c / NB == b)
                            intra[lr].emplace_back(c - r0, idx);
                        else
                            cross[lr].push_back(idx);
                    }
                    if (diag[lr] == SIZE_MAX)
                    {
                        // A row without a diagonal block: its first padding slot (own column,
                        // zero values) stands in for it.
                        if (L < slots)
                            diag[lr] = static_cast<size_t>(sp + L) * W + lane;
                        else
                            fine = false;
                    }
                    len[lr] = 1 + static_cast<int>(cross[lr].size());
                    cnt[lr] = 0;
                }
                if (!fine)
                {
                    ok = false;
                    failed_block = b;
                    break;
                }
                // Pairs (lr < c) in row order; the mirror (c, lr) must exist.
                for (int lr = 0; lr < NB && fine; ++lr)
                    for (const auto &e : intra[lr])
                    {
                        const int c = e.first;
                        if (c < lr)
                            continue;
                        const auto &row = intra[c];
                        const auto it = std::lower_bound(row.begin(), row.end(), std::make_pair(lr, size_t(0)),
                                                         [](const std::pair<int, size_t> &a,
                                                            const std::pair<int, size_t> &v) { return a.first < v.first; });
                        if (it == row.end() || it->first != lr)
                        {
                            fine = false;
                            break;
                        }
                        const bool mine = len[lr] <= len[c];
                        const int owner = mine ? lr : c, other = mine ? c : lr;
                        const size_t src = mine ? e.second : it->second, mir = mine ? it->second : e.second;
                        ++len[owner];
                        owned[owner].push_back({other, src, mir, cnt[other]++});
                    }
                // Every intra entry of a row must have been paired (a structurally symmetric block).
                for (int lr = 0; lr < NB && fine; ++lr)
                {
                    int paired = static_cast<int>(owned[lr].size()) + cnt[lr];
                    if (paired != static_cast<int>(intra[lr].size()) || cnt[lr] > 255)
                        fine = false;
                }
                if (!fine)
                {
                    ok = false;
                    failed_block = b;
                    break;
                }
                // Rows longer than the block's mean length hand blocks of theirs that couple to other
                // node blocks to the shortest rows, which compute those products for them and
                // deliver them through the incoming slots (foreign entries).
                int sum = 0;
                for (int lr = 0; lr < NB; ++lr)
                {
                    sum += len[lr];
                    foreign[lr].clear();
                }
                const int T = (sum + NB - 1) / NB;
                pool.clear();
                for (int lr = 0; lr < NB; ++lr)
                    while (len[lr] > T && !cross[lr].empty() && cnt[lr] < 255)
                    {
                        pool.push_back({lr, cross[lr].back(), cnt[lr]++});
                        cross[lr].pop_back();
                        --len[lr];
                    }
                {
                    int rorder[NB];
                    std::iota(rorder, rorder + NB, 0);
                    std::stable_sort(rorder, rorder + NB, [&](int a, int c) { return len[a] < len[c]; });
                    size_t q = 0;
                    for (int i = 0; i < NB && q < pool.size(); ++i)
                        for (const int lr = rorder[i]; len[lr] < T && q < pool.size(); ++len[lr])
                            foreign[lr].push_back(pool[q++]);
                    for (int i = 0; q < pool.size(); i = (i + 1) % NB) // (sum <= NB * T: not reached)
                    {
                        foreign[rorder[i]].push_back(pool[q++]);
                        ++len[rorder[i]];
                    }
                }
                int total = 0;
                for (int lr = 0; lr < NB; ++lr)
                {
                    beg[lr] = total;
                    total += cnt[lr];
                }
                if (total > static_cast<int>(NONE) - 1)
                {
                    ok = false;
                    failed_block = b;
                    break;
                }
                Block &B = blocks[b];
                B.incoming = total;
                B.ent.clear();
                B.off.assign(NB + 1, 0);
                B.info.assign(NB, 0);
                int order[NB];
                std::iota(order, order + NB, 0);
                std::stable_sort(order, order + NB, [&](int a, int c) { return len[a] > len[c]; });
                for (int p = 0; p < NB; ++p)
                {
                    const int lr = order[p];
                    const uint32_t self = lcols[diag[lr]];
                    B.ent.push_back({r0 + lr, real_diag[lr] ? static_cast<int>(diag[lr]) : -1, -1, self | (NONE << 16)});
                    rest.clear();
                    for (const Own &o : owned[lr])
                        rest.push_back({r0 + o.other, static_cast<int>(o.src), static_cast<int>(o.mirror),
                                        static_cast<uint32_t>(lcols[o.src]) |
                                            (static_cast<uint32_t>(beg[o.other] + o.rank) << 16)});
                    for (size_t idx : cross[lr])
                        rest.push_back({cols[idx], static_cast<int>(idx), -1, lcols[idx] | (NONE << 16)});
                    for (const Foreign &f : foreign[lr])
                        rest.push_back({cols[f.idx], static_cast<int>(f.idx), -1,
                                        static_cast<uint32_t>(lcols[f.idx]) |
                                            (static_cast<uint32_t>(beg[f.row] + f.rank) << 16) | FOREIGN});
                    std::sort(rest.begin(), rest.end(), [](const Ent &a, const Ent &c) { return a.col < c.col; });
                    B.ent.insert(B.ent.end(), rest.begin(), rest.end());
                    B.off[p + 1] = static_cast<int>(B.ent.size());
                    B.info[p] = static_cast<uint32_t>(lr) | (static_cast<uint32_t>(cnt[lr]) << 8) |
                                (static_cast<uint32_t>(beg[lr]) << 16);
                    B.len[p / W] = std::max(B.len[p / W], B.off[p + 1] - B.off[p]);
                }
            }
        };
        {
            const int nt = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
            std::vector<std::thread> th;
            const int chunk = div_up(nblk, nt);
            for (int t = 0; t < nt; ++t)
            {
                const int b = t * chunk, e = std::min(nblk, b + chunk);
                if (b < e)
                    th.emplace_back(work, b, e);
            }
            for (auto &x : th)
                x.join();
        }
        if (!ok)
        {
            out.reason = "node block " + std::to_string(failed_block.load()) +
                         ": pattern not symmetric within the block, a row without a diagonal slot, or too many "
                         "transposed products";
            return false;
        }
        out.slice_ptr.assign(static_cast<size_t>(npad / W) + 1, 0);
        for (int b = 0; b < nblk; ++b)
        {
            out.incoming_max = std::max(out.incoming_max, blocks[b].incoming);
            for (int q = 0; q < SL; ++q)
                out.slice_ptr[b * SL + q + 1] = out.slice_ptr[b * SL + q] + std::max(1, blocks[b].len[q]);
        }
        out.slots = out.slice_ptr.back();
        for (int b = 0; b < nblk; ++b)
            for (const Ent &e : blocks[b].ent)
                out.stored += e.src >= 0 ? 1 : 0;
        const size_t n = static_cast<size_t>(out.slots) * W;
        out.meta.assign(n, 0);
        out.src.assign(n, -1);
        out.mirror.assign(n, -1);
        out.row_info.assign(static_cast<size_t>(npad), 0);
        auto fill = [&](int b0, int b1) {
            for (int b = b0; b < b1; ++b)
            {
                const Block &B = blocks[b];
                for (int p = 0; p < NB; ++p)
                {
                    const int s = b * SL + p / W, lane = p % W;
                    const int sp = out.slice_ptr[s], slots = out.slice_ptr[s + 1] - sp;
                    const Ent *E = B.ent.data() + B.off[p];
                    const int ne = B.off[p + 1] - B.off[p];
                    for (int k = 0; k < slots; ++k)
                    {
                        const size_t at = static_cast<size_t>(sp + k) * W + lane;
                        if (k < ne)
                        {
                            out.meta[at] = E[k].meta;
                            out.src[at] = E[k].src;
                            out.mirror[at] = E[k].mirror;
                        }
                        else
                            out.meta[at] = (E[0].meta & 0xFFFFu) | (NONE << 16); // zero block, own column
                    }
                    out.row_info[static_cast<size_t>(b) * NB + p] = B.info[p];
                }
            }
        };
        {
            const int nt = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
            std::vector<std::thread> th;
            const int chunk = div_up(nblk, nt);
            for (int t = 0; t < nt; ++t)
            {
                const int b = t * chunk, e = std::min(nblk, b + chunk);
                if (b < e)
                    th.emplace_back(fill, b, e);
            }
            for (auto &x : th)
                x.join();
        }
        return true;
    }
} // namespace psmas
