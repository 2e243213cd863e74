// Host-side setup of the MAS solver: input validation, block graph, graph aggregation into
// <=32-node domains, and the permuted, bank-padded BSR matrix.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace psmas
{
    constexpr int BANK = 32; ///< Nodes per MAS domain (bank); fixed by the reference (MASPreconditioner.cu:33).
    constexpr int MAX_COARSE_LEVELS = 4; ///< Coarse levels (MASPreconditioner.hpp:70).

    /// Compressed column-major input, the layout of polysolve's StiffnessMatrix
    /// (src/polysolve/Types.hpp:11-15) with int32 indices.
    struct CscView
    {
        int n = 0;
        long long nnz = 0;
        const int *colptr = nullptr;
        const int *rowidx = nullptr;
        const double *vals = nullptr;
    };

    /// Validate shape and index ranges; throws InvalidArgumentError ("[MAS] invalid A ...") the
    /// way the reference rejects bad input (BSRAdjacency.cu:466-481, BSRMatrix.cu:438-442).
    void validate_csc(const CscView &A);

    /// Symmetric block graph without self loops; weight of {u,v} = sum of |a| over the scalars of
    /// blocks (u,v) and (v,u) (the reference's adjacency weight, BSRAdjacency.cu:60-84, without
    /// the integer quantisation that only its external partitioner needed).
    struct BlockGraph
    {
        int nb = 0;
        std::vector<int> ptr;
        std::vector<int> adj;
        std::vector<double> w;
    };
    BlockGraph build_block_graph(const CscView &A, int bd);

    /// Domain decomposition: domain p holds permuted block ids [offsets[p], offsets[p+1]), at most
    /// 32 nodes; perm maps an input block id to its permuted id.  Mirrors the reference's
    /// part_offsets / permutation contract (MASSolver.cu:86-114,223-242).
    struct Partition
    {
        int nparts = 0;
        std::vector<int> offsets;
        std::vector<int> perm;
    };

    /// Greedy aggregation on the block graph: seeds in breadth-first order from a
    /// pseudo-peripheral node; each aggregate grows by the unassigned neighbour with the most
    /// edges into it (ties: larger weight, then earlier BFS rank) until it holds max_size nodes.
    /// Aggregates smaller than max_size/4 merge into their most strongly connected neighbour
    /// aggregate when the merged size stays <= max_size.
    Partition graph_aggregate(const BlockGraph &g, int max_size);

    /// Partition from a node order (e.g. Morton): consecutive runs of 32.
    Partition chunk_partition(const std::vector<int> &order, int max_size);

    /// Bank-padded BSR in the solver's working space: node slot = 32*part + rank in part; the
    /// part count is rounded up to a multiple of 8 so the system covers whole 256-node blocks
    /// (four SELL slices of 64 rows, each holding two whole banks).
    /// Rows of virtual slots are empty.  Blocks are row-major bd x bd, columns sorted; the
    /// trailing partial block gets 1.0 on its padded diagonal (BSRMatrix.cu:123-134,189-191).
    struct PaddedBsr
    {
        int bd = 0;
        int nb = 0;        ///< input block rows
        int nparts = 0;    ///< multiple of 8
        int npad = 0;      ///< 32 * nparts
        std::vector<int> slot_of_block; ///< input block -> slot
        std::vector<int> block_of_slot; ///< slot -> input block, -1 virtual
        std::vector<int> rows;          ///< npad + 1
        std::vector<int> cols;
        std::vector<double> vals;
    };
    void build_padded_bsr(const CscView &A, int bd, const Partition &part, PaddedBsr &out);
    /// Only the slot assignment of build_padded_bsr (bd, nb, nparts, npad, slot_of_block,
    /// block_of_slot), for the device matrix build.
    void assign_slots(int nb, int bd, const Partition &part, PaddedBsr &out);
    /// Refresh only the values of an existing pattern (same sparsity), for re-factorisation.
    bool refresh_padded_bsr_values(const CscView &A, PaddedBsr &bsr);
    /// Index of every non-zero of A in the SELL-64 FP64 value array of one system's padded BSR
    /// (rows, cols; slot_of_block of its partition; slice_ptr = slot offset of every 64-row slice):
    /// ((slice_ptr[R / 64] + position of the block in row R) * bd*bd + e) * 64 + R % 64, with R the
    /// block row's slot and e the entry's place in the block.  Returns false when two non-zeros
    /// land on one value (duplicate entries), whose values cannot be scattered.
    bool build_value_map(const CscView &A, int bd, const std::vector<int> &slot_of_block, const std::vector<int> &rows,
                         const std::vector<int> &cols, const std::vector<int> &slice_ptr, std::vector<int> &map);
    /// Dirichlet elimination (FEMSolver.cpp:123-160, prefactorize FEMSolver.cpp:269-306) with
    /// N(k) = 1 for a fixed dof k.  The eliminated matrix keeps a_ij when neither i nor j is fixed
    /// and has N(k) added on every diagonal: 1 on a fixed dof's otherwise empty row and column,
    /// a_kk + 0 on a free one (an explicit zero where A has no diagonal).  keep[k] is the position
    /// in it of A's k-th non-zero, or -1 for a dropped one.  The coupling list holds the entries
    /// a_ij of A with i free and j fixed, grouped by row (crow, cptr) in ascending column order
    /// (ccol, and cidx = their index among A's non-zeros): the lifted right-hand side is
    /// g = f - (I - N) A N f, i.e. g_i = f_i - sum over the list of a_ij f_j.
    struct DirichletSystem
    {
        std::vector<int> colptr, rowidx;
        std::vector<double> vals;
        std::vector<int> keep;
        std::vector<int> crow, cptr, ccol, cidx;
    };
    void eliminate_dirichlet(const CscView &A, const std::vector<unsigned char> &fixed, DirichletSystem &out);
    /// a[0..n) == b[0..n), compared in parallel chunks.
    bool same_ints(const int *a, const int *b, size_t n);

    /// Symmetric form of a SELL-64 matrix whose rows come in node blocks of 256 (four slices):
    /// a block coupling rows i != j of the same node block is kept in one of the two rows, its
    /// owner, which applies it to its own row and its transpose to the other one; blocks on the
    /// diagonal and blocks coupling two node blocks stay in their row.  The owner of each pair is
    /// chosen greedily (pairs in row order, the row with fewer stored blocks so far takes it);
    /// rows still longer than the node block's mean length then hand blocks that couple to other
    /// node blocks to the shortest rows, which compute those products for them (foreign
    /// entries); and each node block's rows are placed in its slices by decreasing stored length.
    /// Per slot (slot * 64 + lane of the result): meta = local column | incoming slot << 16
    /// (0x7FFF: none) | foreign << 31, src = the full matrix's slot * 64 + lane whose block it
    /// holds (-1: padding), mirror = the transposed block's (where an owned block's transpose goes
    /// to an incoming slot, else -1).  row_info per position: local row | incoming count << 8 |
    /// first incoming slot << 16.  Fails (reason set) when the pattern is not structurally
    /// symmetric within the node blocks, a row has no diagonal slot, or a node block has more
    /// than 32,766 incoming slots or a row more than 255.
    struct SymSell
    {
        std::vector<int> slice_ptr;
        std::vector<uint32_t> meta;
        std::vector<int> src, mirror;
        std::vector<uint32_t> row_info;
        int incoming_max = 0;
        long long slots = 0;
        long long stored = 0; ///< slots holding a block of the matrix (src >= 0)
        std::string reason;
    };
    /// slice_ptr (npad / 64 + 1), cols and lcols (the full matrix's 32-bit and halo-local
    /// columns, slot * 64 + lane), row_len (real blocks per row, npad).
    bool build_symmetric_sell(int npad, const std::vector<int> &slice_ptr, const std::vector<int> &cols,
                              const std::vector<unsigned short> &lcols, const std::vector<int> &row_len, SymSell &out);
} // namespace psmas
