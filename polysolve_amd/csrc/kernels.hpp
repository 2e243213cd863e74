// Launch interface of the MI355X MAS kernels.  Every launcher enqueues on the given stream and
// returns immediately.
#pragma once

#include "lattice.hpp"

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace psmas
{
    /// Nodes per workgroup of the node-blocked kernels.  Every system of a batch occupies a whole
    /// number of these blocks, so a workgroup never straddles two systems.
    constexpr int NODE_BLOCK = 256;

    /// SELL-64 block matrix: rows (bank slots) in slices of 64; slot k of lane l holds block k of
    /// row 64*s + l.  cols[(slice_ptr[s] + k) * 64 + l]; vals[((slice_ptr[s] + k) * bd*bd + e) * 64 + l].
    /// Short rows are padded with zero blocks whose column is the row itself.
    template <typename TV>
    struct SellView
    {
        int nslices = 0;
        const int *slice_ptr = nullptr;
        const int *cols = nullptr;
        const TV *vals = nullptr;
        int unroll = 4; ///< slots per batch in the SpMV row loop: 2, 4 or 8 (autotuned)
        /// Halo form of the columns: per node block (256 rows, one SpMV workgroup) the distinct
        /// columns its slots read, ascending (halo[halo_ptr[b] .. halo_ptr[b+1])), and every slot's
        /// column as a 16-bit index into its block's list (lcols, same slot layout as cols).
        /// halo_max = the longest list; 0 = not in use (x is gathered through cols).
        const unsigned short *lcols = nullptr;
        const int *halo_ptr = nullptr;
        const int *halo = nullptr;
        int halo_max = 0;
    };

    /// Longest per-workgroup halo the SpMV stages in LDS (nodes); larger sets fall back to
    /// gathering x through the 32-bit columns.
    constexpr int SPMV_HALO_CAP = 2048;

    /// Restriction of fine (bank-slot) values to coarse level 0: slot p that leads its bank-local
    /// connected component carries the component's member mask (lanes of its bank) and the
    /// component's global id.
    struct FineRestrict
    {
        const uint32_t *lead = nullptr;
        const int *gid = nullptr;
    };

    /// Per-system device scalars (double): the division guard set by the host, and the totals
    /// of the PCG reductions published by launch_finalize for the host's convergence checks.
    enum Scalar : int
    {
        S_RZ = 0,      ///< r . z (current)
        S_RZ_OLD = 1,  ///< r . z of the previous iteration
        S_PAP = 2,     ///< p . A p
        S_RR = 3,      ///< r . r of the recurrence residual
        S_RR_TRUE = 4, ///< r . r of an explicitly computed residual b - A x (FP64)
        S_GUARD = 5,   ///< division guard (see safe_div)
        S_COUNT = 8,
    };

    /// Reduction sites (one partials array of node-block length each).
    enum Site : int
    {
        R_PAP = 0,
        R_RR = 1,
        R_RZ = 2,
        R_RR_TRUE = 3,
        R_RZ_OLD = 4,
        R_DLT = 5,    ///< single-reduction PCG: u . A u
        R_EPS = 6,    ///< single-reduction PCG: u . s_prev
        R_PS = 7,     ///< single-reduction PCG: p . s of the update
        R_PS_OLD = 8, ///< single-reduction PCG: the previous p . s
        R_COUNT = 9,
    };

    /// Per-slot lattice data and the coarse coefficients, as the prolongation kernel reads them.
    struct LccApply
    {
        const int4 *cell = nullptr; ///< {corner-0 lattice point (global) or -1, fx, fy, fz as float bits}
        const float *coef = nullptr; ///< [lattice point][bd] coarse correction
        int k = 0;                   ///< lattice points per axis
    };

    /// Flexible-PCG extras for the prolongation kernel: r . z_prev partials.
    template <typename T>
    struct FlexDot
    {
        const T *zprev = nullptr;
        double *rzo = nullptr; ///< per-node-block partials of r . z_prev
    };

    /// The PCG reductions.  Producers write one partial sum per node block (workgroup); every
    /// consumer re-sums the partials of its own system in one fixed order, so all workgroups see
    /// bit-identical alpha / beta without atomics or cross-workgroup fences.  sc holds S_COUNT
    /// doubles per system: the division guard (input) and the totals that launch_finalize
    /// publishes for the host.
    struct Reductions
    {
        double *pap = nullptr;     ///< p . A p
        double *rr = nullptr;      ///< r . r of the recurrence residual
        double *rz = nullptr;      ///< r . z (current)
        double *rz_old = nullptr;  ///< r . z of the previous iteration
        double *rr_true = nullptr; ///< r . r of an explicitly computed residual
        double *sc = nullptr;
    };

    /// Systems of a batch (a single solve is a batch of one): node block -> system, each
    /// system's block range, and per-system flags written by the host between graph launches.
    struct SysMap
    {
        const int *block_sys = nullptr; ///< node block -> system
        const int *blk_begin = nullptr; ///< system -> first node block
        const int *blk_count = nullptr; ///< system -> node blocks
        const int *active = nullptr;    ///< system -> 1 while it iterates
        const int *select = nullptr;    ///< system -> 1 when picked for a reliable update
        const int *bank_sys = nullptr;  ///< fine bank -> system
        /// > 0: one system (not a batch) of this many node blocks.  Its kernels are launched only
        /// while it iterates, so they skip the block -> system and activity lookups (two dependent
        /// loads ahead of all other work) and its block range needs no lookup either.
        int single_blocks = 0;
    };

    // ----- SpMV ---------------------------------------------------------------------------
    /// y = A x and the partials of x . y -> red.pap.  TV matrix value type, TX vector type.
    template <int BD, typename TV, typename TX>
    void launch_spmv_ap(const SellView<TV> &A, const TX *x, TX *y, const Reductions &red, const SysMap &sm,
                        hipStream_t s);

    /// The update of the search direction folded into the SpMV that uses it: the SpMV computes
    /// p = z + beta p_old itself (beta from the partials of r . z as launch_update_p would, or
    /// p = z with init) for every node it reads, multiplies A p and writes p to p_new for its own
    /// rows (p_old and p_new must differ: other workgroups still read p_old).  Flexible PCG (rzo
    /// set): beta = r . (z - z_prev) / r_old . z_old, and zprev receives z.  Bit-identical to
    /// launch_update_p followed by launch_spmv_ap.  Needs the halo form of the matrix.
    template <typename TX>
    struct FusedDirection
    {
        const TX *z = nullptr;
        const TX *p_old = nullptr;
        TX *p_new = nullptr;
        TX *zprev = nullptr;
        const double *rzo = nullptr;
        int init = 0;
    };

    template <int BD, typename TV, typename TX>
    void launch_spmv_ap_fused(const SellView<TV> &A, const FusedDirection<TX> &fd, TX *y, const Reductions &red,
                              const SysMap &sm, hipStream_t s);

    /// Symmetric form of a SELL-64 matrix with halo lists: a block coupling two rows of the same
    /// node block is stored once, in the row chosen as its owner, which also applies its
    /// transpose to the other row through the workgroup's LDS.  Per node block the 256 rows are
    /// placed in the four slices by decreasing stored length (row_info maps a slice position to
    /// its row).  Slot k of position (s, l): meta = local column | incoming slot << 16 (0x7FFF:
    /// none) | foreign << 31 and the block's values, laid out as SellView's; slot 0 of every row
    /// is its diagonal block.  A foreign slot holds a block of another row of the node block
    /// (balancing the rows' lengths) and sends its product there.  row_info = local row |
    /// incoming count << 8 | first incoming slot << 16; the incoming slots of a row hold the
    /// transposed and foreign products addressed to it, in a fixed order.
    template <typename TV>
    struct SymView
    {
        int nslices = 0;
        const int *slice_ptr = nullptr;
        const uint32_t *meta = nullptr;
        const TV *vals = nullptr;
        const uint32_t *row_info = nullptr;
        int incoming_max = 0; ///< longest incoming list of a node block (slots)
        int unroll = 4;
    };

    /// y = A x and the partials of x . y -> red.pap, as launch_spmv_ap, from the symmetric form
    /// S of A (the halo lists are A's).  The sums run in another order than launch_spmv_ap's.
    template <int BD, typename TV, typename TX>
    void launch_spmv_sym(const SellView<TV> &A, const SymView<TV> &S, const TX *x, TX *y, const Reductions &red,
                         const SysMap &sm, hipStream_t s);
    /// Largest dynamic LDS launch_spmv_sym takes (bytes; of the 160 KB of a gfx950 CU).
    constexpr size_t SPMV_SYM_LDS_CAP = 120 * 1024;
    /// launch_spmv_res from the symmetric form S of A (the halo lists are A's).
    template <int BD, typename TV, typename TX, typename TY>
    void launch_spmv_res_sym(const SellView<TV> &A, const SymView<TV> &S, const TX *x, const TX *b, TY *y,
                             FineRestrict fr, TY *yc0, double *rr_partials, const SysMap &sm, bool selected_only,
                             hipStream_t s);
    /// Dynamic LDS of launch_spmv_sym: the halo of x, then BD incoming arrays.
    template <int BD, typename TV, typename TX>
    size_t spmv_sym_lds(int halo_max, int incoming_max);
    /// Values of the symmetric form: entry e of the sym slots takes the full matrix's block
    /// src[e] (an index slot * 64 + lane of vals64; < 0: zero), converted to TV.
    template <typename TV>
    void launch_sym_gather(int bd, long long entries, const int *src, const double *vals64, TV *out, hipStream_t s);
    /// flag = 1 if for some entry e with mirror[e] >= 0 the blocks src[e] and mirror[e] of vals64
    /// are not each other's transpose within rel_tol of the pair's largest entry.  flag must be
    /// zero before the launch.
    void launch_sym_check(int bd, long long entries, const int *src, const int *mirror, const double *vals64,
                          double rel_tol, int *flag, hipStream_t s);

    /// y = b - A x, partials of y . y -> rr_partials; restrict y -> yc0.  TX: x,b; TY: y.
    /// selected_only: only systems with select == 1.
    template <int BD, typename TV, typename TX, typename TY>
    void launch_spmv_res(const SellView<TV> &A, const TX *x, const TX *b, TY *y, FineRestrict fr, TY *yc0,
                         double *rr_partials, const SysMap &sm, bool selected_only, hipStream_t s);

    /// Plain y = A x without reductions (introspection path).
    template <int BD, typename TV, typename TX>
    void launch_spmv_plain(const SellView<TV> &A, const TX *x, TX *y, hipStream_t s);
    /// launch_spmv_ap for a batch of `copies` (<= 8) systems that repeat one matrix of npad1 slots:
    /// each block of copy 0's slots is read once for all copies.
    template <int BD, typename TV, typename TX>
    void launch_spmv_ap_copies(const SellView<TV> &A, int copies, int npad1, const TX *x, TX *y, const Reductions &red,
                               const SysMap &sm, hipStream_t s);

    // ----- single-reduction PCG (Chronopoulos & Gear 1989), mixed path ---------------------
    /// Partials of one iteration of the single-reduction PCG.  The SpMV kernel writes
    /// gamma = r . u (rz), delta = u . A u (dlt), eps = u . s_prev (eps) and, flexible, zeta =
    /// r . u_prev (zeta), after moving its block's previous rz and ps to rz_old and ps_old; the
    /// update kernel writes r . r (rr) and p . s (ps).
    struct CgReductions
    {
        double *rz = nullptr, *rz_old = nullptr, *dlt = nullptr, *eps = nullptr;
        double *ps = nullptr, *ps_old = nullptr, *zeta = nullptr, *rr = nullptr;
        double *sc = nullptr;
    };

    /// The prolongation the single-reduction SpMV performs where it reads its input:
    /// u_p = z_p + sum_l z_l[cmap_p.l] (+ P c), in FP64 and in level order as
    /// launch_prolong_dot, from the multi-level z (mz).  own[b]: position of node block b's
    /// first own node in its halo list (every own node is in it, in order).
    struct CgProlong
    {
        const float *mz = nullptr;
        const int4 *cmap = nullptr;
        const int *lev_off = nullptr;
        int level_num = 0;
        const int *own = nullptr;
        LccApply lcc;
    };

    /// u = prolongation of z (CgProlong) at every node the block's rows read, w = A u (FP32 SELL,
    /// halo form), u written for the own rows; partials gamma = r . u, delta = w . u and, unless
    /// init, eps = u . s_prev and (u_prev set) zeta = r . u_prev.
    template <int BD>
    void launch_spmv_cg(const SellView<float> &A, const CgProlong &pr, float *u, float *w, const float *r,
                        const float *s_prev, const float *u_prev, const CgReductions &red, const SysMap &sm,
                        bool init, hipStream_t s);

    /// beta = gamma / gamma_old (flexible, u_prev set: (gamma - zeta) / gamma_old; init: 0),
    /// p.s = delta + 2 beta eps + beta^2 ps_old (init: delta), alpha = gamma / p.s;
    /// p = u + beta p; s = w + beta s; x += alpha p; r -= alpha s; partials of r . r and p . s;
    /// restrict r -> rc0; u_prev = u when set; with lb the lattice block partials of r.
    template <int BD>
    void launch_update_cg(int nblocks, const float *u, const float *w, float *p, float *s_vec, float *x, float *r,
                          FineRestrict fr, float *rc0, float *u_prev, const CgReductions &red, const SysMap &sm,
                          bool init, hipStream_t s, const LatticeBlocks *lb = nullptr);

    // ----- vectors (node-blocked: nblocks workgroups of NODE_BLOCK nodes) ------------------
    /// alpha = rz / pap; x += alpha p; r -= alpha Ap; partials of r . r -> red.rr; restrict r -> rc0;
    /// with lb, also the lattice block partials of the new r (learned correction).
    template <int BD, typename T>
    void launch_update_xr(int nblocks, const T *p, const T *Ap, T *x, T *r, FineRestrict fr, T *rc0,
                          const Reductions &red, const SysMap &sm, hipStream_t s, const LatticeBlocks *lb = nullptr);

    /// x += alpha p (alpha = rz / pap).
    template <int BD, typename T>
    void launch_axpy_alpha(int nblocks, const T *p, T *x, const Reductions &red, const SysMap &sm, hipStream_t s);

    /// p = z + beta p (beta = rz / rz_old).  Flexible form when rzo is set: beta =
    /// (rz - rzo) / rz_old, and z is copied to zprev.
    template <int BD, typename T>
    void launch_update_p(int nblocks, const T *z, T *p, const Reductions &red, const SysMap &sm, hipStream_t s,
                         const double *rzo = nullptr, T *zprev = nullptr);

    /// out = x64 + x32 (snapshot of the mixed path's iterate).
    void launch_snapshot(long long n, const double *x64, const float *x32, double *out, hipStream_t s);
    /// out = a - b.
    void launch_difference(long long n, const double *a, const double *b, double *out, hipStream_t s);

    /// Per-system totals of every reduction -> sc (S_PAP, S_RR, S_RZ, S_RZ_OLD, S_RR_TRUE) for the host.
    void launch_finalize(int nsys, const Reductions &red, const SysMap &sm, hipStream_t s);

    /// x64 += x32; x32 = 0 for systems selected for a reliable update.
    template <int BD>
    void launch_accumulate_correction(int nblocks, float *x32, double *x64, const SysMap &sm, hipStream_t s);

    // ----- plain element-wise helpers ------------------------------------------------------
    template <typename TS, typename TD>
    void launch_copy_convert(long long n, const TS *src, TD *dst, hipStream_t s);

    /// Fill with a value (used instead of memset inside captured graphs).
    template <typename T>
    void launch_fill(long long n, T *dst, T value, hipStream_t s);

    /// dst[slot*bd + c] = src[block*bd + c] for real slots (block = block_of_slot[slot], scalar < n),
    /// zero elsewhere.
    template <typename T>
    void launch_gather_in(int npad, int bd, int n, const int *block_of_slot, const T *src, T *dst, hipStream_t s);
    /// dst[block*bd + c] = src[slot_of_block[block]*bd + c] for scalar < n.
    template <typename T>
    void launch_scatter_out(int nb, int bd, int n, const int *slot_of_block, const T *src, T *dst, hipStream_t s);

    /// dst[map[k]] = src[k] for k < n with map[k] >= 0 (injective on those; -1 = dropped entry):
    /// a values-only matrix refresh.
    void launch_scatter_values(long long n, const int *map, const double *src, double *dst, hipStream_t s);
    /// dst[k] = src[idx[k]] for k < n.
    void launch_gather_values(int n, const int *idx, const double *src, double *dst, hipStream_t s);
    /// Dirichlet right-hand side lift g = f - (I - N) A N f over the coupling rows (crow[t] with
    /// couplings cptr[t]..cptr[t+1]: values cval at fixed columns ccol); rows without couplings
    /// are left as they are in g.  g may be f (the coupled rows are free, the columns read fixed).
    void launch_dirichlet_lift(int nrows, const int *crow, const int *cptr, const int *ccol, const double *cval,
                               const double *f, double *g, hipStream_t s);

    /// sc[s][slot] = value for systems s < count.
    void launch_set_scalar(double *sc, int count, int slot, double value, hipStream_t s);

    // ----- MAS apply ----------------------------------------------------------------------
    /// y_m = Inv_m x_m for matrices [mat_begin, mat_begin + count) of the multi-level vectors.
    /// FP64 storage: packed upper row-major.  FP16 storage is scaled, Inv = D S D with D per row
    /// (dscale), and S is tiled in 8 x 8 blocks with an implicit unit diagonal (k_make_half).
    /// bank_sys/active (may be null): skip matrices of systems that stopped iterating
    /// (bank_sys[m] = system of matrix m, -1 for coarse matrices, which are never skipped).
    /// Longest descendant list of a restriction task that takes one thread; longer lists are cut
    /// into pieces of at most RESTRICT_PIECE entries that take a wave each.
    constexpr int RESTRICT_SMALL_MAX = 8;
    // Network layer 1 in the second apply launch: at most LIN_WGS_MAX workgroups, each lane holds
    // LIN_CHUNKS 16-byte chunks of a weight row (rows up to 64 * 8 * LIN_CHUNKS wide) and each
    // thread LIN_U_PER_THREAD values of u.
    constexpr int LIN_WGS_MAX = 768;
    constexpr int LIN_CHUNKS = 4;
    constexpr int LIN_U_PER_THREAD = 16;
    constexpr int RESTRICT_PIECE = 512;
    /// Threads per workgroup of the FP16 apply of one right-hand side (a wave per matrix, four
    /// matrices per workgroup, k_mas_apply_w16).
    constexpr int APPLY_W16_BLOCK = 256;
    constexpr int APPLY_W16_WAVES = APPLY_W16_BLOCK / 64;

    /// Work done by the leading `wgs` workgroups of an apply launch beside the bank matrices (the
    /// other workgroups apply the matrices).  Either part may be empty.
    struct ApplySide
    {
        int wgs = 0;
        int apply_cap = 0; ///< > 0: at most this many matrix workgroups (host launch choice only)
        /// FP16 storage and one right-hand side: a wave per matrix (k_mas_apply_w16, workgroups of
        /// APPLY_W16_BLOCK threads) instead of the persistent kernel that unpacks each matrix in
        /// LDS (host launch choice only; apply_threads gives the side workgroups' size).
        bool wave_apply = false;
        // Restriction of coarse segments 2.. from segment 1: task i sums the segment-1 values of
        // rlist[rtask[i].y .. rtask[i].z) (ascending node ids) in FP64.  Tasks [0, nsmall) take a
        // thread each, [nsmall, nsmall + nbig) a wave each (lanes stride the list, fixed-order
        // reduction).  rtask[i].w < 0: the sum is multi-level node rtask[i].x's value (rounded to
        // the vector type, written to out); otherwise it is partial rtask[i].w of a node cut into
        // pieces (written to rpart; the launch that reads the node adds its pieces).
        int nsmall = 0, nbig = 0;
        const int4 *rtask = nullptr;
        const int *rlist = nullptr;
        const void *seg1 = nullptr; ///< segment-1 values, of the launch's vector type
        void *out = nullptr;        ///< multi-level vector the sums go to
        double *rpart = nullptr;    ///< [pieces][bd]
        // Inputs of the launch that reads segments 2..: nodes whose restriction was cut into pieces
        // take the sum of rpart[rcomb[i].y .. rcomb[i].y + rcomb[i].z) (in piece order) with
        // i = split_of[node - node0] >= 0; all other nodes read mr.
        const int *split_of = nullptr;
        const int4 *rcomb = nullptr;
        int node0 = 0;
        int split_count = 0; ///< entries of split_of
        /// Any readable device int (the apply kernels load from it unconditionally, in place of a
        /// split or dense index that does not exist, and ignore the value).
        const int *dummy = nullptr;
        // Dense coarsest solve y = Inv x, one wave per row; its nodes are coarsest-segment nodes
        // dense_seg + dense_ids[i].  Bank outputs of nodes with dense_index[n - dense_seg] >= 0
        // (n - dense_seg < dense_count) are left to it.
        int dense_nd = 0;
        const double *dense_inv = nullptr;
        const int *dense_ids = nullptr;
        int dense_seg = 0;
        const int *dense_index = nullptr;
        int dense_count = 0;
        // Several right-hand sides of one factorisation: copy c's coarsest segment starts at
        // dense_seg + c dense_stride; one pass over each inverse row serves every copy.
        int dense_copies = 1;
        int dense_stride = 0;
        // Layer 1 and the skip layer of the learned correction's network for a single system, done
        // by side workgroups [lin_wg0, wgs) (the others do the work above): each stages
        // x = u sqrt(m) / |u| (zero-padded to lin_md) in LDS; waves take rows (row < lin_h:
        // h1 = relu(W1 x + b1), else skip = S x); workgroup lin_wg0 writes
        // scale = |u| lin_ascale / sqrt(m).  Skipped when lin_active[0] == 0.
        int lin_rows = 0; ///< lin_h + lin_md (0: none)
        int lin_wg0 = 0;
        int lin_m = 0, lin_md = 0, lin_h = 0;
        const double *lin_u = nullptr;
        const unsigned short *lin_w1 = nullptr, *lin_s = nullptr;
        const float *lin_b1 = nullptr;
        float *lin_h1 = nullptr, *lin_skip = nullptr, *lin_scale = nullptr;
        float lin_ascale = 0.0f;
        const int *lin_active = nullptr;
        // Lattice gather of the learned correction: u[J] = sum of lat_partial[lat_src[i]] over
        // i in [lat_ptr[J], lat_ptr[J + 1]), J < lat_npts (wave per point, fixed-order reduction).
        int lat_npts = 0;
        const int *lat_ptr = nullptr;
        const int *lat_src = nullptr;
        const double *lat_partial = nullptr;
        double *lat_u = nullptr;
    };

    /// Right-hand sides that share one factorisation (factorize_copies): the bank matrices are
    /// stored once, in the single system's order (level l's cnt[l] matrices from soff[l]); the
    /// multi-level vectors hold every copy level by level, copy c's banks of level l being
    /// moff[l] + c cnt[l] .. + cnt[l] - 1.  Levels: the fine banks, then each coarse level.
    /// copies == 1: one system, matrices and banks coincide.
    struct CopyLayout
    {
        static constexpr int MAXL = 5; ///< fine + 4 coarse levels
        int copies = 1;
        int levels = 0;
        int moff[MAXL + 1] = {}; ///< first bank of each level (all copies), then the total
        int cnt[MAXL] = {};      ///< banks of each level per copy
        int soff[MAXL + 1] = {}; ///< first stored matrix of each level, then the total
    };

    template <int BD, typename TM, typename T>
    /// cl.copies > 1: the FP16 apply takes [mat_begin, mat_begin + count) as stored matrices and
    /// applies each to every copy's bank (one matrix read for all right-hand sides); the FP64
    /// apply takes them as banks and reads each bank's stored matrix.
    void launch_mas_apply(int mat_begin, int count, const TM *mats, const float *dscale, const T *mr, T *mz,
                          const int *bank_sys, const int *active, hipStream_t s, const ApplySide &side = ApplySide{},
                          const CopyLayout &cl = CopyLayout{});

    /// Best streaming-read rate (GB/s) over `reps` passes of `bytes` of device memory
    /// (16-byte nontemporal loads, grid-stride); allocates and frees its buffer.
    double read_bandwidth_probe(long long bytes, int reps);

    /// Elements of the vector type the FP16 apply kernel's side workgroups can stage in LDS
    /// (wave: the wave-per-matrix kernel of one right-hand side).
    int apply_h16_stage_floats(int bd, bool wave);
    /// Threads per workgroup of an apply launch (side workgroups included).
    int apply_threads(int bd, bool wave);


    /// z_p = mz_p + sum_l mz_l[cmap_p.l] (+ the lattice correction P c when lcc.cell is set);
    /// partials of r . z -> red.rz (the previous ones move to red.rz_old unless init) and, when
    /// flex.zprev is set, of r . z_prev -> flex.rzo.
    template <int BD, typename T>
    void launch_prolong_dot(int nblocks, int level_num, const int4 *cmap, const int *lev_off, const T *r, T *mz,
                            const Reductions &red, bool init, const SysMap &sm, const LccApply &lcc,
                            const FlexDot<T> &flex, hipStream_t s);

    // ----- MAS build ----------------------------------------------------------------------
    /// Bank-local connected components at coarse level 0 (one lane per slot, two banks per wave).
    void launch_cc_level0(int npad, const SellView<double> &A, const int *row_len, const int *block_of_slot,
                          uint32_t *lead, int *local_id, int *bank_count, hipStream_t s);
    /// Neighbour masks among level-(lv-1) ids (atomicOr into nbr).
    void launch_level_masks(int npad, int lv, const SellView<double> &A, const int *row_len, const int *block_of_slot,
                            const int4 *cmap, uint32_t *nbr, hipStream_t s);
    /// Components among `count` level-(lv-1) nodes grouped in banks of 32.
    void launch_cc_coarse(int count, const uint32_t *nbr, uint32_t *lead, int *local_id, int *bank_count,
                          hipStream_t s);
    /// Exclusive scan of n ints (single workgroup); total -> out[n].
    void launch_scan(int n, const int *in, int *out, hipStream_t s);
    /// gid[q] = local[q] + prefix[q / 32] for q < count.
    void launch_make_gid(int count, const int *local_id, const int *prefix, int *gid, hipStream_t s);
    /// cmap[p].lv = (lv == 0) ? gid0[p] : gid_lv[cmap[p].(lv-1)] for real slots.
    void launch_fill_cmap(int npad, int lv, const int *block_of_slot, const int *gid, int4 *cmap, hipStream_t s);
    /// Galerkin bank matrices (packed upper FP64), all levels, from the fine matrix.
    void launch_galerkin(int npad, int bd, int level_num, const int *mat_off, const SellView<double> &A,
                         const int *row_len, const int *block_of_slot, const int4 *cmap, double *mats, hipStream_t s);
    /// In-place batched inverse of packed SPD matrices (zero diagonal -> 1 first).  Sets *fail on
    /// a zero / non-finite pivot or a non-positive inverse diagonal.
    void launch_invert(int bd, int count, double *mats, int *fail, hipStream_t s);
    /// Scaled FP16 form of packed inverses: Inv = D S D, S packed FP16, D FP32 per row.
    void launch_make_half(int bd, int count, const double *mats, __half *out, float *dscale, hipStream_t s);

    /// Morton keys of quantised coordinates (21 bits per axis).
    void launch_morton_keys(int nb, int dim, const double *xyz, const double *lo, const double *inv_extent,
                            uint64_t *keys, int *ids, hipStream_t s);
    /// Radix sort (key, id) pairs ascending by key; temp storage managed by the caller.
    size_t morton_sort_temp_bytes(int nb);
    void morton_sort(int nb, void *temp, size_t temp_bytes, const uint64_t *keys_in, uint64_t *keys_out,
                     const int *ids_in, int *ids_out, hipStream_t s);
} // namespace psmas
