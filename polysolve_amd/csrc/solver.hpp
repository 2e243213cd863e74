// MI355X MAS-preconditioned PCG solver: the object behind the psmas_* C ABI.
//
// Public surface mirrors polysolve::linear::MASSolver (src/polysolve/linear/MASSolver.hpp:35-72):
// set_parameters / get_info / analyze_pattern / factorize / solve / name, with the same
// semantics (MASSolver.cu:134-639), plus device-pointer solves, optional node coordinates for
// Morton aggregation, batches of independent systems solved in one launch sequence, and
// introspection used by the parity tests.
//
// Internally every factorisation is a batch of K >= 1 systems laid out one after another in the
// bank-padded node space; a single solve is K = 1.
#pragma once

#include "common.hpp"
#include "host_setup.hpp"
#include "kernels.hpp"
#include "lcc.hpp"
#include "params.hpp"

#include <functional>
#include <map>
#include <memory>
#include <future>
#include <string>
#include <chrono>
#include <vector>

namespace psmas
{
    enum class Status
    {
        Running = 0,
        ReachRelativeTolerance = 1,
        ReachAbsoluteTolerance = 2,
        ReachMaxIterations = 3,
    };

    /// Strings of MASSolver.hpp:18-33.
    const char *status_string(Status s);

    struct KernelTiming
    {
        std::string name;
        int launches = 0;
        double total_ms = 0.0;
        double bytes_per_launch = 0.0; ///< algorithmic bytes
    };

    /// CSC arrays in device memory.
    struct DeviceCsc
    {
        int n = 0;
        long long nnz = 0;
        const int *colptr = nullptr;
        const int *rowidx = nullptr;
        const double *vals = nullptr;
    };

    class Solver
    {
    public:
        Solver();
        ~Solver();
        Solver(const Solver &) = delete;
        Solver &operator=(const Solver &) = delete;

        void set_parameters_json(const std::string &json);
        std::string get_info_json() const { return system_info_json(0); }
        std::string system_info_json(int s) const;
        std::string parameters_json() const { return params_to_json(params_); }
        void analyze_pattern(const CscView &) {}
        void factorize(const CscView &A) { factorize_batch(std::vector<CscView>{A}); }
        void factorize_batch(const std::vector<CscView> &As);
        void solve_host(int n, const double *b, double *x);
        void solve_device(int n, const double *d_b, double *d_x);
        /// solve_device without any device allocation: the reduction partials and per-system
        /// scalars live in the library's static per-GPU scratch array (shared by every solver on
        /// that GPU: overlapping calls race), kernels are launched one by one, and a state that
        /// would need an allocation (parts built lazily at the first solve, pair collection,
        /// kernel timing, a system above noalloc_max_blocks() node blocks) is refused.
        void solve_device_noalloc(int n, const double *d_b, double *d_x);
        static int noalloc_max_blocks();
        void solve_batch(const std::vector<const double *> &b, const std::vector<double *> &x, bool device);
        void set_coordinates(int sys, int nb, int dim, const double *xyz);
        int num_systems() const { return static_cast<int>(sys_.size()); }

        // Introspection for tests (host arrays, original ordering unless stated; system 0).
        void get_partition(std::vector<int> &offsets, std::vector<int> &perm) const;
        std::vector<int> level_info() const;
        std::vector<int> coarse_map() const;       ///< nb x 4, per input block
        std::vector<double> bank_inverses() const; ///< packed FP64, all levels
        void apply_preconditioner(const double *r, double *z, bool mixed);
        void spmv(const double *x, double *y, bool fp32);
        /// Learned-correction lattice of system 0: number of coarse unknowns m, u = P^T v and
        /// e = P w in the input ordering.
        int lattice_size();
        void lattice_restrict(const double *v, double *u);
        void lattice_prolong(const double *w, double *e);

        // Learned coarse correction (mixed path).
        /// Train the correction network on the pairs collected so far; returns the final loss.
        double learned_train(int epochs, int batch, float lr, unsigned seed);
        /// Drop the trained network (and the collected pairs when drop_pairs).
        void learned_reset(bool drop_pairs);
        /// Replace the network by the exact lattice coarse solve of the current factorisation,
        /// c = (P^T A P)^-1 P^T r (the linear two-level method the network approximates):
        /// P^T A P is assembled column by column on the GPU, inverted on the host (Cholesky) and
        /// held in the skip layer.  Returns the wall time in ms.
        double learned_exact();

        /// Time the tile / schedule candidates on this factorisation and the last right-hand side,
        /// keep the fastest in the parameters, return a JSON report.
        std::string autotune(int reps);

        void set_kernel_timing(bool on) { timing_on_ = on; }
        std::vector<KernelTiming> kernel_timings() const;
        void reset_kernel_timings();
        int block_dim() const { return bd_; }
        int dim() const { return sys_.empty() ? 0 : sys_[0].n; }

        /// Values-only refactorisation from device memory (values in the recorded CSC order of the
        /// last full factorisation with lazy_partitioning).
        void refresh_values_device(long long nnz, const double *d_vals);
        /// Several right-hand sides of one matrix: one factorisation of A (one SELL matrix, one
        /// set of bank inverses and coarsest inverse) shared by `copies` (<= MAX_COPIES) systems
        /// whose vectors, coarse numbering and scalars are per copy; the SpMV and the bank apply
        /// read each matrix once for all of them.  solve_batch then takes the right-hand sides.
        void factorize_copies(const CscView &A, int copies, const DeviceCsc *dev = nullptr);
        static constexpr int MAX_COPIES = 8;
        /// Full factorisation of one system from CSC arrays in device memory (the matrix is built
        /// on the GPU; the host sees the pattern only for a graph aggregation).
        void factorize_device(int n, long long nnz, const int *d_colptr, const int *d_rowidx, const double *d_vals);

        /// Dirichlet elimination on the factorisation (prefactorize, FEMSolver.cpp:269-315): factorise
        /// the matrix with the rows and columns of the fixed dofs replaced by identity; remembers A's
        /// couplings to the fixed dofs for the right-hand side lift.  With lazy_partitioning and the
        /// same pattern and node set as the last such call, only the values are refreshed.
        void dirichlet_factorize(const CscView &A, const int *nodes, int ndir);
        /// The same with A's CSC arrays in device memory (validated there); nodes on the host.
        void dirichlet_factorize_device_input(int n, long long nnz, const int *d_colptr, const int *d_rowidx,
                                              const double *d_vals, const int *nodes, int ndir);
        /// dirichlet_solve_prefactorized (FEMSolver.cpp:318-342): f := g = f - (I - N) A N f with
        /// the A of the last Dirichlet factorisation, then u = solve(g).
        void dirichlet_solve_host(int n, double *f, double *u);
        void dirichlet_solve_device(int n, double *d_f, double *d_u);

    private:
        struct Graphs;

        /// One system of the batch.
        struct System
        {
            int n = 0, nb = 0, nparts = 0, npad = 0;
            int node_off = 0;  ///< first padded node
            int blk_begin = 0; ///< first node block
            long long nnzb = 0;
            Partition part;
            bool have_partition = false;
            std::vector<int> slot_of_block; ///< local slots
            std::vector<double> coords;
            int coord_dim = 0;
            std::string aggregation;
            int sob_off = 0; ///< offset of this system's slot_of_block in the device array
            // last solve
            int iterations = 0;
            double residual_norm = 0.0;
            Status status = Status::Running;
            int reliable_updates = 0;
            double rel_residual = 0.0;
        };

        Params params_;
        std::vector<System> sys_;
        double solve_ms_ = 0.0;

        hipStream_t stream_ = nullptr;
        int device_ = 0;

        int bd_ = 0, npad_ = 0, nparts_ = 0, nblocks_ = 0;
        long long nnzb_ = 0, nnz_scalar_ = 0, n_total_ = 0, nb_total_ = 0;
        bool factorized_ = false;
        bool rhs_ready_ = false; ///< d_b_ holds the last solve's right-hand sides
        std::vector<int> block_of_slot_; ///< global slot -> local block (-1 virtual)

        // SELL matrix
        int nslices_ = 0;
        long long nslots_ = 0;
        DeviceBuffer<int> d_slice_ptr_, d_cols_, d_row_len_, d_block_of_slot_, d_slot_of_block_;
        // Halo form of the SELL columns (see SellView): 16-bit local columns and per-node-block
        // lists; halo_max_ = 0 when some block reads more than SPMV_HALO_CAP distinct nodes.
        DeviceBuffer<unsigned short> d_lcols_;
        DeviceBuffer<int> d_halo_ptr_, d_halo_;
        /// Per node block: position of its first own node in its halo list, when every block's
        /// own nodes appear there in order (halo_own_ok_; the single-reduction SpMV needs it).
        DeviceBuffer<int> d_halo_own_;
        bool halo_own_ok_ = false;
        int halo_max_ = 0;
        long long halo_total_ = 0;
        DeviceBuffer<double> d_vals64_;
        DeviceBuffer<float> d_vals32_;
        // Symmetric form of the matrix (SymView; spmv_symmetric): the structure is built with the
        // halo lists, the values of each type on first use (from d_vals64_), after a check that
        // the blocks it stores once are symmetric (sym_values_ok_: -1 not checked yet).
        DeviceBuffer<int> d_sym_slice_ptr_, d_sym_src_, d_sym_mirror_;
        DeviceBuffer<uint32_t> d_sym_meta_, d_sym_info_;
        DeviceBuffer<float> d_symv32_;
        DeviceBuffer<double> d_symv64_;
        long long sym_slots_ = 0, sym_stored_ = 0;
        int sym_incoming_max_ = 0;
        bool sym_built_ = false;
        int sym_values_ok_ = -1;
        std::string sym_note_;
        std::future<SymSell> sym_future_; ///< the host build of the structure, while it runs
        bool symv_fresh_[2] = {false, false}; ///< d_symv32_ / d_symv64_ hold the current values
        DeviceBuffer<int> d_sym_flag_;

        // systems
        DeviceBuffer<int> d_block_sys_, d_blk_begin_, d_blk_count_, d_active_, d_select_, d_bank_sys_;
        std::vector<int> h_active_, h_select_;

        // MAS
        int level_num_ = 0;
        int cco_[MAX_COARSE_LEVELS] = {0, 0, 0, 0};
        int mat_cnt_[MAX_COARSE_LEVELS + 1] = {0, 0, 0, 0, 0};
        int mat_off_[MAX_COARSE_LEVELS + 1] = {0, 0, 0, 0, 0};
        int lev_off_[MAX_COARSE_LEVELS + 1] = {0, 0, 0, 0, 0};
        int total_mats_ = 0, total_level_nodes_ = 0;
        DeviceBuffer<uint32_t> d_lead0_;
        DeviceBuffer<int> d_gid0_;
        DeviceBuffer<uint32_t> d_leadc_[MAX_COARSE_LEVELS];
        DeviceBuffer<int> d_gidc_[MAX_COARSE_LEVELS];
        DeviceBuffer<int4> d_cmap_;
        DeviceBuffer<int> d_lev_off_, d_mat_off_;
        DeviceBuffer<double> d_mats64_;
        DeviceBuffer<__half> d_mats16_;
        DeviceBuffer<float> d_dscale_;

        // exact coarsest-level solve (coarse_solver = "cholesky")
        std::vector<char> nonisolated_; ///< slot has an off-diagonal block
        bool dense_built_ = false;      ///< tried for this factorisation
        int dense_nd_ = 0;              ///< rows of the dense coarse matrix (0 = not used)
        int dense_nodes_ = 0;
        std::string dense_note_;
        DeviceBuffer<double> d_dense_inv_;
        DeviceBuffer<int> d_dense_ids_;
        DeviceBuffer<int> d_dense_index_; ///< per coarsest-segment node: dense index or -1

        // Values-only refactorisation: the last full factorisation's pattern and the position of
        // each of its non-zeros in the SELL value array.
        std::vector<int> vr_colptr_, vr_rowidx_; ///< host pattern (empty after a device-input factorisation)
        long long vr_nnz_ = -1;                   ///< non-zeros the value map covers (-1: none)
        std::vector<int> h_slice_ptr_;
        DeviceBuffer<int> d_vmap_;
        DeviceBuffer<double> d_csc_vals_;
        const char *factorize_mode_ = "full";
        int copies_ = 1; ///< > 1: right-hand sides sharing one factorisation (factorize_copies)
        CopyLayout cl_;  ///< bank and stored-matrix numbering of those copies
        int stored_mats_ = 0; ///< bank matrices held (one copy's)
        int npad1_ = 0, nblocks1_ = 0; ///< one copy's padded nodes and node blocks
        std::vector<std::pair<std::string, double>> setup_ms_; ///< phases of the last full factorisation

        // Dirichlet elimination: the fixed dofs (sorted, unique) and A's couplings a_ij (i free,
        // j fixed) by row, for the right-hand side lift.
        bool dirichlet_ = false;
        std::vector<int> dir_nodes_;
        int n_crow_ = 0;
        long long n_coupl_ = 0;
        DeviceBuffer<int> d_crow_, d_cptr_, d_ccol_, d_cidx_;
        DeviceBuffer<double> d_cval_, d_dir_f_;

        // Restriction of coarse segments 2.. straight from segment 1: every node's segment-1
        // descendants (ascending), summed by the side workgroups of the first apply launch (a
        // thread per node of at most RESTRICT_SMALL_MAX descendants, a wave per piece of at most
        // RESTRICT_PIECE descendants of a larger node; the launch that reads such a node adds its
        // pieces).
        int nrsmall_ = 0, nrbig_ = 0, nrcomb_ = 0;
        DeviceBuffer<int4> d_rtask_;
        DeviceBuffer<int> d_rlist_;
        DeviceBuffer<int4> d_rcomb_;   ///< nodes cut into pieces: {node, first partial, pieces}
        DeviceBuffer<double> d_rpart_; ///< piece sums of those nodes
        DeviceBuffer<int> d_split_of_; ///< per node of segments 2..: index into d_rcomb_ or -1

        // vectors
        DeviceBuffer<double> d_b_, d_x_, d_mr_, d_mz_, d_p_, d_Ap_, d_io_;
        DeviceBuffer<float> d_x32_, d_mr32_, d_mz32_, d_p32_, d_Ap32_;
        // Second search-direction buffers: with the direction update folded into the SpMV, p
        // alternates between the two (other workgroups still read the previous p).
        DeviceBuffer<double> d_pb_;
        DeviceBuffer<float> d_p32b_;
        // Single-reduction PCG (mixed path): u = M r assembled by the SpMV, and s = A p by recurrence.
        DeviceBuffer<float> d_u32_, d_s32_;
        bool single_reduction() const;
        CgReductions cg_reductions() const;
        int p_cur_ = 0; ///< which buffer holds the current p (0: d_p_ / d_p32_)
        bool fuse_direction() const;
        double *p64(int which) { return which ? d_pb_.data() : d_p_.data(); }
        float *p32(int which) { return which ? d_p32b_.data() : d_p32_.data(); }
        DeviceBuffer<double> d_sc_;
        DeviceBuffer<double> d_partials_;
        int partial_cap_ = 0;
        std::vector<double> h_sc_;

        std::unique_ptr<Graphs> graphs_;

        // Static scratch of solve_device_noalloc: this device's address of the library's array
        // (null when the lookup failed, with the reason), and whether the running solve uses it
        // for its partials, scalars and flexible-PCG partials instead of the buffers above.
        double *scratch_ = nullptr;
        std::string scratch_error_;
        bool use_scratch_ = false;
        double *sc_ptr() const;
        double *rzo_ptr() const;

        // learned coarse correction
        std::unique_ptr<LearnedCoarse> lcc_;
        bool lattice_stale_ = true;
        DeviceBuffer<float> d_zprev32_;
        DeviceBuffer<double> d_rzo_;
        DeviceBuffer<double> d_snap_;
        int nsnap_ = 0;
        bool learned_used_ = false, learned_fallback_ = false;
        int learned_fallback_iter_ = -1;
        std::string learned_note_;

        bool timing_on_ = false;
        std::map<std::string, KernelTiming> timings_;
        std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending_events_;
        std::vector<hipEvent_t> event_pool_;
        size_t event_pool_used_ = 0;

        Reductions reductions() const;
        /// The partials of copy c's node blocks (pointers offset to its first block), for one
        /// launch over that copy's rows with copy 0's matrix.
        Reductions reductions_of_copy(int c) const;
        SysMap sysmap_of_copy(int c) const;
        /// y = A x with the PCG partials, for every system: one launch, the copies launch (matrix
        /// read once for all copies) or one launch per copy (spmv_copies false).
        template <int BD, typename T>
        void spmv_ap_all(const SellView<T> &A, const T *x, T *y);
        /// y = b - A x (FP64 matrix) with its level-0 restriction and r . r partials into rr, for
        /// every system (one launch per copy with copies); y is the multi-level vector.
        template <int BD, typename TX, typename TY>
        void spmv_res_all(const SellView<double> &A, const TX *x, const TX *b, TY *y, double *rr, bool selected);
        void finalize_scalars();
        SysMap sysmap() const;
        FineRestrict fine_restrict() const;
        template <typename TV>
        SellView<TV> sell(const TV *vals) const;
        /// The symmetric form for value type TV when the single-matrix SpMV may use it
        /// (spmv_symmetric, structure built, values symmetric, LDS within 64 KB); builds its
        /// values on first use.  Returns false otherwise.
        template <typename TV>
        bool sym_view(SymView<TV> &out) const;
        /// Checks the values' symmetry (once per values) and builds the TV values when the
        /// symmetric form applies; outside graph capture only (it allocates and synchronizes).
        template <typename TV>
        bool prepare_sym();
        void build_sym(std::vector<int> slice_ptr, std::vector<int> cols, std::vector<unsigned short> lcols);
        void join_sym();
        void reset_sym_values(bool free_buffers = false);

        /// Domains of system S: Morton chunks (coordinates), else the graph aggregation on the
        /// GPU when dcsc (A in device memory) is given, else on the host.
        void build_partition(System &S, const CscView &A, const DeviceCsc *dcsc = nullptr);
        void upload_matrix(const PaddedBsr &bsr);
        /// Per node block, the distinct columns of its SELL slots and the 16-bit local columns.
        void build_halo(const std::vector<int> &slice_ptr, std::vector<int> cols);
        /// Algorithmic bytes of one SpMV: values (val_bytes each), columns, the halo lists when in
        /// use, and vec_passes row vectors of x_bytes per scalar.
        double spmv_bytes(int val_bytes, int x_bytes, int vec_passes) const;
        /// Algorithmic bytes of one y = A p launch of the PCG loop (the symmetric form's when it
        /// is used).
        template <typename TV>
        double spmv_ap_bytes() const;
        void build_mas();
        /// Several right-hand sides of the single-system factorisation just built: per-copy node
        /// maps (coarse map, restriction lists and masks, slot maps, systems), level-major
        /// multi-level numbering, vectors; the matrices stay as built.
        void expand_copies(int copies);
        void build_dense_coarse();
        void build_bank_inverses(); ///< Galerkin bank matrices and their inverses (numeric part)
        /// Values-only refactorisation: one system, lazy_partitioning, the same n, block size and
        /// sparsity pattern as the last full factorisation, and the same fixed dofs (dirichlet_nodes,
        /// sorted and unique) when that was a Dirichlet one (nullptr: a plain one).  Scatters the
        /// new values into the SELL arrays and rebuilds the numeric MAS parts; false when it does
        /// not apply.
        bool try_refresh_values(const CscView &A, const std::vector<int> *dirichlet_nodes);
        /// Full factorisation; for a Dirichlet one, As[0] is the eliminated matrix, raw the matrix
        /// before elimination (whose non-zeros the value map then indexes) and dir the elimination.
        /// d_keep: a Dirichlet elimination made on the GPU (dev = A′, raw = A, dir = nullptr) with
        /// its raw -> A′ map in device memory.
        void factorize_full(const std::vector<CscView> &As, const CscView *raw, const DirichletSystem *dir,
                            const DeviceCsc *dev = nullptr, const DeviceBuffer<int> *d_keep = nullptr);
        /// The SELL matrix built on the GPU (sell_build.hpp) from the CSC arrays in device memory
        /// (dcsc[s] per system: uploaded, or the caller's arrays when device_input), for one
        /// system or every system of a batch; also the halo
        /// lists and, with lazy_partitioning, the value map (and A's host pattern for host input);
        /// for a Dirichlet factorisation (raw, dir as in factorize_full) the map indexes raw's
        /// non-zeros.  phase(name) closes a timed setup phase.
        template <typename Phase>
        void device_matrix_build(const CscView &A, const std::vector<DeviceCsc> &dcsc, bool device_input,
                                 const std::vector<int> &h_slot_of_block, const CscView *raw,
                                 const DirichletSystem *dir, const DeviceBuffer<int> *d_keep, Phase &phase);
        /// New SELL values (and coupling values of a Dirichlet factorisation) from d_vals in the
        /// recorded CSC order.
        void scatter_new_values(const double *d_vals);
        /// The Dirichlet factorisation with the elimination on the GPU (device_setup); list holds
        /// the sorted, unique fixed dofs and is taken over.
        void dirichlet_factorize_device(const CscView &A, std::vector<int> &list);
        /// Sorted, unique fixed dofs from a caller's list (checked against n).
        static std::vector<int> dirichlet_list(int n, const int *nodes, int ndir);
        /// Elimination of the fixed dofs in list (taken over) from dA (A in device memory; A's
        /// host arrays, when not null, give the pattern for values-only refactorisations), then
        /// the factorisation of A′; t0 starts the elimination phase's clock.
        void dirichlet_eliminate_and_factorize(const CscView &A, const DeviceCsc &dA, std::vector<int> &list,
                                               std::chrono::steady_clock::time_point t0);
        void dirichlet_lift_device(const double *d_f, double *d_g);
        void check_dirichlet(int n) const;
        /// The numeric rebuild after the SELL values were refreshed.
        void finish_value_refresh();

        void build_restrict_chunks();
        /// Side work of the first apply launch (restriction of segments 2.. from the segment-1
        /// values seg1 into the multi-level vector out; with lattice_gather also the learned
        /// correction's lattice gather) and of the second (the dense coarsest solve); `block` is
        /// the apply kernel's workgroup size.
        ApplySide restrict_side(const void *seg1, void *out, int block, bool lattice_gather = false) const;
        ApplySide coarse_side(int block) const;
        /// Matrices of the first apply launch: fine banks and segment-1 banks.
        int first_apply_mats() const { return level_num_ >= 2 ? mat_off_[2] : total_mats_; }
        /// The same boundary in stored matrices (differs with copies).
        int first_apply_stored() const { return level_num_ >= 2 ? cl_.soff[2] : stored_mats_; }
        void ensure_mixed_data();
        void ensure_coarse_solver();
        void allocate_vectors();
        void check_ready_single(int n) const;

        void solve_padded();
        void solve_fp64();
        void solve_mixed();
        void ensure_lattice();
        template <int BD>
        void learned_collect_pairs();
        template <int BD>
        void learned_pair_from_coef(const std::vector<double> &w);
        void upload_flags();
        double *scal(int s, int slot) { return &h_sc_[static_cast<size_t>(s) * S_COUNT + slot]; }

        // per-iteration pieces
        template <int BD>
        void fp64_iteration(int k, bool init_direction);
        template <int BD>
        void fp64_precondition(bool init);
        template <int BD>
        void mixed_part_a(bool learned, bool init_direction);
        template <int BD>
        void mixed_part_b(bool init, bool learned = false);
        template <int BD>
        LccApply mixed_apply(bool init, bool learned);
        template <int BD>
        void cg_part_a(bool learned, bool init);
        template <int BD>
        void cg_part_b(bool init, bool learned);
        template <int BD>
        void mixed_reliable_residual(bool learned);

        void run_block(long long key, const std::function<void()> &enqueue);

        // timing helpers
        void tic(const char *name, double bytes);
        void toc();
        void collect_timings();
        std::string pending_name_;
        double pending_bytes_ = 0.0;
        hipEvent_t pending_start_ = nullptr;

        void read_scalars();
    };
} // namespace psmas
