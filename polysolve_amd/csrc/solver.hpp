// MI355X MAS-preconditioned PCG solver: the object behind the psmas_* C ABI.
//
// Public surface mirrors polysolve::linear::MASSolver (src/polysolve/linear/MASSolver.hpp:35-72):
// set_parameters / get_info / analyze_pattern / factorize / solve / name, with the same
// semantics (MASSolver.cu:134-639), plus device-pointer solves, optional node coordinates for
// Morton aggregation, and introspection used by the parity tests.
#pragma once

#include "common.hpp"
#include "host_setup.hpp"
#include "kernels.hpp"
#include "params.hpp"

#include <functional>
#include <map>
#include <memory>
#include <string>
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

    class Solver
    {
    public:
        Solver();
        ~Solver();
        Solver(const Solver &) = delete;
        Solver &operator=(const Solver &) = delete;

        void set_parameters_json(const std::string &json);
        std::string get_info_json() const;
        std::string parameters_json() const { return params_to_json(params_); }
        void analyze_pattern(const CscView &) {}
        void factorize(const CscView &A);
        void solve_host(int n, const double *b, double *x);
        void solve_device(int n, const double *d_b, double *d_x);
        void set_coordinates(int nb, int dim, const double *xyz);

        // Introspection for tests (host arrays, original ordering unless stated).
        void get_partition(std::vector<int> &offsets, std::vector<int> &perm) const;
        std::vector<int> level_info() const;
        std::vector<int> coarse_map() const;                 ///< nb x 4, per input block
        std::vector<double> bank_inverses() const;           ///< packed FP64, all levels
        void apply_preconditioner(const double *r, double *z, bool mixed);
        void spmv(const double *x, double *y, bool fp32);

        void set_kernel_timing(bool on) { timing_on_ = on; }
        std::vector<KernelTiming> kernel_timings() const;
        void reset_kernel_timings();
        int block_dim() const { return bd_; }
        int dim() const { return n_; }

    private:
        struct Graphs;

        Params params_;
        int iterations_ = 0;
        double residual_norm_ = 0.0;
        Status status_ = Status::Running;
        int reliable_updates_ = 0;
        double true_rel_residual_ = 0.0;
        double solve_ms_ = 0.0;

        hipStream_t stream_ = nullptr;
        int device_ = 0;

        int n_ = 0, bd_ = 0, nb_ = 0, npad_ = 0, nparts_ = 0;
        long long nnzb_ = 0;
        bool factorized_ = false;
        std::vector<double> coords_;
        int coord_dim_ = 0;
        Partition part_;
        bool have_partition_ = false;
        std::vector<int> slot_of_block_, block_of_slot_;
        std::string aggregation_used_;

        // SELL matrix
        int nslices_ = 0;
        long long nslots_ = 0;
        DeviceBuffer<int> d_slice_ptr_, d_cols_, d_row_len_, d_block_of_slot_, d_slot_of_block_;
        DeviceBuffer<double> d_vals64_;
        DeviceBuffer<float> d_vals32_;

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

        // vectors
        DeviceBuffer<double> d_b_, d_x_, d_mr_, d_mz_, d_p_, d_Ap_, d_io_;
        DeviceBuffer<float> d_x32_, d_mr32_, d_mz32_, d_p32_, d_Ap32_;
        DeviceBuffer<double> d_sc_;
        DeviceBuffer<double> d_partials_;
        DeviceBuffer<unsigned> d_counters_;
        int partial_cap_ = 0;

        std::unique_ptr<Graphs> graphs_;

        bool timing_on_ = false;
        std::map<std::string, KernelTiming> timings_;
        std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending_events_;
        std::vector<hipEvent_t> event_pool_;
        size_t event_pool_used_ = 0;

        ReduceSite site(int s) const;
        FineRestrict fine_restrict() const;
        CoarseRestrict coarse_restrict() const;
        template <typename TV>
        SellView<TV> sell(const TV *vals) const;

        void build_partition(const CscView &A);
        void upload_matrix(const PaddedBsr &bsr);
        void build_mas();
        void ensure_mixed_data();
        void allocate_vectors();
        void check_ready(int n) const;

        void solve_padded();
        void solve_fp64();
        void solve_mixed();

        // per-iteration pieces
        template <int BD>
        void fp64_iteration(int k);
        template <int BD>
        void fp64_precondition(bool init);
        template <int BD>
        void mixed_part_a();
        template <int BD>
        void mixed_part_b(bool init);
        template <int BD>
        void mixed_reliable_residual();

        void run_block(int key, int count, const std::function<void()> &enqueue);

        // timing helpers
        void tic(const char *name, double bytes);
        void toc();
        void collect_timings();
        std::string pending_name_;
        double pending_bytes_ = 0.0;
        hipEvent_t pending_start_ = nullptr;

        double read_scalar(int slot);
        void read_scalars(double *out);
    };
} // namespace psmas
