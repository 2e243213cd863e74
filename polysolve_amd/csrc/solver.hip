// Host orchestration of the MI355X MAS-preconditioned PCG solver.
//
// Follows the reference's MASSolverImpl (src/polysolve/linear/MASSolver.cu:134-596):
//   factorize  partition (lazy reuse) -> permuted BSR -> MAS build -> PCG buffers  (:292-374)
//   solve      b -> permuted space; x0 = 0; PCG; un-permute                         (:376-409)
//   pcg_solve  x0 = 0, true residual every 4th iteration, convergence every 10th     (:469-595)
// The working space is the bank-padded node space (domain p owns slots [32p, 32p+32)), so the
// fine level of the preconditioner needs no gather at all, and every device scalar stays on the
// device between the host convergence checks.  Iterations between checks are captured once into
// HIP graphs and replayed.
//
// A factorisation holds K >= 1 independent systems laid out one after another, each covering
// whole 256-node blocks.  They share every kernel launch; each has its own CG scalars, its own
// deterministic reductions and its own convergence test, and stops (its workgroups exit at
// once) when it converges.  The preconditioner's coarse levels are built over the whole batch,
// which is the reference construction applied to the block-diagonal batch matrix.
//
// The mixed-precision path ("precision": "mixed") runs an FP32 inner PCG (FP32 SELL matrix,
// FP16 fine-level bank inverses with FP32 scaling, FP64 coarse-level inverses) with reliable
// updates: at a check point where a system's recurrence residual fell by reliable_update_delta
// (or below its target), its FP32 correction is added to the FP64 solution and its FP64
// residual b - A x is recomputed with the FP64 matrix; convergence is decided on that FP64
// residual only.
#include "solver.hpp"
#include "dirichlet.hpp"
#include "sell_build.hpp"
#include "aggregate.hpp"
#include "coarse_solve.hpp"
#include "device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <atomic>
#include <future>
#include <thread>

namespace psmas
{
    // ------------------------------------------------------------------------------ errors
    void throw_hip_error(hipError_t err, const char *what, const char *file, int line)
    {
        std::ostringstream o;
        o << "[MAS] HIP error " << hipGetErrorName(err) << " (" << hipGetErrorString(err) << ") in " << what
          << " at " << file << ":" << line;
        if (err == hipErrorOutOfMemory)
            throw OutOfMemoryError(o.str());
        throw std::runtime_error(o.str());
    }

    void *device_alloc(size_t bytes, const char *context)
    {
        void *p = nullptr;
        const hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess)
            return p;
        (void)hipGetLastError();
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        constexpr size_t mb = 1024ull * 1024ull;
        throw OutOfMemoryError(std::string("[MAS] Failed to allocate ") + std::to_string((bytes + mb - 1) / mb) +
                               " MB for " + (context ? context : "device allocation") + "; only " +
                               std::to_string(free_b / mb) + " MB free.");
    }

    void device_free(void *p) noexcept
    {
        if (p)
            (void)hipFree(p);
    }

    const char *status_string(Status s)
    {
        switch (s)
        {
        case Status::Running:
            return "Running";
        case Status::ReachRelativeTolerance:
            return "Reach relative tolerance";
        case Status::ReachAbsoluteTolerance:
            return "Reach absolute tolerance";
        case Status::ReachMaxIterations:
            return "Reach max iterations";
        }
        return "Unknown";
    }

    namespace
    {
        /// Graph cache keys: precision path, iteration pattern, iterations per graph, and flags
        /// (bit 0: fused direction update, bit 1: direction buffer at the segment's start, bit 2:
        /// the segment's first SpMV starts the direction from z, bit 3: single-reduction iteration).
        long long graph_key(int path, int pattern, int count, int flags = 0)
        {
            return (static_cast<long long>(path) << 56) | (static_cast<long long>(flags) << 48) |
                   (static_cast<long long>(pattern) << 24) | static_cast<long long>(count);
        }

        // Scratch of solve_device_noalloc (1 MiB per GPU, held until the process ends): the
        // per-system scalars in the first SCRATCH_SC_PAD doubles, then the five reduction
        // partial arrays and the flexible-PCG partials, one double per 256-node block each.
        // Host code never passes `scratch` itself to a kernel: it is a host-side stand-in, and
        // the device address comes from hipGetSymbolAddress (scratch_address).
        constexpr size_t SCRATCH_DOUBLES = size_t(1) << 17;
        constexpr size_t SCRATCH_SC_PAD = 32; ///< >= S_COUNT doubles for the one system
        static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

        /// This device's address of `scratch`, looked up once per device with that device current.
        double *scratch_address(int device, std::string &error)
        {
            static std::mutex mtx;
            static std::map<int, std::pair<double *, std::string>> addr;
            std::lock_guard<std::mutex> lock(mtx);
            auto it = addr.find(device);
            if (it == addr.end())
            {
                void *p = nullptr;
                const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
                std::string msg;
                if (e != hipSuccess)
                {
                    (void)hipGetLastError();
                    msg = std::string("hipGetSymbolAddress of the static scratch failed: ") + hipGetErrorString(e);
                    p = nullptr;
                }
                it = addr.emplace(device, std::make_pair(static_cast<double *>(p), msg)).first;
            }
            error = it->second.second;
            return it->second.first;
        }
    } // namespace

    // ------------------------------------------------------------------------------ graphs
    struct Solver::Graphs
    {
        std::map<long long, hipGraphExec_t> execs;
        ~Graphs() { clear(); }
        void clear()
        {
            for (auto &kv : execs)
                (void)hipGraphExecDestroy(kv.second);
            execs.clear();
        }
    };

    // ------------------------------------------------------------------------ construction
    Solver::Solver() : graphs_(new Graphs())
    {
        int count = 0;
        const hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0)
        {
            (void)hipGetLastError();
            throw NoDeviceError("[MAS] No AMD GPU!!");
        }
        device_ = 0;
        PSMAS_HIP(hipSetDevice(device_));
        PSMAS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        scratch_ = scratch_address(device_, scratch_error_);
        sys_.resize(1);
    }

    Solver::~Solver()
    {
        (void)hipSetDevice(device_);
        if (stream_)
            (void)hipStreamSynchronize(stream_);
        graphs_.reset();
        for (hipEvent_t ev : event_pool_)
            (void)hipEventDestroy(ev);
        if (stream_)
            (void)hipStreamDestroy(stream_);
    }

    void Solver::set_parameters_json(const std::string &json)
    {
        const Params before = params_;
        apply_json_parameters(params_, json);
        if (before.use_graphs != params_.use_graphs || before.check_interval != params_.check_interval ||
            before.true_residual_period != params_.true_residual_period || before.precision != params_.precision ||
            before.coarse_solver != params_.coarse_solver || before.spmv_unroll != params_.spmv_unroll ||
            before.spmv_halo != params_.spmv_halo || before.spmv_copies != params_.spmv_copies ||
            before.spmv_fuse_direction != params_.spmv_fuse_direction ||
            before.apply_workgroups != params_.apply_workgroups || before.apply_wave != params_.apply_wave ||
            before.single_reduction != params_.single_reduction || before.spmv_symmetric != params_.spmv_symmetric)
            graphs_->clear();
    }

    std::string Solver::autotune(int reps)
    {
        // Tile and schedule choices, timed on this factorisation and the last right-hand side:
        // (1) the SpMV row-loop batch depth, by timing `reps` SpMVs per candidate with HIP events;
        // (2) the check interval (graph segment length) and the coarsest-level solver, by timing
        // full solves.  The fastest converged candidate is kept in the parameters.
        PSMAS_HIP(hipSetDevice(device_));
        if (!factorized_ || !rhs_ready_)
            throw std::runtime_error("[MAS] autotune needs a factorisation and one solve first.");
        reps = std::max(reps, 1);
        std::ostringstream o;
        o.precision(6);
        const bool mixed = params_.precision == Precision::Mixed;
        if (mixed)
            ensure_mixed_data();
        hipEvent_t e0, e1;
        PSMAS_HIP(hipEventCreate(&e0));
        PSMAS_HIP(hipEventCreate(&e1));
        const long long nvec = static_cast<long long>(npad_) * bd_;
        int best_u = params_.spmv_unroll;
        float best_t = 1e30f;
        o << "{\"spmv_us\": {";
        const int cands[3] = {2, 4, 8};
        for (int ci = 0; ci < 3; ++ci)
        {
            params_.spmv_unroll = cands[ci];
            auto body = [&](auto bdc) {
                constexpr int BD = decltype(bdc)::value;
                if (mixed)
                    launch_spmv_plain<BD, float, float>(sell(d_vals32_.data()), d_p32_.data(), d_Ap32_.data(), stream_);
                else
                    launch_spmv_plain<BD, double, double>(sell(d_vals64_.data()), d_p_.data(), d_Ap_.data(), stream_);
            };
            auto run = [&]() {
                if (bd_ == 1)
                    body(std::integral_constant<int, 1>());
                else if (bd_ == 2)
                    body(std::integral_constant<int, 2>());
                else
                    body(std::integral_constant<int, 3>());
            };
            run(); // warm
            PSMAS_HIP(hipEventRecord(e0, stream_));
            for (int r = 0; r < reps; ++r)
                run();
            PSMAS_HIP(hipEventRecord(e1, stream_));
            PSMAS_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            PSMAS_HIP(hipEventElapsedTime(&ms, e0, e1));
            const float us = ms * 1e3f / reps;
            o << (ci ? ", " : "") << "\"" << cands[ci] << "\": " << json_number(us, 6);
            if (us < best_t)
            {
                best_t = us;
                best_u = cands[ci];
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)nvec;
        params_.spmv_unroll = best_u;
        o << "}, \"spmv_unroll\": " << best_u << ", \"solve_ms\": {";
        const Params base = params_;
        const int intervals[3] = {5, 10, 20};
        const CoarseSolver coarse[2] = {CoarseSolver::Bank, CoarseSolver::Cholesky};
        double best_ms = 1e300;
        Params best = base;
        bool firstc = true;
        for (CoarseSolver cs : coarse)
            for (int iv : intervals)
            {
                params_ = base;
                params_.check_interval = iv;
                params_.coarse_solver = cs;
                graphs_->clear();
                solve_padded(); // captures this variant's graphs
                solve_padded();
                bool ok = true;
                for (const auto &S : sys_)
                    ok = ok && (S.status == Status::ReachRelativeTolerance || S.status == Status::ReachAbsoluteTolerance);
                o << (firstc ? "" : ", ") << "\"" << (cs == CoarseSolver::Bank ? "bank" : "cholesky") << "/" << iv
                  << "\": " << json_number(ok ? solve_ms_ : -1.0, 6);
                firstc = false;
                if (ok && solve_ms_ < best_ms)
                {
                    best_ms = solve_ms_;
                    best = params_;
                }
            }
        params_ = best;
        graphs_->clear();
        o << "}, \"check_interval\": " << params_.check_interval << ", \"coarse_solver\": \""
          << (params_.coarse_solver == CoarseSolver::Bank ? "bank" : "cholesky") << "\", \"best_solve_ms\": "
          << json_number(best_ms < 1e300 ? best_ms : NAN, 6) << "}";
        return o.str();
    }

    std::string Solver::system_info_json(int s) const
    {
        if (s < 0 || s >= static_cast<int>(sys_.size()))
            throw InvalidArgumentError("[MAS] system index out of range.");
        const System &S = sys_[s];
        std::ostringstream o;
        o.precision(17);
        o << "{\"solver_iter\": " << S.iterations << ", \"solver_error\": " << json_number(S.residual_norm)
          << ", \"solver_status\": \"" << status_string(S.status) << "\""
          << ", \"precision\": \"" << (params_.precision == Precision::FP64 ? "fp64" : "mixed") << "\""
          << ", \"reliable_updates\": " << S.reliable_updates << ", \"relative_residual\": " << json_number(S.rel_residual)
          << ", \"solve_ms\": " << json_number(solve_ms_) << ", \"levels\": " << level_num_ << ", \"parts\": " << S.nparts
          << ", \"systems\": " << sys_.size() << ", \"aggregation\": \"" << S.aggregation << "\""
          << ", \"coarse_solver\": \"" << (dense_nd_ > 0 ? dense_note_ : std::string("bank")) << "\""
          << ", \"factorize_mode\": \"" << factorize_mode_ << "\""
          << ", \"copies\": " << copies_
          << ", \"pcg_iteration\": \""
          << (params_.precision == Precision::Mixed && single_reduction() ? "single_reduction" : "classic") << "\""
          << ", \"setup_ms\": {" << [&] {
                 std::ostringstream q;
                 q.precision(6);
                 for (size_t i = 0; i < setup_ms_.size(); ++i)
                     q << (i ? ", " : "") << "\"" << setup_ms_[i].first << "\": " << json_number(setup_ms_[i].second, 6);
                 return q.str();
             }() << "}"
          << ", \"dirichlet\": {\"active\": " << (dirichlet_ ? "true" : "false")
          << ", \"fixed_dofs\": " << dir_nodes_.size() << ", \"coupled_rows\": " << n_crow_
          << ", \"couplings\": " << n_coupl_ << "}"
          << ", \"spmv_halo\": {\"active\": " << (params_.spmv_halo && halo_max_ > 0 ? "true" : "false")
          << ", \"max_nodes\": " << halo_max_ << ", \"total_nodes\": " << halo_total_ << "}"
          << ", \"spmv_symmetric\": {\"fp32\": " << [this] {
                 SymView<float> v;
                 return sym_view(v) ? "true" : "false";
             }() << ", \"fp64\": " << [this] {
                 SymView<double> v;
                 return sym_view(v) ? "true" : "false";
             }() << ", \"slots\": " << sym_slots_ << ", \"full_slots\": " << nslots_
          << ", \"incoming_max\": " << sym_incoming_max_ << ", \"note\": \"" << sym_note_ << "\"}"
          << ", \"coarse_restriction\": {\"thread_nodes\": " << nrsmall_ << ", \"wave_pieces\": " << nrbig_
          << ", \"split_nodes\": " << nrcomb_ << "}"
          << ", \"learned\": {\"used\": " << (learned_used_ ? "true" : "false")
          << ", \"fallback\": " << (learned_fallback_ ? "true" : "false")
          << ", \"fallback_iteration\": " << learned_fallback_iter_
          << ", \"trained\": " << (lcc_ && lcc_->trained() ? "true" : "false")
          << ", \"network\": \"" << (lcc_ && lcc_->trained() ? lcc_->kind() : std::string()) << "\""
          << ", \"pairs\": " << (lcc_ ? lcc_->num_pairs() : 0)
          << ", \"train_loss\": " << json_number(lcc_ ? lcc_->last_loss() : 0.0)
          << ", \"lattice_points\": " << (lcc_ && lcc_->lattice_ready() ? lcc_->k() : 0)
          << ", \"block_restriction\": " << (lcc_ && lcc_->block_restriction() ? "true" : "false")
          << ", \"note\": \"" << learned_note_ << "\"}}";
        return o.str();
    }

    void Solver::set_coordinates(int sys, int nb, int dim, const double *xyz)
    {
        if (nb <= 0 || dim < 1 || dim > 3 || xyz == nullptr)
            throw InvalidArgumentError("[MAS] invalid coordinates: need nb > 0 nodes with 1 to 3 components.");
        if (sys < 0)
            throw InvalidArgumentError("[MAS] system index out of range.");
        if (sys >= static_cast<int>(sys_.size()))
            sys_.resize(sys + 1);
        System &S = sys_[sys];
        S.coords.assign(xyz, xyz + static_cast<size_t>(nb) * dim);
        S.coord_dim = dim;
        S.have_partition = false;
    }

    double *Solver::sc_ptr() const { return use_scratch_ ? scratch_ : d_sc_.data(); }

    double *Solver::rzo_ptr() const
    {
        return use_scratch_ ? scratch_ + SCRATCH_SC_PAD + static_cast<size_t>(R_COUNT) * partial_cap_ : d_rzo_.data();
    }

    int Solver::noalloc_max_blocks()
    {
        return static_cast<int>((SCRATCH_DOUBLES - SCRATCH_SC_PAD) / (R_COUNT + 1));
    }

    Reductions Solver::reductions() const
    {
        Reductions r;
        double *const base = use_scratch_ ? scratch_ + SCRATCH_SC_PAD : d_partials_.data();
        auto site = [&](int i) { return base + static_cast<size_t>(i) * partial_cap_; };
        r.pap = site(R_PAP);
        r.rr = site(R_RR);
        r.rz = site(R_RZ);
        r.rz_old = site(R_RZ_OLD);
        r.rr_true = site(R_RR_TRUE);
        r.sc = sc_ptr();
        return r;
    }

    CgReductions Solver::cg_reductions() const
    {
        CgReductions c;
        double *const base = use_scratch_ ? scratch_ + SCRATCH_SC_PAD : d_partials_.data();
        auto site = [&](int i) { return base + static_cast<size_t>(i) * partial_cap_; };
        c.rz = site(R_RZ);
        c.rz_old = site(R_RZ_OLD);
        c.dlt = site(R_DLT);
        c.eps = site(R_EPS);
        c.ps = site(R_PS);
        c.ps_old = site(R_PS_OLD);
        c.zeta = rzo_ptr();
        c.rr = site(R_RR);
        c.sc = sc_ptr();
        return c;
    }

    bool Solver::single_reduction() const
    {
        return params_.single_reduction && !params_.spmv_fuse_direction && params_.spmv_halo && halo_max_ > 0 &&
               halo_own_ok_ && copies_ == 1 && level_num_ <= 4;
    }

    void Solver::finalize_scalars()
    {
        launch_finalize(static_cast<int>(sys_.size()), reductions(), sysmap(), stream_);
    }

    SysMap Solver::sysmap() const
    {
        SysMap m;
        m.block_sys = d_block_sys_.data();
        m.blk_begin = d_blk_begin_.data();
        m.blk_count = d_blk_count_.data();
        m.active = d_active_.data();
        m.select = d_select_.data();
        m.bank_sys = d_bank_sys_.data();
        // One system, not copies of one: its kernels skip the system lookups.
        m.single_blocks = sys_.size() == 1 && copies_ == 1 ? nblocks_ : 0;
        return m;
    }

    Reductions Solver::reductions_of_copy(int c) const
    {
        Reductions r = reductions();
        const size_t o = static_cast<size_t>(c) * nblocks1_;
        r.pap += o;
        r.rr += o;
        r.rz += o;
        r.rz_old += o;
        r.rr_true += o;
        return r;
    }

    SysMap Solver::sysmap_of_copy(int c) const
    {
        SysMap m = sysmap();
        m.block_sys += static_cast<size_t>(c) * nblocks1_;
        return m;
    }

    FineRestrict Solver::fine_restrict() const
    {
        FineRestrict f;
        f.lead = d_lead0_.data();
        f.gid = d_gid0_.data();
        return f;
    }

    template <typename TV>
    SellView<TV> Solver::sell(const TV *vals) const
    {
        SellView<TV> v;
        v.nslices = nslices_;
        v.slice_ptr = d_slice_ptr_.data();
        v.cols = d_cols_.data();
        v.vals = vals;
        v.unroll = params_.spmv_unroll;
        if (params_.spmv_halo && halo_max_ > 0)
        {
            v.lcols = d_lcols_.data();
            v.halo_ptr = d_halo_ptr_.data();
            v.halo = d_halo_.data();
            v.halo_max = halo_max_;
        }
        return v;
    }

    double Solver::spmv_bytes(int val_bytes, int x_bytes, int vec_passes) const
    {
        const bool halo = params_.spmv_halo && halo_max_ > 0;
        return static_cast<double>(nnzb_) * (bd_ * bd_ * val_bytes + (halo ? 2 : 4)) +
               (halo ? static_cast<double>(halo_total_) * 4.0 + (nblocks_ + 1) * 4.0 : (nb_total_ + 1) * 4.0) +
               static_cast<double>(vec_passes) * n_total_ * x_bytes;
    }

    template <typename TV>
    double Solver::spmv_ap_bytes() const
    {
        SymView<TV> v;
        if (copies_ != 1 || !sym_view(v))
            return spmv_bytes(sizeof(TV), sizeof(TV), 2);
        // Stored blocks with their 32-bit metadata, the halo lists, one row word per row, x and y.
        return static_cast<double>(sym_stored_) * (bd_ * bd_ * sizeof(TV) + 4) +
               static_cast<double>(halo_total_) * 4.0 + (nblocks_ + 1) * 4.0 + static_cast<double>(npad_) * 4.0 +
               2.0 * n_total_ * sizeof(TV);
    }

    void Solver::read_scalars()
    {
        h_sc_.resize(static_cast<size_t>(S_COUNT) * sys_.size());
        PSMAS_HIP(hipMemcpyAsync(h_sc_.data(), sc_ptr(), sizeof(double) * h_sc_.size(), hipMemcpyDeviceToHost,
                                 stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
    }

    void Solver::upload_flags()
    {
        PSMAS_HIP(hipMemcpyAsync(d_active_.data(), h_active_.data(), sizeof(int) * h_active_.size(),
                                 hipMemcpyHostToDevice, stream_));
        PSMAS_HIP(hipMemcpyAsync(d_select_.data(), h_select_.data(), sizeof(int) * h_select_.size(),
                                 hipMemcpyHostToDevice, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
    }

    // --------------------------------------------------------------------------- factorize
    void Solver::factorize_batch(const std::vector<CscView> &As)
    {
        PSMAS_HIP(hipSetDevice(device_));
        if (As.empty())
            throw InvalidArgumentError("[MAS] empty batch.");
        for (const auto &A : As)
            validate_csc(A);
        if (As.size() == 1 && try_refresh_values(As[0], nullptr))
            return;
        factorize_full(As, nullptr, nullptr);
    }

    void Solver::factorize_full(const std::vector<CscView> &As, const CscView *raw, const DirichletSystem *dir,
                                const DeviceCsc *dev, const DeviceBuffer<int> *d_keep)
    {
        const int bd = params_.block_dim;
        factorized_ = false;
        factorize_mode_ = "full";
        copies_ = 1;
        setup_ms_.clear();
        auto t_mark = std::chrono::steady_clock::now();
        auto phase = [&](const char *name) {
            // Host wall time of one setup phase (the stream is drained so GPU work is counted).
            PSMAS_HIP(hipStreamSynchronize(stream_));
            const auto now = std::chrono::steady_clock::now();
            setup_ms_.emplace_back(name, std::chrono::duration<double, std::milli>(now - t_mark).count());
            t_mark = now;
        };
        vr_colptr_.clear();
        vr_rowidx_.clear();
        d_vmap_.reset();
        dirichlet_ = false;
        dir_nodes_.clear();
        n_crow_ = 0;
        graphs_->clear();
        const int K = static_cast<int>(As.size());
        if (static_cast<int>(sys_.size()) != K)
            sys_.resize(K);
        const bool bd_changed = (bd != bd_);
        bd_ = bd;
        // The SELL matrix is built on the GPU from the CSC arrays (uploaded here, or the caller's
        // device arrays): one system or every system of a batch.
        const bool gpu_build = dev != nullptr || params_.device_setup;
        vr_nnz_ = -1;
        // The CSC arrays in device memory for the device aggregation and build (host input is
        // uploaded once here; one system's values stay for values-only refactorisations).
        std::vector<DeviceCsc> dcsc(K);
        std::vector<DeviceBuffer<int>> up_cp(K), up_ri(K);
        std::vector<DeviceBuffer<double>> up_v(K);
        if (gpu_build && dev != nullptr)
            dcsc[0] = *dev; // (one system)
        else if (gpu_build)
        {
            for (int u = 0; u < K; ++u)
            {
                const CscView &A = As[u];
                const size_t nz = static_cast<size_t>(std::max<long long>(A.nnz, 1));
                up_cp[u].allocate(static_cast<size_t>(A.n) + 1, "MAS input pattern");
                up_ri[u].allocate(nz, "MAS input pattern");
                DeviceBuffer<double> &v = (K == 1) ? d_csc_vals_ : up_v[u];
                v.ensure(nz, "MAS input values");
                PSMAS_HIP(hipMemcpyAsync(up_cp[u].data(), A.colptr, sizeof(int) * (A.n + 1), hipMemcpyHostToDevice,
                                         stream_));
                if (A.nnz > 0)
                {
                    PSMAS_HIP(hipMemcpyAsync(up_ri[u].data(), A.rowidx, sizeof(int) * A.nnz, hipMemcpyHostToDevice,
                                             stream_));
                    PSMAS_HIP(hipMemcpyAsync(v.data(), A.vals, sizeof(double) * A.nnz, hipMemcpyHostToDevice, stream_));
                }
                dcsc[u] = DeviceCsc{A.n, A.nnz, up_cp[u].data(), up_ri[u].data(), v.data()};
            }
            phase("csc_upload");
        }

        // Per-system partition and bank-padded BSR, concatenated.
        PaddedBsr all;
        all.bd = bd;
        all.rows.assign(1, 0);
        block_of_slot_.clear();
        int node_off = 0;
        nnzb_ = 0;
        nnz_scalar_ = 0;
        n_total_ = 0;
        nb_total_ = 0;
        int sob_off = 0;
        std::vector<int> h_slot_of_block;
        for (int s = 0; s < K; ++s)
        {
            System &S = sys_[s];
            const CscView &A = As[s];
            const int nb = div_up(A.n, bd);
            const bool rebuild = !params_.lazy_partitioning || !S.have_partition ||
                                 static_cast<int>(S.part.perm.size()) != nb || bd_changed;
            S.n = A.n;
            S.nb = nb;
            if (rebuild)
                build_partition(S, A, gpu_build ? &dcsc[s] : nullptr);
            if (s == 0 && K == 1)
                phase("partition");
            PaddedBsr bsr;
            if (gpu_build)
                assign_slots(nb, bd, S.part, bsr);
            else
                build_padded_bsr(A, bd, S.part, bsr);
            S.nparts = bsr.nparts;
            S.npad = bsr.npad;
            S.node_off = node_off;
            S.blk_begin = node_off / NODE_BLOCK;
            S.nnzb = static_cast<long long>(bsr.cols.size());
            S.slot_of_block = bsr.slot_of_block;
            S.sob_off = sob_off;
            sob_off += nb;
            h_slot_of_block.insert(h_slot_of_block.end(), bsr.slot_of_block.begin(), bsr.slot_of_block.end());
            block_of_slot_.insert(block_of_slot_.end(), bsr.block_of_slot.begin(), bsr.block_of_slot.end());
            if (!gpu_build) // (the device build makes the matrix itself; bsr holds only the slots)
            {
                const int row_base = all.rows.back();
                for (int r = 0; r < bsr.npad; ++r)
                    all.rows.push_back(row_base + bsr.rows[r + 1]);
                for (int c : bsr.cols)
                    all.cols.push_back(c + node_off);
                all.vals.insert(all.vals.end(), bsr.vals.begin(), bsr.vals.end());
            }
            node_off += bsr.npad;
            nnzb_ += S.nnzb;
            nnz_scalar_ += A.nnz;
            n_total_ += A.n;
            nb_total_ += nb;
        }
        npad_ = node_off;
        nparts_ = npad_ / BANK;
        nblocks_ = npad_ / NODE_BLOCK;
        all.npad = npad_;
        all.nparts = nparts_;
        dense_built_ = false;
        dense_nd_ = 0;
        dense_note_.clear();
        d_dense_inv_.reset();
        d_dense_ids_.reset();
        d_dense_index_.reset();

        if (gpu_build)
        {
            if (K > 1)
                phase("partition");
            device_matrix_build(As[0], dcsc, dev != nullptr, h_slot_of_block, raw, dir, d_keep, phase);
        }
        else
        {
            nonisolated_.assign(npad_, 0);
            for (int r = 0; r < npad_; ++r)
                for (int k = all.rows[r]; k < all.rows[r + 1]; ++k)
                    if (all.cols[k] != r)
                    {
                        nonisolated_[r] = 1;
                        break;
                    }
            phase(K == 1 ? "padded_bsr" : "partition_and_bsr");
            upload_matrix(all);
            phase("sell_upload_and_halo");
        }
        if (!gpu_build && K == 1 && params_.lazy_partitioning)
        {
            // Remember the pattern and where each non-zero lives, for values-only refactorisation.
            std::vector<int> map;
            if (build_value_map(As[0], bd, sys_[0].slot_of_block, all.rows, all.cols, h_slice_ptr_, map))
            {
                const CscView &P = raw != nullptr ? *raw : As[0];
                if (dir != nullptr)
                {
                    // Map the non-zeros of the matrix before elimination; dropped ones get -1.
                    std::vector<int> full(dir->keep.size());
                    for (size_t k = 0; k < full.size(); ++k)
                        full[k] = dir->keep[k] >= 0 ? map[static_cast<size_t>(dir->keep[k])] : -1;
                    map.swap(full);
                }
                vr_colptr_.assign(P.colptr, P.colptr + P.n + 1);
                vr_rowidx_.assign(P.rowidx, P.rowidx + P.nnz);
                vr_nnz_ = P.nnz;
                d_vmap_.allocate(map.size(), "MAS value map");
                PSMAS_HIP(hipMemcpy(d_vmap_.data(), map.data(), d_vmap_.bytes(), hipMemcpyHostToDevice));
            }
            phase("value_map");
        }
        all = PaddedBsr();

        // System maps.
        std::vector<int> block_sys(nblocks_), blk_begin(K), blk_count(K), bank_sys(nparts_, -1);
        for (int s = 0; s < K; ++s)
        {
            blk_begin[s] = sys_[s].blk_begin;
            blk_count[s] = sys_[s].npad / NODE_BLOCK;
            for (int b = 0; b < blk_count[s]; ++b)
                block_sys[blk_begin[s] + b] = s;
            for (int p = 0; p < sys_[s].nparts; ++p)
                bank_sys[sys_[s].node_off / BANK + p] = s;
        }
        d_block_sys_.allocate(block_sys.size(), "MAS block systems");
        d_blk_begin_.allocate(K, "MAS system blocks");
        d_blk_count_.allocate(K, "MAS system blocks");
        d_bank_sys_.allocate(bank_sys.size(), "MAS bank systems");
        d_active_.allocate(K, "MAS system flags");
        d_select_.allocate(K, "MAS system flags");
        d_slot_of_block_.allocate(h_slot_of_block.size(), "MAS block map");
        PSMAS_HIP(hipMemcpy(d_block_sys_.data(), block_sys.data(), d_block_sys_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_blk_begin_.data(), blk_begin.data(), d_blk_begin_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_blk_count_.data(), blk_count.data(), d_blk_count_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_bank_sys_.data(), bank_sys.data(), d_bank_sys_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_slot_of_block_.data(), h_slot_of_block.data(), d_slot_of_block_.bytes(),
                            hipMemcpyHostToDevice));
        h_active_.assign(K, 1);
        h_select_.assign(K, 1);

        d_vals32_.reset();
        reset_sym_values();
        d_mats16_.reset();
        d_dscale_.reset();
        lattice_stale_ = true;
        rhs_ready_ = false;
        phase("system_maps");
        build_mas();
        phase("coarse_levels_and_inverses");
        {
            // Matrix -> system for the apply's skip test: fine banks as above, coarse matrices -1.
            std::vector<int> mat_sys(total_mats_, -1);
            std::copy(bank_sys.begin(), bank_sys.end(), mat_sys.begin());
            d_bank_sys_.allocate(mat_sys.size(), "MAS matrix systems");
            PSMAS_HIP(hipMemcpy(d_bank_sys_.data(), mat_sys.data(), d_bank_sys_.bytes(), hipMemcpyHostToDevice));
        }
        allocate_vectors();
        if (params_.precision == Precision::Mixed)
            ensure_mixed_data();
        prepare_sym<double>();
        phase("vectors_and_mixed_data");
        factorized_ = true;
    }

    template <typename Phase>
    void Solver::device_matrix_build(const CscView &A, const std::vector<DeviceCsc> &dcsc, bool device_input,
                                     const std::vector<int> &h_slot_of_block, const CscView *raw,
                                     const DirichletSystem *dir, const DeviceBuffer<int> *d_keep, Phase &phase)
    {
        System &S = sys_[0];
        const long long nnz = dcsc[0].nnz;
        DeviceBuffer<int> d_sob(h_slot_of_block.size(), "MAS block map");
        PSMAS_HIP(hipMemcpyAsync(d_sob.data(), h_slot_of_block.data(), d_sob.bytes(), hipMemcpyHostToDevice, stream_));
        SellDevice sd;
        const int K = static_cast<int>(sys_.size());
        if (K == 1)
        {
            build_sell_device(S.n, nnz, bd_, dcsc[0].colptr, dcsc[0].rowidx, dcsc[0].vals, d_sob.data(), S.npad,
                              params_.lazy_partitioning, stream_, sd);
            S.nnzb = sd.nnzb;
        }
        else
        {
            // Every system on its own (its slot map is its segment of the concatenated one), then
            // one block-diagonal matrix.
            std::vector<SellDevice> parts(K);
            std::vector<int> npads(K);
            for (int s = 0; s < K; ++s)
            {
                System &C = sys_[s];
                build_sell_device(C.n, dcsc[s].nnz, bd_, dcsc[s].colptr, dcsc[s].rowidx, dcsc[s].vals,
                                  d_sob.data() + C.sob_off, C.npad, false, stream_, parts[s]);
                C.nnzb = parts[s].nnzb;
                npads[s] = C.npad;
            }
            concat_sell(parts, npads, stream_, sd);
        }
        nnzb_ = sd.nnzb;
        nonisolated_ = std::move(sd.nonisolated);
        nslices_ = npad_ / WAVE;
        nslots_ = sd.nslots;
        h_slice_ptr_ = sd.slice_ptr;
        d_slice_ptr_ = std::move(sd.d_slice_ptr);
        d_cols_ = std::move(sd.d_cols);
        d_vals64_ = std::move(sd.d_vals);
        d_row_len_ = std::move(sd.d_row_len);
        d_block_of_slot_.allocate(block_of_slot_.size(), "MAS slot map");
        PSMAS_HIP(hipMemcpyAsync(d_block_of_slot_.data(), block_of_slot_.data(), d_block_of_slot_.bytes(),
                                 hipMemcpyHostToDevice, stream_));
        phase("device_matrix_build");
        std::vector<int> cols(static_cast<size_t>(nslots_) * WAVE);
        PSMAS_HIP(hipMemcpyAsync(cols.data(), d_cols_.data(), sizeof(int) * cols.size(), hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        build_halo(h_slice_ptr_, std::move(cols));
        phase("halo");
        if (params_.lazy_partitioning && sd.map_ok)
        {
            d_vmap_ = std::move(sd.d_map);
            vr_nnz_ = nnz;
            const CscView &P = raw != nullptr ? *raw : A;
            if (dir != nullptr || d_keep != nullptr)
            {
                // Index the non-zeros of the matrix before elimination; dropped ones get -1.
                vr_nnz_ = P.nnz;
                d_vmap_ = d_keep != nullptr ? compose_value_map(*d_keep, static_cast<size_t>(P.nnz), d_vmap_, stream_)
                                            : compose_value_map(dir->keep, d_vmap_, stream_);
            }
            if (P.colptr != nullptr) // (host input: the pattern for values-only refactorisations)
            {
                vr_colptr_.assign(P.colptr, P.colptr + P.n + 1);
                vr_rowidx_.assign(P.rowidx, P.rowidx + P.nnz);
            }
            phase("value_map");
        }
    }

    void Solver::factorize_device(int n, long long nnz, const int *d_colptr, const int *d_rowidx, const double *d_vals)
    {
        PSMAS_HIP(hipSetDevice(device_));
        validate_csc_device(n, nnz, d_colptr, d_rowidx, stream_);
        if (nnz > 0 && d_vals == nullptr)
            throw InvalidArgumentError("[MAS] invalid A: empty matrix or null arrays.");
        // The host never sees the arrays: domains come from coordinates or the device aggregation.
        const CscView A{n, nnz, nullptr, nullptr, nullptr};
        const DeviceCsc dv{n, nnz, d_colptr, d_rowidx, d_vals};
        factorize_full(std::vector<CscView>{A}, nullptr, nullptr, &dv);
    }

    void Solver::factorize_copies(const CscView &A, int copies, const DeviceCsc *dev)
    {
        PSMAS_HIP(hipSetDevice(device_));
        if (copies < 1 || copies > MAX_COPIES)
            throw InvalidArgumentError("[MAS] invalid number of right-hand sides (1 .. 8).");
        if (dev != nullptr)
        {
            validate_csc_device(dev->n, dev->nnz, dev->colptr, dev->rowidx, stream_);
            if (dev->nnz > 0 && dev->vals == nullptr)
                throw InvalidArgumentError("[MAS] invalid A: empty matrix or null arrays.");
        }
        else
            validate_csc(A);
        // One factorisation (system 0 keeps the coordinates), then per-copy node maps.
        if (sys_.size() > 1)
            sys_.resize(1);
        factorize_full(std::vector<CscView>{A}, nullptr, nullptr, dev);
        if (copies == 1)
            return;
        const auto t0 = std::chrono::steady_clock::now();
        // The coarsest-level inverse is built now, from the single system's maps, whatever the
        // coarse_solver parameter (it may be switched on later).
        build_dense_coarse();
        expand_copies(copies);
        PSMAS_HIP(hipStreamSynchronize(stream_));
        setup_ms_.emplace_back("expand_copies",
                               std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }

    void Solver::expand_copies(int k)
    {
        const int npad1 = npad_, nblocks1 = nblocks_, levels = cl_.levels;
        const System S0 = sys_[0];
        // Level-major multi-level numbering: level l holds copy 0's cnt[l] banks, then copy 1's...
        CopyLayout cl = cl_;
        cl.copies = k;
        int P[CopyLayout::MAXL] = {}; // one copy's (bank-padded) nodes of each level
        cl.moff[0] = 0;
        for (int l = 0; l < levels; ++l)
        {
            P[l] = BANK * cl.cnt[l];
            cl.moff[l + 1] = cl.moff[l] + k * cl.cnt[l];
        }
        for (int l = 0; l <= MAX_COARSE_LEVELS; ++l)
        {
            mat_off_[l] = l < levels ? cl.moff[l] : cl.moff[levels];
            mat_cnt_[l] = l < levels ? k * cl.cnt[l] : 0;
            lev_off_[l] = BANK * mat_off_[l];
        }
        total_mats_ = cl.moff[levels];
        total_level_nodes_ = BANK * total_mats_;
        cl_ = cl;
        npad_ = k * npad1;
        nparts_ = npad_ / BANK;
        nblocks_ = k * nblocks1;
        n_total_ *= k;
        nb_total_ *= k;
        sys_.assign(k, S0);
        for (int c = 0; c < k; ++c)
        {
            System &S = sys_[c];
            S.node_off = c * npad1;
            S.blk_begin = c * nblocks1;
            S.sob_off = c * S0.nb;
        }

        // Node maps: copy c's coarse ids are copy 0's plus c times the level's padded size.
        std::vector<int4> cm(npad1);
        std::vector<uint32_t> lead(npad1);
        std::vector<int> gid(npad1), sob(S0.slot_of_block);
        PSMAS_HIP(hipMemcpyAsync(cm.data(), d_cmap_.data(), sizeof(int4) * npad1, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipMemcpyAsync(lead.data(), d_lead0_.data(), sizeof(uint32_t) * npad1, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipMemcpyAsync(gid.data(), d_gid0_.data(), sizeof(int) * npad1, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        std::vector<int4> cmk(static_cast<size_t>(npad_));
        std::vector<uint32_t> leadk(static_cast<size_t>(npad_));
        std::vector<int> gidk(static_cast<size_t>(npad_)), bosk(static_cast<size_t>(npad_)),
            sobk(static_cast<size_t>(k) * sob.size());
        std::vector<char> nonisok(static_cast<size_t>(npad_));
        auto shift = [](int id, int by) { return id >= 0 ? id + by : id; };
        for (int c = 0; c < k; ++c)
        {
            const size_t o = static_cast<size_t>(c) * npad1;
            for (int p = 0; p < npad1; ++p)
            {
                const int4 v = cm[p];
                cmk[o + p] = int4{shift(v.x, c * P[1]), shift(v.y, c * P[2]), shift(v.z, c * P[3]), shift(v.w, c * P[4])};
                leadk[o + p] = lead[p];
                gidk[o + p] = shift(gid[p], c * P[1]);
                bosk[o + p] = block_of_slot_[p];
                nonisok[o + p] = nonisolated_[p];
            }
            std::copy(sob.begin(), sob.end(), sobk.begin() + static_cast<size_t>(c) * sob.size());
        }
        block_of_slot_.swap(bosk);
        nonisolated_.swap(nonisok);
        auto upload = [&](auto &buf, const auto &host, const char *what) {
            buf.allocate(host.size(), what);
            PSMAS_HIP(hipMemcpy(buf.data(), host.data(), buf.bytes(), hipMemcpyHostToDevice));
        };
        upload(d_cmap_, cmk, "MAS coarse map");
        upload(d_lead0_, leadk, "MAS level-0 leader masks");
        upload(d_gid0_, gidk, "MAS level-0 ids");
        upload(d_block_of_slot_, block_of_slot_, "MAS slot map");
        upload(d_slot_of_block_, sobk, "MAS block map");

        // Systems: copy c's node blocks and fine banks.
        std::vector<int> block_sys(nblocks_), blk_begin(k), blk_count(k), bank_sys(total_mats_, -1);
        for (int c = 0; c < k; ++c)
        {
            blk_begin[c] = c * nblocks1;
            blk_count[c] = nblocks1;
            std::fill(block_sys.begin() + c * nblocks1, block_sys.begin() + (c + 1) * nblocks1, c);
            std::fill(bank_sys.begin() + c * cl.cnt[0], bank_sys.begin() + (c + 1) * cl.cnt[0], c);
        }
        upload(d_block_sys_, block_sys, "MAS block systems");
        upload(d_blk_begin_, blk_begin, "MAS system blocks");
        upload(d_blk_count_, blk_count, "MAS system blocks");
        upload(d_bank_sys_, bank_sys, "MAS matrix systems");
        d_active_.allocate(k, "MAS system flags");
        d_select_.allocate(k, "MAS system flags");
        h_active_.assign(k, 1);
        h_select_.assign(k, 1);
        PSMAS_HIP(hipMemcpy(d_lev_off_.data(), lev_off_, sizeof(lev_off_), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_mat_off_.data(), mat_off_, sizeof(mat_off_), hipMemcpyHostToDevice));
        build_restrict_chunks(); // from the copies' coarse map: each copy's lists are copy 0's, shifted
        allocate_vectors();
        if (d_vals32_)
            ensure_mixed_data();
        copies_ = k;
        lattice_stale_ = true;
        graphs_->clear();
    }

    bool Solver::try_refresh_values(const CscView &A, const std::vector<int> *dirichlet_nodes)
    {
        if (!params_.lazy_partitioning || !factorized_ || sys_.size() != 1 || !d_vmap_ ||
            params_.block_dim != bd_ || A.n != sys_[0].n || static_cast<size_t>(A.nnz) != vr_rowidx_.size() ||
            (dirichlet_nodes != nullptr) != dirichlet_ ||
            (dirichlet_nodes != nullptr && *dirichlet_nodes != dir_nodes_) ||
            !same_ints(A.colptr, vr_colptr_.data(), vr_colptr_.size()) ||
            !same_ints(A.rowidx, vr_rowidx_.data(), vr_rowidx_.size()))
            return false;
        factorized_ = false;
        graphs_->clear();
        if (d_csc_vals_.size() != static_cast<size_t>(A.nnz))
            d_csc_vals_.allocate(static_cast<size_t>(A.nnz), "MAS input values");
        PSMAS_HIP(hipMemcpyAsync(d_csc_vals_.data(), A.vals, d_csc_vals_.bytes(), hipMemcpyHostToDevice, stream_));
        scatter_new_values(d_csc_vals_.data());
        finish_value_refresh();
        return true;
    }

    void Solver::scatter_new_values(const double *d_vals)
    {
        launch_scatter_values(vr_nnz_, d_vmap_.data(), d_vals, d_vals64_.data(), stream_);
        if (dirichlet_ && n_coupl_ > 0) // (the coupling buffers hold one unused entry when there are none)
            launch_gather_values(static_cast<int>(n_coupl_), d_cidx_.data(), d_vals, d_cval_.data(), stream_);
    }

    void Solver::dirichlet_factorize(const CscView &A, const int *nodes, int ndir)
    {
        PSMAS_HIP(hipSetDevice(device_));
        validate_csc(A);
        std::vector<int> list = dirichlet_list(A.n, nodes, ndir);
        if (try_refresh_values(A, &list))
            return;
        if (params_.device_setup)
        {
            dirichlet_factorize_device(A, list);
            return;
        }
        std::vector<unsigned char> fixed(A.n, 0);
        for (int d : list)
            fixed[d] = 1;
        DirichletSystem D;
        eliminate_dirichlet(A, fixed, D);
        const CscView At{A.n, static_cast<long long>(D.rowidx.size()), D.colptr.data(), D.rowidx.data(),
                         D.vals.data()};
        factorize_full(std::vector<CscView>{At}, &A, &D);
        n_crow_ = static_cast<int>(D.crow.size());
        const size_t nc = D.ccol.size();
        std::vector<double> cval(nc);
        for (size_t k = 0; k < nc; ++k)
            cval[k] = A.vals[D.cidx[k]];
        d_crow_.allocate(std::max<size_t>(1, D.crow.size()), "MAS Dirichlet couplings");
        d_cptr_.allocate(D.cptr.size(), "MAS Dirichlet couplings");
        d_ccol_.allocate(std::max<size_t>(1, nc), "MAS Dirichlet couplings");
        d_cidx_.allocate(std::max<size_t>(1, nc), "MAS Dirichlet couplings");
        d_cval_.allocate(std::max<size_t>(1, nc), "MAS Dirichlet couplings");
        PSMAS_HIP(hipMemcpy(d_cptr_.data(), D.cptr.data(), d_cptr_.bytes(), hipMemcpyHostToDevice));
        if (nc > 0)
        {
            PSMAS_HIP(hipMemcpy(d_crow_.data(), D.crow.data(), sizeof(int) * D.crow.size(), hipMemcpyHostToDevice));
            PSMAS_HIP(hipMemcpy(d_ccol_.data(), D.ccol.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
            PSMAS_HIP(hipMemcpy(d_cidx_.data(), D.cidx.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
            PSMAS_HIP(hipMemcpy(d_cval_.data(), cval.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
        }
        n_coupl_ = static_cast<long long>(nc);
        dir_nodes_.swap(list);
        dirichlet_ = true;
    }

    void Solver::dirichlet_factorize_device(const CscView &A, std::vector<int> &list)
    {
        const auto t0 = std::chrono::steady_clock::now();
        // A once in device memory: its values stay for values-only refactorisations.
        const size_t nz = static_cast<size_t>(std::max<long long>(A.nnz, 1));
        DeviceBuffer<int> cp(static_cast<size_t>(A.n) + 1, "MAS input pattern"), ri(nz, "MAS input pattern");
        d_csc_vals_.ensure(nz, "MAS input values");
        PSMAS_HIP(hipMemcpyAsync(cp.data(), A.colptr, sizeof(int) * (A.n + 1), hipMemcpyHostToDevice, stream_));
        if (A.nnz > 0)
        {
            PSMAS_HIP(hipMemcpyAsync(ri.data(), A.rowidx, sizeof(int) * A.nnz, hipMemcpyHostToDevice, stream_));
            PSMAS_HIP(hipMemcpyAsync(d_csc_vals_.data(), A.vals, sizeof(double) * A.nnz, hipMemcpyHostToDevice,
                                     stream_));
        }
        dirichlet_eliminate_and_factorize(A, DeviceCsc{A.n, A.nnz, cp.data(), ri.data(), d_csc_vals_.data()}, list,
                                          t0);
    }

    void Solver::dirichlet_factorize_device_input(int n, long long nnz, const int *d_colptr, const int *d_rowidx,
                                                  const double *d_vals, const int *nodes, int ndir)
    {
        PSMAS_HIP(hipSetDevice(device_));
        const auto t0 = std::chrono::steady_clock::now();
        validate_csc_device(n, nnz, d_colptr, d_rowidx, stream_);
        if (nnz > 0 && d_vals == nullptr)
            throw InvalidArgumentError("[MAS] invalid A: empty matrix or null arrays.");
        std::vector<int> list = dirichlet_list(n, nodes, ndir);
        // The host never sees A: later values come through psmas_factorize_values_device.
        dirichlet_eliminate_and_factorize(CscView{n, nnz, nullptr, nullptr, nullptr},
                                          DeviceCsc{n, nnz, d_colptr, d_rowidx, d_vals}, list, t0);
    }

    std::vector<int> Solver::dirichlet_list(int n, const int *nodes, int ndir)
    {
        if (ndir < 0 || (ndir > 0 && nodes == nullptr))
            throw InvalidArgumentError("[MAS] invalid Dirichlet node list.");
        std::vector<int> list(nodes, nodes + ndir);
        for (int d : list)
            if (d < 0 || d >= n)
                throw InvalidArgumentError("[MAS] Dirichlet node index out of range.");
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
        return list;
    }

    void Solver::dirichlet_eliminate_and_factorize(const CscView &A, const DeviceCsc &dA, std::vector<int> &list,
                                                   std::chrono::steady_clock::time_point t0)
    {
        DeviceBuffer<int> dn(std::max<size_t>(list.size(), 1), "MAS Dirichlet nodes");
        if (!list.empty())
            PSMAS_HIP(hipMemcpyAsync(dn.data(), list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice,
                                     stream_));
        DirichletDevice D;
        eliminate_dirichlet_device(dA.n, dA.nnz, dA.colptr, dA.rowidx, dA.vals, dn.data(),
                                   static_cast<int>(list.size()), stream_, D);
        const double elim_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const CscView At{A.n, D.nnz, nullptr, nullptr, nullptr};
        const DeviceCsc dv{A.n, D.nnz, D.colptr.data(), D.rowidx.data(), D.vals.data()};
        factorize_full(std::vector<CscView>{At}, &A, nullptr, &dv, &D.keep);
        setup_ms_.insert(setup_ms_.begin(), {"dirichlet_elimination", elim_ms});
        n_crow_ = D.nrows;
        n_coupl_ = D.ncoupl;
        d_crow_ = std::move(D.crow);
        d_cptr_ = std::move(D.cptr);
        d_ccol_ = std::move(D.ccol);
        d_cidx_ = std::move(D.cidx);
        d_cval_.allocate(static_cast<size_t>(std::max<long long>(n_coupl_, 1)), "MAS Dirichlet couplings");
        if (n_coupl_ > 0)
            launch_gather_values(static_cast<int>(n_coupl_), d_cidx_.data(), dA.vals, d_cval_.data(), stream_);
        PSMAS_HIP(hipStreamSynchronize(stream_));
        dir_nodes_.swap(list);
        dirichlet_ = true;
    }

    void Solver::dirichlet_lift_device(const double *d_f, double *d_g)
    {
        const int n = sys_[0].n;
        if (d_g != d_f)
            PSMAS_HIP(hipMemcpyAsync(d_g, d_f, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
        tic("dirichlet_lift", static_cast<double>(n_coupl_) * 20.0 + static_cast<double>(n_crow_) * 24.0);
        launch_dirichlet_lift(n_crow_, d_crow_.data(), d_cptr_.data(), d_ccol_.data(), d_cval_.data(), d_f, d_g,
                              stream_);
        toc();
    }

    void Solver::check_dirichlet(int n) const
    {
        if (!factorized_ || !dirichlet_)
            throw InvalidArgumentError("[MAS] no Dirichlet factorisation: call the Dirichlet factorize first.");
        if (n != sys_[0].n)
            throw InvalidArgumentError("[MAS] invalid f: size does not match the factorised matrix.");
    }

    void Solver::dirichlet_solve_host(int n, double *f, double *u)
    {
        PSMAS_HIP(hipSetDevice(device_));
        check_dirichlet(n);
        if (f == nullptr || u == nullptr)
            throw InvalidArgumentError("[MAS] invalid f or u: null pointer.");
        if (d_dir_f_.size() != static_cast<size_t>(n))
            d_dir_f_.allocate(static_cast<size_t>(n), "MAS Dirichlet right-hand side");
        PSMAS_HIP(hipMemcpyAsync(d_dir_f_.data(), f, sizeof(double) * n, hipMemcpyHostToDevice, stream_));
        dirichlet_lift_device(d_dir_f_.data(), d_dir_f_.data());
        PSMAS_HIP(hipMemcpyAsync(f, d_dir_f_.data(), sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        solve_host(n, f, u);
    }

    void Solver::dirichlet_solve_device(int n, double *d_f, double *d_u)
    {
        PSMAS_HIP(hipSetDevice(device_));
        check_dirichlet(n);
        if (d_f == nullptr || d_u == nullptr)
            throw InvalidArgumentError("[MAS] invalid f or u: null pointer.");
        dirichlet_lift_device(d_f, d_f);
        solve_device(n, d_f, d_u);
    }

    void Solver::refresh_values_device(long long nnz, const double *d_vals)
    {
        PSMAS_HIP(hipSetDevice(device_));
        if (!factorized_ || sys_.size() != 1 || !d_vmap_ || d_vals == nullptr || nnz != vr_nnz_)
            throw InvalidArgumentError(
                "[MAS] device values need a recorded pattern: a full factorisation of one system with "
                "lazy_partitioning, no duplicate entries, and the same number of non-zeros.");
        factorized_ = false;
        graphs_->clear();
        scatter_new_values(d_vals);
        finish_value_refresh();
    }

    void Solver::finish_value_refresh()
    {
        // Numeric parts only: the partition, SELL pattern, coarse levels, restriction lists and
        // lattice depend on the pattern (and coordinates), which are unchanged.
        dense_built_ = false;
        dense_nd_ = 0;
        dense_note_.clear();
        d_dense_inv_.reset();
        d_dense_ids_.reset();
        d_dense_index_.reset();
        d_vals32_.reset();
        reset_sym_values();
        d_mats16_.reset();
        d_dscale_.reset();
        rhs_ready_ = false;
        build_bank_inverses();
        if (params_.precision == Precision::Mixed)
            ensure_mixed_data();
        prepare_sym<double>();
        PSMAS_HIP(hipStreamSynchronize(stream_));
        factorize_mode_ = "values";
        factorized_ = true;
    }

    void Solver::build_partition(System &S, const CscView &A, const DeviceCsc *dcsc)
    {
        const int nb = S.nb;
        Aggregation agg = params_.aggregation;
        const bool have_coords = S.coord_dim > 0 && static_cast<int>(S.coords.size()) == nb * S.coord_dim;
        if (agg == Aggregation::Auto)
            agg = have_coords ? Aggregation::Morton : Aggregation::Graph;
        if (agg == Aggregation::Morton && !have_coords)
            throw InvalidArgumentError(
                "[MAS] Morton aggregation needs one coordinate tuple per block node (psmas_set_coordinates).");
        if (agg == Aggregation::Morton)
        {
            const int dim = S.coord_dim;
            double lo[3] = {0, 0, 0}, inv[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d)
            {
                double mn = S.coords[d], mx = S.coords[d];
                for (int i = 0; i < nb; ++i)
                {
                    mn = std::min(mn, S.coords[static_cast<size_t>(i) * dim + d]);
                    mx = std::max(mx, S.coords[static_cast<size_t>(i) * dim + d]);
                }
                lo[d] = mn;
                inv[d] = (mx > mn) ? 1.0 / (mx - mn) : 0.0;
            }
            DeviceBuffer<double> d_xyz(S.coords.size(), "MAS morton coordinates");
            DeviceBuffer<double> d_lo(3, "MAS morton box"), d_inv(3, "MAS morton box");
            DeviceBuffer<uint64_t> keys(nb, "MAS morton keys"), keys_out(nb, "MAS morton keys");
            DeviceBuffer<int> ids(nb, "MAS morton ids"), ids_out(nb, "MAS morton ids");
            PSMAS_HIP(hipMemcpyAsync(d_xyz.data(), S.coords.data(), d_xyz.bytes(), hipMemcpyHostToDevice, stream_));
            PSMAS_HIP(hipMemcpyAsync(d_lo.data(), lo, sizeof(lo), hipMemcpyHostToDevice, stream_));
            PSMAS_HIP(hipMemcpyAsync(d_inv.data(), inv, sizeof(inv), hipMemcpyHostToDevice, stream_));
            launch_morton_keys(nb, dim, d_xyz.data(), d_lo.data(), d_inv.data(), keys.data(), ids.data(), stream_);
            const size_t tb = morton_sort_temp_bytes(nb);
            DeviceBuffer<unsigned char> temp(std::max<size_t>(tb, 1), "MAS morton sort");
            morton_sort(nb, temp.data(), tb, keys.data(), keys_out.data(), ids.data(), ids_out.data(), stream_);
            std::vector<int> order(nb);
            PSMAS_HIP(hipMemcpyAsync(order.data(), ids_out.data(), sizeof(int) * nb, hipMemcpyDeviceToHost, stream_));
            PSMAS_HIP(hipStreamSynchronize(stream_));
            S.part = chunk_partition(order, BANK);
            S.aggregation = "morton";
        }
        else if (dcsc != nullptr)
        {
            S.part = gpu_aggregate(dcsc->n, dcsc->nnz, bd_, dcsc->colptr, dcsc->rowidx, dcsc->vals, BANK, stream_);
            S.aggregation = "graph_gpu";
        }
        else
        {
            const BlockGraph g = build_block_graph(A, bd_);
            S.part = graph_aggregate(g, BANK);
            S.aggregation = "graph";
        }
        S.have_partition = true;
    }

    void Solver::upload_matrix(const PaddedBsr &bsr)
    {
        const int bs = bd_ * bd_;
        nslices_ = npad_ / WAVE;
        std::vector<int> slice_ptr(nslices_ + 1, 0);
        std::vector<int> row_len(npad_, 0);
        for (int r = 0; r < npad_; ++r)
            row_len[r] = bsr.rows[r + 1] - bsr.rows[r];
        for (int s = 0; s < nslices_; ++s)
        {
            int mx = 1; // every slice at least one slot (as the device build): each row reads its own node
            for (int l = 0; l < WAVE; ++l)
                mx = std::max(mx, row_len[s * WAVE + l]);
            slice_ptr[s + 1] = slice_ptr[s] + mx;
        }
        nslots_ = slice_ptr[nslices_];
        h_slice_ptr_ = slice_ptr;
        std::vector<int> cols(static_cast<size_t>(nslots_) * WAVE);
        std::vector<double> vals(static_cast<size_t>(nslots_) * WAVE * bs);
        auto fill = [&](int s0, int s1) {
            for (int s = s0; s < s1; ++s)
            {
                const int len = slice_ptr[s + 1] - slice_ptr[s];
                for (int l = 0; l < WAVE; ++l)
                {
                    const int row = s * WAVE + l;
                    const int rl = row_len[row];
                    for (int k = 0; k < len; ++k)
                    {
                        const size_t slot = static_cast<size_t>(slice_ptr[s] + k);
                        if (k < rl)
                        {
                            const size_t src = static_cast<size_t>(bsr.rows[row] + k);
                            cols[slot * WAVE + l] = bsr.cols[src];
                            for (int e = 0; e < bs; ++e)
                                vals[(slot * bs + e) * WAVE + l] = bsr.vals[src * bs + e];
                        }
                        else
                        {
                            cols[slot * WAVE + l] = row;
                            for (int e = 0; e < bs; ++e)
                                vals[(slot * bs + e) * WAVE + l] = 0.0;
                        }
                    }
                }
            }
        };
        {
            const int nt = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
            std::vector<std::thread> th;
            const int chunk = div_up(nslices_, nt);
            for (int t = 0; t < nt; ++t)
            {
                const int b = t * chunk, e = std::min(nslices_, b + chunk);
                if (b < e)
                    th.emplace_back(fill, b, e);
            }
            for (auto &x : th)
                x.join();
        }
        d_slice_ptr_.allocate(slice_ptr.size(), "MAS SELL slice pointers");
        d_cols_.allocate(cols.size(), "MAS SELL columns");
        d_vals64_.allocate(vals.size(), "MAS SELL values (FP64)");
        d_row_len_.allocate(row_len.size(), "MAS SELL row lengths");
        d_block_of_slot_.allocate(block_of_slot_.size(), "MAS slot map");
        PSMAS_HIP(hipMemcpy(d_slice_ptr_.data(), slice_ptr.data(), d_slice_ptr_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_cols_.data(), cols.data(), d_cols_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_vals64_.data(), vals.data(), d_vals64_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_row_len_.data(), row_len.data(), d_row_len_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_block_of_slot_.data(), block_of_slot_.data(), d_block_of_slot_.bytes(),
                            hipMemcpyHostToDevice));
        build_halo(slice_ptr, std::move(cols));
    }

    void Solver::build_halo(const std::vector<int> &slice_ptr, std::vector<int> cols)
    {
        halo_max_ = 0;
        halo_total_ = 0;
        halo_own_ok_ = false;
        if (sym_future_.valid())
            sym_future_.wait(); // (a previous pattern's build; its result is dropped)
        sym_future_ = std::future<SymSell>();
        sym_built_ = false;
        sym_slots_ = 0;
        sym_stored_ = 0;
        sym_incoming_max_ = 0;
        sym_note_ = "no halo lists";
        d_sym_slice_ptr_.reset();
        d_sym_src_.reset();
        d_sym_mirror_.reset();
        d_sym_meta_.reset();
        d_sym_info_.reset();
        reset_sym_values(true);
        d_lcols_.reset();
        d_halo_ptr_.reset();
        d_halo_.reset();
        d_halo_own_.reset();
        constexpr int SLICES = NODE_BLOCK / WAVE;
        const int nblk = npad_ / NODE_BLOCK;
        std::vector<std::vector<int>> lists(nblk);
        std::vector<unsigned short> lcols(cols.size());
        std::atomic<bool> fits{true};
        auto work = [&](int b0, int b1) {
            std::vector<int> local(npad_, -1);
            for (int b = b0; b < b1 && fits; ++b)
            {
                std::vector<int> &u = lists[b];
                const size_t i0 = static_cast<size_t>(slice_ptr[b * SLICES]) * WAVE;
                const size_t i1 = static_cast<size_t>(slice_ptr[(b + 1) * SLICES]) * WAVE;
                for (size_t i = i0; i < i1; ++i)
                    if (local[cols[i]] < 0)
                    {
                        local[cols[i]] = 0;
                        u.push_back(cols[i]);
                    }
                std::sort(u.begin(), u.end());
                const bool ok = static_cast<int>(u.size()) <= SPMV_HALO_CAP;
                if (ok)
                {
                    for (size_t k = 0; k < u.size(); ++k)
                        local[u[k]] = static_cast<int>(k);
                    for (size_t i = i0; i < i1; ++i)
                        lcols[i] = static_cast<unsigned short>(local[cols[i]]);
                }
                else
                {
                    fits = false;
                }
                for (int c : u)
                    local[c] = -1;
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
        if (!fits)
            return;
        std::vector<int> ptr(static_cast<size_t>(nblk) + 1, 0);
        int mx = 0;
        for (int b = 0; b < nblk; ++b)
        {
            ptr[b + 1] = ptr[b] + static_cast<int>(lists[b].size());
            mx = std::max(mx, static_cast<int>(lists[b].size()));
        }
        std::vector<int> all(static_cast<size_t>(ptr[nblk]));
        for (int b = 0; b < nblk; ++b)
            std::copy(lists[b].begin(), lists[b].end(), all.begin() + ptr[b]);
        // Where each block's own nodes sit in its (ascending) list: every row reads its own node
        // through its diagonal block (padded rows point at themselves), so they form one run.
        std::vector<int> own(static_cast<size_t>(std::max(nblk, 1)), 0);
        bool own_ok = true;
        for (int b = 0; b < nblk && own_ok; ++b)
        {
            const std::vector<int> &u = lists[b];
            const int first = b * NODE_BLOCK;
            const auto it = std::lower_bound(u.begin(), u.end(), first);
            const int pos = static_cast<int>(it - u.begin());
            own[b] = pos;
            if (pos + NODE_BLOCK > static_cast<int>(u.size()))
                own_ok = false;
            else
                for (int t = 0; t < NODE_BLOCK && own_ok; ++t)
                    own_ok = u[pos + t] == first + t;
        }
        d_lcols_.allocate(lcols.size(), "MAS SELL local columns");
        d_halo_ptr_.allocate(ptr.size(), "MAS SpMV halo lists");
        d_halo_.allocate(std::max<size_t>(1, all.size()), "MAS SpMV halo lists");
        PSMAS_HIP(hipMemcpy(d_lcols_.data(), lcols.data(), d_lcols_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_halo_ptr_.data(), ptr.data(), d_halo_ptr_.bytes(), hipMemcpyHostToDevice));
        if (!all.empty())
            PSMAS_HIP(hipMemcpy(d_halo_.data(), all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice));
        if (own_ok)
        {
            d_halo_own_.allocate(own.size(), "MAS SpMV halo own positions");
            PSMAS_HIP(hipMemcpy(d_halo_own_.data(), own.data(), d_halo_own_.bytes(), hipMemcpyHostToDevice));
        }
        halo_own_ok_ = own_ok;
        halo_max_ = mx;
        halo_total_ = static_cast<long long>(all.size());
        build_sym(std::vector<int>(slice_ptr), std::move(cols), std::move(lcols));
    }

    void Solver::build_sym(std::vector<int> slice_ptr, std::vector<int> cols, std::vector<unsigned short> lcols)
    {
        // The host build runs beside the rest of the factorisation (GPU work); join_sym() takes
        // its result before the first use.
        std::vector<int> row_len(static_cast<size_t>(npad_));
        PSMAS_HIP(hipMemcpy(row_len.data(), d_row_len_.data(), sizeof(int) * row_len.size(), hipMemcpyDeviceToHost));
        sym_note_ = "being built";
        sym_future_ = std::async(std::launch::async,
                                 [npad = npad_, sp = std::move(slice_ptr), c = std::move(cols), lc = std::move(lcols),
                                  rl = std::move(row_len)]() {
                                     SymSell S;
                                     build_symmetric_sell(npad, sp, c, lc, rl, S);
                                     return S;
                                 });
    }

    void Solver::join_sym()
    {
        if (!sym_future_.valid())
            return;
        SymSell S = sym_future_.get();
        if (S.slots == 0)
        {
            sym_note_ = S.reason;
            return;
        }
        auto up = [](auto &buf, const auto &v, const char *what) {
            buf.allocate(std::max<size_t>(1, v.size()), what);
            if (!v.empty())
                PSMAS_HIP(hipMemcpy(buf.data(), v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice));
        };
        up(d_sym_slice_ptr_, S.slice_ptr, "MAS symmetric SELL slice pointers");
        up(d_sym_meta_, S.meta, "MAS symmetric SELL columns");
        up(d_sym_src_, S.src, "MAS symmetric SELL value map");
        up(d_sym_mirror_, S.mirror, "MAS symmetric SELL value map");
        up(d_sym_info_, S.row_info, "MAS symmetric SELL rows");
        sym_slots_ = S.slots;
        sym_stored_ = S.stored;
        sym_incoming_max_ = S.incoming_max;
        sym_built_ = true;
        sym_note_.clear();
    }

    void Solver::reset_sym_values(bool free_buffers)
    {
        // New values: the check and the gathers run again at the next prepare_sym, into the same
        // buffers (freeing and re-allocating 0.2-0.4 GB per refresh would cost more than both).
        if (free_buffers)
        {
            d_symv32_.reset();
            d_symv64_.reset();
        }
        symv_fresh_[0] = symv_fresh_[1] = false;
        sym_values_ok_ = -1;
    }

    template <typename TV>
    bool Solver::prepare_sym()
    {
        if (!use_scratch_)
            join_sym();
        if (!params_.spmv_symmetric || !sym_built_ || copies_ != 1 || !params_.spmv_halo || halo_max_ < NODE_BLOCK)
            return false;
        const size_t lds = (static_cast<size_t>(halo_max_) * bd_ * sizeof(TV) + 15) / 16 * 16 +
                           static_cast<size_t>(sym_incoming_max_) * bd_ * sizeof(TV);
        if (lds > SPMV_SYM_LDS_CAP)
            return false;
        const long long entries = sym_slots_ * WAVE;
        if (use_scratch_) // (the allocation-free solve uses what exists)
        {
            SymView<TV> v;
            return sym_view(v);
        }
        if (sym_values_ok_ < 0)
        {
            if (!d_sym_flag_)
                d_sym_flag_.allocate(1, "MAS symmetry check");
            launch_fill<int>(1, d_sym_flag_.data(), 0, stream_);
            launch_sym_check(bd_, entries, d_sym_src_.data(), d_sym_mirror_.data(), d_vals64_.data(), 1e-12,
                             d_sym_flag_.data(), stream_);
            int h = 0;
            PSMAS_HIP(hipMemcpyAsync(&h, d_sym_flag_.data(), sizeof(int), hipMemcpyDeviceToHost, stream_));
            PSMAS_HIP(hipStreamSynchronize(stream_));
            sym_values_ok_ = h == 0 ? 1 : 0;
            sym_note_ = h == 0 ? "" : "the matrix is not symmetric to 1e-12 of its largest entries";
        }
        if (sym_values_ok_ == 0)
            return false;
        DeviceBuffer<TV> &buf = [this]() -> DeviceBuffer<TV> & {
            if constexpr (std::is_same_v<TV, float>)
                return d_symv32_;
            else
                return d_symv64_;
        }();
        bool &fresh = symv_fresh_[std::is_same_v<TV, float> ? 0 : 1];
        if (!buf || !fresh)
        {
            if (!buf)
                buf.allocate(static_cast<size_t>(entries) * bd_ * bd_, "MAS symmetric SELL values");
            launch_sym_gather<TV>(bd_, entries, d_sym_src_.data(), d_vals64_.data(), buf.data(), stream_);
            fresh = true;
        }
        return true;
    }

    template <typename TV>
    bool Solver::sym_view(SymView<TV> &out) const
    {
        const DeviceBuffer<TV> &buf = [this]() -> const DeviceBuffer<TV> & {
            if constexpr (std::is_same_v<TV, float>)
                return d_symv32_;
            else
                return d_symv64_;
        }();
        if (!params_.spmv_symmetric || !sym_built_ || sym_values_ok_ != 1 || copies_ != 1 || !params_.spmv_halo ||
            halo_max_ < NODE_BLOCK || !buf || !symv_fresh_[std::is_same_v<TV, float> ? 0 : 1])
            return false;
        out.nslices = nslices_;
        out.slice_ptr = d_sym_slice_ptr_.data();
        out.meta = d_sym_meta_.data();
        out.vals = buf.data();
        out.row_info = d_sym_info_.data();
        out.incoming_max = sym_incoming_max_;
        out.unroll = params_.spmv_unroll;
        return true;
    }

    void Solver::build_mas()
    {
        const SellView<double> A = sell(d_vals64_.data());
        const int banks = npad_ / BANK;
        d_lead0_.allocate(npad_, "MAS level-0 leader masks");
        d_gid0_.allocate(npad_, "MAS level-0 ids");
        DeviceBuffer<int> local(npad_, "MAS local component ids");
        DeviceBuffer<int> bank_count(banks, "MAS component counts");
        DeviceBuffer<int> prefix(banks + 1, "MAS component prefix");
        launch_cc_level0(npad_, A, d_row_len_.data(), d_block_of_slot_.data(), d_lead0_.data(), local.data(),
                         bank_count.data(), stream_);
        launch_scan(banks, bank_count.data(), prefix.data(), stream_);
        int total = 0;
        PSMAS_HIP(hipMemcpyAsync(&total, prefix.data() + banks, sizeof(int), hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        for (int l = 0; l < MAX_COARSE_LEVELS; ++l)
        {
            cco_[l] = 0;
            d_leadc_[l].reset();
            d_gidc_[l].reset();
        }
        cco_[0] = total;
        launch_make_gid(npad_, local.data(), prefix.data(), d_gid0_.data(), stream_);
        d_cmap_.allocate(npad_, "MAS coarse map");
        launch_fill<int>(static_cast<long long>(npad_) * 4, reinterpret_cast<int *>(d_cmap_.data()), -1, stream_);
        launch_fill_cmap(npad_, 0, d_block_of_slot_.data(), d_gid0_.data(), d_cmap_.data(), stream_);
        level_num_ = 1;

        for (int lv = 1; lv < MAX_COARSE_LEVELS; ++lv)
        {
            const int count = cco_[lv - 1];
            const int nbank = div_up(count, BANK);
            DeviceBuffer<uint32_t> nbr(count, "MAS coarse neighbour masks");
            launch_fill<unsigned>(count, nbr.data(), 0u, stream_);
            launch_level_masks(npad_, lv, A, d_row_len_.data(), d_block_of_slot_.data(), d_cmap_.data(), nbr.data(),
                               stream_);
            DeviceBuffer<uint32_t> lead(count, "MAS coarse leader masks");
            DeviceBuffer<int> loc(count, "MAS coarse local ids");
            DeviceBuffer<int> bc(nbank, "MAS coarse counts");
            DeviceBuffer<int> pre(nbank + 1, "MAS coarse prefix");
            launch_cc_coarse(count, nbr.data(), lead.data(), loc.data(), bc.data(), stream_);
            launch_scan(nbank, bc.data(), pre.data(), stream_);
            int next = 0;
            PSMAS_HIP(hipMemcpyAsync(&next, pre.data() + nbank, sizeof(int), hipMemcpyDeviceToHost, stream_));
            PSMAS_HIP(hipStreamSynchronize(stream_));
            if (next == count)
                break; // nothing merged (MASPreconditioner.cu:464-469)
            DeviceBuffer<int> gid(count, "MAS coarse ids");
            launch_make_gid(count, loc.data(), pre.data(), gid.data(), stream_);
            launch_fill_cmap(npad_, lv, d_block_of_slot_.data(), gid.data(), d_cmap_.data(), stream_);
            d_leadc_[lv] = std::move(lead);
            d_gidc_[lv] = std::move(gid);
            cco_[lv] = next;
            level_num_ = lv + 1;
            PSMAS_HIP(hipStreamSynchronize(stream_));
        }

        mat_cnt_[0] = banks;
        mat_off_[0] = 0;
        total_mats_ = banks;
        lev_off_[0] = 0;
        total_level_nodes_ = npad_;
        for (int l = 0; l < MAX_COARSE_LEVELS; ++l)
        {
            mat_off_[l + 1] = total_mats_;
            mat_cnt_[l + 1] = (l < level_num_) ? div_up(cco_[l], BANK) : 0;
            total_mats_ += mat_cnt_[l + 1];
            lev_off_[l + 1] = total_level_nodes_;
            total_level_nodes_ += BANK * mat_cnt_[l + 1];
        }
        d_lev_off_.allocate(MAX_COARSE_LEVELS + 1, "MAS level offsets");
        d_mat_off_.allocate(MAX_COARSE_LEVELS + 1, "MAS matrix offsets");
        PSMAS_HIP(hipMemcpyAsync(d_lev_off_.data(), lev_off_, sizeof(lev_off_), hipMemcpyHostToDevice, stream_));
        PSMAS_HIP(hipMemcpyAsync(d_mat_off_.data(), mat_off_, sizeof(mat_off_), hipMemcpyHostToDevice, stream_));
        // One copy: banks and stored matrices coincide.
        stored_mats_ = total_mats_;
        npad1_ = npad_;
        nblocks1_ = nblocks_;
        cl_ = CopyLayout{};
        cl_.levels = level_num_ + 1;
        for (int l = 0; l < cl_.levels; ++l)
        {
            cl_.moff[l] = cl_.soff[l] = mat_off_[l];
            cl_.cnt[l] = mat_cnt_[l];
        }
        cl_.moff[cl_.levels] = cl_.soff[cl_.levels] = total_mats_;
        build_restrict_chunks();
        build_bank_inverses();
    }

    void Solver::build_bank_inverses()
    {
        const SellView<double> A = sell(d_vals64_.data());
        const int dim = BANK * bd_;
        const size_t L = static_cast<size_t>(dim) * (dim + 1) / 2;
        if (d_mats64_.size() != static_cast<size_t>(total_mats_) * L)
            d_mats64_.allocate(static_cast<size_t>(total_mats_) * L, "MAS coarse_matrices");
        d_mats64_.zero(stream_);
        launch_galerkin(npad_, bd_, level_num_, d_mat_off_.data(), A, d_row_len_.data(), d_block_of_slot_.data(),
                        d_cmap_.data(), d_mats64_.data(), stream_);
        DeviceBuffer<int> fail(1, "MAS invert flag");
        fail.zero(stream_);
        launch_invert(bd_, total_mats_, d_mats64_.data(), fail.data(), stream_);
        int h_fail = 0;
        PSMAS_HIP(hipMemcpyAsync(&h_fail, fail.data(), sizeof(int), hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        if (h_fail)
            throw std::runtime_error("[MAS] MAS packed inverse failed.");
    }

    void Solver::build_restrict_chunks()
    {
        nrsmall_ = nrbig_ = nrcomb_ = 0;
        d_rtask_.reset();
        d_rlist_.reset();
        d_rcomb_.reset();
        d_rpart_.reset();
        d_split_of_.reset();
        if (level_num_ < 2)
            return;
        // Ancestors of every segment-1 node in segments 2.. (the same for all its fine slots).
        std::vector<int4> cm(npad_);
        PSMAS_HIP(hipMemcpyAsync(cm.data(), d_cmap_.data(), sizeof(int4) * npad_, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        const int n1 = lev_off_[2] - lev_off_[1]; // segment-1 slots (every copy's, bank-padded)
        std::vector<int4> anc(n1, int4{-1, -1, -1, -1});
        for (int p = 0; p < npad_; ++p)
            if (cm[p].x >= 0)
                anc[cm[p].x] = cm[p];
        const int node0 = lev_off_[2];
        const int nrel = total_level_nodes_ - node0;
        auto ancestor = [&](int q, int seg) { // seg >= 2
            const int4 &a = anc[q];
            return seg == 2 ? a.y : (seg == 3 ? a.z : a.w);
        };
        std::vector<int> eoff(static_cast<size_t>(nrel) + 1, 0);
        for (int seg = 2; seg <= level_num_; ++seg)
            for (int q = 0; q < n1; ++q)
            {
                const int a = ancestor(q, seg);
                if (a >= 0)
                    ++eoff[static_cast<size_t>(lev_off_[seg] - node0 + a) + 1];
            }
        for (int i = 0; i < nrel; ++i)
            eoff[i + 1] += eoff[i];
        std::vector<int> list(eoff[nrel]);
        std::vector<int> fill(eoff.begin(), eoff.end() - 1);
        for (int seg = 2; seg <= level_num_; ++seg)
            for (int q = 0; q < n1; ++q) // ascending: each node's list is sorted
            {
                const int a = ancestor(q, seg);
                if (a >= 0)
                    list[fill[lev_off_[seg] - node0 + a]++] = q;
            }
        // Every node of segments 2.. gets a task (padding nodes: an empty list, i.e. zero); long
        // lists are cut into pieces whose sums a combine step adds.
        std::vector<int4> small, big, comb;
        int nparts = 0;
        for (int i = 0; i < nrel; ++i)
        {
            const int b = eoff[i], e = eoff[i + 1];
            if (e - b <= RESTRICT_SMALL_MAX)
                small.push_back(int4{node0 + i, b, e, -1});
            else if (e - b <= params_.restrict_piece)
                big.push_back(int4{node0 + i, b, e, -1});
            else
            {
                const int first = nparts;
                for (int k = b; k < e; k += params_.restrict_piece)
                    big.push_back(int4{node0 + i, k, std::min(k + params_.restrict_piece, e), nparts++});
                comb.push_back(int4{node0 + i, first, nparts - first, 0});
            }
        }
        nrsmall_ = static_cast<int>(small.size());
        nrbig_ = static_cast<int>(big.size());
        nrcomb_ = static_cast<int>(comb.size());
        if (nrcomb_ > 0)
        {
            d_rcomb_.allocate(comb.size(), "MAS restriction combines");
            PSMAS_HIP(hipMemcpy(d_rcomb_.data(), comb.data(), d_rcomb_.bytes(), hipMemcpyHostToDevice));
            d_rpart_.allocate(static_cast<size_t>(nparts) * bd_, "MAS restriction piece sums");
            std::vector<int> split_of(nrel, -1);
            for (int i = 0; i < nrcomb_; ++i)
                split_of[comb[i].x - node0] = i;
            d_split_of_.allocate(split_of.size(), "MAS restriction split nodes");
            PSMAS_HIP(hipMemcpy(d_split_of_.data(), split_of.data(), d_split_of_.bytes(), hipMemcpyHostToDevice));
        }
        small.insert(small.end(), big.begin(), big.end());
        d_rtask_.allocate(small.size(), "MAS restriction tasks");
        PSMAS_HIP(hipMemcpy(d_rtask_.data(), small.data(), d_rtask_.bytes(), hipMemcpyHostToDevice));
        if (!list.empty())
        {
            d_rlist_.allocate(list.size(), "MAS restriction lists");
            PSMAS_HIP(hipMemcpy(d_rlist_.data(), list.data(), d_rlist_.bytes(), hipMemcpyHostToDevice));
        }
    }

    ApplySide Solver::restrict_side(const void *seg1, void *out, int block, bool lattice_gather) const
    {
        ApplySide sd;
        sd.dummy = d_blk_begin_.data();
        sd.apply_cap = params_.apply_workgroups;
        if (nrsmall_ + nrbig_ > 0)
        {
            sd.nsmall = nrsmall_;
            sd.nbig = nrbig_;
            sd.rtask = d_rtask_.data();
            sd.rlist = d_rlist_.data();
            sd.seg1 = seg1;
            sd.out = out;
            sd.rpart = const_cast<double *>(d_rpart_.data());
        }
        if (lattice_gather)
            lcc_->gather_into(sd);
        if (sd.nsmall + sd.nbig + sd.lat_npts == 0)
            return sd;
        // A thread per small node, a wave per piece of a large node and per lattice point; at
        // most 256 workgroups.
        const int want = std::max(div_up(sd.nsmall, block), div_up(sd.nbig + sd.lat_npts, block / 64));
        sd.wgs = std::min(256, std::max(1, want));
        return sd;
    }

    ApplySide Solver::coarse_side(int block) const
    {
        ApplySide sd;
        sd.dummy = d_blk_begin_.data();
        sd.apply_cap = params_.apply_workgroups;
        if (nrcomb_ > 0)
        {
            sd.split_of = d_split_of_.data();
            sd.split_count = static_cast<int>(d_split_of_.size());
            sd.rcomb = d_rcomb_.data();
            sd.rpart = const_cast<double *>(d_rpart_.data());
            sd.node0 = lev_off_[2];
        }
        if (params_.coarse_solver == CoarseSolver::Cholesky && dense_nd_ > 0)
        {
            sd.dense_nd = dense_nd_;
            sd.dense_inv = d_dense_inv_.data();
            sd.dense_ids = d_dense_ids_.data();
            sd.dense_seg = lev_off_[level_num_];
            sd.dense_index = d_dense_index_.data();
            sd.dense_count = cco_[level_num_ - 1];
            sd.dense_copies = copies_;
            sd.dense_stride = BANK * cl_.cnt[level_num_];
            sd.wgs = std::min(256, div_up(dense_nd_, block / 64)); // wave per row
        }
        return sd;
    }

    void Solver::allocate_vectors()
    {
        const size_t nvec = static_cast<size_t>(npad_) * bd_;
        const size_t nml = static_cast<size_t>(total_level_nodes_) * bd_;
        int nmax = 0;
        for (const auto &S : sys_)
            nmax = std::max(nmax, S.n);
        d_b_.allocate(nvec, "device_buffers b");
        d_x_.allocate(nvec, "device_buffers x");
        d_p_.allocate(nvec, "device_buffers p");
        d_pb_.allocate(nvec, "device_buffers p (second)");
        d_Ap_.allocate(nvec, "device_buffers Ap");
        d_mr_.allocate(nml, "MAS coarse_vectors r");
        d_mz_.allocate(nml, "MAS coarse_vectors z");
        d_io_.allocate(static_cast<size_t>(nmax), "device_buffers io");
        d_mr_.zero(stream_);
        d_mz_.zero(stream_);
        d_x32_.reset();
        d_mr32_.reset();
        d_mz32_.reset();
        d_p32_.reset();
        d_p32b_.reset();
        d_Ap32_.reset();
        d_u32_.reset();
        d_s32_.reset();
        const int K = static_cast<int>(sys_.size());
        d_sc_.allocate(static_cast<size_t>(S_COUNT) * K, "MAS device scalars");
        d_sc_.zero(stream_);
        partial_cap_ = nblocks_;
        d_partials_.allocate(static_cast<size_t>(R_COUNT) * partial_cap_, "MAS reduction partials");
        d_partials_.zero(stream_);
    }

    void Solver::ensure_mixed_data()
    {
        const int bs = bd_ * bd_;
        if (!d_vals32_)
        {
            d_vals32_.allocate(static_cast<size_t>(nslots_) * WAVE * bs, "MAS SELL values (FP32)");
            launch_copy_convert<double, float>(static_cast<long long>(d_vals32_.size()), d_vals64_.data(),
                                               d_vals32_.data(), stream_);
        }
        prepare_sym<float>();
        const int dim = BANK * bd_;
        if (!d_mats16_)
        {
            // Every level's inverses in scaled FP16 (S = D^-1 Inv D^-1, unit diagonal): one apply
            // launch covers all matrices.
            d_mats16_.allocate(static_cast<size_t>(stored_mats_) * (dim * dim / 2), "MAS bank inverses (FP16)");
            d_dscale_.allocate(static_cast<size_t>(stored_mats_) * dim, "MAS bank scaling");
            launch_make_half(bd_, stored_mats_, d_mats64_.data(), d_mats16_.data(), d_dscale_.data(), stream_);
        }
        const size_t nvec = static_cast<size_t>(npad_) * bd_;
        const size_t nml = static_cast<size_t>(total_level_nodes_) * bd_;
        if (!d_x32_)
        {
            d_x32_.allocate(nvec, "device_buffers x32");
            d_p32_.allocate(nvec, "device_buffers p32");
            d_p32b_.allocate(nvec, "device_buffers p32 (second)");
            d_Ap32_.allocate(nvec, "device_buffers Ap32");
            d_mr32_.allocate(nml, "MAS coarse_vectors r32");
            d_mz32_.allocate(nml, "MAS coarse_vectors z32");
            d_mr32_.zero(stream_);
            d_mz32_.zero(stream_);
        }
        if (!d_u32_)
        {
            d_u32_.allocate(nvec, "device_buffers u32");
            d_s32_.allocate(nvec, "device_buffers s32");
        }
    }

    void Solver::ensure_coarse_solver()
    {
        if (params_.coarse_solver != CoarseSolver::Cholesky || dense_built_)
            return;
        build_dense_coarse();
    }

    void Solver::build_dense_coarse()
    {
        if (copies_ > 1)
            throw std::logic_error("[MAS] internal error: the coarsest inverse is built before the copies' maps.");
        dense_built_ = true;
        dense_nd_ = 0;
        const int level = level_num_ - 1; // coarsest coarse level (segment level_num_)
        const int count = cco_[level];
        std::vector<int4> cm(npad_);
        PSMAS_HIP(hipMemcpy(cm.data(), d_cmap_.data(), sizeof(int4) * npad_, hipMemcpyDeviceToHost));
        std::vector<int> active_index(count, -1);
        for (int p = 0; p < npad_; ++p)
        {
            if (block_of_slot_[p] < 0 || !nonisolated_[p])
                continue;
            const int4 c = cm[p];
            const int id = level == 0 ? c.x : (level == 1 ? c.y : (level == 2 ? c.z : c.w));
            active_index[id] = 1;
        }
        std::vector<int> ids;
        for (int c = 0; c < count; ++c)
            if (active_index[c] > 0)
            {
                active_index[c] = static_cast<int>(ids.size());
                ids.push_back(c);
            }
        dense_nodes_ = static_cast<int>(ids.size());
        if (ids.empty() || dense_nodes_ > DENSE_COARSE_MAX_NODES)
        {
            dense_note_ = ids.empty() ? "no active coarsest nodes" : "too many active coarsest nodes";
            return;
        }
        const int nd = bd_ * dense_nodes_;
        DeviceBuffer<int> d_index(count, "MAS coarsest active index");
        PSMAS_HIP(hipMemcpy(d_index.data(), active_index.data(), sizeof(int) * count, hipMemcpyHostToDevice));
        d_dense_index_.allocate(count, "MAS coarsest dense index");
        PSMAS_HIP(hipMemcpy(d_dense_index_.data(), active_index.data(), sizeof(int) * count, hipMemcpyHostToDevice));
        DeviceBuffer<double> dense(static_cast<size_t>(nd) * nd, "MAS coarsest Galerkin matrix");
        dense.zero(stream_);
        launch_dense_galerkin(bd_, npad_, level, sell(d_vals64_.data()), d_row_len_.data(), d_block_of_slot_.data(),
                              d_cmap_.data(), d_index.data(), nd, dense.data(), stream_);
        std::vector<double> h(static_cast<size_t>(nd) * nd);
        PSMAS_HIP(hipMemcpyAsync(h.data(), dense.data(), dense.bytes(), hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        for (int i = 0; i < nd; ++i) // atomics leave round-off asymmetry: use the symmetric part
            for (int j = i + 1; j < nd; ++j)
            {
                const double v = 0.5 * (h[static_cast<size_t>(i) * nd + j] + h[static_cast<size_t>(j) * nd + i]);
                h[static_cast<size_t>(i) * nd + j] = v;
                h[static_cast<size_t>(j) * nd + i] = v;
            }
        if (!cholesky_inverse(nd, h))
            throw std::runtime_error("[MAS] coarsest-level Cholesky failed (coarse matrix not positive definite).");
        d_dense_inv_.allocate(h.size(), "MAS coarsest inverse");
        d_dense_ids_.allocate(ids.size(), "MAS coarsest active ids");
        PSMAS_HIP(hipMemcpy(d_dense_inv_.data(), h.data(), d_dense_inv_.bytes(), hipMemcpyHostToDevice));
        PSMAS_HIP(hipMemcpy(d_dense_ids_.data(), ids.data(), d_dense_ids_.bytes(), hipMemcpyHostToDevice));
        dense_nd_ = nd;
        dense_note_ = "dense FP64 Cholesky on " + std::to_string(dense_nodes_) + " coarsest nodes";
    }

    void Solver::check_ready_single(int n) const
    {
        if (!factorized_ || sys_.size() != 1 || n <= 0 || n != sys_[0].n)
            throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
    }

    // ------------------------------------------------------------------------------- solve
    void Solver::solve_host(int n, const double *b, double *x)
    {
        sys_[0].status = Status::Running;
        check_ready_single(n);
        solve_batch({b}, {x}, false);
    }

    void Solver::solve_device(int n, const double *d_b, double *d_x)
    {
        sys_[0].status = Status::Running;
        check_ready_single(n);
        solve_batch({d_b}, {d_x}, true);
    }

    void Solver::solve_device_noalloc(int n, const double *d_b, double *d_x)
    {
        PSMAS_HIP(hipSetDevice(device_));
        sys_[0].status = Status::Running;
        check_ready_single(n);
        if (scratch_ == nullptr)
            throw std::runtime_error("[MAS] solve_device_noalloc: " + scratch_error_);
        if (d_b == nullptr || d_x == nullptr)
            throw InvalidArgumentError("[MAS] null b or x.");
        if (nblocks_ > noalloc_max_blocks())
            throw InvalidArgumentError("[MAS] solve_device_noalloc: " + std::to_string(nblocks_) +
                                       " node blocks of 256 padded nodes exceed the static scratch (1 MiB: at most " +
                                       std::to_string(noalloc_max_blocks()) + " blocks); use psmas_solve_device.");
        // Everything the solve would otherwise build or allocate on the way must exist already.
        std::string missing;
        if (params_.coarse_solver == CoarseSolver::Cholesky && !dense_built_)
            missing = "the dense coarsest-level solve";
        if (params_.precision == Precision::Mixed && (!d_vals32_ || !d_mats16_ || !d_x32_ || !d_u32_))
            missing = "the FP32/FP16 data of the mixed path";
        if (params_.precision == Precision::Mixed && params_.learned_correction && lcc_ && lcc_->trained() &&
            (lattice_stale_ || !lcc_->lattice_ready() || lcc_->k() != params_.lattice_points ||
             d_zprev32_.size() < static_cast<size_t>(npad_) * bd_))
            missing = "the learned correction's lattice and flexible-PCG buffers";
        if (!missing.empty())
            throw std::runtime_error("[MAS] solve_device_noalloc: " + missing +
                                     " is built at the first solve of a factorisation; call psmas_solve_device once "
                                     "after factorize and these parameters.");
        if (params_.learned_collect)
            throw std::runtime_error("[MAS] solve_device_noalloc: learned_collect allocates snapshots; not available.");
        if (timing_on_)
            throw std::runtime_error("[MAS] solve_device_noalloc: kernel timing is on; not available.");
        use_scratch_ = true;
        try
        {
            // The scratch keeps the previous call's contents: clear the part this system uses.
            const long long used = static_cast<long long>(SCRATCH_SC_PAD) + static_cast<long long>(R_COUNT + 1) * partial_cap_;
            launch_fill<double>(used, scratch_, 0.0, stream_);
            solve_batch({d_b}, {d_x}, true);
        }
        catch (...)
        {
            use_scratch_ = false;
            throw;
        }
        use_scratch_ = false;
    }

    void Solver::solve_batch(const std::vector<const double *> &b, const std::vector<double *> &x, bool device)
    {
        PSMAS_HIP(hipSetDevice(device_));
        for (auto &S : sys_)
            S.status = Status::Running;
        if (!factorized_ || b.size() != sys_.size() || x.size() != sys_.size())
            throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
        for (size_t s = 0; s < sys_.size(); ++s)
        {
            const System &S = sys_[s];
            if (b[s] == nullptr || x[s] == nullptr)
                throw InvalidArgumentError("[MAS] null b or x.");
            const double *src = b[s];
            if (!device)
            {
                PSMAS_HIP(hipMemcpyAsync(d_io_.data(), b[s], sizeof(double) * S.n, hipMemcpyHostToDevice, stream_));
                src = d_io_.data();
            }
            launch_gather_in<double>(S.npad, bd_, S.n, d_block_of_slot_.data() + S.node_off, src,
                                     d_b_.data() + static_cast<size_t>(S.node_off) * bd_, stream_);
        }
        rhs_ready_ = true;
        solve_padded();
        for (size_t s = 0; s < sys_.size(); ++s)
        {
            const System &S = sys_[s];
            double *dst = device ? x[s] : d_io_.data();
            launch_scatter_out<double>(S.nb, bd_, S.n, d_slot_of_block_.data() + S.sob_off,
                                       d_x_.data() + static_cast<size_t>(S.node_off) * bd_, dst, stream_);
            if (!device)
                PSMAS_HIP(hipMemcpyAsync(x[s], d_io_.data(), sizeof(double) * S.n, hipMemcpyDeviceToHost, stream_));
        }
        PSMAS_HIP(hipStreamSynchronize(stream_));
    }

    void Solver::solve_padded()
    {
        const auto t0 = std::chrono::steady_clock::now();
        const int K = static_cast<int>(sys_.size());
        for (auto &S : sys_)
        {
            S.iterations = 0;
            S.residual_norm = 0.0;
            S.reliable_updates = 0;
            S.rel_residual = 0.0;
        }
        h_active_.assign(K, 1);
        h_select_.assign(K, 1);
        upload_flags();
        launch_set_scalar(sc_ptr(), K, S_GUARD, params_.division_guard, stream_);
        ensure_coarse_solver();
        if (params_.precision == Precision::Mixed)
        {
            ensure_mixed_data();
            solve_mixed();
        }
        else
        {
            prepare_sym<double>(); // (built with the factorisation unless the parameters changed since)
            solve_fp64();
        }
        PSMAS_HIP(hipStreamSynchronize(stream_));
        collect_timings();
        solve_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void Solver::run_block(long long key, const std::function<void()> &enqueue)
    {
        // Captured graphs hold the allocated partial buffers; the scratch solve launches directly.
        if (!params_.use_graphs || timing_on_ || use_scratch_)
        {
            enqueue();
            return;
        }
        auto it = graphs_->execs.find(key);
        if (it == graphs_->execs.end())
        {
            hipGraph_t g = nullptr;
            PSMAS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try
            {
                enqueue();
            }
            catch (...)
            {
                hipGraph_t dead = nullptr;
                (void)hipStreamEndCapture(stream_, &dead);
                if (dead)
                    (void)hipGraphDestroy(dead);
                throw;
            }
            PSMAS_HIP(hipStreamEndCapture(stream_, &g));
            hipGraphExec_t ge = nullptr;
            const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess)
                throw_hip_error(e, "hipGraphInstantiate", __FILE__, __LINE__);
            it = graphs_->execs.emplace(key, ge).first;
        }
        PSMAS_HIP(hipGraphLaunch(it->second, stream_));
    }

    template <int BD, typename T>
    void Solver::spmv_ap_all(const SellView<T> &A, const T *x, T *y)
    {
        if (copies_ > 1 && params_.spmv_copies)
            launch_spmv_ap_copies<BD, T, T>(A, copies_, npad1_, x, y, reductions(), sysmap(), stream_);
        else if (copies_ > 1)
            for (int c = 0; c < copies_; ++c)
            {
                const size_t o = static_cast<size_t>(c) * npad1_ * BD;
                launch_spmv_ap<BD, T, T>(A, x + o, y + o, reductions_of_copy(c), sysmap_of_copy(c), stream_);
            }
        else
        {
            SymView<T> S;
            if (sym_view(S))
                launch_spmv_sym<BD, T, T>(A, S, x, y, reductions(), sysmap(), stream_);
            else
                launch_spmv_ap<BD, T, T>(A, x, y, reductions(), sysmap(), stream_);
        }
    }

    template <int BD, typename TX, typename TY>
    void Solver::spmv_res_all(const SellView<double> &A, const TX *x, const TX *b, TY *y, double *rr, bool selected)
    {
        TY *const level1 = y + static_cast<size_t>(npad_) * BD; // every copy's (copy-shifted ids)
        SymView<double> S;
        if (copies_ == 1 && sym_view(S))
        {
            launch_spmv_res_sym<BD, double, TX, TY>(A, S, x, b, y, fine_restrict(), level1, rr, sysmap(), selected,
                                                    stream_);
            return;
        }
        for (int c = 0; c < copies_; ++c)
        {
            const size_t o = static_cast<size_t>(c) * npad1_;
            FineRestrict fr = fine_restrict();
            fr.lead += o;
            fr.gid += o;
            launch_spmv_res<BD, double, TX, TY>(A, x + o * BD, b + o * BD, y + o * BD, fr, level1,
                                                rr + static_cast<size_t>(c) * nblocks1_, sysmap_of_copy(c), selected,
                                                stream_);
        }
    }

    // ---- FP64 path: the reference schedule -------------------------------------------------
    template <int BD>
    void Solver::fp64_precondition(bool init)
    {
        // Launch 1: fine and segment-1 banks, with the restriction of segments 2.. from segment 1
        // in its leading workgroups; launch 2: banks of segments 2.. (inputs from the restriction
        // sums) and the dense coarsest solve.
        constexpr int block = 64 * BD;
        const int n1 = first_apply_mats();
        const double per_mat = (BANK * BD) * (BANK * BD + 1) / 2 * 8.0 + 2.0 * BANK * BD * 8.0;
        tic("mas_apply_fp64", static_cast<double>(n1) * per_mat);
        launch_mas_apply<BD, double, double>(0, n1, d_mats64_.data(), nullptr, d_mr_.data(), d_mz_.data(), nullptr,
                                             nullptr, stream_,
                                             restrict_side(d_mr_.data() + static_cast<size_t>(lev_off_[1]) * BD,
                                                           d_mr_.data(), block),
                                             cl_);
        toc();
        const ApplySide cs = coarse_side(block);
        if (total_mats_ > n1 || cs.wgs > 0)
        {
            tic("mas_apply_fp64_coarse", static_cast<double>(total_mats_ - n1) * per_mat +
                                             static_cast<double>(dense_nd_) * dense_nd_ * 8.0);
            launch_mas_apply<BD, double, double>(n1, total_mats_ - n1, d_mats64_.data(), nullptr, d_mr_.data(),
                                                 d_mz_.data(), nullptr, nullptr, stream_, cs, cl_);
            toc();
        }
        launch_prolong_dot<BD, double>(nblocks_, level_num_, d_cmap_.data(), d_lev_off_.data(), d_mr_.data(),
                                       d_mz_.data(), reductions(), init, sysmap(), LccApply{}, FlexDot<double>{},
                                       stream_);
    }

    bool Solver::fuse_direction() const
    {
        return params_.spmv_fuse_direction && params_.spmv_halo && halo_max_ > 0 && copies_ == 1;
    }

    template <int BD>
    void Solver::fp64_iteration(int k, bool init_direction)
    {
        const SellView<double> A = sell(d_vals64_.data());
        const bool fuse = fuse_direction();
        if (fuse)
        {
            // p = z + beta p (or z) formed by the SpMV; the new p lands in the other buffer.
            FusedDirection<double> fd;
            fd.z = d_mz_.data();
            fd.p_old = p64(p_cur_);
            fd.p_new = p64(p_cur_ ^ 1);
            fd.init = init_direction ? 1 : 0;
            tic("spmv_dir_fp64", spmv_bytes(8, 8, 4));
            launch_spmv_ap_fused<BD, double, double>(A, fd, d_Ap_.data(), reductions(), sysmap(), stream_);
            toc();
            p_cur_ ^= 1;
        }
        else
        {
            tic("spmv_fp64", spmv_ap_bytes<double>());
            spmv_ap_all<BD, double>(A, d_p_.data(), d_Ap_.data());
            toc();
        }
        double *p = fuse ? p64(p_cur_) : d_p_.data();
        if (k % params_.true_residual_period == 0)
        {
            launch_axpy_alpha<BD, double>(nblocks_, p, d_x_.data(), reductions(), sysmap(), stream_);
            tic("spmv_res_fp64", spmv_ap_bytes<double>() + n_total_ * 8.0);
            spmv_res_all<BD, double, double>(A, d_x_.data(), d_b_.data(), d_mr_.data(), reductions().rr, false);
            toc();
        }
        else
        {
            launch_update_xr<BD, double>(nblocks_, p, d_Ap_.data(), d_x_.data(), d_mr_.data(), fine_restrict(),
                                         d_mr_.data() + static_cast<size_t>(npad_) * BD, reductions(), sysmap(),
                                         stream_);
        }
        fp64_precondition<BD>(false);
        if (!fuse)
            launch_update_p<BD, double>(nblocks_, d_mz_.data(), d_p_.data(), reductions(), sysmap(), stream_);
    }

    void Solver::solve_fp64()
    {
        auto body = [&](auto bdc) {
            constexpr int BD = decltype(bdc)::value;
            const int K = static_cast<int>(sys_.size());
            const SellView<double> A = sell(d_vals64_.data());
            const long long nvec = static_cast<long long>(npad_) * BD;
            launch_fill<double>(nvec, d_x_.data(), 0.0, stream_);
            // r = b - A x (x = 0), rr, restriction.
            spmv_res_all<BD, double, double>(A, d_x_.data(), d_b_.data(), d_mr_.data(), reductions().rr, false);
            fp64_precondition<BD>(true);
            const bool fuse = fuse_direction();
            p_cur_ = 0;
            if (!fuse) // otherwise the first SpMV starts the direction from z
                launch_copy_convert<double, double>(nvec, d_mz_.data(), d_p_.data(), stream_);
            finalize_scalars();
            read_scalars();
            std::vector<double> rz0(K), rr0(K), bb(K);
            for (int s = 0; s < K; ++s)
            {
                rz0[s] = *scal(s, S_RZ);
                if (std::isnan(rz0[s]) || !std::isfinite(rz0[s]))
                    throw std::runtime_error(K == 1 ? std::string("[MAS] Invalid initial residual.")
                                                    : "[MAS] Invalid initial residual (system " +
                                                          std::to_string(s) + ").");
                bb[s] = *scal(s, S_RR);
                rr0[s] = params_.use_preconditioned_residual_norm ? 0.0 : bb[s];
            }
            const double rel2 = params_.rel_tol * params_.rel_tol;
            const double abs2 = params_.abs_tol * params_.abs_tol;
            const int ci = params_.check_interval;
            const int trp = params_.true_residual_period;
            int k = 1;
            int running = K;
            while (k <= params_.max_iter && running > 0)
            {
                const int next_check = ((k - 1) / ci + 1) * ci;
                const int last = std::min(next_check, params_.max_iter);
                const int count = last - k + 1;
                const int k0 = k;
                // The lambda runs only when a graph is captured: the direction buffer parity is
                // set from the segment's start either way.
                const int par0 = p_cur_;
                run_block(graph_key(0, k % trp, count, (fuse ? 1 : 0) | (par0 << 1) | (k0 == 1 ? 4 : 0)), [&]() {
                    p_cur_ = par0;
                    for (int i = k0; i <= last; ++i)
                        fp64_iteration<BD>(i, i == 1);
                    finalize_scalars();
                });
                p_cur_ = fuse ? par0 ^ (count & 1) : 0;
                for (int s = 0; s < K; ++s)
                    if (h_active_[s])
                        sys_[s].iterations = last;
                if (last % ci == 0)
                {
                    read_scalars();
                    bool changed = false;
                    for (int s = 0; s < K; ++s)
                    {
                        if (!h_active_[s])
                            continue;
                        System &S = sys_[s];
                        const double v = params_.use_preconditioned_residual_norm ? *scal(s, S_RZ) : *scal(s, S_RR);
                        const double v0 = params_.use_preconditioned_residual_norm ? rz0[s] : rr0[s];
                        S.residual_norm = std::sqrt(v);
                        S.rel_residual = bb[s] > 0.0 ? std::sqrt(*scal(s, S_RR) / bb[s]) : 0.0;
                        if (v <= rel2 * v0 || v <= abs2)
                        {
                            S.status = (v <= abs2) ? Status::ReachAbsoluteTolerance : Status::ReachRelativeTolerance;
                            h_active_[s] = 0;
                            --running;
                            changed = true;
                        }
                    }
                    if (changed && running > 0)
                        upload_flags();
                }
                k = last + 1;
            }
            for (int s = 0; s < K; ++s)
                if (sys_[s].iterations == params_.max_iter)
                    sys_[s].status = Status::ReachMaxIterations;
        };
        if (bd_ == 1)
            body(std::integral_constant<int, 1>());
        else if (bd_ == 2)
            body(std::integral_constant<int, 2>());
        else
            body(std::integral_constant<int, 3>());
    }

    // ---- mixed path: FP32 inner PCG with FP64 reliable updates ------------------------------
    template <int BD>
    void Solver::mixed_part_a(bool learned, bool init_direction)
    {
        const SellView<float> A = sell(d_vals32_.data());
        const bool fuse = fuse_direction();
        if (fuse)
        {
            // p = z + beta p (flexible beta with the learned correction; p = z at the start) formed
            // by the SpMV; the new p lands in the other buffer.
            FusedDirection<float> fd;
            fd.z = d_mz32_.data();
            fd.p_old = p32(p_cur_);
            fd.p_new = p32(p_cur_ ^ 1);
            fd.init = init_direction ? 1 : 0;
            if (learned)
            {
                fd.zprev = d_zprev32_.data();
                fd.rzo = init_direction ? nullptr : rzo_ptr();
            }
            tic("spmv_dir_fp32", spmv_bytes(4, 4, 4) + (learned ? n_total_ * 4.0 : 0.0));
            launch_spmv_ap_fused<BD, float, float>(A, fd, d_Ap32_.data(), reductions(), sysmap(), stream_);
            toc();
            p_cur_ ^= 1;
        }
        else
        {
            tic("spmv_fp32", spmv_ap_bytes<float>());
            spmv_ap_all<BD, float>(A, d_p32_.data(), d_Ap32_.data());
            toc();
        }
        // With the learned correction on the block path, the update also leaves the lattice block
        // partials of the new r.
        const LatticeBlocks lb = learned && lcc_ ? lcc_->block_view() : LatticeBlocks{};
        tic(lb.cell ? "update_xr_lattice" : "update_xr", static_cast<double>(n_total_) * 4.0 * 6.0);
        launch_update_xr<BD, float>(nblocks_, fuse ? p32(p_cur_) : d_p32_.data(), d_Ap32_.data(), d_x32_.data(),
                                    d_mr32_.data(), fine_restrict(), d_mr32_.data() + static_cast<size_t>(npad_) * BD,
                                    reductions(), sysmap(), stream_, &lb);
        toc();
    }

    template <int BD>
    void Solver::mixed_part_b(bool init, bool learned)
    {
        const long long nvec = static_cast<long long>(npad_) * BD;
        const LccApply lv = mixed_apply<BD>(init, learned);
        FlexDot<float> flex;
        if (learned && !init)
        {
            flex.zprev = d_zprev32_.data();
            flex.rzo = rzo_ptr();
        }
        tic(learned ? "prolong_dot_learned" : "prolong_dot", static_cast<double>(n_total_) * 4.0 * 3 + nblocks_ * 256.0 * 16);
        launch_prolong_dot<BD, float>(nblocks_, level_num_, d_cmap_.data(), d_lev_off_.data(), d_mr32_.data(),
                                      d_mz32_.data(), reductions(), init, sysmap(), lv, flex, stream_);
        toc();
        if (fuse_direction())
            return; // the next SpMV forms p (and z_prev)
        if (init)
        {
            launch_copy_convert<float, float>(nvec, d_mz32_.data(), d_p32_.data(), stream_);
            if (learned)
                launch_copy_convert<float, float>(nvec, d_mz32_.data(), d_zprev32_.data(), stream_);
        }
        else if (learned)
        {
            launch_update_p<BD, float>(nblocks_, d_mz32_.data(), d_p32_.data(), reductions(), sysmap(), stream_,
                                       rzo_ptr(), d_zprev32_.data());
        }
        else
        {
            launch_update_p<BD, float>(nblocks_, d_mz32_.data(), d_p32_.data(), reductions(), sysmap(), stream_);
        }
    }

    // ---- mixed path, single-reduction iteration ------------------------------------------
    // One iteration: update (alpha, beta from one set of partials; p, s, x, r) | apply launches
    // (and the network) | SpMV of u = M r, which assembles u from the levels where it reads it.
    template <int BD>
    void Solver::cg_part_b(bool init, bool learned)
    {
        const LccApply lv = mixed_apply<BD>(init, learned);
        CgProlong pr;
        pr.mz = d_mz32_.data();
        pr.cmap = d_cmap_.data();
        pr.lev_off = d_lev_off_.data();
        pr.level_num = level_num_;
        pr.own = d_halo_own_.data();
        pr.lcc = lv;
        const SellView<float> A = sell(d_vals32_.data());
        tic(learned ? "spmv_cg_learned" : "spmv_cg",
            spmv_bytes(4, 4, 2) + static_cast<double>(n_total_) * 4.0 * (init ? 2.0 : (learned ? 4.0 : 3.0)));
        launch_spmv_cg<BD>(A, pr, d_u32_.data(), d_Ap32_.data(), d_mr32_.data(), d_s32_.data(),
                           learned ? d_zprev32_.data() : nullptr, cg_reductions(), sysmap(), init, stream_);
        toc();
    }

    template <int BD>
    void Solver::cg_part_a(bool learned, bool init)
    {
        const LatticeBlocks lb = learned && lcc_ ? lcc_->block_view() : LatticeBlocks{};
        tic(lb.cell ? "update_cg_lattice" : "update_cg",
            static_cast<double>(n_total_) * 4.0 * (init ? 8.0 : 10.0) + (learned ? n_total_ * 4.0 : 0.0));
        launch_update_cg<BD>(nblocks_, d_u32_.data(), d_Ap32_.data(), d_p32_.data(), d_s32_.data(), d_x32_.data(),
                             d_mr32_.data(), fine_restrict(), d_mr32_.data() + static_cast<size_t>(npad_) * BD,
                             learned ? d_zprev32_.data() : nullptr, cg_reductions(), sysmap(), init, stream_, &lb);
        toc();
    }

    /// Both apply launches (and the network) on the current r: mz holds every level's z, and the
    /// returned view carries the learned correction's coefficients when it is used.
    template <int BD>
    LccApply Solver::mixed_apply(bool init, bool learned)
    {
        const int dim = BANK * BD;
        // Apply launch 1: fine and segment-1 banks, with the restriction of segments 2.. (and, for
        // the learned correction on the block path, the lattice gather of the block partials the
        // residual update left) in its leading workgroups; launch 2: banks of segments 2.. and the
        // dense coarsest solve; then the network.
        // One right-hand side: a wave per bank matrix (k_mas_apply_w16); copies of one
        // factorisation: the persistent kernel that unpacks each stored matrix once for all.
        const bool wave = params_.apply_wave && copies_ == 1;
        const int block = apply_threads(BD, wave);
        const int n1 = first_apply_stored(); // (the FP16 apply walks stored matrices)
        const double per_mat = dim * dim * 1.0 + dim * 4.0 + 2.0 * dim * 4.0 * copies_;
        const bool lat_blocks = learned && lcc_->block_restriction();
        LccApply lv;
        if (learned)
        {
            if (lat_blocks)
                lcc_->set_output_factor(static_cast<float>(params_.learned_output_scale));
            else
            {
                // Entry path: per coupled slot 8 {slot, weight} entries (8 B each) and 8 gathers of r.
                tic("lcc_restrict", static_cast<double>(lcc_->restrict_entries()) * (8.0 + BD * 4.0));
                lcc_->restrict_to_lattice<BD, float>(d_mr32_.data(), d_active_.data(), stream_,
                                                     static_cast<float>(params_.learned_output_scale));
                toc();
            }
            lv = lcc_->apply_view();
        }
        (void)init;
        tic("mas_apply_fp16", n1 * per_mat);
        // A single system is active for the whole apply: no per-matrix system lookups (two dependent
        // loads per matrix on each workgroup's critical path).  Copies of one factorisation apply
        // every stored matrix to all of them.
        const bool batch = sys_.size() > 1 && copies_ == 1;
        const int *bsys = batch ? d_bank_sys_.data() : nullptr;
        const int *bact = batch ? d_active_.data() : nullptr;
        ApplySide rs = restrict_side(d_mr32_.data() + static_cast<size_t>(lev_off_[1]) * BD, d_mr32_.data(), block,
                                     lat_blocks);
        rs.wave_apply = wave;
        launch_mas_apply<BD, __half, float>(0, n1, d_mats16_.data(), d_dscale_.data(), d_mr32_.data(),
                                            d_mz32_.data(), bsys, bact, stream_, rs, cl_);
        toc();
        // The network's first layer and skip layer join the second launch's side work when they
        // fit; the output layer follows it.
        ApplySide cs = coarse_side(block);
        cs.wave_apply = wave;
        const bool lin_side =
            learned && lcc_->in_side(cs, apply_h16_stage_floats(BD, wave), block, d_active_.data());
        if (stored_mats_ > n1 || cs.wgs > 0)
        {
            tic("mas_apply_fp16_coarse", (stored_mats_ - n1) * per_mat + static_cast<double>(dense_nd_) * dense_nd_ * 8.0);
            launch_mas_apply<BD, __half, float>(n1, stored_mats_ - n1, d_mats16_.data(), d_dscale_.data(),
                                                d_mr32_.data(), d_mz32_.data(), bsys, bact, stream_, cs, cl_);
            toc();
        }
        if (learned)
        {
            tic(lin_side ? "lcc_infer_out" : "lcc_infer", static_cast<double>(lcc_->inference_weight_bytes()));
            if (lin_side)
                lcc_->infer_out(d_active_.data(), stream_);
            else
                lcc_->infer(d_active_.data(), stream_);
            toc();
        }
        return lv;
    }

    template <int BD>
    void Solver::mixed_reliable_residual(bool learned)
    {
        const SellView<double> A = sell(d_vals64_.data());
        tic("spmv_res_fp64", spmv_ap_bytes<double>() + n_total_ * 8.0);
        spmv_res_all<BD, double, float>(A, d_x_.data(), d_b_.data(), d_mr32_.data(), reductions().rr_true, true);
        toc();
        // Block path of the learned correction: the preconditioner expects the lattice block
        // partials of the current r (the residual update leaves them; here r was recomputed).
        if (learned && lcc_ && lcc_->block_restriction())
            lcc_->restrict_blocks<BD, float>(d_mr32_.data(), stream_);
    }

    void Solver::solve_mixed()
    {
        auto body = [&](auto bdc) {
            constexpr int BD = decltype(bdc)::value;
            const int K = static_cast<int>(sys_.size());
            const long long nvec = static_cast<long long>(npad_) * BD;
            // Learned coarse correction: used when requested and a trained network matches the
            // lattice; snapshots for training pairs when collecting (single systems only).
            learned_used_ = false;
            learned_fallback_ = false;
            learned_fallback_iter_ = -1;
            learned_note_.clear();
            bool learned = false;
            if (params_.learned_correction)
            {
                if (!lcc_ || !lcc_->trained())
                    learned_note_ = "requested but no trained network";
                else
                {
                    ensure_lattice();
                    if (lcc_->trained())
                        learned = true;
                    else
                        learned_note_ = "network dropped: lattice size changed";
                }
            }
            const bool collect = params_.learned_collect && K == 1 && params_.learned_snapshots > 0;
            if (params_.learned_collect && K != 1)
                learned_note_ = "pair collection needs a single system";
            nsnap_ = 0;
            if (collect)
            {
                ensure_lattice();
                const size_t need = static_cast<size_t>(params_.learned_snapshots) * nvec;
                if (d_snap_.size() < need)
                    d_snap_.allocate(need, "MAS learned snapshots");
            }
            if (learned)
            {
                if (d_zprev32_.size() < static_cast<size_t>(nvec))
                    d_zprev32_.allocate(nvec, "MAS flexible z_prev");
                if (!use_scratch_ && d_rzo_.size() < static_cast<size_t>(nblocks_))
                    d_rzo_.allocate(nblocks_, "MAS flexible partials");
                learned_used_ = true;
            }
            launch_fill<double>(nvec, d_x_.data(), 0.0, stream_);
            launch_fill<float>(nvec, d_x32_.data(), 0.0f, stream_);
            mixed_reliable_residual<BD>(learned); // every system selected: r = b
            p_cur_ = 0;
            const bool cg = single_reduction();
            auto part_a = [&](bool lrn, bool init) {
                if (cg)
                    cg_part_a<BD>(lrn, init);
                else
                    mixed_part_a<BD>(lrn, init);
            };
            auto part_b = [&](bool init, bool lrn) {
                if (cg)
                    cg_part_b<BD>(init, lrn);
                else
                    mixed_part_b<BD>(init, lrn);
            };
            part_b(true, learned);
            finalize_scalars();
            read_scalars();
            std::vector<double> bnorm(K), target(K), ref(K), best(K), seg_best(K);
            std::vector<int> stalls(K, 0), seg_stall(K, 0);
            int running = K;
            for (int s = 0; s < K; ++s)
            {
                System &S = sys_[s];
                const double rz0 = *scal(s, S_RZ);
                if (std::isnan(rz0) || !std::isfinite(rz0))
                    throw std::runtime_error(K == 1 ? std::string("[MAS] Invalid initial residual.")
                                                    : "[MAS] Invalid initial residual (system " +
                                                          std::to_string(s) + ").");
                bnorm[s] = std::sqrt(*scal(s, S_RR_TRUE));
                target[s] = std::max(params_.rel_tol * bnorm[s], params_.abs_tol);
                ref[s] = best[s] = seg_best[s] = bnorm[s];
                S.residual_norm = bnorm[s];
                S.rel_residual = bnorm[s] > 0.0 ? 1.0 : 0.0;
                if (bnorm[s] <= params_.abs_tol)
                {
                    S.status = Status::ReachAbsoluteTolerance;
                    h_active_[s] = 0;
                    --running;
                }
            }
            std::fill(h_select_.begin(), h_select_.end(), 0);
            upload_flags();
            bool first = true, restart_plain = false;
            int k = 0;
            const int ci = params_.check_interval;
            while (k < params_.max_iter && running > 0)
            {
                if (restart_plain)
                {
                    // Restart from the current residual with plain MAS: z = M r, p = z.
                    part_b(true, false);
                    first = true;
                    restart_plain = false;
                }
                const int count = std::min(ci, params_.max_iter - k);
                const bool lead_b = !first;
                // The lambda runs only when a graph is captured: the direction buffer parity is
                // set from the segment's start either way.
                const bool fuse = fuse_direction();
                const int par0 = p_cur_;
                run_block(graph_key(learned ? 2 : 1, first ? 1 : 2, count, (fuse ? 1 : 0) | (par0 << 1) | (cg ? 8 : 0)),
                          [&]() {
                              p_cur_ = par0;
                              if (lead_b)
                                  part_b(false, learned);
                              for (int i = 0; i < count; ++i)
                              {
                                  part_a(learned, first && i == 0);
                                  if (i + 1 < count)
                                      part_b(false, learned);
                              }
                              finalize_scalars();
                          });
                p_cur_ = fuse ? par0 ^ (count & 1) : 0;
                first = false;
                k += count;
                read_scalars();
                if (collect && nsnap_ < params_.learned_snapshots)
                {
                    launch_snapshot(nvec, d_x_.data(), d_x32_.data(), d_snap_.data() + static_cast<size_t>(nsnap_) * nvec,
                                    stream_);
                    ++nsnap_;
                }
                // Plain MAS judges the learned correction: when some system's recurrence residual
                // has not reached a new minimum for learned_stall_segments check intervals, the
                // correction is switched off and PCG restarts from the current residual.
                bool stalled = false;
                for (int s = 0; s < K; ++s)
                {
                    if (!h_active_[s])
                        continue;
                    const double rn = std::sqrt(std::max(*scal(s, S_RR), 0.0));
                    if (rn < seg_best[s])
                    {
                        seg_best[s] = rn;
                        seg_stall[s] = 0;
                    }
                    else if (++seg_stall[s] >= params_.learned_stall_segments)
                        stalled = true;
                }
                if (learned && stalled)
                {
                    learned = false;
                    learned_fallback_ = true;
                    learned_fallback_iter_ = k;
                    std::fill(seg_stall.begin(), seg_stall.end(), 0);
                    restart_plain = true; // after any reliable update below has replaced r
                }
                int nsel = 0;
                for (int s = 0; s < K; ++s)
                {
                    h_select_[s] = 0;
                    if (!h_active_[s])
                        continue;
                    sys_[s].iterations = k;
                    const double rn = std::sqrt(std::max(*scal(s, S_RR), 0.0));
                    if (rn <= params_.reliable_update_delta * ref[s] || rn <= target[s] || k >= params_.max_iter ||
                        !std::isfinite(rn))
                    {
                        h_select_[s] = 1;
                        ++nsel;
                    }
                }
                if (nsel == 0)
                    continue;
                // Reliable update: fold the FP32 correction into x and recompute r in FP64.
                upload_flags();
                launch_accumulate_correction<BD>(nblocks_, d_x32_.data(), d_x_.data(), sysmap(), stream_);
                mixed_reliable_residual<BD>(learned);
                finalize_scalars();
                read_scalars();
                for (int s = 0; s < K; ++s)
                {
                    if (!h_select_[s])
                        continue;
                    System &S = sys_[s];
                    ++S.reliable_updates;
                    const double rt = std::sqrt(std::max(*scal(s, S_RR_TRUE), 0.0));
                    S.residual_norm = rt;
                    S.rel_residual = bnorm[s] > 0.0 ? rt / bnorm[s] : 0.0;
                    bool stop = false;
                    if (rt <= target[s])
                    {
                        S.status = (rt <= params_.abs_tol) ? Status::ReachAbsoluteTolerance
                                                           : Status::ReachRelativeTolerance;
                        stop = true;
                    }
                    else
                    {
                        stalls[s] = (rt < 0.9 * best[s]) ? 0 : stalls[s] + 1;
                        if (stalls[s] >= 8 || !std::isfinite(rt))
                        {
                            S.status = Status::ReachMaxIterations;
                            stop = true;
                        }
                        best[s] = std::min(best[s], rt);
                        ref[s] = rt;
                    }
                    if (stop)
                    {
                        h_active_[s] = 0;
                        --running;
                    }
                    h_select_[s] = 0;
                }
                upload_flags();
            }
            for (int s = 0; s < K; ++s)
                if (sys_[s].status == Status::Running)
                    sys_[s].status = Status::ReachMaxIterations;
            if (collect && (sys_[0].status == Status::ReachRelativeTolerance ||
                            sys_[0].status == Status::ReachAbsoluteTolerance))
                learned_collect_pairs<BD>();
        };
        if (bd_ == 1)
            body(std::integral_constant<int, 1>());
        else if (bd_ == 2)
            body(std::integral_constant<int, 2>());
        else
            body(std::integral_constant<int, 3>());
    }

    // ------------------------------------------------------------ learned coarse correction
    void Solver::ensure_lattice()
    {
        if (!lcc_)
            lcc_ = std::make_unique<LearnedCoarse>();
        if (!lattice_stale_ && lcc_->lattice_ready() && lcc_->k() == params_.lattice_points)
            return;
        std::vector<int> offs, pads;
        std::vector<const std::vector<double> *> coords;
        for (const auto &S : sys_)
        {
            if (S.coord_dim != 3 || static_cast<int>(S.coords.size()) != 3 * S.nb)
                throw InvalidArgumentError(
                    "[MAS] learned correction needs 3-D coordinates for every block node (psmas_set_coordinates).");
            offs.push_back(S.node_off);
            pads.push_back(S.npad);
            coords.push_back(&S.coords);
        }
        const int old_width = lcc_->lattice_ready() ? lcc_->width() : -1;
        lcc_->build_lattice(params_.lattice_points, bd_, offs, pads, coords, block_of_slot_, nonisolated_, stream_);
        if (old_width >= 0 && old_width != lcc_->width() && lcc_->trained())
            lcc_->reset_network();
        lattice_stale_ = false;
        graphs_->clear(); // captured launches hold the lattice buffers
    }

    template <int BD>
    void Solver::learned_pair_from_coef(const std::vector<double> &w)
    {
        // e = P w, r = A e (FP64), u = P^T r: an exact pair (u, w) of the lattice Galerkin map.
        const long long nvec = static_cast<long long>(npad_) * BD;
        lcc_->prolong_host_coef<BD>(w, d_p_.data(), nvec, stream_);
        launch_spmv_plain<BD, double, double>(sell(d_vals64_.data()), d_p_.data(), d_Ap_.data(), stream_);
        lcc_->restrict_to_lattice<BD, double>(d_Ap_.data(), nullptr, stream_);
        const std::vector<double> u = lcc_->download_u(stream_);
        double nu = 0.0;
        for (double v : u)
            nu += v * v;
        if (nu > 0.0 && std::isfinite(nu))
            lcc_->add_pair(u, w);
    }

    template <int BD>
    void Solver::learned_collect_pairs()
    {
        // Snapshot errors e_i = x* - x_i, projected onto the lattice in the least-squares sense.
        const long long nvec = static_cast<long long>(npad_) * BD;
        for (int i = 0; i < nsnap_; ++i)
        {
            launch_difference(nvec, d_x_.data(), d_snap_.data() + static_cast<size_t>(i) * nvec, d_Ap_.data(), stream_);
            lcc_->restrict_to_lattice<BD, double>(d_Ap_.data(), nullptr, stream_);
            const std::vector<double> v = lcc_->download_u(stream_);
            double nv = 0.0;
            for (double x : v)
                nv += x * x;
            if (!(nv > 0.0) || !std::isfinite(nv))
                continue;
            learned_pair_from_coef<BD>(lcc_->ls_project(v));
        }
        std::normal_distribution<double> N01(0.0, 1.0);
        for (int j = 0; j < params_.learned_random_pairs; ++j)
        {
            std::vector<double> w(lcc_->m());
            for (double &x : w)
                x = N01(lcc_->rng());
            learned_pair_from_coef<BD>(w);
        }
        PSMAS_HIP(hipStreamSynchronize(stream_));
    }

    double Solver::learned_train(int epochs, int batch, float lr, unsigned seed)
    {
        if (!lcc_)
            throw std::runtime_error("[MAS] learned correction: no training pairs collected.");
        if (epochs < 1 || batch < 1 || !(lr > 0.0f))
            throw InvalidArgumentError("[MAS] learned correction: epochs, batch and lr must be positive.");
        const double loss = lcc_->train(params_.learned_hidden, epochs, batch, lr, seed, params_.learned_linear);
        graphs_->clear(); // captured launches hold the network's buffers
        return loss;
    }

    double Solver::learned_exact()
    {
        const auto t0 = std::chrono::steady_clock::now();
        PSMAS_HIP(hipSetDevice(device_));
        check_ready_single(dim());
        ensure_lattice();
        const int m = lcc_->m();
        std::vector<double> G(static_cast<size_t>(m) * m);
        auto body = [&](auto bdc) {
            constexpr int BD = decltype(bdc)::value;
            const long long nvec = static_cast<long long>(npad_) * BD;
            std::vector<double> w(m, 0.0);
            for (int j = 0; j < m; ++j)
            {
                // column j: P^T A P e_j
                w[j] = 1.0;
                lcc_->prolong_host_coef<BD>(w, d_p_.data(), nvec, stream_);
                launch_spmv_plain<BD, double, double>(sell(d_vals64_.data()), d_p_.data(), d_Ap_.data(), stream_);
                lcc_->restrict_to_lattice<BD, double>(d_Ap_.data(), nullptr, stream_);
                const std::vector<double> u = lcc_->download_u(stream_);
                for (int i = 0; i < m; ++i)
                    G[static_cast<size_t>(i) * m + j] = u[i];
                w[j] = 0.0;
            }
        };
        if (bd_ == 1)
            body(std::integral_constant<int, 1>());
        else if (bd_ == 2)
            body(std::integral_constant<int, 2>());
        else
            body(std::integral_constant<int, 3>());
        for (int i = 0; i < m; ++i)
        {
            for (int j = i + 1; j < m; ++j) // symmetric part (the sums differ in rounding only)
            {
                const double v = 0.5 * (G[static_cast<size_t>(i) * m + j] + G[static_cast<size_t>(j) * m + i]);
                G[static_cast<size_t>(i) * m + j] = G[static_cast<size_t>(j) * m + i] = v;
            }
            // A lattice unknown no coupled node touches has a zero row: its u entry is always 0.
            if (G[static_cast<size_t>(i) * m + i] == 0.0)
                G[static_cast<size_t>(i) * m + i] = 1.0;
        }
        if (!cholesky_inverse(m, G))
            throw std::runtime_error("[MAS] learned correction: lattice Galerkin matrix not positive definite.");
        lcc_->install_linear(G, params_.learned_hidden, 1u, "galerkin");
        graphs_->clear(); // captured launches hold the network's buffers
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void Solver::learned_reset(bool drop_pairs)
    {
        if (!lcc_)
            return;
        lcc_->reset_network();
        if (drop_pairs)
            lcc_->clear_pairs();
        graphs_->clear();
    }

    // --------------------------------------------------------------------- introspection
    void Solver::get_partition(std::vector<int> &offsets, std::vector<int> &perm) const
    {
        if (!factorized_)
            throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
        const System &S = sys_[0];
        offsets.assign(S.nparts + 1, 0);
        for (int p = 0; p <= S.part.nparts; ++p)
            offsets[p] = S.part.offsets[p];
        for (int p = S.part.nparts + 1; p <= S.nparts; ++p)
            offsets[p] = S.part.offsets[S.part.nparts];
        perm = S.part.perm;
    }

    std::vector<int> Solver::level_info() const
    {
        const System &S = sys_[0];
        return {level_num_, cco_[0], cco_[1], cco_[2], cco_[3], S.nparts, S.npad, stored_mats_,
                static_cast<int>(std::min<long long>(S.nnzb, 2147483647LL)), static_cast<int>(nslots_), S.nb, S.n};
    }

    std::vector<int> Solver::coarse_map() const
    {
        if (!factorized_)
            throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
        std::vector<int4> cm(npad_);
        PSMAS_HIP(hipMemcpy(cm.data(), d_cmap_.data(), sizeof(int4) * npad_, hipMemcpyDeviceToHost));
        const System &S = sys_[0];
        std::vector<int> out(static_cast<size_t>(S.nb) * 4);
        for (int b = 0; b < S.nb; ++b)
        {
            const int4 c = cm[S.node_off + S.slot_of_block[b]];
            out[4 * b + 0] = c.x;
            out[4 * b + 1] = c.y;
            out[4 * b + 2] = c.z;
            out[4 * b + 3] = c.w;
        }
        return out;
    }

    std::vector<double> Solver::bank_inverses() const
    {
        if (!factorized_)
            throw std::runtime_error("[MAS] Size mismatch. Did you forget to call factorize?");
        std::vector<double> out(d_mats64_.size());
        PSMAS_HIP(hipMemcpy(out.data(), d_mats64_.data(), d_mats64_.bytes(), hipMemcpyDeviceToHost));
        return out;
    }

    void Solver::apply_preconditioner(const double *r, double *z, bool mixed)
    {
        PSMAS_HIP(hipSetDevice(device_));
        check_ready_single(dim());
        const System &S = sys_[0];
        const long long nvec = static_cast<long long>(npad_) * bd_;
        h_active_.assign(1, 1);
        h_select_.assign(1, 1);
        upload_flags();
        PSMAS_HIP(hipMemcpyAsync(d_io_.data(), r, sizeof(double) * S.n, hipMemcpyHostToDevice, stream_));
        launch_gather_in<double>(npad_, bd_, S.n, d_block_of_slot_.data(), d_io_.data(), d_p_.data(), stream_);
        launch_fill<double>(nvec, d_Ap_.data(), 0.0, stream_);
        launch_set_scalar(d_sc_.data(), 1, S_GUARD, params_.division_guard, stream_);
        ensure_coarse_solver();
        if (mixed)
            ensure_mixed_data();
        auto body = [&](auto bdc) {
            constexpr int BD = decltype(bdc)::value;
            const SellView<double> A = sell(d_vals64_.data());
            if (!mixed)
            {
                // r - A*0 = r, restricted to coarse level 0 by the same fused kernel the solve uses.
                launch_spmv_res<BD, double, double, double>(A, d_Ap_.data(), d_p_.data(), d_mr_.data(),
                                                            fine_restrict(),
                                                            d_mr_.data() + static_cast<size_t>(npad_) * BD,
                                                            reductions().rr_true, sysmap(), false, stream_);
                fp64_precondition<BD>(true);
                launch_scatter_out<double>(S.nb, BD, S.n, d_slot_of_block_.data(), d_mz_.data(), d_io_.data(),
                                           stream_);
            }
            else
            {
                launch_spmv_res<BD, double, double, float>(A, d_Ap_.data(), d_p_.data(), d_mr32_.data(),
                                                           fine_restrict(),
                                                           d_mr32_.data() + static_cast<size_t>(npad_) * BD,
                                                           reductions().rr_true, sysmap(), false, stream_);
                mixed_part_b<BD>(true);
                launch_copy_convert<float, double>(nvec, d_mz32_.data(), d_Ap_.data(), stream_);
                launch_scatter_out<double>(S.nb, BD, S.n, d_slot_of_block_.data(), d_Ap_.data(), d_io_.data(),
                                           stream_);
            }
        };
        if (bd_ == 1)
            body(std::integral_constant<int, 1>());
        else if (bd_ == 2)
            body(std::integral_constant<int, 2>());
        else
            body(std::integral_constant<int, 3>());
        PSMAS_HIP(hipMemcpyAsync(z, d_io_.data(), sizeof(double) * S.n, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
        collect_timings();
    }

    void Solver::spmv(const double *x, double *y, bool fp32)
    {
        PSMAS_HIP(hipSetDevice(device_));
        check_ready_single(dim());
        const System &S = sys_[0];
        const long long nvec = static_cast<long long>(npad_) * bd_;
        PSMAS_HIP(hipMemcpyAsync(d_io_.data(), x, sizeof(double) * S.n, hipMemcpyHostToDevice, stream_));
        launch_gather_in<double>(npad_, bd_, S.n, d_block_of_slot_.data(), d_io_.data(), d_p_.data(), stream_);
        if (fp32)
            ensure_mixed_data();
        else
            prepare_sym<double>();
        // The kernel the PCG loop uses for y = A p (the symmetric form when it applies).
        auto body = [&](auto bdc) {
            constexpr int BD = decltype(bdc)::value;
            if (!fp32)
            {
                SymView<double> S;
                if (sym_view(S))
                    launch_spmv_sym<BD, double, double>(sell(d_vals64_.data()), S, d_p_.data(), d_Ap_.data(),
                                                        reductions(), sysmap(), stream_);
                else
                    launch_spmv_plain<BD, double, double>(sell(d_vals64_.data()), d_p_.data(), d_Ap_.data(), stream_);
            }
            else
            {
                launch_copy_convert<double, float>(nvec, d_p_.data(), d_p32_.data(), stream_);
                SymView<float> S;
                if (sym_view(S))
                    launch_spmv_sym<BD, float, float>(sell(d_vals32_.data()), S, d_p32_.data(), d_Ap32_.data(),
                                                      reductions(), sysmap(), stream_);
                else
                    launch_spmv_plain<BD, float, float>(sell(d_vals32_.data()), d_p32_.data(), d_Ap32_.data(), stream_);
                launch_copy_convert<float, double>(nvec, d_Ap32_.data(), d_Ap_.data(), stream_);
            }
        };
        if (bd_ == 1)
            body(std::integral_constant<int, 1>());
        else if (bd_ == 2)
            body(std::integral_constant<int, 2>());
        else
            body(std::integral_constant<int, 3>());
        launch_scatter_out<double>(S.nb, bd_, S.n, d_slot_of_block_.data(), d_Ap_.data(), d_io_.data(), stream_);
        PSMAS_HIP(hipMemcpyAsync(y, d_io_.data(), sizeof(double) * S.n, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
    }

    int Solver::lattice_size()
    {
        PSMAS_HIP(hipSetDevice(device_));
        check_ready_single(dim());
        ensure_lattice();
        return lcc_->m();
    }

    void Solver::lattice_restrict(const double *v, double *u)
    {
        const int m = lattice_size();
        const System &S = sys_[0];
        PSMAS_HIP(hipMemcpyAsync(d_io_.data(), v, sizeof(double) * S.n, hipMemcpyHostToDevice, stream_));
        launch_gather_in<double>(npad_, bd_, S.n, d_block_of_slot_.data(), d_io_.data(), d_p_.data(), stream_);
        auto body = [&](auto bdc) {
            constexpr int BD = decltype(bdc)::value;
            lcc_->restrict_to_lattice<BD, double>(d_p_.data(), nullptr, stream_);
        };
        if (bd_ == 1)
            body(std::integral_constant<int, 1>());
        else if (bd_ == 2)
            body(std::integral_constant<int, 2>());
        else
            body(std::integral_constant<int, 3>());
        const std::vector<double> h = lcc_->download_u(stream_);
        std::copy(h.begin(), h.begin() + m, u);
    }

    void Solver::lattice_prolong(const double *w, double *e)
    {
        const int m = lattice_size();
        const System &S = sys_[0];
        const std::vector<double> hw(w, w + m);
        auto body = [&](auto bdc) {
            constexpr int BD = decltype(bdc)::value;
            lcc_->prolong_host_coef<BD>(hw, d_p_.data(), static_cast<long long>(npad_) * BD, stream_);
        };
        if (bd_ == 1)
            body(std::integral_constant<int, 1>());
        else if (bd_ == 2)
            body(std::integral_constant<int, 2>());
        else
            body(std::integral_constant<int, 3>());
        launch_scatter_out<double>(S.nb, bd_, S.n, d_slot_of_block_.data(), d_p_.data(), d_io_.data(), stream_);
        PSMAS_HIP(hipMemcpyAsync(e, d_io_.data(), sizeof(double) * S.n, hipMemcpyDeviceToHost, stream_));
        PSMAS_HIP(hipStreamSynchronize(stream_));
    }

    // ------------------------------------------------------------------------------ timing
    void Solver::tic(const char *name, double bytes)
    {
        if (!timing_on_)
            return;
        if (event_pool_used_ + 2 > event_pool_.size())
        {
            for (int i = 0; i < 64; ++i)
            {
                hipEvent_t ev;
                PSMAS_HIP(hipEventCreate(&ev));
                event_pool_.push_back(ev);
            }
        }
        pending_start_ = event_pool_[event_pool_used_++];
        pending_name_ = name;
        pending_bytes_ = bytes;
        PSMAS_HIP(hipEventRecord(pending_start_, stream_));
    }

    void Solver::toc()
    {
        if (!timing_on_ || pending_start_ == nullptr)
            return;
        hipEvent_t stop = event_pool_[event_pool_used_++];
        PSMAS_HIP(hipEventRecord(stop, stream_));
        pending_events_.push_back({pending_name_, {pending_start_, stop}});
        auto &t = timings_[pending_name_];
        t.name = pending_name_;
        t.bytes_per_launch = pending_bytes_;
        pending_start_ = nullptr;
    }

    void Solver::collect_timings()
    {
        if (pending_events_.empty())
            return;
        PSMAS_HIP(hipStreamSynchronize(stream_));
        for (auto &pe : pending_events_)
        {
            float ms = 0.0f;
            PSMAS_HIP(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
            auto &t = timings_[pe.first];
            t.launches += 1;
            t.total_ms += ms;
        }
        pending_events_.clear();
        event_pool_used_ = 0;
    }

    std::vector<KernelTiming> Solver::kernel_timings() const
    {
        std::vector<KernelTiming> out;
        for (const auto &kv : timings_)
            out.push_back(kv.second);
        return out;
    }

    void Solver::reset_kernel_timings() { timings_.clear(); }
} // namespace psmas
