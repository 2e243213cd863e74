// Solver parameters and their JSON reader.
//
// Keys, types and meaning follow the reference's MAS settings (linear-solver-spec.json:456-509,
// read by MASSolverImpl::set_parameters at src/polysolve/linear/MASSolver.cu:198-212, which only
// looks at params["MAS"] when called through MASSolver::set_parameters, MASSolver.cu:605-614).
// Defaults are the reference's in-code defaults (MASSolver.cu:137-143), which are what a
// MASSolver created by name and configured only through set_parameters uses.
//
// Keys this implementation adds (none collides with the reference spec):
//   precision            "fp64" (reference algorithm) | "mixed" (FP32 inner PCG, FP16 bank
//                        inverses, FP64 reliable updates)
//   aggregation          "auto" | "morton" | "graph"; auto = morton when coordinates were given
//   coarse_solver        "bank" (reference: every level is bank-block-diagonal) | "cholesky"
//                        (coarsest level solved exactly by a dense FP64 Cholesky)
//   reliable_update_delta   residual reduction that triggers an FP64 reliable update (mixed)
//   check_interval       iterations between host convergence checks (reference: 10)
//   true_residual_period reference: 4 (fp64 path only)
//   use_graphs           capture the iteration sequence in HIP graphs
//   learned_correction   mixed path: add the trained lattice coarse correction (flexible PCG;
//                        plain MAS takes over when the residual stops decreasing)
//   learned_collect      mixed path: after each converged solve, turn snapshot iterates into
//                        training pairs for the coarse-correction network
//   lattice_points       points per axis of the correction's coarse lattice (k; k^3 points)
//   learned_hidden       hidden width of the correction network (multiple of 8)
//   learned_random_pairs Gaussian training pairs added per collected solve
//   learned_snapshots    at most this many snapshot iterates per collected solve
//   learned_stall_segments  check intervals without a new smallest residual before the learned
//                        correction is switched off for the rest of the solve
//   learned_output_scale factor on the learned correction (1 = as trained)
//   learned_network      "mlp" (default: hidden layers plus the linear skip layer) | "linear"
//                        (only the skip layer is trained: a learned linear coarse map)
//   spmv_unroll          SpMV row-loop batch depth 2 | 4 | 8 (chosen by psmas_autotune)
//   spmv_copies          after factorize_copies, one SpMV launch reads the matrix once for every
//                        right-hand side (default true; false: one SpMV launch per right-hand side)
//   device_setup         build the SELL matrix on the GPU from the CSC arrays, for one system, a
//                        batch or the copies of factorize_copies, and aggregate there when no
//                        coordinates are given (default true; false: the host build)
//   spmv_halo            SpMV gathers x from a per-workgroup copy of the entries its rows read,
//                        staged in LDS, with 16-bit local columns (default true; used when every
//                        workgroup's set fits in SPMV_HALO_CAP nodes)
//   spmv_fuse_direction  the SpMV forms the search direction p = z + beta p itself while staging
//                        its halo, instead of a separate update kernel (bit-identical iterates;
//                        used with the halo form and without copies).  Default false: on MI355X
//                        the longer halo stage (two gathers and beta's reduction ahead of the
//                        row loop) costs more than the update kernel it removes
//   restrict_piece       longest level-1 descendant list one wave sums when restricting coarse
//                        levels 2.. (longer lists are cut into pieces added where they are read);
//                        9..512, default 512; takes effect at the next factorize
//   apply_workgroups     at most this many workgroups apply bank matrices in one apply launch
//                        (0, the default: as many as are resident at once); each walks its
//                        matrices grid-stride, so a small cap makes every workgroup take many
//                        matrices (parity tests of that loop on small systems)
//   single_reduction     the mixed path's PCG iteration with one reduction phase (Chronopoulos &
//                        Gear 1989): the SpMV multiplies u = M r, which it assembles itself from
//                        the preconditioner's levels where it reads it, and one update kernel
//                        forms p, s = A p (by recurrence), x and r (default false; used for one
//                        right-hand side or a batch, with the halo form of the matrix; false: the
//                        classic iteration with separate prolongation and direction kernels)
//   apply_wave           the FP16 apply of one right-hand side takes a wave per bank matrix with
//                        the matrix's tile rows in registers (default true); false: the persistent
//                        kernel that unpacks each matrix into LDS (the one copies of a
//                        factorisation always use)
//   spmv_symmetric       the single-matrix SpMV of the PCG loop (y = A p) reads the symmetric form
//                        of the matrix: a block coupling two rows of the same 256-node block is
//                        stored once and applied to both rows (default true; used when the
//                        pattern is symmetric within the node blocks and the blocks stored once
//                        are symmetric to 1e-12 of their largest entry, see get_info
//                        "spmv_symmetric"); false: every block is read in its own row
//   division_guard       alpha/beta are set to 0 when |denominator| < division_guard; 0 (default)
//                        sets them to 0 only for a non-finite quotient; 1e-20 is the reference's
//                        absolute rule (MASSolver.cu:43-56)
#pragma once

#include <string>

namespace psmas
{
    enum class Precision
    {
        FP64,
        Mixed,
    };

    enum class Aggregation
    {
        Auto,
        Morton,
        Graph,
    };

    enum class CoarseSolver
    {
        Bank,
        Cholesky,
    };

    struct Params
    {
        int block_dim = 1;
        int max_iter = 5000;
        double rel_tol = 1e-5;
        double abs_tol = 1e-8;
        bool lazy_partitioning = false;
        bool use_preconditioned_residual_norm = false;
        int true_residual_period = 4;

        Precision precision = Precision::FP64;
        Aggregation aggregation = Aggregation::Auto;
        CoarseSolver coarse_solver = CoarseSolver::Bank;
        double reliable_update_delta = 0.1;
        int check_interval = 10;
        bool use_graphs = true;
        bool learned_correction = false;
        double division_guard = 0.0;
        bool learned_collect = false;
        int lattice_points = 7;
        int learned_hidden = 128;
        int learned_random_pairs = 8;
        int learned_snapshots = 64;
        int learned_stall_segments = 3;
        double learned_output_scale = 1.0;
        int spmv_unroll = 4;
        bool spmv_halo = true;
        bool device_setup = true;
        bool spmv_copies = true;
        bool spmv_fuse_direction = false;
        int restrict_piece = 512;
        int apply_workgroups = 0;
        bool apply_wave = true;
        bool single_reduction = false;
        bool spmv_symmetric = true;
        bool learned_linear = false;
    };

    /// Apply a JSON object to params.  Accepts either {"MAS": {...}} (only the "MAS" member is
    /// read, as MASSolver::set_parameters does) or the inner object itself.  Unknown keys are
    /// ignored.  Throws InvalidArgumentError on malformed JSON or a bad value.
    void apply_json_parameters(Params &p, const std::string &json);

    /// Canonical JSON dump of all parameters.
    std::string params_to_json(const Params &p);
} // namespace psmas
