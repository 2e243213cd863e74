/*
 * psmas — C ABI of the MI355X-native MAS-preconditioned PCG linear solver.
 *
 * Drop-in boundary for polysolve's "MAS" linear solver.  The reference exposes the solver as
 * polysolve::linear::MASSolver (src/polysolve/linear/MASSolver.hpp:35-72), created by name through
 * polysolve::linear::Solver::create("MAS", ...) (src/polysolve/linear/Solver.cpp:403-408) and used
 * through the Solver virtuals (src/polysolve/linear/Solver.hpp:90-131).  Each entry point below
 * names the reference interface it replaces.  Signatures use plain pointers and sizes only;
 * matrices are compressed column-major with int32 indices, the layout of polysolve's
 * StiffnessMatrix (src/polysolve/Types.hpp:11-15).
 *
 * Every call returns a status code.  PSMAS_OK is success; every other code corresponds to the
 * exception the reference throws: PSMAS_ERROR_OUT_OF_MEMORY <-> CudaOutOfMemoryError
 * (src/polysolve/linear/mas_utils/CudaUtils.cuh:33-80), all others <-> std::runtime_error.  The
 * message is available from psmas_last_error.  A handle is not thread-safe; all calls are
 * synchronous on return (MASSolver.cu:408).
 */
#ifndef PSMAS_H
#define PSMAS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psmas_solver *psmas_handle;

enum psmas_status
{
    PSMAS_OK = 0,
    PSMAS_ERROR_RUNTIME = 1,          /* std::runtime_error */
    PSMAS_ERROR_OUT_OF_MEMORY = 2,    /* CudaOutOfMemoryError */
    PSMAS_ERROR_INVALID_ARGUMENT = 3, /* std::runtime_error ("[MAS] invalid ...") */
    PSMAS_ERROR_NO_DEVICE = 4,        /* std::runtime_error from the constructor (MASSolver.cu:188-191) */
    PSMAS_ERROR_BUFFER_TOO_SMALL = 5  /* output string truncated; *needed holds the size */
};

/* ---- Solver lifecycle ------------------------------------------------------------------- */

/* Replaces Solver::create("MAS", precond) -> std::make_unique<MASSolver>()
 * (Solver.cpp:403-408; constructor MASSolver.cu:186-196,598-601).  Fails with
 * PSMAS_ERROR_NO_DEVICE when no GPU is visible. */
int psmas_create(psmas_handle *out);

/* Replaces ~MASSolver (MASSolver.cu:603). */
int psmas_destroy(psmas_handle h);

/* Replaces MASSolver::name() == "MAS" (MASSolver.cu:636-639).  Static string. */
const char *psmas_name(void);

/* Replaces MASSolver::set_parameters(const json&) (MASSolver.cu:198-212,605-614).  json is a
 * JSON object text; its "MAS" member is read with the keys of linear-solver-spec.json:456-509
 * (block_dim, max_iter, relative_tolerance, absolute_tolerance, lazy_partitioning,
 * use_preconditioned_residual_norm) plus this implementation's extension keys (precision,
 * aggregation, coarse_solver, reliable_update_delta, check_interval, true_residual_period,
 * use_graphs, learned_correction).  Without a "MAS" member the text is read as the inner
 * object when it holds any of those keys, and ignored otherwise. */
int psmas_set_parameters(psmas_handle h, const char *json);

/* Current parameters as JSON (no reference counterpart; for inspection). */
int psmas_get_parameters(psmas_handle h, char *buf, size_t cap, size_t *needed);

/* Replaces MASSolver::get_info(json&) (MASSolver.cu:214-219,616-619): writes a JSON object with
 * solver_iter, solver_error, solver_status (strings of MASSolver.hpp:18-33) and extra keys. */
int psmas_get_info(psmas_handle h, char *buf, size_t cap, size_t *needed);

/* Replaces MASSolver::analyze_pattern(A, precond_num) — a no-op (MASSolver.cu:221,621-624). */
int psmas_analyze_pattern(psmas_handle h, int n, const int *colptr, const int *rowidx, int precond_num);

/* Replaces MASSolver::factorize(const StiffnessMatrix&) (MASSolver.cu:292-374,626-629).
 * A is n x n, symmetric, compressed column-major: colptr[n+1], rowidx[nnz], vals[nnz].  The
 * matrix is copied; the caller may free it on return. */
int psmas_factorize_csc(psmas_handle h, int n, long long nnz, const int *colptr, const int *rowidx,
                        const double *vals);

/* Replaces MASSolver::solve(b, x) (MASSolver.cu:376-409,631-634).  Host arrays of length n.
 * The initial x is ignored (the reference starts from zero, MASSolver.cu:395-398). */
int psmas_solve(psmas_handle h, int n, const double *b, double *x);

/* Error message of the last failing call on h (or of the last failing psmas_create when h is
 * NULL).  Valid until the next call on h. */
const char *psmas_last_error(psmas_handle h);

/* ---- Dirichlet boundary conditions (src/polysolve/linear/FEMSolver.hpp:35-47) ------------ */

/* prefactorize (FEMSolver.cpp:269-315): factorise the matrix A' obtained from A by dropping every
 * entry in the row or column of a fixed dof and putting N(k) on each diagonal (1 for a fixed dof,
 * a_kk otherwise; FEMSolver.cpp:283-306).  dirichlet_nodes: ndir scalar dof indices in [0, n)
 * (duplicates allowed).  The reference overwrites A with A'; here A is left as it is, and the
 * couplings a_ij (i free, j fixed) are kept on the device for the right-hand side lift.  With
 * "lazy_partitioning": true, a later call with the same pattern and the same fixed dofs only
 * refreshes the values (on the GPU), as does psmas_factorize_values_device with A's values. */
int psmas_dirichlet_factorize_csc(psmas_handle h, int n, long long nnz, const int *colptr, const int *rowidx,
                                  const double *vals, int ndir, const int *dirichlet_nodes);

/* The same with A's CSC arrays in device memory (HIP pointers on the solver's device; validated
 * there, as in psmas_factorize_csc_device); dirichlet_nodes stays a host array.  The elimination
 * and the matrix build run on the GPU and the host never receives A; the values of later Newton
 * steps follow through psmas_factorize_values_device ("lazy_partitioning": true).  Replaces
 * prefactorize (FEMSolver.cpp:269-315) for callers that assemble on the GPU. */
int psmas_dirichlet_factorize_csc_device(psmas_handle h, int n, long long nnz, const int *d_colptr,
                                         const int *d_rowidx, const double *d_vals, int ndir,
                                         const int *dirichlet_nodes);

/* dirichlet_solve_prefactorized (FEMSolver.cpp:318-342): f (length n, host) is replaced by the
 * lifted right-hand side g = f - (I - N) A N f, computed on the GPU with the A of the last
 * psmas_dirichlet_factorize_csc, and x receives the solution of A' x = g (x[i] = f[i] at a fixed
 * dof up to the solver tolerance).  The one-call dirichlet_solve (FEMSolver.cpp:97-266) is
 * psmas_dirichlet_factorize_csc followed by this call.  Fails without a Dirichlet
 * factorisation. */
int psmas_dirichlet_solve(psmas_handle h, int n, double *f, double *x);
/* Same with f and x in device memory (HIP pointers on the solver's device); f is overwritten
 * by g in place. */
int psmas_dirichlet_solve_device(psmas_handle h, int n, double *d_f, double *d_x);

/* ---- Extensions (no counterpart in the reference) --------------------------------------- */

/* Same as psmas_solve with b and x in device memory (HIP pointers on the solver's device). */
int psmas_solve_device(psmas_handle h, int n, const double *d_b, double *d_x);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same solve as
 * psmas_solve_device (bit-identical results), but it allocates nothing: the solve's reduction
 * partials and per-system scalars live in a static 1 MiB array of the library, one per GPU, held
 * until the process ends, and kernels are launched one by one (no graph replay).  Limit: systems
 * of at most 13,104 node blocks (256 padded nodes each, about 3.35 M block nodes); larger ones are
 * refused with PSMAS_ERROR_INVALID_ARGUMENT.  Parts that a factorisation builds at its first
 * solve (the mixed path's FP32/FP16 copies, the dense coarsest solve, the learned correction's
 * lattice) must exist: call psmas_solve_device once after factorize, otherwise the call fails
 * with PSMAS_ERROR_RUNTIME; learned_collect and kernel timing are refused.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (several threads or handles) race. */
int psmas_solve_device_noalloc(psmas_handle h, int n, const double *d_b, double *d_x);

/* Same as psmas_factorize_csc with the CSC arrays in device memory (HIP pointers on the solver's
 * device): d_colptr n+1, d_rowidx and d_vals nnz entries.  The arrays are validated on the GPU
 * first (the same errors as psmas_factorize_csc); the domains (Morton chunks with coordinates,
 * the GPU graph aggregation without) and the matrix are built there, and the host never copies
 * A back.  With "lazy_partitioning"
 * the values of later Newton steps can follow through psmas_factorize_values_device.  Replaces
 * the reference's two host-to-device uploads of the CSC arrays (BSRAdjacency.cu:180-197,
 * BSRMatrix.cu:214-231) for callers that assemble on the GPU. */
int psmas_factorize_csc_device(psmas_handle h, int n, long long nnz, const int *d_colptr, const int *d_rowidx,
                               const double *d_vals);

/* Values-only refactorisation from device memory: d_vals holds nnz doubles (HIP pointer on the
 * solver's device) in the order of the column-compressed pattern of the last full factorisation,
 * which must have run with "lazy_partitioning": true on one system without duplicate entries.
 * The values are scattered into the solver's matrix and the numeric MAS parts rebuilt, with no
 * host transfer (for a caller that assembles A on the GPU).  Fails when no such pattern is
 * recorded or nnz differs.  After psmas_dirichlet_factorize_csc the values are those of the
 * matrix before elimination (its pattern): the eliminated entries are dropped and the couplings
 * refreshed on the device.  A refinement of the reference's lazy re-factorisation
 * (MASSolver.cu:304-321), which has no device-value entry. */
int psmas_factorize_values_device(psmas_handle h, long long nnz, const double *d_vals);

/* Several right-hand sides of one matrix (no counterpart in the reference; for Newton or IPC
 * callers that solve one Hessian for several vectors): factorise A once (one SELL matrix, one set
 * of bank inverses, one coarsest inverse) for nrhs (1..8) right-hand sides whose vectors, coarse
 * numbering and CG scalars are per copy; the SpMV (parameter spmv_copies) and every bank apply
 * read each matrix once for all of them.  Then solve all right-hand sides in one launch sequence,
 * each with its own convergence test and status (psmas_batch_get_info(h, c, ...) for copy c);
 * per copy the arithmetic is that of the single system.  B and X hold nrhs vectors of length n
 * one after the other (column-major n x nrhs), host arrays or device pointers (_device). */
int psmas_factorize_multi_csc(psmas_handle h, int nrhs, int n, long long nnz, const int *colptr, const int *rowidx,
                              const double *vals);
/* The same with A's CSC arrays in device memory (HIP pointers on the solver's device, validated
 * there); the matrix is built on the GPU. */
int psmas_factorize_multi_csc_device(psmas_handle h, int nrhs, int n, long long nnz, const int *d_colptr,
                                     const int *d_rowidx, const double *d_vals);
int psmas_solve_multi(psmas_handle h, int n, int nrhs, const double *B, double *X);
int psmas_solve_multi_device(psmas_handle h, int n, int nrhs, const double *d_B, double *d_X);

/* Node coordinates for Morton-ordered aggregation into 32-node domains: nb tuples of dim
 * (1..3) doubles, one per block node (block_dim scalars).  Takes effect at the next factorize. */
int psmas_set_coordinates(psmas_handle h, int nb, int dim, const double *xyz);

/* Batches of independent systems (no counterpart in the reference): nsys systems are factorised
 * together and solved in one launch sequence; each keeps its own CG scalars, convergence test and
 * status.  Arrays of nsys pointers / sizes; each system as in psmas_factorize_csc.  After a
 * batch factorisation psmas_solve / psmas_get_info refer to the batch only when nsys == 1. */
int psmas_batch_factorize_csc(psmas_handle h, int nsys, const int *n, const long long *nnz,
                              const int *const *colptr, const int *const *rowidx, const double *const *vals);
/* b[s], x[s]: host arrays (device == 0) or device pointers (device != 0) of length n[s]. */
int psmas_batch_solve(psmas_handle h, int nsys, const double *const *b, double *const *x, int device);
/* get_info of system sys of the last batch solve. */
int psmas_batch_get_info(psmas_handle h, int sys, char *buf, size_t cap, size_t *needed);
/* Coordinates of system sys (before psmas_batch_factorize_csc). */
int psmas_batch_set_coordinates(psmas_handle h, int sys, int nb, int dim, const double *xyz);

/* Learned coarse correction of the mixed path (no counterpart in the reference).  With the
 * parameter learned_collect, every converged single-system mixed solve turns snapshot iterates
 * into training pairs (u = P^T A P w, w) on a lattice coarse space P built from the coordinates;
 * psmas_learned_train fits the correction network to all pairs collected so far (Adam on the GPU,
 * cosine learning-rate decay) and *loss receives the final epoch's mean loss; with the parameter
 * learned_correction the mixed solve then adds the network's coarse correction inside a flexible
 * PCG and falls back to plain MAS when the residual stops decreasing.  psmas_learned_reset drops
 * the network (and the pairs when drop_pairs != 0). */
int psmas_learned_train(psmas_handle h, int epochs, int batch, float lr, unsigned seed, double *loss);
int psmas_learned_reset(psmas_handle h, int drop_pairs);
/* Ablation of the learned correction (no counterpart in the reference): replace the network by
 * the exact lattice coarse solve of the current single-system factorisation,
 * c = (P^T A P)^-1 P^T r, the linear two-level method the network learns to approximate.
 * P^T A P is assembled on the GPU one column per lattice unknown and inverted on the host; the
 * inverse is held (BF16) in the network's skip layer with the nonlinear branch zero, so the solve
 * uses it exactly as a trained network (parameter learned_correction).  *ms receives the time it
 * took.  With the parameter learned_network = "linear", psmas_learned_train fits the skip layer
 * alone instead (a learned linear map). */
int psmas_learned_exact(psmas_handle h, double *ms);

/* Autotuner (no counterpart in the reference): on the current factorisation (single system or
 * batch) and the right-hand sides of the last solve, times the SpMV row-loop batch depth (2, 4, 8;
 * reps SpMVs each) and, by full solves, the check interval (5, 10, 20) with each coarsest-level
 * solver; keeps the fastest converged choice in the parameters (spmv_unroll, check_interval,
 * coarse_solver) and writes a JSON report of every timing.  Re-runs the solves, so call it once
 * per workload shape with a buffer large enough for the report (4 KB). */
int psmas_autotune(psmas_handle h, int reps, char *buf, size_t cap, size_t *needed);

/* Number of visible GPUs (0 on a machine without one). */
int psmas_device_count(void);

/* Device memory for callers without their own GPU allocator (benchmarks, tests): plain
 * hipMalloc / hipFree / hipMemcpy on the solver's device.  kind: 1 host->device, 2 device->host,
 * 3 device->device. */
int psmas_device_alloc(size_t bytes, void **ptr);
int psmas_device_free(void *ptr);
int psmas_memcpy(void *dst, const void *src, size_t bytes, int kind);
int psmas_device_synchronize(void);

/* Streaming-read probe (no counterpart in the reference; measurement only): reads `bytes` of
 * device memory once per repetition with 16-byte nontemporal loads and writes the best rate in
 * GB/s to *gbps — the read ceiling the SpMV and the bank apply are compared with, beside the
 * 8 TB/s specification.  bytes >= 1 MiB. */
int psmas_read_bandwidth(long long bytes, int reps, double *gbps);

/* ---- Learned coarse-correction network (no counterpart in the reference) ------------------ */

/* MLP d_in -> hidden -> hidden -> d_out (ReLU) with an optional linear skip layer S (skip != 0),
 * BF16-operand MFMA GEMMs with FP32 accumulation, FP32 master weights trained by Adam on the GPU.
 * Parameter vector (FP32): [W1 (hidden x d_in) | b1 | W2 (hidden x hidden) | b2 | W3 (d_out x hidden)
 * | b3 | S (d_out x d_in, skip only)], row-major; y = W3 relu(W2 relu(W1 x + b1) + b2) + b3 (+ S x). */
typedef struct psmas_mlp *psmas_mlp_handle;
int psmas_mlp_create(int d_in, int d_out, int hidden, int skip, unsigned seed, psmas_mlp_handle *out);
int psmas_mlp_destroy(psmas_mlp_handle h);
const char *psmas_mlp_last_error(psmas_mlp_handle h);
long long psmas_mlp_num_parameters(psmas_mlp_handle h);
int psmas_mlp_set_parameters(psmas_mlp_handle h, const float *theta);
int psmas_mlp_get_parameters(psmas_mlp_handle h, float *theta);
/* Y[batch x d_out] = f(X[batch x d_in]) (host arrays). */
int psmas_mlp_forward(psmas_mlp_handle h, int batch, const float *X, float *Y);
/* Mean squared error against T and its gradient w.r.t. the parameter vector (host arrays). */
int psmas_mlp_loss_grad(psmas_mlp_handle h, int batch, const float *X, const float *T, double *loss, float *grad);
/* Adam training on n samples (host arrays); *final_loss = mean loss of the last epoch. */
int psmas_mlp_train(psmas_mlp_handle h, int n, const float *X, const float *T, int epochs, int batch, float lr,
                    unsigned seed, double *final_loss);
/* C[M x N] = op(A) B with BF16-rounded operands and FP32 accumulation (the MLP's GEMM);
 * trans_a: A stored K x M. */
int psmas_gemm_bf16(int M, int N, int K, int trans_a, const float *A, const float *B, float *C);

/* Parse parameters without a device (checks the JSON contract): writes the resulting parameter
 * set as JSON. */
int psmas_parse_parameters(const char *json, char *buf, size_t cap, size_t *needed);

/* ---- Introspection used by the parity tests ---------------------------------------------- */

/* The block graph of a CSC matrix with block size bd (no device needed): the undirected pairs of
 * distinct block nodes joined by a stored entry, in both directions, each row's neighbours in
 * ascending order (ptr[nb+1], adj[edges]), with the L1 weight sum of |a_ij| over the entries of
 * both blocks (w[edges]).  The graph and weights of build_adjacency_from_csc
 * (BSRAdjacency.cu:184-460, keys and weights :60-84) before its quantisation (:131-162); the host
 * graph aggregation matches on these weights.  *nedges receives the edge count; a call with
 * edge_cap smaller than it writes ptr only and returns PSMAS_ERROR_BUFFER_TOO_SMALL. */
int psmas_block_graph(int n, long long nnz, const int *colptr, const int *rowidx, const double *vals, int bd,
                      int *ptr, int ptr_cap, int *adj, double *w, long long edge_cap, long long *nedges);

/* info[0..11] = level_num, cco_nums[4], parts (even), padded nodes, bank matrices, nnz blocks,
 * SELL slots, block nodes, n.  cap >= 12. */
int psmas_get_level_info(psmas_handle h, int *info, int cap);

/* Domain partition in the reference's part_offsets / permutation form (MASSolver.cu:86-114):
 * offsets[parts+1] (parts even; a trailing empty part pads to an even count), perm[nb] maps an
 * input block node to its position. */
int psmas_get_partition(psmas_handle h, int *nparts, int *offsets, int offsets_cap, int *perm, int perm_cap);

/* Coarse-space ids of each input block node for levels 0..3 (-1 past the last level): map[4*nb]. */
int psmas_get_coarse_map(psmas_handle h, int *map, long long cap);

/* Packed upper FP64 bank inverses of all levels, (bank matrices) x (32*bd)(32*bd+1)/2. */
int psmas_get_bank_inverses(psmas_handle h, double *out, long long cap);

/* z = M^-1 r in the input ordering (mixed != 0: the FP16/FP32 preconditioner of the mixed path). */
int psmas_apply_preconditioner(psmas_handle h, int n, const double *r, double *z, int mixed);

/* y = A x in the input ordering with the solver's stored matrix (fp32 != 0: FP32 copy). */
int psmas_spmv(psmas_handle h, int n, const double *x, double *y, int fp32);

/* Lattice coarse space of the learned correction for the factorised single system (built from
 * its coordinates with the current lattice_points): *m = coarse unknowns; u[m] = P^T v and
 * e[n] = P w with v, e in the input ordering. */
int psmas_lattice_size(psmas_handle h, int *m);
int psmas_lattice_restrict(psmas_handle h, int n, const double *v, double *u);
int psmas_lattice_prolong(psmas_handle h, int n, const double *w, double *e);

/* Per-kernel HIP-event timing of the solve's dominant kernels (disables graph replay while on).
 * psmas_get_kernel_timing writes {"name": {"launches": k, "total_ms": t, "bytes": b}, ...}. */
int psmas_set_kernel_timing(psmas_handle h, int on);
int psmas_get_kernel_timing(psmas_handle h, char *buf, size_t cap, size_t *needed);

#ifdef __cplusplus
}
#endif

#endif /* PSMAS_H */
