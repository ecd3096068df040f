/*
 * raptor_amd.h -- C ABI of the MI355X (gfx950) AMG solve path.
 *
 * One context owns one GPU and one HIP stream.  Matrices are CSR with int32
 * row pointers and columns and FP64 values, the layout of
 * raptor::CSRMatrix (core/matrix.hpp:392-394 of the reference); at one rank
 * a raptor::ParCSRMatrix is exactly its on_proc block with identity column
 * maps (core/par_matrix.cpp:111-157), so the same arrays describe both.
 *
 * Conventions
 *   - Every function returns RAMD_OK (0) or a negative RAMD_E* code; the
 *     text of the last error of the calling thread is in ramd_last_error().
 *   - Pointers named *_d are device pointers (from ramd_malloc or any HIP
 *     allocation on the context's device); all others are host pointers.
 *   - Calls on one context are serialised on its stream.  Calls that return
 *     host data synchronise; the rest are stream-ordered.  Handles are not
 *     thread-safe.
 *   - The caller keeps ownership of every array it passes; handles returned
 *     through out-parameters are owned by the caller and freed with the
 *     matching *_destroy / ramd_free call.
 *   - No entry point falls back to a CPU path: without a usable gfx950
 *     device every device entry point fails with RAMD_E_DEVICE.
 */
#ifndef RAPTOR_AMD_H
#define RAPTOR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status codes                                                         */
/* ------------------------------------------------------------------ */
#define RAMD_OK               0
#define RAMD_E_ARG          (-1)  /* invalid argument / shape mismatch          */
#define RAMD_E_DEVICE       (-2)  /* HIP runtime error or no usable device      */
#define RAMD_E_ALLOC        (-3)  /* device or host allocation failed           */
#define RAMD_E_INDEFINITE   (-4)  /* p.Ap < 0 in CG/PCG (reference: exit(-1))   */
#define RAMD_E_NOT_CONVERGED (-5) /* iteration limit reached (info, x is valid) */
#define RAMD_E_STATE        (-6)  /* handle in the wrong state (e.g. no setup)  */
#define RAMD_E_NUMERIC      (-7)  /* NaN/Inf detected or singular coarse matrix */

typedef struct ramd_ctx_s*  ramd_ctx;
typedef struct ramd_csr_s*  ramd_csr;
typedef struct ramd_hier_s* ramd_hier;

/* Text of the calling thread's last error ("" when none). */
int ramd_last_error(char* buf, size_t len);

/* ------------------------------------------------------------------ */
/* context and device memory                                            */
/* ------------------------------------------------------------------ */
/* Opens device `device` and creates the context's stream.  No reference
 * counterpart: the reference's equivalent state is MPI_COMM_WORLD
 * (core/mpi_types.hpp:52-131). */
int ramd_init(int device, ramd_ctx* out);
int ramd_finalize(ramd_ctx ctx);
int ramd_sync(ramd_ctx ctx);
int ramd_malloc(ramd_ctx ctx, size_t bytes, void** out_d);
int ramd_free(ramd_ctx ctx, void* p_d);
int ramd_memcpy_h2d(ramd_ctx ctx, void* dst_d, const void* src, size_t bytes);
int ramd_memcpy_d2h(ramd_ctx ctx, void* dst, const void* src_d, size_t bytes);
int ramd_memcpy_d2d(ramd_ctx ctx, void* dst_d, const void* src_d, size_t bytes);
int ramd_memset_zero(ramd_ctx ctx, void* p_d, size_t bytes);

/* ------------------------------------------------------------------ */
/* device-resident CSR (ParCSRMatrix at one rank)                       */
/* ------------------------------------------------------------------ */
/* Copies host CSR arrays to the device.  Replaces constructing a
 * ParCSRMatrix and its finalize() (core/par_matrix.hpp:481-611,
 * core/par_matrix.cpp:111-157); the rows are stored as given. */
int ramd_csr_create(ramd_ctx ctx, int32_t n_rows, int32_t n_cols, int64_t nnz,
                    const int32_t* rowptr, const int32_t* col, const double* val,
                    ramd_csr* out);
int ramd_csr_info(ramd_csr A, int32_t* n_rows, int32_t* n_cols, int64_t* nnz);
/* Copies the device arrays back (any output pointer may be NULL). */
int ramd_csr_download(ramd_csr A, int32_t* rowptr, int32_t* col, double* val);
int ramd_csr_destroy(ramd_csr A);
/* New values (device array, nnz, A's stored order) for A's pattern; the
 * derived copies (FP32 values, diagonal planes, transpose) follow.  Refused
 * for matrices owned by a hierarchy.  No reference counterpart beyond
 * writing ParCSRMatrix::on_proc->vals. */
int ramd_csr_set_values(ramd_csr A, const double* val_d);
/* Number of stored diagonals when the matrix also has the stencil-diagonal
 * form the SpMV family uses for stencil-structured operators (values per
 * diagonal, a per-row presence mask, no column indices; built automatically
 * for square matrices of >= 4096 rows with <= 32 distinct column offsets,
 * <= 15% padding and one common per-row order); 0 when it is CSR only. */
int ramd_csr_diagonals(ramd_csr A, int* nd);
/* Kernel variant of the stencil-diagonal form for the precision pair pp
 * (0 FP64 values and vectors, 1 FP32/FP32, 2 FP32 values/FP64 vectors):
 * mode 0 gathers x from memory, 1 stages each workgroup's x window in LDS,
 * 2 marches a workgroup's tile of a box-structured grid through `planes`
 * planes (4..256; 0 keeps the current count) with the x planes in LDS
 * (only when every offset is dz*P + dy*L + dx, |d| <= 1, of an L x (P/L) x
 * (n/P) grid), -1 lets the autotuner time them at the next hierarchy setup
 * or solve.  Every variant gives identical bits for the operator's output;
 * the flexible PCG's p.Ap, summed per workgroup in the same launch, follows
 * the workgroup shape in its last bits.  No reference counterpart. */
int ramd_csr_set_dia_staging(ramd_csr A, int pp, int mode, int planes);
/* The variant in use for the precision pair pp (mode as above, -1 not timed
 * yet) and, for the plane-march kernel, its planes per workgroup and the
 * grid's line and plane lengths L and P (0 when the offsets have no box
 * geometry).  No reference counterpart. */
int ramd_csr_dia_staging(ramd_csr A, int pp, int* mode, int* planes, int* L, int* P);
/* CSR kernel of the operator for FP64 (prec 0) or FP32 (prec 1) vectors:
 * on = 1 builds and selects the sliced-row form (64 rows per slice, one
 * lane per row), 2 the same with 16-bit column offsets (col - row; square
 * operators whose offsets fit, RAMD_E_ARG otherwise), 0 returns to the
 * row-block kernel.  All give identical bits; the autotuner picks one per
 * operator at hierarchy setup and at FP32 enable.  No reference
 * counterpart. */
int ramd_csr_set_sliced(ramd_csr A, int prec, int on);
/* The same choice for the lane-per-row kernel over the CSR arrays (one lane
 * walks its own row; no extra storage): on = 1 selects it (and leaves the
 * sliced form), 0 returns to the row-block kernel.  Identical bits.  No
 * reference counterpart. */
int ramd_csr_set_lane_rows(ramd_csr A, int prec, int on);
/* The same choice for the wavefront-per-row kernel (one 64-lane wavefront
 * streams a row's entries with coalesced loads and stages the products in
 * LDS; one lane adds them in stored order): on = 1 selects it, 0 returns to
 * the row-block kernel.  Identical bits; the autotuner tries it for
 * operators averaging >= 24 entries per row.  It is CSR_spmv's row loop
 * (util/linalg/spmv.cpp:52-67) with the loads of a row spread over a
 * wavefront, for operators with few, long rows.  No reference counterpart. */
int ramd_csr_set_wave_rows(ramd_csr A, int prec, int on);
/* Sorted-gather order for the row-block kernel's SpMV at vector precision
 * prec (0 FP64, 1 FP32): within each row block the entries are read in
 * column order, so a wavefront's gathers of x touch few cache lines, and
 * each product is placed at its stored position; rows still add their
 * products in stored order.  on = 1 builds it for the current row blocks
 * (extra column, position and value arrays), 0 releases it.  Identical
 * bits; the autotuner tries it for operators with more columns than rows
 * (restrictions R = P^T, CSR_spmv over P^T: util/linalg/spmv.cpp:52-67).
 * No reference counterpart. */
int ramd_csr_set_sorted_gather(ramd_csr A, int prec, int on);
/* Whether the flexible PCG's FP64 products with A (its outer operator and
 * true residuals, ramd_fpcg and ramd_adjoint) may read A's FP32 copy of the
 * values when every value is exactly representable in FP32: the same FP64
 * products from half the value bytes.  on = 1 (default) allows it, 0 reads
 * the FP64 values as the reference stores them (ParCSRMatrix::mult,
 * util/linalg/par_spmv.cpp:25-59).  y = A p is bit-identical either way;
 * p.Ap, summed per workgroup in the same launch, follows the workgroup shape
 * the autotuner chose for each value precision, so the iterates agree to
 * rounding, and bit for bit when both launches have the same shape
 * (ramd_csr_set_dia_staging).  No reference counterpart. */
int ramd_csr_set_fp32_outer(ramd_csr A, int on);

/* ------------------------------------------------------------------ */
/* SpMV family (FP64; each row summed left to right in stored order, as  */
/* util/linalg/spmv.cpp:52-163 does, so results are bit-identical)       */
/* ------------------------------------------------------------------ */
/* y = A x.        ParCSRMatrix::mult, util/linalg/par_spmv.cpp:25-59,292-295 */
int ramd_spmv(ramd_csr A, const double* x_d, double* y_d);
/* y += A x.       ParMatrix::mult_append, util/linalg/par_spmv.cpp:91-125 */
int ramd_spmv_append(ramd_csr A, const double* x_d, double* y_d);
/* r = b - A x.    ParMatrix::residual, util/linalg/par_spmv.cpp:215-252 */
int ramd_residual(ramd_csr A, const double* x_d, const double* b_d, double* r_d);
/* y = A^T x (y has n_cols entries).  ParMatrix::mult_T,
 * util/linalg/par_spmv.cpp:157-187.  Computed as a gather over a device
 * transpose built on first use, summing in ascending row order like the
 * reference's scatter. */
int ramd_spmv_t(ramd_csr A, const double* x_d, double* y_d);
/* `sweeps` sweeps of weighted Jacobi, x <- (1-w) x + w (b - sum_{j!=i} a_ij x_j)/a_ii.
 * The formula of util/linalg/relax.cpp:10-43, which the reference's
 * parallel jacobi (util/linalg/par_relax.cpp:196-220) intends; tmp_d is n
 * doubles of workspace, as in the reference's signature. */
int ramd_jacobi(ramd_csr A, double* x_d, const double* b_d, double* tmp_d,
                int sweeps, double omega);

/* ------------------------------------------------------------------ */
/* vectors (core/par_vector.cpp)                                         */
/* ------------------------------------------------------------------ */
/* ParVector::inner_product, core/par_vector.cpp:101-121 (deterministic:
 * fixed partition and fixed combination order). */
int ramd_dot(ramd_ctx ctx, int64_t n, const double* x_d, const double* y_d, double* out);
/* ParVector::norm(2), core/par_vector.cpp:88-98 (entries |v|<=1e-16 skipped). */
int ramd_norm2(ramd_ctx ctx, int64_t n, const double* x_d, double* out);
/* y += alpha x.   ParVector::axpy, core/par_vector.cpp:21-27 */
int ramd_axpy(ramd_ctx ctx, int64_t n, double alpha, const double* x_d, double* y_d);

/* ------------------------------------------------------------------ */
/* kernel timing (replaces the reference's track_times buckets,          */
/* multilevel/par_multilevel.hpp:133-211, with device-side events)       */
/* ------------------------------------------------------------------ */
/* From now on, every row-block kernel (SpMV, append, residual, Jacobi) on a
 * matrix with at least min_nnz stored entries is bracketed by hipEvents on
 * the context's stream; min_nnz < 0 turns timing off.  Resets the totals. */
int ramd_prof_enable(ramd_ctx ctx, int64_t min_nnz);
/* Totals for op (0 SpMV, 1 append, 2 residual, 3 Jacobi, 4 x+Pe into a second
 * vector) and precision (RAMD_PREC_FP64 / RAMD_PREC_FP32): launches, of which
 * in the stencil-diagonal form, summed kernel milliseconds and summed
 * algorithmic bytes (matrix once, x once per column, vectors once).
 * Synchronises the stream. */
int ramd_prof_read(ramd_ctx ctx, int op, int prec, int64_t* launches, int64_t* dia_launches, double* ms,
                   double* bytes);

/* ------------------------------------------------------------------ */
/* Galerkin product (two-phase SpGEMM on the device)                     */
/* ------------------------------------------------------------------ */
/* C = A B.  ParCSRMatrix::mult(ParCSRMatrix*), util/linalg/par_matmult.cpp:
 * 73-107 -> spgemm_helper util/linalg/matmult.cpp:71-137 + add_append
 * util/linalg/add.cpp:39-78.  Entries with |c_ij| <= 1e-16 are dropped,
 * every entry is accumulated in the reference's product order, and the
 * rows come out sorted by column. */
int ramd_spgemm(ramd_csr A, ramd_csr B, ramd_csr* C);
/* C = P^T B.  ParCSRMatrix::mult_T(ParCSRMatrix*),
 * util/linalg/par_matmult.cpp:141-196,387-570 -> spgemm_T_helper
 * util/linalg/matmult.cpp:139-204.  Same drop rule and row order. */
int ramd_spgemm_tn(ramd_csr P, ramd_csr B, ramd_csr* C);
/* Ac = P^T (A P) with the diagonal moved to the front of each row: the two
 * calls at ruge_stuben/par_ruge_stuben_solver.hpp:139-143 (and
 * aggregation/par_smoothed_aggregation_solver.hpp:99-101). */
int ramd_galerkin(ramd_csr A, ramd_csr P, ramd_csr* Ac);

/* ------------------------------------------------------------------ */
/* host hierarchy setup (bit-identical to the reference's setup)         */
/* ------------------------------------------------------------------ */
/* coarsen_t, interp_t, strength_t, relax_t, agg_t of core/types.hpp:158-164 */
enum { RAMD_RS = 0, RAMD_CLJP = 1, RAMD_FALGOUT = 2, RAMD_PMIS = 3, RAMD_HMIS = 4 };
enum { RAMD_DIRECT = 0, RAMD_MOD_CLASSICAL = 1, RAMD_EXTENDED = 2 };
enum { RAMD_CLASSICAL = 0, RAMD_SYMMETRIC = 1 };
/* relax_t (Jacobi, SOR, SSOR) plus this build's multicolour forms: the
 * device cycle smooths with RAMD_JACOBI (the intended formula of
 * util/linalg/relax.cpp:10-43) or with RAMD_MC_SOR / RAMD_MC_SSOR, sweeps
 * over a greedy colouring of A + A^T with the reference's SOR_forward /
 * SOR_backward row updates (util/linalg/par_relax.cpp:29-104).  The
 * reference's sequential SOR/SSOR (row order 0..n-1) is host-only there
 * and not offered on the device: the multicolour order gives different
 * results (a labelled non-reference smoother). */
enum { RAMD_JACOBI = 0, RAMD_SOR = 1, RAMD_SSOR = 2, RAMD_MC_SOR = 3, RAMD_MC_SSOR = 4 };
enum { RAMD_PREC_FP64 = 0, RAMD_PREC_FP32 = 1 };
enum { RAMD_SOLVER_RS = 0, RAMD_SOLVER_SA = 1 };

/* Options of ParMultilevel / ParRugeStubenSolver /
 * ParSmoothedAggregationSolver (multilevel/par_multilevel.hpp:79-100,
 * ruge_stuben/par_ruge_stuben_solver.hpp:15-25,
 * aggregation/par_smoothed_aggregation_solver.hpp:17-32). */
typedef struct {
    int solver;             /* RAMD_SOLVER_RS or RAMD_SOLVER_SA                 */
    double strong_threshold;/* RS default 0.25, SA 0.0                          */
    int coarsen_type;       /* RS only, RAMD_RS..RAMD_HMIS                      */
    int interp_type;        /* RS only, RAMD_DIRECT..RAMD_EXTENDED              */
    int strength_type;      /* RAMD_CLASSICAL or RAMD_SYMMETRIC                 */
    double interp_filter;   /* extended interpolation filter (default 0.3)      */
    double prolong_weight;  /* SA Jacobi prolongation weight (default 4/3)      */
    int num_smooth_sweeps;  /* default 1                                        */
    double relax_weight;    /* default 1.0 (configs use 2/3)                    */
    int max_coarse;         /* default 50                                       */
    int max_levels;         /* default 25                                       */
    const double* weights;  /* optional CLJP/PMIS/MIS-2 random weights (n
                               entries); NULL = the reference's
                               srand(2448422)+rand() stream
                               (multilevel/par_multilevel.hpp:215-225)          */
    int prolong_smooth_steps;/* SA Jacobi prolongation steps (default 1)       */
    int relax_type;         /* RAMD_JACOBI (default), RAMD_MC_SOR, RAMD_MC_SSOR */
    int device_setup;       /* 1 (default): the setup steps run on the
                               device (strength, splitting, interpolation;
                               SA: strength, MIS-2, aggregation and the
                               prolongation smoothing; same results);
                               0: on the host                             */
} ramd_amg_opts;

/* Fills opts with the reference's defaults for `solver`. */
int ramd_amg_opts_default(int solver, ramd_amg_opts* opts);

/* ParMultilevel::setup (multilevel/par_multilevel.hpp:126-212): builds
 * strength, splitting/aggregation and interpolation on the host exactly as
 * the reference does, forms every Galerkin product on the device, factors
 * the coarsest matrix and keeps every level device-resident.  A is the
 * caller's fine matrix (host arrays, copied). */
int ramd_hier_setup(ramd_ctx ctx, int32_t n, int64_t nnz, const int32_t* rowptr,
                    const int32_t* col, const double* val, const ramd_amg_opts* opts,
                    ramd_hier* out);
int ramd_hier_destroy(ramd_hier h);
/* num_levels and per-level sizes (arrays of length >= num_levels, may be NULL). */
int ramd_hier_info(ramd_hier h, int* num_levels, int32_t* n_rows, int64_t* nnz_A,
                   int64_t* nnz_P);
/* Borrowed handles of level l's A (diagonal first) and P (NULL at the
 * coarsest level); valid until ramd_hier_destroy. */
int ramd_hier_level(ramd_hier h, int level, ramd_csr* A, ramd_csr* P);
/* Level l's restriction R = P_l^T as the cycle stores it (the gather form
 * of P->mult_T, util/linalg/par_spmv.cpp:157-187), NULL on the coarsest
 * level; a borrowed handle, for reading and kernel overrides. */
int ramd_hier_restriction(ramd_hier h, int level, ramd_csr* R);
/* C/F states (1 = C, 0 = F, -2 = no neighbours) of level l, or for SA the
 * aggregate of each row (its root row, -1 for none). */
int ramd_hier_states(ramd_hier h, int level, int32_t* states);
/* Row-block caps (products per workgroup) the autotuner chose for the
 * CSR kernels: six entries per level, [A, P, R] x [FP64, FP32 vectors];
 * -1 where the sliced-row kernel (64 rows per slice, one lane per row) was
 * measured faster, -2 the sliced-row kernel with 16-bit column offsets,
 * 1 where the lane-per-row kernel over the CSR arrays was, 2 where the
 * wavefront-per-row kernel was, minus the cap (-2048, -4096, -8192) where
 * the row-block kernel in sorted-gather order was,
 * 0 for a stencil-diagonal operator or an absent one.  No
 * reference counterpart. */
int ramd_hier_row_caps(ramd_hier h, int32_t* caps);
/* Host-seconds spent per setup phase: [strength, coarsen, interp, galerkin, other]. */
int ramd_hier_setup_times(ramd_hier h, double* times5);
/* Per-level timing buckets in the reference's layout: times[5 * l + 0] is
 * level l's total (ParMultilevel::setup_times / solve_times with
 * track_times = true, multilevel/par_multilevel.hpp:133-211,343-456; the
 * collective / p2p / vector / matrix communication buckets 1-4 are 0 at one
 * rank).  phase 0: host seconds of each level's setup step (level l's
 * extend_hierarchy, the coarsest level's dense factorisation); phase 1:
 * device seconds of the V-cycle work done on each level since tracking was
 * switched on.  times holds 5 * num_levels entries. */
int ramd_hier_level_times(ramd_hier h, int phase, double* times);
/* Switches the per-level solve timing on (clearing the buckets) or off.
 * Replaces setting ParMultilevel::track_times before a solve. */
int ramd_hier_track_times(ramd_hier h, int on);
/* Size of the coarsest (dense) system and dgetrf_'s info for it (> 0: the
 * info-th pivot is exactly zero).  As in the reference
 * (multilevel/par_multilevel.hpp:330), a zero pivot does not fail the setup;
 * a cycle then divides by it and the solvers report RAMD_E_NUMERIC. */
int ramd_hier_coarse_info(ramd_hier h, int32_t* n_coarse, int32_t* info);
/* Writes the hierarchy (options, every level's A in its stored order, P,
 * the splitting or aggregates, the coarse LU factors and pivots) to a file,
 * and rebuilds a device hierarchy from one without redoing the setup; the
 * loaded hierarchy cycles bit-identically.  The file is this library's own
 * (native byte order); the reference keeps its hierarchy in memory only. */
int ramd_hier_save(ramd_hier h, const char* path);
int ramd_hier_load(ramd_ctx ctx, const char* path, ramd_hier* out);
/* Builds the FP32 copies of every level used by RAMD_PREC_FP32 cycles. */
int ramd_hier_enable_fp32(ramd_hier h);
/* Coarse tail: the cycle of the small coarsest levels (up to 2048 rows and
 * 16384 entries of A and P each, coarse system of at most 64 rows), with
 * the restriction into them and the interpolation out, as one workgroup in
 * one launch instead of one launch per operation.  Same bits either way.
 * level -1: off; 0 (default): automatic -- the first cycle in each
 * precision times a V-cycle without a tail and with every level that can
 * start one, and keeps the fastest; >= 1: the tail starts at that level
 * (RAMD_E_ARG if it cannot).  No reference counterpart
 * (ParMultilevel::cycle, multilevel/par_multilevel.hpp:341-457, is one
 * routine per level); a tuning switch. */
int ramd_hier_set_tail(ramd_hier h, int32_t level);
/* The tail's first level for the next cycle in precision prec (-1 none, -2
 * automatic and not timed yet), and the smallest level that can start one
 * (-1 none). */
int ramd_hier_tail_level(ramd_hier h, int32_t prec, int32_t* level, int32_t* first);
/* on != 0: ramd_fpcg and ramd_adjoint replay each iteration as two
 * captured HIP graphs (operator step, preconditioner step; kept per
 * hierarchy for up to four operator / precision combinations and rebuilt
 * when a level's buffers or kernel choices change; the solution vector's
 * address is read from the device, so any x replays the same capture);
 * 0 (the default): every kernel is launched eagerly, which measured faster
 * on MI355X (the host sets up the next launches while the GPU runs, where
 * a graph launch is prepared before its first kernel starts).  Same results
 * either way.  While kernel timing (ramd_prof_enable) or per-level tracking
 * (ramd_hier_track_times) is on, iterations run eagerly.  No reference
 * counterpart. */
int ramd_set_graphs(ramd_ctx ctx, int on);

/* ------------------------------------------------------------------ */
/* solve path                                                          */
/* ------------------------------------------------------------------ */
/* ParMultilevel::cycle(x, b, 0), multilevel/par_multilevel.hpp:341-457:
 * one V(nu,nu) cycle on the device.  x_is_zero != 0 promises x == 0 on
 * entry (as in every PCG call), which lets the first pre-smoothing sweep
 * skip its matrix read without changing a bit of the result. */
int ramd_vcycle(ramd_hier h, double* x_d, const double* b_d, int x_is_zero);
/* The same cycle from x = 0 on FP32 copies of every level (values and
 * vectors), the coarsest solve kept in FP64: the preconditioner of
 * ramd_fpcg(..., RAMD_PREC_FP32).  No reference counterpart (the reference
 * is FP64 only). */
int ramd_vcycle_fp32(ramd_hier h, float* x_d, const float* b_d);

/* ParMultilevel::solve, multilevel/par_multilevel.hpp:459-538: stationary
 * V-cycles until ||b-Ax||/||b|| <= tol or max_iter cycles; residuals has
 * max_iter+1 entries (may be NULL).  Returns the iteration count in *iters. */
int ramd_amg_solve(ramd_hier h, double* x_d, const double* b_d, double tol,
                   int max_iter, double* residuals, int32_t* iters);

/* PCG(A, ml, x, b, res, tol, max_iter), krylov/par_cg.cpp:121-263, with the
 * reference's rules: an extra V-cycle on b sets the scale (tol *= sqrt(b.z0)),
 * res[0] = sqrt(r0.z0) and later res[k] = (r_k.z_k)/(b.z0), stop when
 * r.z < tol, true residual and p = z restart every 8th iteration.  A is the
 * caller's matrix (the one PCG multiplies by; typically the sorted fine
 * matrix).  res must hold max_res entries; *n_res gets the count written.
 * Returns RAMD_E_INDEFINITE where the reference calls exit(-1). */
int ramd_pcg(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double tol,
             int max_iter, double* res, int32_t max_res, int32_t* n_res, int32_t* iters);
/* The same PCG with the reference's two timing outputs (PCG's precond_t and
 * comm_t, krylov/par_cg.hpp:14-16, krylov/par_cg.cpp:150-228): when not
 * NULL, *precond_t gets the device seconds of the V-cycles added and
 * *comm_t those of the inner products (the reference's Allreduce time). */
int ramd_pcg_timed(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double tol,
                   int max_iter, double* res, int32_t max_res, int32_t* n_res, int32_t* iters,
                   double* precond_t, double* comm_t);

/* CG(A, x, b, res, tol, max_iter), krylov/par_cg.cpp:8-118: unpreconditioned
 * CG with the reference's rules (tol *= ||r0||, res[k] = ||r_k|| / ||b||,
 * true residual every 8th iteration).  res must hold max_res entries. */
int ramd_cg(ramd_csr A, double* x_d, const double* b_d, double tol, int max_iter, double* res,
            int32_t max_res, int32_t* n_res, int32_t* iters);

/* BiCGStab(A, x, b, res, tol, max_iter), krylov/par_bicgstab.cpp:11-110:
 * res[k] = ||r_k||_2 (ParVector::norm) from k = 0, stop when ||r|| <=
 * tol * ||r_0|| or after max_iter iterations (max_iter <= 0: 1.3 n + 2).
 * Vector updates bit-identical to the reference's passes; inner products
 * are deterministic tree sums.  RAMD_E_NUMERIC when a breakdown makes the
 * residual norm non-finite (the reference then stops silently). */
int ramd_bicgstab(ramd_csr A, double* x_d, const double* b_d, double tol, int32_t max_iter, double* res,
                  int32_t max_res, int32_t* n_res, int32_t* iters);
/* Pre_BiCGStab(A, x, b, ml, res, tol, max_iter), krylov/par_bicgstab.cpp:
 * 240-360: the same with p_hat = M^-1 p and s_hat = M^-1 s, M one FP64
 * V-cycle of h from zero.  The reference's routine reads an uninitialised
 * iteration counter and unsized p_hat / s_hat; this is the algorithm it
 * spells out (counter from 0, vectors of n entries). */
int ramd_pre_bicgstab(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double tol, int32_t max_iter,
                      double* res, int32_t max_res, int32_t* n_res, int32_t* iters);

/* Mixed-precision solve: flexible PCG in FP64 (true residual b-Ax, inner
 * products and the coarse LU in FP64) preconditioned by one V-cycle in the
 * precision `cycle_prec`, iterated until ||b-Ax||_2/||b||_2 <= rtol (true,
 * FP64) or max_iter.  No reference counterpart (the reference is FP64 only);
 * it is the precision form of PCG above.  relres_hist (may be NULL) receives
 * the true relative residual after each iteration. */
int ramd_fpcg(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double rtol,
              int max_iter, int cycle_prec, double* relres_hist, int32_t* iters,
              double* final_relres);

/* k right-hand sides at once (1 <= k <= 64).  Blocks are n x k row-major
 * ("row-interleaved": entry (i, c) at i*k + c), FP64 unless stated.
 * No reference counterpart (the reference solves one vector at a time). */
/* Lockstep flexible PCG: column c is solved as ramd_fpcg would solve it
 * (own scalars, own stopping test on its true relative residual), all
 * columns sharing each operator and V-cycle launch.  use_graph != 0
 * replays the iteration as two captured HIP graphs.  iters_col and
 * relres_col (k entries, may be NULL) receive each column's iteration count
 * and final relative residual; *iters_total the number of lockstep
 * iterations. */
/* ramd_fpcg that allocates no device memory: the same solve with the same
 * kernels and arguments (eager launches, so bit-identical results to
 * ramd_fpcg on the same hierarchy and operator), with its partial sums in
 * a fixed per-GPU scratch array of 65536 doubles (512 KiB, held until the
 * process ends) instead of memory it would allocate.  Needs a hierarchy
 * whose FP32 copies exist for an FP32 cycle (ramd_hier_enable_fp32) and,
 * for an operator with the stencil-diagonal form, one earlier ramd_fpcg
 * call with that operator (its FP32 values and kernel choice); refuses
 * (RAMD_E_ARG) an operator needing more than 65536 partial sums (about
 * 16.7M rows) and (RAMD_E_STATE) per-level timing or kernel profiling.
 * Calls that overlap on one GPU share the scratch and race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc). */
int ramd_fpcg_noalloc(ramd_hier h, ramd_csr A, double* x_d, const double* b_d, double rtol,
                      int max_iter, int cycle_prec, double* relres_hist, int32_t* iters,
                      double* final_relres);
int ramd_fpcg_block(ramd_hier h, ramd_csr A, int32_t k, double* X_d, const double* B_d, double rtol,
                    int max_iter, int cycle_prec, int use_graph, int32_t* iters_col, double* relres_col,
                    int32_t* iters_total);
/* One V-cycle from X = 0 for k columns in precision prec (X, B float* for
 * RAMD_PREC_FP32); column c is bit-identical to ramd_vcycle (FP64) or
 * ramd_vcycle_fp32 applied to column c. */
int ramd_vcycle_block(ramd_hier h, int32_t k, int prec, void* X_d, const void* B_d);
/* Kernel choice of the multi-vector autotuner for k columns and vector
 * precision prec: three entries per level (A, P, R; 0 where absent).  A
 * positive entry is the lanes per row of the one-column-per-lane-step kernel,
 * a negative one -(lanes per row) of the kernel reading 16-byte units of
 * consecutive columns (k a multiple of 16 / sizeof(element)).  No reference
 * counterpart. */
int ramd_block_tuning(ramd_hier h, int32_t k, int32_t prec, int32_t* codes);
/* Sets those kernel choices (same encoding; entries of the coarsest level
 * are ignored), e.g. to compare kernels; every column stays bit-identical
 * whichever kernel runs.  No reference counterpart. */
int ramd_block_set_tuning(ramd_hier h, int32_t k, int32_t prec, const int32_t* codes);

/* Adjoint of x = A^-1 b for a symmetric positive definite A (no reference
 * counterpart: the reference has no differentiation).  Given the solution
 * x and g = dL/dx, solves A^T lambda = g by the same flexible PCG and
 * hierarchy (A^T = A), then writes dL/db = lambda into db (n) and, when dA
 * is not NULL, dL/dA_k = -lambda_row(k) * x_col(k) for every stored entry k
 * of A in A's stored order (nnz).  All arrays are device arrays. */
int ramd_adjoint(ramd_hier h, ramd_csr A, const double* x_d, const double* g_d, double rtol,
                 int max_iter, int cycle_prec, double* dA_d, double* db_d, int32_t* iters,
                 double* final_relres);

/* ------------------------------------------------------------------ */
/* host setup pieces, exposed for parity tests                          */
/* ------------------------------------------------------------------ */
/* Results are malloc'd by the library and released with ramd_host_free. */
void ramd_host_free(void* p);
/* strength(Classical|Symmetric, theta): par_strength.cpp:8-547. */
int ramd_host_strength(int32_t n, const int32_t* rowptr, const int32_t* col,
                       const double* val, int strength_type, double theta,
                       int32_t** s_rowptr, int32_t** s_col, int64_t* s_nnz);
/* C/F splitting of S (split_* ruge_stuben/par_cf_splitting.cpp:7-112) with
 * optional weights (NULL = the reference's rand() stream). */
int ramd_host_split(int32_t n, const int32_t* s_rowptr, const int32_t* s_col,
                    int coarsen_type, const double* weights, int32_t* states);
/* Serial Ruge-Stuben splitting split_rs(CSRMatrix*, states, false, second_pass),
 * ruge_stuben/cf_splitting.cpp:277-322 (the building block of RS, Falgout
 * and HMIS coarsening). */
int ramd_host_split_rs_serial(int32_t n, const int32_t* s_rowptr, const int32_t* s_col,
                              int second_pass, int32_t* states);
/* Interpolation P (ruge_stuben/par_interpolation.cpp): direct :1453-1754,
 * mod_classical :991-1451, extended :280-989 (+ filter_interp :175-278). */
int ramd_host_interp(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                     const int32_t* s_rowptr, const int32_t* s_col, const int32_t* states,
                     int interp_type, double filter, int32_t* n_coarse,
                     int32_t** p_rowptr, int32_t** p_col, double** p_val, int64_t* p_nnz);
/* Stencil-grid generators (gallery/par_stencil.hpp:16-236 with
 * gallery/laplacian27pt.hpp:27-39 / gallery/diffusion.hpp:58-85); kind 0 =
 * 27-point Laplacian on m^3, kind 1 = rotated anisotropic 9-point diffusion
 * (eps, theta) on m^2, kind 2 = 7-point finite-volume variable-coefficient
 * diffusion on m^3 with log-normal cell coefficients (seed). */
int ramd_host_gallery(int kind, int32_t m, double eps, double theta, uint64_t seed,
                      int32_t* n, int32_t** rowptr, int32_t** col, double** val, int64_t* nnz);
/* readMatrix(filename), gallery/matrix_IO.cpp:24-110: a PETSc binary matrix
 * file (big-endian as PETSc writes it; a little-endian file is detected by
 * its class id).  Arrays are malloc'd; release them with ramd_host_free. */
int ramd_host_read_petsc(const char* path, int32_t* n_rows, int32_t* n_cols, int32_t** rowptr,
                         int32_t** col, double** val, int64_t* nnz);
/* The matching writer (the reference has none): big-endian PETSc binary. */
int ramd_host_write_petsc(const char* path, int32_t n_rows, int32_t n_cols, const int32_t* rowptr,
                          const int32_t* col, const double* val);
/* Device forms of the setup steps of the 27-point benchmarks' RS path,
 * bit-identical to the host ones (ramd_host_strength / ramd_host_split /
 * ramd_host_interp) on a matrix stored diagonal first:
 * classical strength (par_strength.cpp:8-340) -> S on the device;
 * PMIS (ruge_stuben/par_cf_splitting.cpp:75-89, 1229-1381) on S with the
 * given random weights (NULL: the reference's srand(2448422) stream) ->
 * host states; extended interpolation with filter
 * (ruge_stuben/par_interpolation.cpp:175-989) -> P on the device
 * (RAMD_E_ARG when a row would need more than 256 columns). */
int ramd_strength_gpu(ramd_csr A, double theta, ramd_csr* S);
/* Symmetric strength (par_strength.cpp:342-534) on the device, bit-identical
 * to ramd_host_strength(..., RAMD_SYMMETRIC, ...). */
int ramd_symmetric_strength_gpu(ramd_csr A, double theta, ramd_csr* S);
int ramd_split_pmis_gpu(ramd_csr S, const double* weights, int32_t* states);
/* HMIS (ruge_stuben/par_cf_splitting.cpp:91-112): the Ruge-Stuben first
 * pass on the host (sequential), then the PMIS loop on the device from its
 * states; bit-identical to ramd_host_split(..., RAMD_HMIS, ...). */
int ramd_split_hmis_gpu(ramd_csr S, const double* weights, int32_t* states);
int ramd_interp_extended_gpu(ramd_csr A, ramd_csr S, const int32_t* states, double filter, ramd_csr* P);
/* CLJP (ruge_stuben/par_cf_splitting.cpp:32-48, 1383-1607) and modified
 * classical interpolation (ruge_stuben/par_interpolation.cpp:991-1451) on
 * the device, bit-identical to ramd_host_split(..., RAMD_CLJP, ...) and
 * ramd_host_interp(..., RAMD_MOD_CLASSICAL, ...) (RAMD_E_ARG when a row has
 * more than 256 strong C neighbours). */
int ramd_split_cljp_gpu(ramd_csr S, const double* weights, int32_t* states);
int ramd_interp_mod_classical_gpu(ramd_csr A, ramd_csr S, const int32_t* states, ramd_csr* P);
/* Direct interpolation (ruge_stuben/par_interpolation.cpp:1453-1754) on the
 * device, bit-identical to ramd_host_interp(..., RAMD_DIRECT, ...). */
int ramd_interp_direct_gpu(ramd_csr A, ramd_csr S, const int32_t* states, ramd_csr* P);
/* MIS-2 (aggregation/par_mis.cpp:204-645) and aggregation
 * (aggregation/par_aggregate.cpp:5-185) on the device, bit-identical to
 * ramd_host_mis2 and ramd_host_aggregate; A and S stored sorted with the
 * diagonal first (the hierarchy's levels and the device strength give them
 * so; ramd_aggregate_gpu returns RAMD_E_ARG for other orders).  weights NULL = the srand(2448422) stream; rand_vals NULL = zeros,
 * as the solver calls it.  states, rand_vals and aggregates are host
 * arrays of n. */
int ramd_mis2_gpu(ramd_csr S, const double* weights, int32_t* states);
int ramd_aggregate_gpu(ramd_csr A, ramd_csr S, const int32_t* states, const double* rand_vals, int32_t* aggregates,
                       int32_t* n_aggs);
/* fit_candidates (aggregation/par_candidates.cpp:6-216, one candidate) on
 * the device of ctx, bit-identical to ramd_host_fit_candidates: T (n x
 * n_aggs, a device matrix) and the aggregate norms R (host, n_aggs) from
 * the aggregates and B (host, n). */
int ramd_fit_candidates_gpu(ramd_ctx ctx, int32_t n, const int32_t* aggregates, int32_t n_aggs, const double* B,
                            double* R, ramd_csr* T);

/* Greedy colouring of the graph of A + A^T without the diagonal (rows in
 * ascending order take the smallest colour no coloured neighbour holds),
 * the colouring the multicolour smoothers use.  No reference counterpart. */
int ramd_host_greedy_colors(int32_t n, const int32_t* rowptr, const int32_t* col, int32_t* colors,
                            int32_t* n_colors);
/* mis2(S, states, off_proc_states, false, weights), aggregation/par_mis.cpp:
 * 204-645.  weights NULL = the srand(2448422) stream.  states: 1 selected,
 * 0 not selected. */
int ramd_host_mis2(int32_t n, const int32_t* s_rowptr, const int32_t* s_col, const double* weights,
                   int32_t* states);
/* aggregate(A, S, states, off, aggregates, false, rand_vals),
 * aggregation/par_aggregate.cpp:5-185.  aggregates[i] = root row or -1;
 * rand_vals may be NULL (zeros, as the solver calls it). */
int ramd_host_aggregate(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                        const int32_t* s_rowptr, const int32_t* s_col, const int32_t* states,
                        const double* rand_vals, int32_t* aggregates, int32_t* n_aggs);
/* fit_candidates(A, n_aggs, aggregates, B, R, 1, false, tol),
 * aggregation/par_candidates.cpp:6-216: tentative prolongator T (n x n_aggs)
 * and the aggregate norms R (n_aggs). */
int ramd_host_fit_candidates(int32_t n, const int32_t* aggregates, int32_t n_aggs, const double* B, double* R,
                             int32_t** t_rowptr, int32_t** t_col, double** t_val, int64_t* t_nnz);
/* jacobi_prolongation(A, T, false, omega, steps),
 * aggregation/par_prolongation.cpp:7-103: P = T - (omega D^-1 A) T per step
 * (D = absolute row sums), the sparse products on the device of ctx.  A is
 * taken in its stored order (the solver passes it sorted, diagonal first). */
int ramd_sa_prolongation(ramd_ctx ctx, int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                         int32_t n_aggs, const int32_t* t_rowptr, const int32_t* t_col, const double* t_val,
                         double omega, int steps, int32_t** p_rowptr, int32_t** p_col, double** p_val,
                         int64_t* p_nnz);

#ifdef __cplusplus
}
#endif
#endif /* RAPTOR_AMD_H */
