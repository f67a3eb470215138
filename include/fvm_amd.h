/*
 * fvm_amd -- MI355X (gfx950) algebraic-multigrid linear solver for finite-volume
 * CRMatrix systems. C ABI: plain pointers and sizes, int status codes, no C++
 * types and no exceptions across the boundary.
 *
 * Reference interfaces replaced (paths relative to src/fvm/src/modules/fvmbase/
 * of c-PRIMED/fvm):
 *   class LinearSolver  LinearSolver.h:15-35   (solve / smooth / cleanup, tolerances)
 *   class AMG           AMG.h:27-110, AMG.cpp:15-337 (parameters, cycle, hierarchy)
 *   SWIG AMG binding    AMG.i:1-34
 *   class BCGStab       BCGStab.h / BCGStab.cpp:27-171 (Krylov loop around AMG::smooth)
 *
 * Conventions (same as the reference):
 *   - The system is stored as CSR with the diagonal kept apart: row[n_total+1],
 *     col[nnz] (never contains the row itself), off[nnz], diag[n_total]
 *     (CRMatrix.h:73-84).
 *   - Rows [0, n_self) are unknowns; rows [n_self, n_total) are ghost rows that
 *     the solver never updates (StorageSite self/ghost counts).
 *   - Delta form: the residual is r = b + A x and the solver drives it to zero,
 *     i.e. it solves A x = -b (CRMatrix.h:407-426).
 *   - Convergence uses the 1-norm of r over self rows; solve() returns the
 *     initial norm r0 through *rnorm0 (AMG.cpp:235,281).
 *
 * Every entry point returns FVM_OK (0) on success or a negative FVM_E_* code;
 * fvm_last_error() returns a thread-local message for the last failure.
 * A handle is not thread-safe; it owns one HIP stream and all device memory of
 * its hierarchy until fvm_amg_cleanup() or fvm_amg_destroy(). Host pointers
 * passed in are only read (or written) during the call. The handles' streams are
 * blocking streams: work the caller queued on the legacy default stream (e.g.
 * PyTorch's default stream) finishes before a call reads device arrays, and
 * every call has finished its device work when it returns.
 */
#ifndef FVM_AMD_H
#define FVM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVM_AMD_VERSION 10000 /* 1.0.0 */

enum {
    FVM_OK = 0,
    FVM_E_INVALID = -1,     /* bad argument or unsupported parameter value */
    FVM_E_HIP = -2,         /* a HIP runtime call failed */
    FVM_E_NOT_SETUP = -3,   /* solve/smooth before setup */
    FVM_E_NUMERIC = -4,     /* zero / non-finite diagonal, breakdown */
    FVM_E_NOMEM = -5
};

/* AMG::CycleType, AMG.h:31-36 */
enum { FVM_V_CYCLE = 0, FVM_W_CYCLE = 1, FVM_F_CYCLE = 2 };
/* AMG::SmootherType, AMG.h:38-42 */
enum { FVM_GAUSS_SEIDEL = 0, FVM_JACOBI = 1 };
/* Gauss-Seidel row order (no counterpart in the reference, which is sequential) */
enum {
    FVM_GS_EXACT = 0,      /* reference row order (dependency-ordered dataflow): bit-identical to the reference */
    FVM_GS_MULTICOLOR = 1  /* graph-coloured order: a different (parallel) Gauss-Seidel */
};
/* arithmetic of the V-cycle (no counterpart in the reference, which is FP64 only) */
enum {
    FVM_PREC_FP64 = 0,     /* FP64 values and vectors on every level */
    FVM_PREC_MIXED = 1     /* FP32 vectors, FP32 finest values, BF16 coarse off-diagonals
                              with an FP32 diagonal that keeps each row's FP64 sum;
                              the outer residual stays FP64 (iterative refinement) */
};

/*
 * Parameters. Names, meanings and defaults are those of LinearSolver
 * (LinearSolver.h:19-24) and AMG (AMG.cpp:15-36); the fields after
 * scaleCorrections are extensions.
 */
typedef struct fvm_amg_params {
    int maxCoarseLevels;          /* 30 */
    int nPreSweeps;               /* 0 */
    int nPostSweeps;              /* 1 */
    int coarseGroupSize;          /* 2 (the only value the GPU coarsening accepts) */
    double weightRatioThreshold;  /* 0.65 */
    int cycleType;                /* FVM_V_CYCLE */
    int smootherType;             /* FVM_GAUSS_SEIDEL */
    int scaleCorrections;         /* 1; active only when isSymmetric (AMG.cpp:119) */
    int nMaxIterations;           /* 100: at most nMaxIterations-1 cycles (AMG.cpp:250) */
    int verbosity;                /* 2 in the reference; >0 prints "i: norm" lines */
    double relativeTolerance;     /* 1e-8 */
    double absoluteTolerance;     /* 1e-50 */
    /* extensions */
    int isSymmetric;              /* LinearSystem::isSymmetric (MMReader sets it for symmetric files) */
    int parallelStopRule;         /* 1: keep a coarse level of size <= 3 then stop (FVM_PARALLEL
                                     build, AMG.cpp:170-181); 0: serial build rule (AMG.cpp:199-204) */
    int gsOrdering;               /* FVM_GS_EXACT */
    int precision;                /* FVM_PREC_FP64 */
} fvm_amg_params;

typedef struct fvm_amg fvm_amg;

/* Version as MAJOR*10000 + MINOR*100 + PATCH. */
int fvm_version(void);
/* Thread-local text of the last error ("" if none). */
const char *fvm_last_error(void);

/* Fill *p with the reference defaults. */
int fvm_amg_default_params(fvm_amg_params *p);

/* AMG::AMG() (AMG.cpp:15-36). device = HIP device ordinal. */
int fvm_amg_create(fvm_amg **out, int device);
/* AMG::~AMG() */
int fvm_amg_destroy(fvm_amg *h);
int fvm_amg_set_params(fvm_amg *h, const fvm_amg_params *p);
int fvm_amg_get_params(const fvm_amg *h, fvm_amg_params *p);

/*
 * Build the hierarchy (AMG::createCoarseLevels, AMG.cpp:149-210) from a host CSR
 * matrix. is_boundary may be NULL (no boundary rows). Replaces the first-call
 * branch of AMG::solve / AMG::smooth (AMG.cpp:222-226).
 */
int fvm_amg_setup(fvm_amg *h, int n_self, int n_total, const int32_t *row, const int32_t *col,
                  const double *diag, const double *off, const uint8_t *is_boundary);
/* Same, with all arrays already in device memory of this handle's device. */
int fvm_amg_setup_device(fvm_amg *h, int n_self, int n_total, const int32_t *d_row,
                         const int32_t *d_col, const double *d_diag, const double *d_off,
                         const uint8_t *d_is_boundary);

/*
 * AMG::solve (AMG.cpp:219-282): b and x have n_total entries (host memory);
 * x holds the initial delta and receives the result on self rows (ghost
 * entries are left unchanged). *rnorm0 receives the initial 1-norm (the
 * reference's return value), *rnorm_final the last computed one, *cycles the
 * number of cycles run. Non-convergence is not an error.
 */
int fvm_amg_solve(fvm_amg *h, const double *b, double *x, double *rnorm0, double *rnorm_final,
                  int *cycles);
/* Same with device pointers; runs on the handle's stream and synchronizes it. */
int fvm_amg_solve_device(fvm_amg *h, const double *d_b, double *d_x, double *rnorm0,
                         double *rnorm_final, int *cycles);
/* AMG::smooth (AMG.cpp:285-298): exactly one cycle from the given x. */
int fvm_amg_smooth(fvm_amg *h, const double *b, double *x);
int fvm_amg_smooth_device(fvm_amg *h, const double *d_b, double *d_x);
/* AMG::cleanup (AMG.cpp:212-217): drop the hierarchy. */
int fvm_amg_cleanup(fvm_amg *h);
/* AMG::getTotalIterations (AMG.h:69) */
long long fvm_amg_total_iterations(const fvm_amg *h);
/* Convergence lines printed at verbosity > 0, as the reference prints them
 * (AMG.cpp:237-272, MultiFieldReduction.cpp:199-211, BCGStab.cpp:55-164):
 * "i: [field_name : norm]". path: AMG::redirectPrintToFile (a new path is
 * truncated, the same path keeps appending; NULL = AMG::redirectPrintToScreen,
 * stdout); style 0 prints every
 * cycle (serial build), 1 only the first line and the one where the loop ends
 * (parallel build, as in the reference's convergence goldens). field_name NULL:
 * "x". */
int fvm_amg_set_output(fvm_amg *h, const char *field_name, const char *path, int style);
/* Residual 1-norms of the last solve: hist[0] = r0, hist[i] after cycle i.
 * Returns the number of entries written (<= cap). */
int fvm_amg_history(const fvm_amg *h, double *hist, int cap);

/* Hierarchy introspection (the reference prints level sizes at verbosity > 1,
 * AMG.cpp:176-178). Level 0 is the finest (self rows only, ghost columns removed). */
int fvm_amg_num_levels(const fvm_amg *h);
int fvm_amg_level_info(const fvm_amg *h, int level, int *n_rows, int *nnz);
/* Copy a level to host arrays (any may be NULL): CSR in reference order and the
 * fine-to-coarse map ci[n_rows] (-1 for rows that join no group; all -1 on the
 * coarsest level). */
int fvm_amg_level_get(const fvm_amg *h, int level, int32_t *row, int32_t *col, double *diag,
                      double *off, int32_t *ci);
/* Colours of a level's coloured smoother (no reference counterpart; for tests and
 * diagnostics): colour[i] of every row in reference order (may be NULL; -1 on a
 * level without colouring), *ncolour the number of colours (0 without). Rows of
 * one colour share no matrix entry. */
int fvm_amg_level_colours(const fvm_amg *h, int level, int32_t *colour, int *ncolour);
/* Coarsest dense solve of the mixed-precision cycle (no reference counterpart):
 * the first coarse level with at most 256 rows is inverted in FP64 by Gauss-Jordan
 * with partial pivoting; a singular level keeps the smoother instead (level -1)
 * and is not factored again until the next setup. *level, *n_rows: the inverted
 * level and its rows (-1 / 0 if none); *attempts: factorizations run by this
 * handle; inv (may be NULL): host copy of the n_rows x n_rows inverse (row-major,
 * rows in the level's reference order as returned by fvm_amg_level_get). */
int fvm_amg_direct_level(const fvm_amg *h, int *level, int *n_rows, int *attempts, double *inv);
/* x = -A^{-1} b on that level with the cycle's kernel (host vectors of n_rows). */
int fvm_amg_direct_apply(fvm_amg *h, const double *b, double *x);
/* Live timing of finest-level kernels with HIP events recorded around their
 * launches on the handle's stream (no reference counterpart; for roofline
 * reporting). kind 0: FP64 residual / SpMV launches; kind 1: smoother sweeps
 * (one direction over all colours, or one Jacobi pass). Enabling resets the
 * counters. */
enum { FVM_KT_RESIDUAL64 = 0, FVM_KT_SWEEP = 1 };
int fvm_amg_set_kernel_timing(fvm_amg *h, int enable);
int fvm_amg_kernel_timing(const fvm_amg *h, int kind, long long *launches, double *total_ms);
/* Algorithmic HBM bytes of the same timed launches, summed: every operand read or
 * written once (per row: the diagonal, the vectors, a 4-byte row offset; per
 * off-diagonal entry: a 4-byte column and its value, at the precision the level
 * is stored in). A smoother direction counts the rows of the colours it updates,
 * so a backward direction that skips a colour counts fewer bytes. */
int fvm_amg_kernel_bytes(const fvm_amg *h, int kind, double *total_bytes);
/* Performance knobs that never change results (no reference counterpart; set
 * by the autotuner): "tail_max_rows" -- V-cycle levels with at most this many
 * rows run in one single-workgroup launch (0 disables); "tail_lds" -- 1
 * (default): there, a level whose matrix and vectors fit in the workgroup's LDS
 * is swept from LDS; "sell_block" -- threads
 * per block of the SELL-64 smoother kernels, "spmv_block" -- of the SELL-64
 * residual and SpMV kernels (64/128/256/512 each); "sell_xcd" / "spmv_xcd" --
 * 1: blocks of those kernels are renumbered so that each of the chip's 8 XCDs
 * works on a contiguous eighth of the rows, 0: hardware order; "cycle_graph" --
 * 1 (default): levels >= 1 of the coloured V-cycle are captured once into a HIP
 * graph (first cycle after setup, parameter or tuning change) and replayed by
 * every later cycle, 0: launched one by one; "nt_levels" -- the matrices of
 * this many finest levels are read with non-temporal loads (default 2);
 * "coarsen_lanes" -- lanes per row in the agglomeration rounds (8/16/32/64; 0,
 * the default: chosen per level from its row width); "coarsen_small_rows" --
 * levels with at most this many rows are agglomerated by one workgroup
 * (default 2048);
 * "coarsen_small_lanes" -- lanes per row in those single-workgroup rounds
 * (default 16; 0: as coarsen_lanes); "coarsen_grid" -- blocks per launched
 * agglomeration round (0: 2048 for narrow rows, 1024 otherwise, per 256 threads
 * of a block); "coarsen_block" -- threads per block of those rounds (256, 512 or
 * 1024; default 512: each block appends its part of the next frontier with one
 * atomic); "coarsen_lds_entries" -- that part is staged in LDS, this many
 * entries per thread (0..16, default 8; at most 8192 per block), and what does not
 * fit is appended at once; "bipartite_levels" -- this many finest
 * levels are first coloured by breadth-first parity (red-black order, two
 * colours) and keep it when no edge joins equal colours, else the greedy
 * colouring (default 1; takes effect at the next setup); "gal_lanes" -- lanes
 * per coarse row of the coarse operator (16/32/64; 0, the default: 16 up to 40
 * flattened entries per coarse row on average, more above); "colour_grid" --
 * blocks per colouring round (16..4096, default 256): the colouring runs on a
 * second stream beside the agglomeration rounds, and fewer blocks leave those
 * latency-bound rounds more of the chip; "coarsen_batch" -- launched agglomeration
 * rounds queued between two reads of the frontier size (4..256, default 32). Only
 * the colouring
 * keys change the coloured smoother's arithmetic (another Gauss-Seidel order);
 * the others never change a result; "reuse_setup" -- opt-in: a setup whose
 * matrix arrays and hierarchy parameters have the signature of the hierarchy
 * in place (64-bit sums of mixed elements) keeps it, also across
 * fvm_amg_cleanup, which then keeps the hierarchy for such a setup (default 0:
 * every setup rebuilds, as the reference does after cleanup); "setups_reused"
 * (read only) counts them; "exact_flow" -- reference-order sweeps as
 * one dataflow launch per direction (default 1; 0: level by level);
 * "exact_flow_spin_limit" -- waits per wavefront after which such a sweep gives
 * up and the call fails with FVM_E_NUMERIC (default 0: no bound; a sweep that
 * sees no progress for two seconds gives up regardless);
 * "exact_wg_width" -- in level-by-level reference-order sweeps, each run of
 * consecutive dependency levels of at most this many rows is one
 * single-workgroup launch, every wider level a launch of its own (default 256);
 * "coarsen_pair_filter" -- 1 (default): an agglomeration row does not wait for
 * a seed that can never pair its neighbour, nor for a neighbour whose decision
 * ignores it (its strongest lower neighbour dominates); 0: every dependency of
 * the reference order is waited for; "coarsen_lower_filter" -- 1 (default): a
 * row waits only for its strongest lower neighbours and those that may pair it;
 * "coarsen_wait_lists" -- 2 (default): a blocked join waits in a list of the row
 * it waits for and returns to the frontier when that row decides; 1: only joins
 * waiting for the group they join do; 0: blocked joins retry every round. These
 * three leave the aggregates unchanged. */
int fvm_amg_set_tuning(fvm_amg *h, const char *key, long long value);
int fvm_amg_get_tuning(const fvm_amg *h, const char *key, long long *value);
/* Autotuner (no counterpart in the reference): times every candidate launch
 * shape that cannot change a result -- "sell_block" (threads per block of the
 * coloured SELL-64 smoother kernels: 128/256/512) x "tail_max_rows"
 * (0/1000/2500/1e4/4e4) x "sell_xcd" (0/1) x "cycle_graph" (0/1) -- with one warm-up and `reps` timed
 * cycles each on the handle's current hierarchy and parameters, keeps the
 * fastest, then "nt_levels" (0..3) with those shapes, then times the
 * finest-level residual alone for "spmv_block"
 * (128/256/512) x "spmv_xcd" (0/1). Writes a JSON report of all
 * candidates and the choice into report (cap bytes, may be NULL); returns the
 * report length or an FVM_E_* code. Needs a completed setup. */
int fvm_amg_autotune(fvm_amg *h, int reps, char *report, int cap);
/* Setup statistics: seconds spent in setup, number of coarsening rounds of the finest level. */
int fvm_amg_setup_stats(const fvm_amg *h, double *setup_seconds, int *finest_rounds);

/* ---------------------------------------------------------------------------
 * FP64 flexible GMRES(m) with the AMG handle's cycle as right preconditioner.
 * No reference counterpart for FGMRES itself; it takes the place of the
 * reference's Krylov loops around AMG::smooth (BCGStab.cpp:27-171) and keeps
 * LinearSolver's fields and stop rule (1-norm of the FP64 true residual
 * r = b + A x relative to r0).
 * ------------------------------------------------------------------------- */
typedef struct fvm_fgmres_params {
    int restart;                  /* 30 */
    int nMaxIterations;           /* 1000 (total inner iterations) */
    int verbosity;                /* 0 */
    double relativeTolerance;     /* 1e-8 */
    double absoluteTolerance;     /* 1e-50 */
} fvm_fgmres_params;

typedef struct fvm_fgmres fvm_fgmres;

int fvm_fgmres_default_params(fvm_fgmres_params *p);
/* The preconditioner handle must have been set up; it is borrowed, not owned: do
 * not solve after it is destroyed. Destroying the solver never touches it, so the
 * two may be destroyed in either order. */
int fvm_fgmres_create(fvm_fgmres **out, fvm_amg *preconditioner);
int fvm_fgmres_destroy(fvm_fgmres *s);
int fvm_fgmres_set_params(fvm_fgmres *s, const fvm_fgmres_params *p);
/* Solve A x = -b (delta form) on device pointers of n_self... n_total entries. */
int fvm_fgmres_solve_device(fvm_fgmres *s, const double *d_b, double *d_x, double *rnorm0,
                            double *rnorm_final, int *iterations);
int fvm_fgmres_solve(fvm_fgmres *s, const double *b, double *x, double *rnorm0,
                     double *rnorm_final, int *iterations);
/* True-residual 1-norms of the last solve: hist[0] = r0, then one per restart
 * cycle. Returns the number of entries written (<= cap). */
int fvm_fgmres_history(const fvm_fgmres *s, double *hist, int cap);

/* ---------------------------------------------------------------------------
 * The reference's Krylov loops with the AMG handle's cycle as preconditioner,
 * in FP64 on the device; AMG::smooth (one cycle from a zero delta) is the
 * preconditioner application:
 *   fvm_bcgstab_solve_device  BCGStab::solve (BCGStab.cpp:27-171)
 *   fvm_cg_solve_device       CG::solve (CG.cpp:25-137)
 * Fields and defaults are LinearSolver's (LinearSolver.h:19-24; verbosity 0
 * here). b, x: device vectors of n_total entries, delta form (r = b + A x);
 * x holds the initial delta and receives the solution on the n_self rows.
 * hist (may be NULL): r0, then the 1-norm of r after every completed iteration,
 * as the reference prints them; returns the number of entries written (<= cap)
 * or an FVM_E_* code. rnorm0 = the initial 1-norm (the reference's return
 * value); rnorm_final = the 1-norm of the recomputed true residual.
 * ------------------------------------------------------------------------- */
typedef struct fvm_linsolver_params {
    int nMaxIterations;           /* 100 */
    int verbosity;                /* 0 */
    double relativeTolerance;     /* 1e-8 */
    double absoluteTolerance;     /* 1e-50 */
} fvm_linsolver_params;

int fvm_linsolver_default_params(fvm_linsolver_params *p);
int fvm_bcgstab_solve_device(fvm_amg *pc, const fvm_linsolver_params *p, const double *d_b, double *d_x,
                             double *rnorm0, double *rnorm_final, int *iterations, double *hist, int cap);
int fvm_cg_solve_device(fvm_amg *pc, const fvm_linsolver_params *p, const double *d_b, double *d_x,
                        double *rnorm0, double *rnorm_final, int *iterations, double *hist, int cap);


/* ---------------------------------------------------------------------------
 * Batched solve of many independent scalar systems (the per-direction systems
 * of PhononModel::advance, PhononModel.h:530-600, which calls AMG::solve once
 * per direction). One handle builds every system's hierarchy at once (a
 * block-diagonal hierarchy; each system keeps its own maps, levels and stop
 * rule, AMG.cpp:149-210) and cycles all systems together, each with its own
 * stop test (AMG.cpp:219-282): per system the result equals that system's own
 * fvm_amg solve with the same parameters. One cycle of all systems is captured
 * into a HIP graph and replayed. Supported: both Gauss-Seidel orderings
 * (the coloured one colours each system as it is coloured alone), precision
 * FVM_PREC_FP64, no scaled corrections.
 * All array arguments are host arrays of nsys entries; the pointers they hold
 * are device pointers (CSR as in fvm_amg_setup_device; is_boundary may be NULL,
 * or hold NULL entries). rnorm0/rnorm_final/cycles: host arrays of nsys.
 * ------------------------------------------------------------------------- */
typedef struct fvm_batch fvm_batch;
int fvm_batch_create(fvm_batch **out, int device);
int fvm_batch_destroy(fvm_batch *b);
int fvm_batch_set_params(fvm_batch *b, const fvm_amg_params *p);
int fvm_batch_setup_device(fvm_batch *b, int nsys, const int32_t *n_self, const int32_t *n_total,
                           const int32_t *const *d_row, const int32_t *const *d_col,
                           const double *const *d_diag, const double *const *d_off,
                           const uint8_t *const *d_is_boundary);
int fvm_batch_solve_device(fvm_batch *b, const double *const *d_b, double *const *d_x,
                           double *rnorm0, double *rnorm_final, int *cycles);
/* rows of each level of system sys; returns that system's number of levels */
int fvm_batch_level_sizes(const fvm_batch *b, int sys, int32_t *sizes, int cap);
/* "graph" (1: replay the captured cycle, default; 0: plain launches) and the
 * fvm_amg_set_tuning keys */
int fvm_batch_set_tuning(fvm_batch *b, const char *key, long long value);
int fvm_batch_setup_stats(const fvm_batch *b, double *setup_seconds, int *total_rows, int *levels);


/* ---------------------------------------------------------------------------
 * Learned correction (no counterpart in the reference): a 3-layer message-passing
 * network over the matrix graph (hidden width 64, bf16 weights, FP32
 * accumulation on the matrix cores) that maps a residual r to a correction e
 * (delta form, A e ~ -r). Edge features are the off-diagonals scaled by the
 * row's diagonal; node features the Jacobi correction -r/d (scaled by its RMS),
 * its neighbour average, and the normalized log|d|. Trained on the GPU with the
 * library's own forward and backward kernels on the residual loss
 * ||r + A e||^2/||r||^2 or, for residual/error pairs, the energy-norm error
 * (A symmetric). Attached to an FGMRES handle it corrects every preconditioner
 * application z = M(v) by z += NN(A z - v), and is switched off (plain AMG) for
 * the rest of a solve when the residual estimate stalls or the output is not
 * finite.
 * Parameter vector layout (FP32, fvm_mpnn_num_params() entries): per layer
 * k = 0..2: Ws_k (64 x fin_k, row-major), Wn_k (64 x fin_k), b_k (64), fin_0 = 16,
 * fin_1 = fin_2 = 64; then wo (64), bo (1).
 * ------------------------------------------------------------------------- */
typedef struct fvm_mpnn fvm_mpnn;
int fvm_mpnn_num_params(void);
int fvm_mpnn_create(fvm_mpnn **out, int device, uint64_t seed);
int fvm_mpnn_destroy(fvm_mpnn *m);
int fvm_mpnn_set_params(fvm_mpnn *m, const float *theta);     /* host, num_params entries */
int fvm_mpnn_get_params(const fvm_mpnn *m, float *theta);
/* graph of a system: device CSR of its self rows (off-diagonals only, diagonal
 * apart), structurally symmetric; the arrays are borrowed until the next call */
int fvm_mpnn_set_graph(fvm_mpnn *m, int n, const int32_t *d_row, const int32_t *d_col,
                       const double *d_off, const double *d_diag);
/* d_e = NN(d_r) (accumulate: d_e += NN(d_r)); device vectors of n entries */
int fvm_mpnn_forward(fvm_mpnn *m, const double *d_r, double *d_e, int accumulate);
/* forward + backward on one residual: adds the gradient of the loss to the
 * handle's gradient buffer and returns the loss */
int fvm_mpnn_loss_grad(fvm_mpnn *m, const double *d_r, double *loss);
/* the same for a residual/error pair (A e = -r, device): energy loss
 * (NN(r) - e)^T (-A) (NN(r) - e) / e^T (-A) e, -A positive definite */
int fvm_mpnn_loss_grad_pair(fvm_mpnn *m, const double *d_r, const double *d_e, double *loss);
/* host copy of the accumulated gradient (may be NULL); reset != 0 clears it */
int fvm_mpnn_get_grad(fvm_mpnn *m, float *grad, int reset);
/* one Adam step (beta 0.9/0.999) with the mean of the accumulated gradients */
int fvm_mpnn_adam_step(fvm_mpnn *m, float lr);
/* attach (m != NULL) or detach the network to an FGMRES handle; it is switched off
 * for the rest of a solve when the 2-norm residual estimate falls by less than
 * the factor stall_ratio over 5 iterations */
int fvm_fgmres_set_learned(fvm_fgmres *s, fvm_mpnn *m, double stall_ratio);
/* iterations of the last solve that applied the network, and the iteration at
 * which it was switched off (-1: never) */
int fvm_fgmres_learned_stats(const fvm_fgmres *s, int *iterations_used, int *fallback_at);

/* ---------------------------------------------------------------------------
 * Finite-volume assembly of the thermal (diffusion) equation on the device.
 * Reference: ThermalModel linearization (ThermalModel_impl.h:236-398) with
 * MeshMetricsCalculator (MeshMetricsCalculator_impl.h:59-462),
 * DiffusionDiscretization (DiffusionDiscretization.h:65-232),
 * SourceDiscretization (SourceDiscretization.h:40-60), GenericBCS Dirichlet /
 * Neumann (GenericBCS.h:77-157) and LinearSystem::initSolve's boundary-row
 * elimination (LinearSystem.cpp:40-64, CRMatrix.h:899-1085). The assembled
 * system is the CRMatrix the AMG consumes: interior cells first, one ghost row
 * per boundary face after them, columns in CRConnectivity::multiply order.
 * The non-orthogonal secondary flux uses the least-squares gradient of the
 * current field (GradientModel.h:128-520), as ThermalModel::linearize does.
 * Boundary groups are walls (no symmetry or immersed-boundary faces); the
 * gradient needs a conformal mesh (two cells share at most one face).
 * ------------------------------------------------------------------------- */
enum { FVM_BC_DIRICHLET = 0, FVM_BC_NEUMANN = 1 };

typedef struct fvm_mesh fvm_mesh;
typedef struct fvm_system fvm_system;

/* Jittered n^3 hexahedral mesh of the unit cube built on the device: interior
 * nodes moved by U[-jitter h, jitter h] per coordinate (walls stay planar).
 * Faces: interior x-, y-, z-faces (lexicographic), then the boundary groups
 * x=0, x=1, y=0, y=1, z=0, z=1 (group ids 1..6). Synthetic input: stands in
 * for a mesh read by the reference's importers (FluentReader). */
int fvm_mesh_create_hex(fvm_mesh **out, int device, int n, double jitter, uint64_t seed);
/* Quad-face mesh from host arrays: face_nodes[4 f], face_cells[2 f] = (c0, c1),
 * groups [group_ptr[g], group_ptr[g+1]) with group 0 the interior faces and
 * ghost cell of boundary face f = n_cells + (f - group_ptr[1]). */
int fvm_mesh_create(fvm_mesh **out, int device, int n_nodes, const double *nodes, int n_faces,
                    const int32_t *face_nodes, const int32_t *face_cells, int n_cells, int n_ghost,
                    int n_groups, const int32_t *group_ptr);
/* General unstructured mesh (the meshes the reference's importers build,
 * FluentReader.cpp:696-950): dim 2 (faces of 2 nodes, z = 0) or 3 (faces of 3..64
 * nodes), faces as a node CSR face_nptr[n_faces + 1] / face_nodes, the same
 * face_cells, ghost and group conventions as fvm_mesh_create. */
int fvm_mesh_create_general(fvm_mesh **out, int device, int dim, int n_nodes, const double *nodes, int n_faces,
                            const int32_t *face_nptr, const int32_t *face_nodes, const int32_t *face_cells,
                            int n_cells, int n_ghost, int n_groups, const int32_t *group_ptr);
int fvm_mesh_destroy(fvm_mesh *m);
/* dimension, total face-node count and (may be NULL) a host copy of face_nptr */
int fvm_mesh_topology(const fvm_mesh *m, int *dim, long long *n_face_nodes, int32_t *face_nptr);
/* GradientModel::compute (GradientModel.h:472-520): least-squares gradient of the
 * device cell field d_x (n_cells + n_ghost) into d_grad (3 per cell, ghosts take
 * their face's cell's gradient) */
int fvm_mesh_gradient(fvm_mesh *m, const double *d_x, double *d_grad);
int fvm_mesh_info(const fvm_mesh *m, int *n_cells, int *n_ghost, int *n_faces, int *n_nodes,
                  int *n_groups);
/* host copies of the mesh arrays (any may be NULL); face_nodes holds the
 * n_face_nodes entries of fvm_mesh_topology */
int fvm_mesh_get(const fvm_mesh *m, double *nodes, int32_t *face_nodes, int32_t *face_cells,
                 int32_t *group_ptr);
/* MeshMetricsCalculator: face area vectors, |A|, face centroids, cell centroids
 * and volumes (ghost cells included); host copies, any may be NULL. */
int fvm_mesh_metrics(fvm_mesh *m, double *fA, double *fAm, double *fc, double *cc, double *vol);
/* Cell fields on the device (n_cells + n_ghost entries, ghosts copy their
 * interior cell): exp(sigma z), z standard normal, constant per block^3 cells
 * on a hex mesh (block 1: per cell); and U[lo, hi). */
int fvm_mesh_lognormal_field(fvm_mesh *m, double sigma, int block, uint64_t seed, double *d_out);
int fvm_mesh_uniform_field(fvm_mesh *m, double lo, double hi, uint64_t seed, double *d_out);
/* Assemble conductivity d_k, source d_source and current field d_x (device,
 * n_cells + n_ghost entries) with boundary conditions per face group
 * (bc_type/bc_value indexed by group id, entry 0 unused) into a new system. */
int fvm_thermal_assemble(fvm_mesh *m, const double *d_k, const double *d_source, const double *d_x,
                         const int32_t *bc_type, const double *bc_value, fvm_system **out);
/* The same with the convection term of ThermalModel::linearize
 * (ConvectionDiscretization.h:53-200): d_flux, the convecting face flux (device,
 * n_faces, positive from c0 to c1), upwind, or with central != 0 the reference's
 * central-difference branch (face value 0.5 (x[c0] + x[c0]), upwind
 * linearization); d_cont_res (device, n_cells, may be NULL = 0) is added to the
 * diagonal; specified-temperature faces with an outflow flux extrapolate
 * (ThermalModel_impl.h:306-318). d_flux NULL: no convection term (as
 * fvm_thermal_assemble). */
int fvm_convdiff_assemble(fvm_mesh *m, const double *d_k, const double *d_source, const double *d_x,
                          const double *d_flux, const double *d_cont_res, int central, const int32_t *bc_type,
                          const double *bc_value, fvm_system **out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight time-stepping or serving loops).
 * fvm_convdiff_assemble's linearization (same kernels, same results) into an
 * existing system S that fvm_convdiff_assemble / fvm_thermal_assemble built from
 * this mesh, with no allocation: its scratch (per-face coefficients and fluxes,
 * cell gradients) is a per-device static array of 4,194,304 doubles (32 MiB, held
 * until the process ends), so 2 n_faces + 3 (n_cells + n_ghost) must not exceed
 * that (FVM_E_INVALID otherwise), and the mesh must have been assembled once
 * (its metrics and gradient coefficients exist). Every call on a device shares
 * that device's scratch: calls that overlap on one device (other threads) race. */
int fvm_convdiff_reassemble_noalloc(fvm_mesh *m, const double *d_k, const double *d_source, const double *d_x,
                                    const double *d_flux, const double *d_cont_res, int central,
                                    const int32_t *bc_type, const double *bc_value, fvm_system *S);
/* One discrete-ordinate direction of a gray phonon BTE (scalar system per
 * direction, PhononModel::linearizePhononModel, PhononModel.h:437-523):
 * relaxation -V (x - e0)/tau (PhononCollisionDiscretization.h:54-68), upwind
 * convection with face flux F = A . vg (PhononConvectionDiscretization.h:150-175),
 * per boundary face Dirichlet with the ghost's current value where vg points
 * into the domain and extrapolation elsewhere (PhononModel.h:498-523,
 * GenericPhononBCS.h:67-124), then initSolve's elimination of the extrapolated
 * ghost rows. vg: host double[3]; d_e0: device, n_cells entries; d_x: device,
 * n_cells + n_ghost entries (ghosts hold the boundary values). */
int fvm_phonon_assemble(fvm_mesh *m, const double *vg, double tau, const double *d_e0,
                        const double *d_x, fvm_system **out);
int fvm_system_info(const fvm_system *S, int *n_self, int *n_total, long long *nnz);
int fvm_system_device(const fvm_system *S, const int32_t **row, const int32_t **col,
                      const double **diag, const double **off, const double **b,
                      const uint8_t **is_boundary);
int fvm_system_get(const fvm_system *S, int32_t *row, int32_t *col, double *diag, double *off,
                   double *b, uint8_t *is_boundary);
/* LinearSystem::postSolve (LinearSystem.cpp:250-252) -> CRMatrix::solveBoundary
 * (CRMatrix.h:433-454): after a solve for the self rows, every boundary row
 * nr >= n_self gets d_delta[nr] = -(b[nr] + sum_j off * d_delta[col]) / diag[nr]
 * (device vector of n_total entries). */
int fvm_system_post_solve(const fvm_system *S, double *d_delta);
int fvm_system_destroy(fvm_system *S);

/* ---------------------------------------------------------------------------
 * Gray phonon BTE model (PhononModel, phononbase/PhononModel.h) on one mesh:
 * K discrete directions (modes) with group velocity vg[3k..3k+2], weight dk3[k],
 * relaxation time tau[k] and the gray equilibrium e0_k(T) = cp[k] (T - tref[k])
 * (pmode.h:146-165); every boundary face group is a temperature wall at
 * wall_T[g] (PhononBoundary.h:165-210; entry 0 unused). init (PhononModel.h:
 * 115-160): TL = T_init and e_k = e0_k(T_init) on every cell, ghosts included.
 * The mesh is borrowed: keep it alive while the model is used.
 * ------------------------------------------------------------------------- */
typedef struct fvm_phonon fvm_phonon;
int fvm_phonon_create(fvm_phonon **out, fvm_mesh *mesh, int K, const double *vg, const double *dk3,
                      const double *tau, const double *cp, const double *tref, const double *wall_T,
                      double T_init);
int fvm_phonon_destroy(fvm_phonon *p);
/* parameters of the per-direction AMG (the phonon linear solver); precision
 * FVM_PREC_FP64, no scaled corrections */
int fvm_phonon_set_params(fvm_phonon *p, const fvm_amg_params *prm);
/* PhononModel::callBoundaryConditions: the temperature walls set the ghosts */
int fvm_phonon_apply_walls(fvm_phonon *p);
/* PhononModel::advance (PhononModel.h:530-600): up to niter outer iterations;
 * every iteration assembles all directions (fvm_phonon_assemble), solves them
 * together with one batched AMG (fvm_batch_*: each direction's result is its own
 * AMG::solve's), sums the directions' initial 1-norms in direction order,
 * postSolve + update, then -- unless the sum is below abs_tol or its ratio to
 * the initial norm (the largest of the first five iterations') below rel_tol --
 * updateTL (Newton, Kspace.h:425-441), updatee0 and the walls. norms (may be
 * NULL when cap is 0): the iterations' norms; *converged: 1 if the stop test
 * ended the call. */
int fvm_phonon_advance(fvm_phonon *p, int niter, double rel_tol, double abs_tol, int *iterations,
                       double *norms, int cap, int *converged);
/* host copies (any may be NULL): TL (n_cells + n_ghost), e and e0 (K blocks of
 * n_cells + n_ghost, direction-major) */
int fvm_phonon_get(const fvm_phonon *p, double *TL, double *e, double *e0);
/* replace every direction's field (host, K blocks of n_cells + n_ghost) */
int fvm_phonon_set_field(fvm_phonon *p, const double *e);
/* One discrete velocity of the BGK kinetic model (KineticModel::
 * linearizeKineticModel, esbgkbase/KineticModel.h:1505-1674): the collision term
 * -V nu (f - feq) (CollisionTermDiscretization.h:46-62), then upwind convection
 * with face flux F = A . c (ConvectionDiscretization_Kmodel.h:171-197), then per
 * boundary face Dirichlet with the ghost's current value for incoming directions
 * and extrapolation otherwise (GenericKineticBCS.h:76-173): bc_kind[g] = 1 (zero
 * gradient) takes every direction as incoming, 0 (diffuse and specular walls)
 * when c.en - uw.en < eps with uw = wall_u[3g..3g+2] (KineticModel.h:1622-1661);
 * then initSolve's elimination of the extrapolated ghost rows. c: host double[3];
 * d_nue: device, n_cells entries; d_feq: device, n_cells; d_x: device, n_cells +
 * n_ghost (ghosts hold the boundary values); bc_kind, wall_u: host, per group
 * (entry 0 unused). */
int fvm_kinetic_assemble(fvm_mesh *m, const double *c, const double *d_nue, const double *d_feq,
                         const double *d_x, const int32_t *bc_kind, const double *wall_u, double eps,
                         fvm_system **out);

/* ---------------------------------------------------------------------------
 * BGK kinetic model (KineticModel, esbgkbase/KineticModel.h) on one mesh, on a
 * discrete velocity quadrature.
 * ------------------------------------------------------------------------- */
/* Quadrature(N1, N2, N3, clim, T2), Cartesian type (Quadrature.h:83-181): N1 N2 N3
 * velocities on the uniform grid [-clim sqrt(T2/2), clim sqrt(T2/2)]^3, x slowest,
 * z fastest; w the product of the spacings. Host arrays of N1 N2 N3 entries. */
int fvm_quadrature_cartesian(int N1, int N2, int N3, double clim, double T2, double *cx, double *cy,
                             double *cz, double *w);
/* KineticModelOptions (esbgkbase/KineticBC.h:61-113) used by the loop */
typedef struct fvm_kinetic_options {
    double rho_init, T_init, nonDimLt, molecularWeight, Tmuref, muref, mu_w, Prandtl;
    int fgamma;          /* 0: Maxwellian of the stored coefficients, 1: BGK, 2: ES-BGK (Newton fits) */
    int ktrial;          /* NewtonsMethod_ktrial */
    double ToleranceX, ToleranceF;
    double epsilon;      /* epsilon_ES: incoming test c.n - uw.n < epsilon */
} fvm_kinetic_options;
/* the reference's defaults (fgamma 2) */
int fvm_kinetic_default_options(fvm_kinetic_options *o);
typedef struct fvm_kinetic fvm_kinetic;
/* K velocities (cx, cy, cz, w: host); per face group g >= 1 (entry 0 unused):
 * bc_kind[g] 0 diffuse wall at velocity wall_u[3g..3g+2] and temperature
 * wall_T[g] ("wall", KineticBoundaryConditions.h:65-140), 1 zero gradient
 * ("ZeroGradBC", :385-406), 2 specular symmetry ("symmetry", :287-330, reflection
 * table from the group's first face, KineticModel.h:290-318), 3 a group whose type
 * has no boundary-value rule (the reference's Couette script sets "CopyBC",
 * test/ESBGK-tests/Couette/Par_Couette.py:199-206): its ghosts keep their values and
 * it is linearized like a wall at rest (KineticModel.h:1647-1661). The constructor's
 * state (KineticModel.h:81-120): density 1, velocity 0, temperature 1, f the
 * Maxwellian of it, then its moments, collision frequency and equilibrium. The
 * mesh is borrowed. */
int fvm_kinetic_create(fvm_kinetic **out, fvm_mesh *mesh, int K, const double *cx, const double *cy,
                       const double *cz, const double *w, const fvm_kinetic_options *opt,
                       const int32_t *bc_kind, const double *wall_u, const double *wall_T);
int fvm_kinetic_destroy(fvm_kinetic *p);
/* parameters of the per-direction AMG (the kinetic linear solver); FVM_PREC_FP64 */
int fvm_kinetic_set_params(fvm_kinetic *p, const fvm_amg_params *prm);
/* KineticModel::advance (KineticModel.h:3360-3462): up to niter outer iterations;
 * each assembles every direction (fvm_kinetic_assemble), solves them together with
 * one batched AMG, sums the directions' initial 1-norms in direction order,
 * postSolve + update, then the boundary values, moments, collision frequency and
 * equilibrium, then stops if the sum is below abs_tol or its ratio to the initial
 * norm (the largest of the first five iterations') below rel_tol. norms (may be
 * NULL when cap is 0): the iterations' norms; *converged: 1 if the stop test ended
 * the call. */
int fvm_kinetic_advance(fvm_kinetic *p, int niter, double rel_tol, double abs_tol, int *iterations,
                        double *norms, int cap, int *converged);
/* host copies (any may be NULL): f and feq (the equilibrium the collision term
 * relaxes to: ES-BGK's for fgamma 2) (K blocks of n_cells + n_ghost,
 * direction-major), density, temperature, collision frequency (n_cells + n_ghost),
 * velocity (3 blocks of n_cells + n_ghost: x, y, z) */
int fvm_kinetic_get(const fvm_kinetic *p, double *f, double *feq, double *density, double *velocity,
                    double *temperature, double *nu);

/* fvm_amg_setup_device on an assembled system */
int fvm_amg_setup_system(fvm_amg *h, const fvm_system *S);

#ifdef __cplusplus
}
#endif
#endif /* FVM_AMD_H */
