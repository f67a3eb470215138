// Device-side AMG hierarchy of fvm_amd.
//
// Level storage in HBM (n = rows of the level, nnz = stored off-diagonals):
//   reference order ("exact" layout, used for parity and by the setup):
//     row[n+1], col[nnz] int32; off[nnz], diag[n] FP64 -- CSR with the diagonal
//     kept apart, column order of the reference (CRMatrix.h:73-84)
//     ci[n]            fine->coarse map (-1: row joins no group)
//     c2f_row[nc+1], c2f_col[n]  coarse->fine rows in ascending fine index
//     fwd/bwd level schedules of the Gauss-Seidel dependency DAG
//   coloured order ("fast" layout):
//     perm[n] (new->old), rows grouped by colour, ascending old index inside a colour
//     SELL-64: slices of 64 consecutive rows, column-major inside a slice,
//       sell_off[nslice+1] element offsets, sell_col[], sell_val[] (FP64),
//       sell_val32 (FP32, finest level) or sell_bf16 (BF16, coarse levels) for the
//       mixed-precision cycle; padding entries have value 0 and col = own row
//     pci[n] fine->coarse map in coloured indices of both levels; pc2f_* likewise
#pragma once

#include "common.hpp"
#include "../../include/fvm_amd.h"

namespace fvm {

struct Level {
    int n = 0;
    long long nnz = 0;
    // reference-order CSR
    DArr<int> row, col;
    DArr<double> diag, off;
    DArr<uint8_t> bnd;      // boundary flags (finest level only; coarse levels have none)
    // transpose pattern rowT(c) = {i : c in row(i)}; only stored when the pattern
    // is not structurally symmetric (sym == 0); sym == -1: not yet known
    int sym = -1;
    DArr<int> rowT, colT;
    DArr<int> tposT;        // with rowT/colT: position of the entry (colT[q], c) in row colT[q]
    DArr<int> ci;           // to the next level (absent on the coarsest)
    int nc = 0;
    DArr<int> c2f_row, c2f_col;  // coarse->fine of ci (on this, the fine, level)
    // exact Gauss-Seidel schedule (device level pointers; host knows the count)
    bool sched_ready = false;
    DArr<int> fwd_ptr, fwd_rows, bwd_ptr, bwd_rows;
    int fwd_nlev = 0, bwd_nlev = 0, fwd_maxw = 0, bwd_maxw = 0;
    std::vector<int> h_fwd_ptr, h_bwd_ptr;
    // launch plan of a sweep: maximal runs of consecutive dependency levels of at most
    // exact_wg_width rows share one single-workgroup launch, wider levels get a launch
    // each (built for runs_width; -1: not built)
    struct SweepRun { int lvl0, nlev, width; bool wg; };
    std::vector<SweepRun> fwd_runs, bwd_runs;
    int runs_width = -1;
    // coloured SELL layout
    bool fast_ready = false;
    int sell_u = 8;              // load depth of the SELL-64 row kernels (mean slice width, rounded up to 8/16/32)
    DArr<int> perm, iperm;
    int ncolor = 0;
    std::vector<int> color_ptr;        // host: row ranges of each colour (coloured index)
    std::vector<long long> color_nnz;  // host: off-diagonal entries of each colour's rows
    DArr<int> d_color_ptr;             // the same on the device
    DArr<int> sell_off, sell_col;
    int sell_total = 0;                // SELL-64 elements (padding included) = sell_off[nslices]
    DArr<double> sell_val, pdiag;
    DArr<int> pci, pc2f_row, pc2f_col;
    // mixed-precision values
    bool mixed_ready_exact = false, mixed_ready_fast = false;
    DArr<float> off32, diag32, sell_val32, pdiag32;
    DArr<uint16_t> off16, sell_bf16;
    // work vectors
    DArr<double> b, x, r, t;
    DArr<float> b32, x32, r32, t32;
};

// Several independent systems solved as one block-diagonal hierarchy: the
// agglomeration, Galerkin product and Gauss-Seidel schedules are local to a
// system's rows, so level l of the combined hierarchy is the concatenation of
// every system's own level l (systems in input order). Systems stop coarsening
// independently; a system whose hierarchy has ended keeps its coarsest rows but
// is excluded from further coarsening.
struct BatchSys {
    const int *row, *col;
    const double *diag, *off;
    const uint8_t *bnd;
    int n_self, n_total;
};

struct BatchInfo {
    int K = 0;
    std::vector<BatchSys> sys;
    std::vector<std::vector<int>> off;   // per level: K+1 row offsets of the systems
    std::vector<int> depth;              // per system: levels of its hierarchy
    DArr<BatchSys> d_sys;
    DArr<int> d_off0;                    // finest-level offsets on the device
    std::vector<DArr<int>> d_off;        // every level's offsets on the device
    DArr<int> sysperm;                   // coloured layout: finest rows grouped by system
    // solve state
    DArr<uint8_t> act;                   // per system: still cycling
    DArr<uint8_t> mask0;                 // per finest row: act of its system
    DArr<int> blk_off;                   // norm partial blocks per system (K+1)
    int nblk = 0;
    DArr<double> norm, r0, rn;           // per system
    DArr<int> cyc, counters;             // cycles per system; [0] iteration, [1] active count
    DArr<const double *> d_bptr;
    DArr<double *> d_xptr;
    DArr<double> partials;
    hipGraphExec_t graph = nullptr;
    bool graph_ok = false;
    bool use_graph = true;
    int *h_active = nullptr;             // pinned: active count after an iteration
    ~BatchInfo() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (h_active) (void)hipHostFree(h_active);
    }
};

struct Workspace {
    Scanner scan;
    DArr<double> partials;   // norm / dot partial sums
    DArr<double> scal;       // device scalars (norms, dots, scale)
    DArr<unsigned char> sort_tmp;
};

}  // namespace fvm

struct fvm_amg {
    int device = 0;
    hipStream_t stream = nullptr;
    fvm::Pool pool;              // declared first: destroyed after every array it lends
    // second stream of the setup (colouring and SELL layout of finished levels
    // while later levels are agglomerated), with its own pool and workspace so
    // that a block is only handed out again to work ordered on the same stream
    fvm::Pool pool2;
    hipStream_t stream2 = nullptr;
    fvm_amg_params prm{};
    std::vector<fvm::Level> lev;
    fvm::Workspace ws;
    fvm::Workspace ws2;          // of stream2
    bool ready = false;
    int n_self = 0, n_total = 0;
    long long total_iterations = 0;
    std::vector<double> hist;
    double setup_seconds = 0;
    int finest_rounds = 0;
    // finest-level staging for host entry points
    fvm::DArr<double> h_stage_b, h_stage_x;
    // mixed-precision coarsest dense inverse (FP64)
    int direct_level = -1;
    bool direct_tried = false;         // factorization attempted since the last setup
    int direct_attempts = 0;           // Gauss-Jordan runs since the handle was created
    fvm::DArr<double> dense_inv;
    // tuning knobs (fvm_amg_set_tuning): rows below which a level runs in the
    // single-workgroup tail of the V-cycle / is coarsened by one workgroup
    long long tail_max_rows = 40000;
    int exact_wg_width = 256;          // runs of consecutive dependency levels of at most this many rows
                                       // sweep in one workgroup launch (no result changes)
    int reuse_setup = 0;               // opt-in: a setup of an unchanged matrix keeps the hierarchy
    unsigned long long setup_sig = 0;  // signature of the hierarchy's inputs (when reuse_setup)
    bool setup_sig_valid = false;
    long long setup_reused = 0;        // setups skipped that way
    bool parked = false;               // cleanup kept the hierarchy for such a setup
    int exact_flow = 1;                // reference-order sweeps as one dataflow launch per direction
    fvm::DArr<int> flow_ctl;           // its block tickets, finished blocks and error flag
    bool flow_used = false;            // a dataflow sweep ran since the last check
    long long flow_ticks = 0;          // two seconds of the device's wall clock
    int flow_spin_limit = 0;           // waits per wavefront before a sweep gives up (0: no bound)
    long long coarsen_small_rows = 2048;
    int coarsen_lanes = 0;             // lanes per row of the agglomeration rounds (0: by row width)
    int coarsen_small_lanes = 16;      // the same for the single-workgroup rounds (0: as coarsen_lanes)
    int coarsen_grid = 0;              // blocks per launched agglomeration round (0: 2048 / 1024 per 256 threads)
    int coarsen_block = 512;           // threads per block of a launched agglomeration round
    int gal_lanes = 0;                 // lanes per coarse row of the coarse operator (0: by row length)
    int colour_grid = 256;             // blocks per colouring round (second setup stream)
    int coarsen_batch = 32;            // launched agglomeration rounds queued between frontier checks
    int coarsen_lds_entries = 8;       // entries per thread of such a block's next-frontier queue in LDS
    int bipartite_levels = 1;          // levels tried with a breadth-first two-colouring first (red-black)
    int coarsen_pair_filter = 1;       // rows do not wait on seeds that can never pair their neighbour
    int coarsen_lower_filter = 1;      // ... nor on lower neighbours other than the strongest ones
    int coarsen_wait_lists = 2;        // blocked joins wait in their blocker's list (1: only those that
                                       // wait for the group they join; 0: they retry every round)
    int sell_block = 256;              // threads per block of the SELL-64 smoother kernels
    int spmv_block = 256;              // threads per block of the SELL-64 residual / SpMV kernels
    int sell_xcd = 1, spmv_xcd = 0;    // XCD-contiguous block order for those kernel classes
    int nt_levels = 2;                 // finest levels read with non-temporal matrix loads
    fvm::BatchInfo *batch = nullptr;   // set while the hierarchy is a batched one
    // levels >= 1 of the coloured V-cycle replayed from a captured graph (per precision)
    int cycle_graph = 1;
    hipGraphExec_t cg_exec[2] = {nullptr, nullptr};
    std::string tune_report;           // last autotuner result (JSON)
    fvm::CounterProbe probe;           // pipelined termination checks of the setup loops
    // single-workgroup tail of the V-cycle: levels >= tail_level run in one launch
    int tail_level = -1;
    int tail_prec = -1;                // precision the descriptors were built for
    int tail_lds = 1;                  // tail levels that fit sweep from LDS
    int tail_lds_lim = -1;             // LDS bytes per workgroup of the device (-1: not queried)
    std::vector<int> tail_lds_bytes[2];   // LDS bytes per level (FP64 / mixed descriptors)
    fvm::DArr<unsigned char> tail_desc;
    // live timing of finest-level kernels with HIP events on the handle's stream
    bool ktime = false;
    std::vector<hipEvent_t> ev_pool;
    struct TimedLaunch { int kind; hipEvent_t start, stop; double bytes; };
    std::vector<TimedLaunch> ev_open;
    long long k_launches[4] = {0, 0, 0, 0};
    double k_ms[4] = {0, 0, 0, 0};
    double k_bytes[4] = {0, 0, 0, 0};  // algorithmic bytes of the timed launches
    // double-buffer used by pointer-swapped Jacobi
    // convergence lines at verbosity > 0 (AMG.cpp:237-272): "i: [field : norm]"
    std::string field_name = "x";
    FILE *out = nullptr;               // redirectPrintToFile target (nullptr: stdout)
    std::string out_path;
    int print_style = 0;               // 0: every cycle (serial build), 1: first and last (parallel build)
};

namespace fvm {
inline void invalidate_cycle_graphs(fvm_amg *h) {
    for (auto &g : h->cg_exec)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}
// one convergence line "label: [field : value]" (MultiFieldReduction.cpp:199-211)
inline void print_norm(const fvm_amg *h, const std::string &label, double v) {
    FILE *f = h->out ? h->out : stdout;
    fprintf(f, "%s: [%s : %g]\n", label.c_str(), h->field_name.c_str(), v);
    fflush(f);
}
// kernel timing kinds
enum { KT_RESID64 = 0, KT_SWEEP = 1 };
hipEvent_t timing_begin(fvm_amg *h);
// bytes: the algorithmic HBM bytes of the timed work (every operand read or written once)
void timing_end(fvm_amg *h, int kind, hipEvent_t start, double bytes);
void timing_harvest(fvm_amg *h);
// setup.hip
void build_hierarchy(fvm_amg *h, int n_self, int n_total, const int *d_row, const int *d_col,
                     const double *d_diag, const double *d_off, const uint8_t *d_bnd);
void prepare_exact_schedule(fvm_amg *h, Level &L);
void prepare_fast(fvm_amg *h);
void prepare_fast_level(fvm_amg *h, Level &L, int l, hipStream_t s, Workspace &ws, const int *seg_off, int nseg);
void prepare_fast_maps(fvm_amg *h);
void prepare_mixed(fvm_amg *h, bool fast);
void prepare_direct(fvm_amg *h);
void build_hierarchy_batched(fvm_amg *h, BatchInfo &B);
// cycle.hip
void solve_device(fvm_amg *h, const double *d_b, double *d_x, double *rnorm0, double *rnorm_final,
                  int *cycles);
void smooth_device(fvm_amg *h, const double *d_b, double *d_x);
void direct_apply(fvm_amg *h, const double *d_b, double *d_x);
void bcgstab_solve(fvm_amg *h, const fvm_linsolver_params &p, const double *d_b, double *d_x, double *rnorm0,
                   double *rnorm_final, int *iterations, std::vector<double> &hist);
void cg_solve(fvm_amg *h, const fvm_linsolver_params &p, const double *d_b, double *d_x, double *rnorm0,
              double *rnorm_final, int *iterations, std::vector<double> &hist);
void batch_solve(fvm_amg *h, BatchInfo &B, const double *const *d_b, double *const *d_x, double *rnorm0,
                 double *rnorm_final, int *cycles);
void apply_precond(fvm_amg *h, const double *d_v, double *d_z);   // z = -M^{-1}... see cycle.hip
void spmv_residual64(fvm_amg *h, const double *d_x, const double *d_b, double *d_r, double *norm1);
double norm1_device(fvm_amg *h, const double *d_r, int n);
std::string autotune(fvm_amg *h, int reps);
// assembly.hip: device of a mesh; its face cells, area vectors and |A| on the
// device (metrics computed if needed)
int mesh_device(const fvm_mesh *m);
void mesh_device_metrics(fvm_mesh *m, const int **fcells, const double **fA, const double **fAm);
hipStream_t mesh_stream(fvm_mesh *m);
fvm_system *system_for_mesh(fvm_mesh *m);
void kinetic_assemble_into(fvm_mesh *m, const double *c, const double *d_nue, const double *d_feq,
                           const double *d_x, const int *grp, const double *wall, double eps, fvm_system *S,
                           double *F, uint8_t *inc);
void system_post_solve_async(const fvm_system *S, double *d_delta, hipStream_t s);
// learned correction (mpnn.hip)
void mpnn_bind(fvm_mpnn *m, int n, const int *row, const int *col, const double *off, const double *diag);
void mpnn_correct(fvm_mpnn *m, hipStream_t st, const int *perm, const double *Az, const double *v, double *z,
                  int n, int *bad_out);
struct LearnedHook {
    fvm_mpnn *m = nullptr;
    double stall_ratio = 0.9;   // off when the 2-norm estimate falls by less over 5 iterations
    int *used = nullptr;        // iterations that applied the network
    int *fallback_at = nullptr; // iteration at which it was switched off (-1: never)
};
}  // namespace fvm
