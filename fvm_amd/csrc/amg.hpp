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
    DArr<int> ci;           // to the next level (absent on the coarsest)
    int nc = 0;
    DArr<int> c2f_row, c2f_col;  // coarse->fine of ci (on this, the fine, level)
    // exact Gauss-Seidel schedule (device level pointers; host knows the count)
    bool sched_ready = false;
    DArr<int> fwd_ptr, fwd_rows, bwd_ptr, bwd_rows;
    int fwd_nlev = 0, bwd_nlev = 0, fwd_maxw = 0, bwd_maxw = 0;
    std::vector<int> h_fwd_ptr, h_bwd_ptr;
    // coloured SELL layout
    bool fast_ready = false;
    DArr<int> perm, iperm;
    int ncolor = 0;
    std::vector<int> color_ptr;        // host: row ranges of each colour (coloured index)
    DArr<int> d_color_ptr;             // the same on the device
    DArr<int> sell_off, sell_col;
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
    fvm_amg_params prm{};
    std::vector<fvm::Level> lev;
    fvm::Workspace ws;
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
    fvm::DArr<double> dense_inv;
    // tuning knobs (fvm_amg_set_tuning): rows below which a level runs in the
    // single-workgroup tail of the V-cycle / is coarsened by one workgroup
    long long tail_max_rows = 40000;
    long long coarsen_small_rows = 16384;
    // single-workgroup tail of the V-cycle: levels >= tail_level run in one launch
    int tail_level = -1;
    int tail_prec = -1;                // precision the descriptors were built for
    fvm::DArr<unsigned char> tail_desc;
    // live timing of finest-level kernels with HIP events on the handle's stream
    bool ktime = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open;
    long long k_launches[4] = {0, 0, 0, 0};
    double k_ms[4] = {0, 0, 0, 0};
    // double-buffer used by pointer-swapped Jacobi
};

namespace fvm {
// kernel timing kinds
enum { KT_RESID64 = 0, KT_SWEEP = 1 };
hipEvent_t timing_begin(fvm_amg *h);
void timing_end(fvm_amg *h, int kind, hipEvent_t start);
void timing_harvest(fvm_amg *h);
// setup.hip
void build_hierarchy(fvm_amg *h, int n_self, int n_total, const int *d_row, const int *d_col,
                     const double *d_diag, const double *d_off, const uint8_t *d_bnd);
void prepare_exact_schedule(fvm_amg *h, Level &L);
void prepare_fast(fvm_amg *h);
void prepare_mixed(fvm_amg *h, bool fast);
void prepare_direct(fvm_amg *h);
// cycle.hip
void solve_device(fvm_amg *h, const double *d_b, double *d_x, double *rnorm0, double *rnorm_final,
                  int *cycles);
void smooth_device(fvm_amg *h, const double *d_b, double *d_x);
void apply_precond(fvm_amg *h, const double *d_v, double *d_z);   // z = -M^{-1}... see cycle.hip
void spmv_residual64(fvm_amg *h, const double *d_x, const double *d_b, double *d_r, double *norm1);
double norm1_device(fvm_amg *h, const double *d_r, int n);
}  // namespace fvm
