// C ABI of fvm_amd (declared in include/fvm_amd.h). No exception crosses it:
// every entry point returns FVM_OK or a negative FVM_E_* code and records the
// message for fvm_last_error().
#include "amg.hpp"

#include <climits>

#include <algorithm>
#include <chrono>

namespace fvm {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

struct KrylovWork;
void fgmres_solve(fvm_amg *h, KrylovWork &W, const fvm_fgmres_params &p, const double *d_b, double *d_x,
                  double *rnorm0, double *rnorm_final, int *iterations, std::vector<double> &hist,
                  const LearnedHook &nn);

}  // namespace fvm

using fvm::DArr;

#define FVM_API_BEGIN try {
#define FVM_API_END                                                    \
    }                                                                  \
    catch (const fvm::Error &e) {                                      \
        fvm::set_last_error(e.what());                                 \
        return e.code;                                                 \
    }                                                                  \
    catch (const std::exception &e) {                                  \
        fvm::set_last_error(e.what());                                 \
        return FVM_E_INVALID;                                          \
    }                                                                  \
    catch (...) {                                                      \
        fvm::set_last_error("unknown error");                          \
        return FVM_E_INVALID;                                          \
    }

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        FVM_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// device of the handle + its allocation pool for the duration of a call
struct HandleScope {
    DeviceGuard dg;
    fvm::PoolScope ps;
    explicit HandleScope(fvm_amg *h) : dg(h->device), ps(&h->pool) {}
};

void check_params(const fvm_amg_params *p) {
    FVM_CHECK(p, FVM_E_INVALID, "null parameter struct");
    FVM_CHECK(p->maxCoarseLevels >= 0, FVM_E_INVALID, "maxCoarseLevels must be >= 0");
    FVM_CHECK(p->nPreSweeps >= 0 && p->nPostSweeps >= 0, FVM_E_INVALID, "sweep counts must be >= 0");
    FVM_CHECK(p->cycleType >= FVM_V_CYCLE && p->cycleType <= FVM_F_CYCLE, FVM_E_INVALID, "bad cycleType");
    FVM_CHECK(p->smootherType == FVM_GAUSS_SEIDEL || p->smootherType == FVM_JACOBI, FVM_E_INVALID,
              "bad smootherType");
    FVM_CHECK(p->gsOrdering == FVM_GS_EXACT || p->gsOrdering == FVM_GS_MULTICOLOR, FVM_E_INVALID,
              "bad gsOrdering");
    FVM_CHECK(p->precision == FVM_PREC_FP64 || p->precision == FVM_PREC_MIXED, FVM_E_INVALID,
              "bad precision");
    FVM_CHECK(p->coarseGroupSize >= 1, FVM_E_INVALID, "coarseGroupSize must be >= 1");
}

}  // namespace

struct fvm_fgmres {
    fvm_amg *pc = nullptr;
    int device = 0;               // of pc, kept so that destroy never reads pc (it may be gone)
    fvm_fgmres_params prm{};
    fvm::KrylovWork *work = nullptr;
    std::vector<double> hist;
    fvm_mpnn *nn = nullptr;       // learned correction of the preconditioner (borrowed)
    double stall_ratio = 0.9;
    int nn_used = 0, nn_fallback_at = -1;
};

struct fvm_batch {
    fvm_amg *h = nullptr;     // owns the combined hierarchy, stream and pool
    fvm::BatchInfo B;
};

namespace fvm {
struct KrylovWork;
KrylovWork *new_krylov_work();
void delete_krylov_work(KrylovWork *w);
}  // namespace fvm

static fvm::LearnedHook hook_of(fvm_fgmres *s) {
    fvm::LearnedHook k;
    k.m = s->nn;
    k.stall_ratio = s->stall_ratio;
    k.used = &s->nn_used;
    k.fallback_at = &s->nn_fallback_at;
    return k;
}

extern "C" {

int fvm_version(void) { return FVM_AMD_VERSION; }

const char *fvm_last_error(void) { return fvm::g_last_error.c_str(); }

int fvm_amg_default_params(fvm_amg_params *p) {
    if (!p) return FVM_E_INVALID;
    // AMG.cpp:15-36 and LinearSolver.h:19-24
    p->maxCoarseLevels = 30;
    p->nPreSweeps = 0;
    p->nPostSweeps = 1;
    p->coarseGroupSize = 2;
    p->weightRatioThreshold = 0.65;
    p->cycleType = FVM_V_CYCLE;
    p->smootherType = FVM_GAUSS_SEIDEL;
    p->scaleCorrections = 1;
    p->nMaxIterations = 100;
    p->verbosity = 2;
    p->relativeTolerance = 1e-8;
    p->absoluteTolerance = 1e-50;
    p->isSymmetric = 0;
    p->parallelStopRule = 1;
    p->gsOrdering = FVM_GS_EXACT;
    p->precision = FVM_PREC_FP64;
    return FVM_OK;
}

int fvm_amg_create(fvm_amg **out, int device) {
    FVM_API_BEGIN
    FVM_CHECK(out, FVM_E_INVALID, "null output handle");
    int ndev = 0;
    FVM_HIP(hipGetDeviceCount(&ndev));
    FVM_CHECK(device >= 0 && device < ndev, FVM_E_INVALID, "device ordinal out of range");
    DeviceGuard g(device);
    fvm_amg *h = new fvm_amg();
    h->device = device;
    fvm_amg_default_params(&h->prm);
    // a blocking stream: work queued on the legacy default stream (where PyTorch and
    // other callers produce the device arrays they pass in) completes before the
    // handle's work starts
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamDefault);
    if (e != hipSuccess) {
        delete h;
        throw fvm::Error(FVM_E_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    *out = h;
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_destroy(fvm_amg *h) {
    FVM_API_BEGIN
    if (!h) return FVM_OK;
    {
        HandleScope g(h);
        (void)hipStreamSynchronize(h->stream);
        fvm::invalidate_cycle_graphs(h);
        if (h->out) { fclose(h->out); h->out = nullptr; }
        h->probe.release();
        h->lev.clear();
        h->ws = fvm::Workspace();
        h->ws2 = fvm::Workspace();
        if (h->stream2) (void)hipStreamDestroy(h->stream2);
        h->pool2.trim();
        h->dense_inv.release();
        h->h_stage_b.release();
        h->h_stage_x.release();
        fvm::timing_harvest(h);
        for (auto e : h->ev_pool) (void)hipEventDestroy(e);
        h->ev_pool.clear();
        h->pool.trim();
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_set_params(fvm_amg *h, const fvm_amg_params *p) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    check_params(p);
    const bool rebuild = h->ready && (p->coarseGroupSize != h->prm.coarseGroupSize ||
                                      p->weightRatioThreshold != h->prm.weightRatioThreshold ||
                                      p->maxCoarseLevels != h->prm.maxCoarseLevels ||
                                      p->parallelStopRule != h->prm.parallelStopRule);
    h->prm = *p;
    {
        HandleScope g(h);
        fvm::invalidate_cycle_graphs(h);   // captured cycles hold the old sweep counts and kernels
    }
    if (rebuild) {
        // hierarchy parameters changed: the next solve needs a new setup (AMG rebuilds per solve)
        HandleScope g(h);
        h->lev.clear();
        h->ready = false;
        h->parked = false;
        h->setup_sig_valid = false;
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_get_params(const fvm_amg *h, fvm_amg_params *p) {
    if (!h || !p) return FVM_E_INVALID;
    *p = h->prm;
    return FVM_OK;
}

int fvm_amg_setup_device(fvm_amg *h, int n_self, int n_total, const int32_t *d_row, const int32_t *d_col,
                         const double *d_diag, const double *d_off, const uint8_t *d_is_boundary) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    FVM_CHECK(n_self >= 0 && n_total >= n_self, FVM_E_INVALID, "need 0 <= n_self <= n_total");
    FVM_CHECK(d_row && (n_total == 0 || (d_diag)), FVM_E_INVALID, "null matrix array");
    HandleScope g(h);
    fvm::build_hierarchy(h, n_self, n_total, d_row, d_col, d_diag, d_off, d_is_boundary);
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_setup(fvm_amg *h, int n_self, int n_total, const int32_t *row, const int32_t *col,
                  const double *diag, const double *off, const uint8_t *is_boundary) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    FVM_CHECK(n_self >= 0 && n_total >= n_self, FVM_E_INVALID, "need 0 <= n_self <= n_total");
    FVM_CHECK(row && (n_total == 0 || diag), FVM_E_INVALID, "null matrix array");
    const long long nnz = n_total > 0 ? row[n_total] : 0;
    FVM_CHECK(nnz >= 0 && row[0] == 0, FVM_E_INVALID, "row pointers must start at 0");
    FVM_CHECK(nnz == 0 || (col && off), FVM_E_INVALID, "null col/off arrays");
    HandleScope g(h);
    DArr<int> d_row((size_t)n_total + 1), d_col(std::max<long long>(nnz, 1));
    DArr<double> d_diag(std::max(n_total, 1)), d_off(std::max<long long>(nnz, 1));
    DArr<uint8_t> d_bnd;
    hipStream_t s = h->stream;
    FVM_HIP(hipMemcpyAsync(d_row.p, row, sizeof(int) * ((size_t)n_total + 1), hipMemcpyHostToDevice, s));
    if (nnz > 0) {
        FVM_HIP(hipMemcpyAsync(d_col.p, col, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
        FVM_HIP(hipMemcpyAsync(d_off.p, off, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    }
    if (n_total > 0)
        FVM_HIP(hipMemcpyAsync(d_diag.p, diag, sizeof(double) * n_total, hipMemcpyHostToDevice, s));
    if (is_boundary && n_total > 0) {
        d_bnd.alloc(n_total);
        FVM_HIP(hipMemcpyAsync(d_bnd.p, is_boundary, n_total, hipMemcpyHostToDevice, s));
    }
    FVM_HIP(hipStreamSynchronize(s));
    fvm::build_hierarchy(h, n_self, n_total, d_row.p, d_col.p, d_diag.p, d_off.p, d_bnd.p);
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_solve_device(fvm_amg *h, const double *d_b, double *d_x, double *rnorm0, double *rnorm_final,
                         int *cycles) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "solve called before setup");
    FVM_CHECK(h->n_self == 0 || (d_b && d_x), FVM_E_INVALID, "null vector");
    HandleScope g(h);
    fvm::solve_device(h, d_b, d_x, rnorm0, rnorm_final, cycles);
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_solve(fvm_amg *h, const double *b, double *x, double *rnorm0, double *rnorm_final, int *cycles) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "solve called before setup");
    FVM_CHECK(h->n_total == 0 || (b && x), FVM_E_INVALID, "null vector");
    HandleScope g(h);
    const size_t n = std::max(h->n_total, 1);
    h->h_stage_b.ensure(n);
    h->h_stage_x.ensure(n);
    hipStream_t s = h->stream;
    if (h->n_total > 0) {
        FVM_HIP(hipMemcpyAsync(h->h_stage_b.p, b, sizeof(double) * h->n_total, hipMemcpyHostToDevice, s));
        FVM_HIP(hipMemcpyAsync(h->h_stage_x.p, x, sizeof(double) * h->n_total, hipMemcpyHostToDevice, s));
    }
    fvm::solve_device(h, h->h_stage_b.p, h->h_stage_x.p, rnorm0, rnorm_final, cycles);
    if (h->n_self > 0) {
        FVM_HIP(hipMemcpyAsync(x, h->h_stage_x.p, sizeof(double) * h->n_self, hipMemcpyDeviceToHost, s));
        FVM_HIP(hipStreamSynchronize(s));
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_smooth_device(fvm_amg *h, const double *d_b, double *d_x) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "smooth called before setup");
    HandleScope g(h);
    fvm::smooth_device(h, d_b, d_x);
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_smooth(fvm_amg *h, const double *b, double *x) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "smooth called before setup");
    FVM_CHECK(h->n_total == 0 || (b && x), FVM_E_INVALID, "null vector");
    HandleScope g(h);
    const size_t n = std::max(h->n_total, 1);
    h->h_stage_b.ensure(n);
    h->h_stage_x.ensure(n);
    hipStream_t s = h->stream;
    if (h->n_total > 0) {
        FVM_HIP(hipMemcpyAsync(h->h_stage_b.p, b, sizeof(double) * h->n_total, hipMemcpyHostToDevice, s));
        FVM_HIP(hipMemcpyAsync(h->h_stage_x.p, x, sizeof(double) * h->n_total, hipMemcpyHostToDevice, s));
    }
    fvm::smooth_device(h, h->h_stage_b.p, h->h_stage_x.p);
    if (h->n_self > 0) {
        FVM_HIP(hipMemcpyAsync(x, h->h_stage_x.p, sizeof(double) * h->n_self, hipMemcpyDeviceToHost, s));
        FVM_HIP(hipStreamSynchronize(s));
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_cleanup(fvm_amg *h) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    HandleScope g(h);
    FVM_HIP(hipStreamSynchronize(h->stream));
    h->ready = false;
    if (h->reuse_setup && h->setup_sig_valid && !h->batch) {
        // kept, but unusable until a setup of the same matrix takes it up again
        h->parked = true;
        return FVM_OK;
    }
    h->parked = false;
    h->setup_sig_valid = false;
    h->lev.clear();
    h->dense_inv.release();
    h->direct_level = -1;
    h->direct_tried = false;
    return FVM_OK;
    FVM_API_END
}

long long fvm_amg_total_iterations(const fvm_amg *h) { return h ? h->total_iterations : 0; }

int fvm_amg_history(const fvm_amg *h, double *hist, int cap) {
    if (!h || (cap > 0 && !hist)) return FVM_E_INVALID;
    const int k = std::min<int>(cap, (int)h->hist.size());
    for (int i = 0; i < k; i++) hist[i] = h->hist[i];
    return k;
}

int fvm_amg_num_levels(const fvm_amg *h) { return (h && h->ready) ? (int)h->lev.size() : 0; }

int fvm_amg_level_info(const fvm_amg *h, int level, int *n_rows, int *nnz) {
    if (!h || !h->ready || level < 0 || level >= (int)h->lev.size()) return FVM_E_INVALID;
    if (n_rows) *n_rows = h->lev[level].n;
    if (nnz) *nnz = (int)h->lev[level].nnz;
    return FVM_OK;
}

int fvm_amg_level_get(const fvm_amg *h, int level, int32_t *row, int32_t *col, double *diag, double *off,
                      int32_t *ci) {
    FVM_API_BEGIN
    FVM_CHECK(h && h->ready && level >= 0 && level < (int)h->lev.size(), FVM_E_INVALID, "bad level");
    DeviceGuard g(h->device);
    const fvm::Level &L = h->lev[level];
    hipStream_t s = h->stream;
    if (row) FVM_HIP(hipMemcpyAsync(row, L.row.p, sizeof(int) * ((size_t)L.n + 1), hipMemcpyDeviceToHost, s));
    if (col && L.nnz) FVM_HIP(hipMemcpyAsync(col, L.col.p, sizeof(int) * L.nnz, hipMemcpyDeviceToHost, s));
    if (diag && L.n) FVM_HIP(hipMemcpyAsync(diag, L.diag.p, sizeof(double) * L.n, hipMemcpyDeviceToHost, s));
    if (off && L.nnz) FVM_HIP(hipMemcpyAsync(off, L.off.p, sizeof(double) * L.nnz, hipMemcpyDeviceToHost, s));
    if (ci) {
        if (L.ci.p && level + 1 < (int)h->lev.size())
            FVM_HIP(hipMemcpyAsync(ci, L.ci.p, sizeof(int) * L.n, hipMemcpyDeviceToHost, s));
        else
            for (int i = 0; i < L.n; i++) ci[i] = -1;
    }
    FVM_HIP(hipStreamSynchronize(s));
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_level_colours(const fvm_amg *h, int level, int32_t *colour, int *ncolour) {
    FVM_API_BEGIN
    FVM_CHECK(h && h->ready && level >= 0 && level < (int)h->lev.size(), FVM_E_INVALID, "bad level");
    DeviceGuard g(h->device);
    const fvm::Level &L = h->lev[level];
    const bool have = L.ncolor > 0 && L.iperm.p && (int)L.color_ptr.size() == L.ncolor + 1;
    if (ncolour) *ncolour = have ? L.ncolor : 0;
    if (colour && L.n > 0) {
        if (!have) {
            for (int i = 0; i < L.n; i++) colour[i] = -1;
        } else {
            // colour of row i: the colour range holding its position iperm[i] in the
            // coloured order
            std::vector<int> ip(L.n);
            FVM_HIP(hipMemcpyAsync(ip.data(), L.iperm.p, sizeof(int) * L.n, hipMemcpyDeviceToHost, h->stream));
            FVM_HIP(hipStreamSynchronize(h->stream));
            for (int i = 0; i < L.n; i++) {
                const int p = ip[i];
                int c = (int)(std::upper_bound(L.color_ptr.begin(), L.color_ptr.end(), p) - L.color_ptr.begin()) - 1;
                colour[i] = (p >= 0 && p < L.n && c >= 0 && c < L.ncolor) ? c : -1;
            }
        }
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_set_output(fvm_amg *h, const char *field_name, const char *path, int style) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    FVM_CHECK(style == 0 || style == 1, FVM_E_INVALID, "style must be 0 (every cycle) or 1 (first and last)");
    if (!path || h->out_path != path) {      // the file stays open (and keeps its lines) while the path is the same
        FILE *f = nullptr;
        if (path) {
            f = fopen(path, "w");
            FVM_CHECK(f, FVM_E_INVALID, std::string("cannot open ") + path);
        }
        if (h->out) fclose(h->out);
        h->out = f;
        h->out_path = path ? path : "";
    }
    h->field_name = field_name ? field_name : "x";
    h->print_style = style;
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_direct_level(const fvm_amg *h, int *level, int *n_rows, int *attempts, double *inv) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    const int lv = h->ready ? h->direct_level : -1;
    const int n = lv >= 0 ? h->lev[lv].n : 0;
    if (level) *level = lv;
    if (n_rows) *n_rows = n;
    if (attempts) *attempts = h->direct_attempts;
    if (inv && n > 0) {
        DeviceGuard g(h->device);
        FVM_HIP(hipMemcpyAsync(inv, h->dense_inv.p, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost,
                               h->stream));
        FVM_HIP(hipStreamSynchronize(h->stream));
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_direct_apply(fvm_amg *h, const double *b, double *x) {
    FVM_API_BEGIN
    FVM_CHECK(h && b && x, FVM_E_INVALID, "null argument");
    FVM_CHECK(h->ready && h->direct_level >= 0, FVM_E_INVALID, "the hierarchy has no direct level");
    HandleScope g(h);
    const int n = h->lev[h->direct_level].n;
    DArr<double> db(n), dx(n);
    FVM_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    fvm::direct_apply(h, db.p, dx.p);
    FVM_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    FVM_HIP(hipStreamSynchronize(h->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_set_kernel_timing(fvm_amg *h, int enable) {
    FVM_API_BEGIN
    FVM_CHECK(h, FVM_E_INVALID, "null handle");
    HandleScope g(h);
    FVM_HIP(hipStreamSynchronize(h->stream));
    fvm::timing_harvest(h);
    h->ktime = enable != 0;
    for (int k = 0; k < 4; k++) { h->k_launches[k] = 0; h->k_ms[k] = 0.0; h->k_bytes[k] = 0.0; }
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_kernel_timing(const fvm_amg *h, int kind, long long *launches, double *total_ms) {
    if (!h || kind < 0 || kind > 3) return FVM_E_INVALID;
    if (launches) *launches = h->k_launches[kind];
    if (total_ms) *total_ms = h->k_ms[kind];
    return FVM_OK;
}

int fvm_amg_kernel_bytes(const fvm_amg *h, int kind, double *total_bytes) {
    if (!h || kind < 0 || kind > 3 || !total_bytes) return FVM_E_INVALID;
    *total_bytes = h->k_bytes[kind];
    return FVM_OK;
}

int fvm_amg_set_tuning(fvm_amg *h, const char *key, long long value) {
    FVM_API_BEGIN
    FVM_CHECK(h && key, FVM_E_INVALID, "null argument");
    const std::string k(key);
    if (k == "tail_max_rows") {
        FVM_CHECK(value >= 0, FVM_E_INVALID, "tail_max_rows must be >= 0");
        h->tail_max_rows = value;
    } else if (k == "gal_lanes") {
        FVM_CHECK(value == 0 || value == 16 || value == 32 || value == 64, FVM_E_INVALID,
                  "gal_lanes must be 0, 16, 32 or 64");
        h->gal_lanes = (int)value;
    } else if (k == "colour_grid") {
        FVM_CHECK(value >= 16 && value <= 4096, FVM_E_INVALID, "colour_grid must be in [16, 4096]");
        h->colour_grid = (int)value;
    } else if (k == "coarsen_batch") {
        FVM_CHECK(value >= 4 && value <= 256, FVM_E_INVALID, "coarsen_batch must be in [4, 256]");
        h->coarsen_batch = (int)value;
    } else if (k == "tail_lds") {
        h->tail_lds = value != 0;
    } else if (k == "coarsen_small_rows") {
        FVM_CHECK(value >= 0, FVM_E_INVALID, "coarsen_small_rows must be >= 0");
        h->coarsen_small_rows = value;
    } else if (k == "sell_block" || k == "spmv_block") {
        FVM_CHECK(value == 64 || value == 128 || value == 256 || value == 512, FVM_E_INVALID,
                  k + " must be 64, 128, 256 or 512");
        (k == "sell_block" ? h->sell_block : h->spmv_block) = (int)value;
    } else if (k == "sell_xcd" || k == "spmv_xcd") {
        (k == "sell_xcd" ? h->sell_xcd : h->spmv_xcd) = value != 0;
    } else if (k == "cycle_graph") {
        h->cycle_graph = value != 0;
    } else if (k == "coarsen_lanes") {
        FVM_CHECK(value == 0 || value == 8 || value == 16 || value == 32 || value == 64, FVM_E_INVALID,
                  "coarsen_lanes must be 0, 8, 16, 32 or 64");
        h->coarsen_lanes = (int)value;
    } else if (k == "exact_wg_width") {
        FVM_CHECK(value >= 0 && value <= INT_MAX, FVM_E_INVALID, "exact_wg_width must be in [0, INT_MAX]");
        h->exact_wg_width = (int)value;
        fvm::invalidate_cycle_graphs(h);
    } else if (k == "reuse_setup") {
        h->reuse_setup = value != 0;
        h->setup_sig_valid = false;   // the signature is taken from the next setup on
        if (h->parked) {              // a hierarchy kept by cleanup is dropped
            h->parked = false;
            h->lev.clear();
            h->dense_inv.release();
            h->direct_level = -1;
            h->direct_tried = false;
        }
    } else if (k == "coarsen_pair_filter") {
        h->coarsen_pair_filter = value != 0;
    } else if (k == "coarsen_lower_filter") {
        h->coarsen_lower_filter = value != 0;
    } else if (k == "coarsen_block") {
        FVM_CHECK(value == 256 || value == 512 || value == 1024, FVM_E_INVALID, "coarsen_block must be 256, 512 or 1024");
        h->coarsen_block = (int)value;
    } else if (k == "coarsen_lds_entries") {
        FVM_CHECK(value >= 0 && value <= 16, FVM_E_INVALID, "coarsen_lds_entries must be in [0, 16]");
        h->coarsen_lds_entries = (int)value;
    } else if (k == "coarsen_wait_lists") {
        FVM_CHECK(value >= 0 && value <= 2, FVM_E_INVALID, "coarsen_wait_lists must be 0, 1 or 2");
        h->coarsen_wait_lists = (int)value;
    } else if (k == "exact_flow") {
        h->exact_flow = value != 0;
    } else if (k == "exact_flow_spin_limit") {
        FVM_CHECK(value >= 0 && value <= INT_MAX, FVM_E_INVALID, "exact_flow_spin_limit must be in [0, INT_MAX]");
        h->flow_spin_limit = (int)value;

    } else if (k == "bipartite_levels") {
        FVM_CHECK(value >= 0 && value <= 64, FVM_E_INVALID, "bipartite_levels must be in [0, 64]");
        h->bipartite_levels = (int)value;
    } else if (k == "coarsen_grid") {
        FVM_CHECK(value >= 0 && value <= 65536, FVM_E_INVALID, "coarsen_grid must be in [0, 65536]");
        h->coarsen_grid = (int)value;
    } else if (k == "coarsen_small_lanes") {
        FVM_CHECK(value == 0 || value == 8 || value == 16 || value == 32 || value == 64, FVM_E_INVALID,
                  "coarsen_small_lanes must be 0, 8, 16, 32 or 64");
        h->coarsen_small_lanes = (int)value;
    } else if (k == "nt_levels") {
        FVM_CHECK(value >= 0 && value <= 64, FVM_E_INVALID, "nt_levels must be in [0, 64]");
        h->nt_levels = (int)value;
    } else {
        throw fvm::Error(FVM_E_INVALID, "unknown tuning key: " + k);
    }
    HandleScope g(h);
    fvm::invalidate_cycle_graphs(h);
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_get_tuning(const fvm_amg *h, const char *key, long long *value) {
    if (!h || !key || !value) return FVM_E_INVALID;
    const std::string k(key);
    if (k == "tail_max_rows") *value = h->tail_max_rows;
    else if (k == "coarsen_small_rows") *value = h->coarsen_small_rows;
    else if (k == "sell_block") *value = h->sell_block;
    else if (k == "spmv_block") *value = h->spmv_block;
    else if (k == "sell_xcd") *value = h->sell_xcd;
    else if (k == "spmv_xcd") *value = h->spmv_xcd;
    else if (k == "cycle_graph") *value = h->cycle_graph;
    else if (k == "nt_levels") *value = h->nt_levels;
    else if (k == "coarsen_lanes") *value = h->coarsen_lanes;
    else if (k == "coarsen_small_lanes") *value = h->coarsen_small_lanes;
    else if (k == "coarsen_grid") *value = h->coarsen_grid;
    else if (k == "bipartite_levels") *value = h->bipartite_levels;
    else if (k == "exact_wg_width") *value = h->exact_wg_width;
    else if (k == "exact_flow") *value = h->exact_flow;
    else if (k == "coarsen_pair_filter") *value = h->coarsen_pair_filter;
    else if (k == "coarsen_lower_filter") *value = h->coarsen_lower_filter;
    else if (k == "coarsen_wait_lists") *value = h->coarsen_wait_lists;
    else if (k == "coarsen_block") *value = h->coarsen_block;
    else if (k == "tail_lds") *value = h->tail_lds;
    else if (k == "gal_lanes") *value = h->gal_lanes;
    else if (k == "colour_grid") *value = h->colour_grid;
    else if (k == "coarsen_batch") *value = h->coarsen_batch;
    else if (k == "coarsen_lds_entries") *value = h->coarsen_lds_entries;
    else if (k == "reuse_setup") *value = h->reuse_setup;
    else if (k == "setups_reused") *value = h->setup_reused;
    else if (k == "exact_flow_spin_limit") *value = h->flow_spin_limit;
    else return FVM_E_INVALID;
    return FVM_OK;
}

int fvm_amg_autotune(fvm_amg *h, int reps, char *report, int cap) {
    FVM_API_BEGIN
    FVM_CHECK(h && reps >= 1, FVM_E_INVALID, "null handle or reps < 1");
    HandleScope g(h);
    const std::string r = fvm::autotune(h, reps);
    if (report && cap > 0) {
        const size_t k = std::min(r.size(), (size_t)cap - 1);
        memcpy(report, r.data(), k);
        report[k] = 0;
    }
    return (int)r.size();
    FVM_API_END
}

int fvm_amg_setup_stats(const fvm_amg *h, double *setup_seconds, int *finest_rounds) {
    if (!h) return FVM_E_INVALID;
    if (setup_seconds) *setup_seconds = h->setup_seconds;
    if (finest_rounds) *finest_rounds = h->finest_rounds;
    return FVM_OK;
}

// ----------------------------------------------------------------------- batched systems
static void batch_check_params(const fvm_amg_params *p) {
    check_params(p);
    FVM_CHECK(p->precision == FVM_PREC_FP64, FVM_E_INVALID, "batched solve: precision must be FVM_PREC_FP64");
    FVM_CHECK(!(p->isSymmetric && p->scaleCorrections), FVM_E_INVALID,
              "batched solve: scaled corrections (isSymmetric && scaleCorrections) are not supported");
}

static void batch_invalidate(fvm_batch *b) {
    if (b->B.graph) { (void)hipGraphExecDestroy(b->B.graph); b->B.graph = nullptr; }
    b->B.graph_ok = false;
    b->B.blk_off.release();
    b->B.sysperm.release();
}

int fvm_batch_create(fvm_batch **out, int device) {
    FVM_API_BEGIN
    FVM_CHECK(out, FVM_E_INVALID, "null output handle");
    fvm_amg *h = nullptr;
    const int rc = fvm_amg_create(&h, device);
    if (rc) return rc;
    fvm_batch *b = new fvm_batch();
    b->h = h;
    b->h->prm.verbosity = 0;
    *out = b;
    return FVM_OK;
    FVM_API_END
}

int fvm_batch_destroy(fvm_batch *b) {
    FVM_API_BEGIN
    if (!b) return FVM_OK;
    fvm_amg *h = b->h;
    {
        HandleScope g(h);
        batch_invalidate(b);
        delete b;   // batch arrays go back to the pool of h, destroyed next
    }
    return fvm_amg_destroy(h);
    FVM_API_END
}

int fvm_batch_set_params(fvm_batch *b, const fvm_amg_params *p) {
    FVM_API_BEGIN
    FVM_CHECK(b, FVM_E_INVALID, "null handle");
    batch_check_params(p);
    const fvm_amg_params old = b->h->prm;
    const bool rebuild = b->h->ready && (p->coarseGroupSize != old.coarseGroupSize ||
                                         p->weightRatioThreshold != old.weightRatioThreshold ||
                                         p->maxCoarseLevels != old.maxCoarseLevels ||
                                         p->parallelStopRule != old.parallelStopRule ||
                                         p->smootherType != old.smootherType ||
                                         p->gsOrdering != old.gsOrdering);
    HandleScope g(b->h);
    b->h->prm = *p;
    batch_invalidate(b);   // the captured iteration depends on every cycle parameter
    if (rebuild) {
        b->h->lev.clear();
        b->h->ready = false;
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_batch_setup_device(fvm_batch *b, int nsys, const int32_t *n_self, const int32_t *n_total,
                           const int32_t *const *d_row, const int32_t *const *d_col, const double *const *d_diag,
                           const double *const *d_off, const uint8_t *const *d_is_boundary) {
    FVM_API_BEGIN
    FVM_CHECK(b && nsys > 0 && n_self && n_total && d_row && d_col && d_diag && d_off, FVM_E_INVALID,
              "null argument or empty batch");
    batch_check_params(&b->h->prm);
    HandleScope g(b->h);
    batch_invalidate(b);
    b->B.K = nsys;
    b->B.sys.resize(nsys);
    for (int k = 0; k < nsys; k++) {
        fvm::BatchSys &S = b->B.sys[k];
        S.row = d_row[k]; S.col = d_col[k]; S.diag = d_diag[k]; S.off = d_off[k];
        S.bnd = d_is_boundary ? d_is_boundary[k] : nullptr;
        S.n_self = n_self[k]; S.n_total = n_total[k];
    }
    fvm::build_hierarchy_batched(b->h, b->B);
    return FVM_OK;
    FVM_API_END
}

int fvm_batch_solve_device(fvm_batch *b, const double *const *d_b, double *const *d_x, double *rnorm0,
                           double *rnorm_final, int *cycles) {
    FVM_API_BEGIN
    FVM_CHECK(b && d_b && d_x, FVM_E_INVALID, "null argument");
    FVM_CHECK(b->h->ready, FVM_E_NOT_SETUP, "batched solve called before setup");
    batch_check_params(&b->h->prm);
    HandleScope g(b->h);
    fvm::batch_solve(b->h, b->B, d_b, d_x, rnorm0, rnorm_final, cycles);
    return FVM_OK;
    FVM_API_END
}

int fvm_batch_level_sizes(const fvm_batch *b, int sys, int32_t *sizes, int cap) {
    if (!b || sys < 0 || sys >= b->B.K || !b->h->ready) return FVM_E_INVALID;
    const int d = b->B.depth[sys];
    for (int l = 0; l < d && l < cap; l++) sizes[l] = b->B.off[l][sys + 1] - b->B.off[l][sys];
    return d;
}

int fvm_batch_set_tuning(fvm_batch *b, const char *key, long long value) {
    FVM_API_BEGIN
    FVM_CHECK(b && key, FVM_E_INVALID, "null argument");
    if (std::string(key) == "graph") {
        HandleScope g(b->h);
        batch_invalidate(b);
        b->B.use_graph = value != 0;
        return FVM_OK;
    }
    return fvm_amg_set_tuning(b->h, key, value);
    FVM_API_END
}

int fvm_batch_setup_stats(const fvm_batch *b, double *setup_seconds, int *total_rows, int *levels) {
    if (!b) return FVM_E_INVALID;
    if (setup_seconds) *setup_seconds = b->h->setup_seconds;
    if (total_rows) *total_rows = b->h->ready ? b->h->lev[0].n : 0;
    if (levels) *levels = (int)b->h->lev.size();
    return FVM_OK;
}

int fvm_fgmres_set_learned(fvm_fgmres *s, fvm_mpnn *m, double stall_ratio) {
    FVM_API_BEGIN
    FVM_CHECK(s, FVM_E_INVALID, "null handle");
    FVM_CHECK(!m || (stall_ratio > 0 && stall_ratio <= 1), FVM_E_INVALID, "stall_ratio must be in (0, 1]");
    s->nn = m;
    s->stall_ratio = stall_ratio;
    return FVM_OK;
    FVM_API_END
}

int fvm_fgmres_learned_stats(const fvm_fgmres *s, int *iterations_used, int *fallback_at) {
    if (!s) return FVM_E_INVALID;
    if (iterations_used) *iterations_used = s->nn_used;
    if (fallback_at) *fallback_at = s->nn_fallback_at;
    return FVM_OK;
}

// ----------------------------------------------------------------------- FGMRES
int fvm_fgmres_default_params(fvm_fgmres_params *p) {
    if (!p) return FVM_E_INVALID;
    p->restart = 30;
    p->nMaxIterations = 1000;
    p->verbosity = 0;
    p->relativeTolerance = 1e-8;
    p->absoluteTolerance = 1e-50;
    return FVM_OK;
}

int fvm_fgmres_create(fvm_fgmres **out, fvm_amg *pc) {
    FVM_API_BEGIN
    FVM_CHECK(out && pc, FVM_E_INVALID, "null argument");
    fvm_fgmres *s = new fvm_fgmres();
    s->pc = pc;
    s->device = pc->device;
    fvm_fgmres_default_params(&s->prm);
    s->work = fvm::new_krylov_work();
    *out = s;
    return FVM_OK;
    FVM_API_END
}

int fvm_fgmres_destroy(fvm_fgmres *s) {
    FVM_API_BEGIN
    if (!s) return FVM_OK;
    {
        // the Krylov arrays are plain allocations of s->device (not from the
        // preconditioner's pool): freeing them needs no access to pc
        DeviceGuard g(s->device);
        fvm::delete_krylov_work(s->work);
    }
    delete s;
    return FVM_OK;
    FVM_API_END
}

int fvm_fgmres_set_params(fvm_fgmres *s, const fvm_fgmres_params *p) {
    if (!s || !p || p->restart < 1 || p->nMaxIterations < 0) return FVM_E_INVALID;
    s->prm = *p;
    return FVM_OK;
}

int fvm_fgmres_solve_device(fvm_fgmres *s, const double *d_b, double *d_x, double *rnorm0, double *rnorm_final,
                            int *iterations) {
    FVM_API_BEGIN
    FVM_CHECK(s && s->pc, FVM_E_INVALID, "null solver");
    FVM_CHECK(s->pc->ready, FVM_E_NOT_SETUP, "preconditioner not set up");
    HandleScope g(s->pc);
    fvm::fgmres_solve(s->pc, *s->work, s->prm, d_b, d_x, rnorm0, rnorm_final, iterations, s->hist, hook_of(s));
    return FVM_OK;
    FVM_API_END
}

int fvm_fgmres_solve(fvm_fgmres *s, const double *b, double *x, double *rnorm0, double *rnorm_final,
                     int *iterations) {
    FVM_API_BEGIN
    FVM_CHECK(s && s->pc, FVM_E_INVALID, "null solver");
    fvm_amg *h = s->pc;
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "preconditioner not set up");
    HandleScope g(h);
    const size_t n = std::max(h->n_total, 1);
    h->h_stage_b.ensure(n);
    h->h_stage_x.ensure(n);
    hipStream_t st = h->stream;
    if (h->n_total > 0) {
        FVM_HIP(hipMemcpyAsync(h->h_stage_b.p, b, sizeof(double) * h->n_total, hipMemcpyHostToDevice, st));
        FVM_HIP(hipMemcpyAsync(h->h_stage_x.p, x, sizeof(double) * h->n_total, hipMemcpyHostToDevice, st));
    }
    fvm::fgmres_solve(h, *s->work, s->prm, h->h_stage_b.p, h->h_stage_x.p, rnorm0, rnorm_final, iterations,
                      s->hist, hook_of(s));
    if (h->n_self > 0) {
        FVM_HIP(hipMemcpyAsync(x, h->h_stage_x.p, sizeof(double) * h->n_self, hipMemcpyDeviceToHost, st));
        FVM_HIP(hipStreamSynchronize(st));
    }
    return FVM_OK;
    FVM_API_END
}

// ---------------------------------------------------------------------- BCGStab
int fvm_linsolver_default_params(fvm_linsolver_params *p) {
    if (!p) return FVM_E_INVALID;
    p->nMaxIterations = 100;
    p->verbosity = 0;
    p->relativeTolerance = 1e-8;
    p->absoluteTolerance = 1e-50;
    return FVM_OK;
}

namespace {
using KrylovFn = void (*)(fvm_amg *, const fvm_linsolver_params &, const double *, double *, double *, double *,
                          int *, std::vector<double> &);

int krylov_device(KrylovFn fn, fvm_amg *pc, const fvm_linsolver_params *p, const double *d_b, double *d_x,
                  double *rnorm0, double *rnorm_final, int *iterations, double *hist, int cap) {
    FVM_API_BEGIN
    FVM_CHECK(pc && p, FVM_E_INVALID, "null argument");
    FVM_CHECK(p->nMaxIterations >= 0 && cap >= 0 && (cap == 0 || hist), FVM_E_INVALID, "invalid argument");
    FVM_CHECK(pc->ready, FVM_E_NOT_SETUP, "preconditioner not set up");
    FVM_CHECK(pc->n_total == 0 || (d_b && d_x), FVM_E_INVALID, "null vector");
    HandleScope g(pc);
    std::vector<double> h;
    fn(pc, *p, d_b, d_x, rnorm0, rnorm_final, iterations, h);
    const int k = std::min<int>(cap, (int)h.size());
    for (int i = 0; i < k; i++) hist[i] = h[i];
    return k;
    FVM_API_END
}
}  // namespace

int fvm_bcgstab_solve_device(fvm_amg *pc, const fvm_linsolver_params *p, const double *d_b, double *d_x,
                             double *rnorm0, double *rnorm_final, int *iterations, double *hist, int cap) {
    return krylov_device(fvm::bcgstab_solve, pc, p, d_b, d_x, rnorm0, rnorm_final, iterations, hist, cap);
}

int fvm_cg_solve_device(fvm_amg *pc, const fvm_linsolver_params *p, const double *d_b, double *d_x,
                        double *rnorm0, double *rnorm_final, int *iterations, double *hist, int cap) {
    return krylov_device(fvm::cg_solve, pc, p, d_b, d_x, rnorm0, rnorm_final, iterations, hist, cap);
}

int fvm_fgmres_history(const fvm_fgmres *s, double *hist, int cap) {
    if (!s || (cap > 0 && !hist)) return FVM_E_INVALID;
    const int k = std::min<int>(cap, (int)s->hist.size());
    for (int i = 0; i < k; i++) hist[i] = s->hist[i];
    return k;
}

}  // extern "C"
