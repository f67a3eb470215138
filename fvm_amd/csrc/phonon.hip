// Gray phonon BTE model loop on the device: PhononModel::advance
// (src/fvm/src/modules/phononbase/PhononModel.h:530-600) around the batched AMG.
//
// Per outer iteration the reference loops over the modes (directions): it
// linearizes each one (collision + upwind convection + boundary conditions,
// PhononModel.h:446-523), solves it with the phonon linear solver (AMG::solve),
// adds the returned initial norm to the iteration's norm (PhononModel.h:557-572),
// then postSolve / updateSolution; after all directions: the initial norm (the
// largest of the first five iterations' norms), the stop test, then updateTL
// (Newton on sum_k e_k dk3 / tau_k = sum_k e0_k(TL) dk3 / tau_k, Kspace.h:425-441,
// :546-562), updatee0 (gray modes: e0 = cp (TL - Tref), pmode.h:146-165) and the
// temperature walls (PhononBoundary.h:165-210: ghost e = the cell's e for outgoing
// directions, e0(Twall) for incoming ones).
//
// The directions are independent within an iteration, so here they are
// assembled one after another on the device and solved together by one batched
// hierarchy (each system's result is the one its own AMG solve gives); the
// norm is summed in direction order on the host.
#include "amg.hpp"

#include <cmath>
#include <vector>

struct fvm_phonon {
    fvm_mesh *mesh = nullptr;          // borrowed
    int device = 0;
    hipStream_t stream = nullptr;
    int K = 0, nc = 0, nct = 0, ngroups = 0;
    std::vector<double> vg, dk3, tau, cp, tref;   // per direction (vg: 3 each)
    std::vector<double> wall_T;                    // per face group (entry 0 unused)
    std::vector<int> gptr;
    fvm::DArr<double> e, e0, delta, TL, coef;      // e, e0, delta: K x nct; coef: K x 4 (dk3, tau, cp, Tref)
    fvm::DArr<double> d_vg, d_wall;
    fvm::DArr<int> d_gptr;
    fvm_batch *batch = nullptr;
    fvm_amg_params prm{};
    int niters = 0;
    bool have_initial = false;
    double initial_norm = 0.0;
    std::vector<double> norms;
};

namespace fvm {
namespace {

__global__ void k_axpy(long long n, const double *d, double *x) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] += d[i];
}

// updateTL (PhononModel.h:302-330) for every cell, ghosts included: e_sum over the
// directions in order, then Kspace::NewtonSolve from the current TL
__global__ void k_update_TL(int nct, int K, const double *e, const double *coef, double *TL) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        double esum = 0.0;
        for (int k = 0; k < K; k++) esum += e[(size_t)k * nct + c] * coef[4 * k] / coef[4 * k + 1];
        double guess = TL[c];
        double deltaT = 1.;
        int iters = 0;
        while ((deltaT > 1e-6) && (iters < 10)) {
            double e0tau = 0., de0taudT = 0.;
            for (int k = 0; k < K; k++) {
                const double dk3 = coef[4 * k], tau = coef[4 * k + 1], cp = coef[4 * k + 2], tr = coef[4 * k + 3];
                e0tau += (guess - tr) * cp / tau * dk3;
                de0taudT += cp / tau * dk3;
            }
            deltaT = (e0tau - esum) / de0taudT;
            const double newguess = guess - deltaT;
            deltaT = fabs(deltaT / guess);
            guess = newguess;
            iters++;
        }
        TL[c] = guess;
    }
}

// updatee0 (PhononModel.h:357-382): e0_k = cp_k (TL - Tref_k)
__global__ void k_update_e0(int nct, int K, const double *TL, const double *coef, double *e0) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)K * nct;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nct), c = (int)(i % nct);
        e0[i] = (TL[c] - coef[4 * k + 3]) * coef[4 * k + 2];
    }
}

// temperature walls (PhononBoundary.h:165-202) on every boundary face and direction
__global__ void k_temperature_walls(int nc, int nct, int nf_int, int nf, int K, int ngroups, const int *gptr,
                                    const double *wall_T, const int *fcells, const double *fA, const double *fAm,
                                    const double *vg, const double *coef, double *e) {
    const long long nb = (long long)(nf - nf_int) * K;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nb; i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / (nf - nf_int));
        const int f = nf_int + (int)(i % (nf - nf_int));
        int g = 1;
        while (g + 1 < ngroups && f >= gptr[g + 1]) g++;
        const int c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
        const double en0 = fA[3 * f] / fAm[f], en1 = fA[3 * f + 1] / fAm[f], en2 = fA[3 * f + 2] / fAm[f];
        const double vdn = vg[3 * k] * en0 + vg[3 * k + 1] * en1 + vg[3 * k + 2] * en2;
        double *ek = e + (size_t)k * nct;
        if (vdn > 0.0) ek[c1] = ek[c0];
        else ek[c1] = (wall_T[g] - coef[4 * k + 3]) * coef[4 * k + 2];
    }
}

}  // namespace
}  // namespace fvm

#define FVM_API_BEGIN try {
#define FVM_API_END                                  \
    }                                                \
    catch (const fvm::Error &e) {                    \
        fvm::set_last_error(e.what());               \
        return e.code;                               \
    }                                                \
    catch (const std::exception &e) {                \
        fvm::set_last_error(e.what());               \
        return FVM_E_INVALID;                        \
    }

namespace {
struct DevGuard {
    int prev = -1;
    explicit DevGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        FVM_HIP(hipSetDevice(d));
    }
    ~DevGuard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

void rc_check(int rc) {
    if (rc != FVM_OK) throw fvm::Error(rc, fvm_last_error());
}
}  // namespace

extern "C" {

int fvm_phonon_create(fvm_phonon **out, fvm_mesh *mesh, int K, const double *vg, const double *dk3, const double *tau,
                      const double *cp, const double *tref, const double *wall_T, double T_init) {
    FVM_API_BEGIN
    FVM_CHECK(out && mesh && vg && dk3 && tau && cp && tref && wall_T, FVM_E_INVALID, "null argument");
    FVM_CHECK(K >= 1, FVM_E_INVALID, "need at least one direction");
    int nc, ng, nf, nn, ngr;
    rc_check(fvm_mesh_info(mesh, &nc, &ng, &nf, &nn, &ngr));
    for (int k = 0; k < K; k++)
        FVM_CHECK(std::isfinite(tau[k]) && tau[k] != 0.0 && std::isfinite(dk3[k]) && std::isfinite(cp[k]) &&
                      std::isfinite(tref[k]), FVM_E_INVALID, "direction parameters must be finite, tau non-zero");
    for (int g = 1; g < ngr; g++)
        FVM_CHECK(std::isfinite(wall_T[g]), FVM_E_INVALID, "every boundary group needs a wall temperature");
    FVM_CHECK(ngr >= 1 && ngr <= 64, FVM_E_INVALID, "bad face groups");
    const int dev = fvm::mesh_device(mesh);
    fvm_phonon *p = new fvm_phonon();
    try {
        p->mesh = mesh;
        p->K = K; p->nc = nc; p->nct = nc + ng; p->ngroups = ngr;
        p->vg.assign(vg, vg + 3 * (size_t)K);
        p->dk3.assign(dk3, dk3 + K); p->tau.assign(tau, tau + K);
        p->cp.assign(cp, cp + K); p->tref.assign(tref, tref + K);
        p->wall_T.assign(wall_T, wall_T + ngr);
        p->gptr.resize(ngr + 1);
        rc_check(fvm_mesh_get(mesh, nullptr, nullptr, nullptr, p->gptr.data()));
        p->device = dev;
        DevGuard d(dev);
        FVM_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamDefault));
        const size_t tot = (size_t)K * p->nct;
        p->e.alloc(std::max<size_t>(tot, 1)); p->e0.alloc(std::max<size_t>(tot, 1));
        p->delta.alloc(std::max<size_t>(tot, 1));
        p->TL.alloc(std::max(p->nct, 1));
        std::vector<double> coef(4 * (size_t)K);
        for (int k = 0; k < K; k++) {
            coef[4 * k] = dk3[k]; coef[4 * k + 1] = tau[k]; coef[4 * k + 2] = cp[k]; coef[4 * k + 3] = tref[k];
        }
        p->coef.alloc(4 * (size_t)K);
        p->d_vg.alloc(3 * (size_t)K);
        p->d_wall.alloc(std::max(ngr, 1));
        p->d_gptr.alloc((size_t)ngr + 1);
        FVM_HIP(hipMemcpy(p->coef.p, coef.data(), sizeof(double) * 4 * K, hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(p->d_vg.p, vg, sizeof(double) * 3 * K, hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(p->d_wall.p, wall_T, sizeof(double) * ngr, hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(p->d_gptr.p, p->gptr.data(), sizeof(int) * (ngr + 1), hipMemcpyHostToDevice));
        // PhononModel::init (PhononModel.h:115-160): TL = T_init and e = e0 = e0(T_init)
        // on every cell, ghosts included; the walls act from the first
        // callBoundaryConditions (fvm_phonon_apply_walls, or after an iteration)
        std::vector<double> tl(p->nct, T_init);
        FVM_HIP(hipMemcpy(p->TL.p, tl.data(), sizeof(double) * p->nct, hipMemcpyHostToDevice));
        const unsigned g = fvm::grid_for((long long)tot, 256);
        fvm::k_update_e0<<<g, 256, 0, p->stream>>>(p->nct, K, p->TL.p, p->coef.p, p->e0.p);
        FVM_HIP(hipMemcpyAsync(p->e.p, p->e0.p, sizeof(double) * tot, hipMemcpyDeviceToDevice, p->stream));
        FVM_HIP(hipStreamSynchronize(p->stream));
        rc_check(fvm_amg_default_params(&p->prm));
        p->prm.verbosity = 0;
        rc_check(fvm_batch_create(&p->batch, dev));
    } catch (...) {
        if (p->batch) fvm_batch_destroy(p->batch);
        if (p->stream) (void)hipStreamDestroy(p->stream);
        delete p;
        throw;
    }
    *out = p;
    return FVM_OK;
    FVM_API_END
}

int fvm_phonon_destroy(fvm_phonon *p) {
    FVM_API_BEGIN
    if (!p) return FVM_OK;
    {
        DevGuard d(p->device);
        (void)hipStreamSynchronize(p->stream);
        if (p->batch) fvm_batch_destroy(p->batch);
        p->e.release(); p->e0.release(); p->delta.release(); p->TL.release(); p->coef.release();
        p->d_vg.release(); p->d_wall.release(); p->d_gptr.release();
        (void)hipStreamDestroy(p->stream);
    }
    delete p;
    return FVM_OK;
    FVM_API_END
}

int fvm_phonon_set_params(fvm_phonon *p, const fvm_amg_params *prm) {
    FVM_API_BEGIN
    FVM_CHECK(p && prm, FVM_E_INVALID, "null argument");
    p->prm = *prm;
    return FVM_OK;
    FVM_API_END
}

// apply the temperature walls to the ghost cells of every direction
static void phonon_walls(fvm_phonon *p) {
    int nc, ng, nf, nn, ngr;
    rc_check(fvm_mesh_info(p->mesh, &nc, &ng, &nf, &nn, &ngr));
    const double *fA, *fAm;
    const int *fcells;
    fvm::mesh_device_metrics(p->mesh, &fcells, &fA, &fAm);
    const int nf_int = p->gptr[1];
    const long long nb = (long long)(nf - nf_int) * p->K;
    if (nb > 0)
        fvm::k_temperature_walls<<<fvm::grid_for(nb, 256), 256, 0, p->stream>>>(
            p->nc, p->nct, nf_int, nf, p->K, p->ngroups, p->d_gptr.p, p->d_wall.p, fcells, fA, fAm, p->d_vg.p,
            p->coef.p, p->e.p);
    FVM_HIP(hipGetLastError());
}

int fvm_phonon_apply_walls(fvm_phonon *p) {
    FVM_API_BEGIN
    FVM_CHECK(p, FVM_E_INVALID, "null handle");
    DevGuard d(p->device);
    phonon_walls(p);
    FVM_HIP(hipStreamSynchronize(p->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_phonon_advance(fvm_phonon *p, int niter, double rel_tol, double abs_tol, int *iterations, double *norms,
                       int cap, int *converged) {
    FVM_API_BEGIN
    FVM_CHECK(p && niter >= 0 && (cap == 0 || norms), FVM_E_INVALID, "bad argument");
    DevGuard d(p->device);
    const int K = p->K, nct = p->nct;
    rc_check(fvm_batch_set_params(p->batch, &p->prm));
    std::vector<fvm_system *> sys(K, nullptr);
    auto free_systems = [&]() {
        for (auto &S : sys)
            if (S) { fvm_system_destroy(S); S = nullptr; }
    };
    int done = 0, conv = 0;
    try {
        std::vector<int32_t> ns(K), nt(K);
        std::vector<const int32_t *> rows(K), cols(K);
        std::vector<const double *> diags(K), offs(K), bs(K);
        std::vector<const uint8_t *> bnds(K);
        std::vector<double *> xs(K);
        std::vector<double> r0(K), rf(K);
        std::vector<int> cyc(K);
        for (int n = 0; n < niter; n++) {
            FVM_HIP(hipStreamSynchronize(p->stream));   // e, e0 of the previous iteration are final
            for (int k = 0; k < K; k++) {
                // PhononModel::linearizePhononModel + LinearSystem::initSolve for direction k
                rc_check(fvm_phonon_assemble(p->mesh, &p->vg[3 * (size_t)k], p->tau[k], p->e0.p + (size_t)k * nct,
                                             p->e.p + (size_t)k * nct, &sys[k]));
                int s_n, s_t;
                long long nz;
                rc_check(fvm_system_info(sys[k], &s_n, &s_t, &nz));
                ns[k] = s_n; nt[k] = s_t;
                const double *bb;
                rc_check(fvm_system_device(sys[k], &rows[k], &cols[k], &diags[k], &offs[k], &bb, &bnds[k]));
                bs[k] = bb;
                xs[k] = p->delta.p + (size_t)k * nct;
            }
            FVM_HIP(hipMemsetAsync(p->delta.p, 0, sizeof(double) * (size_t)K * nct, p->stream));
            FVM_HIP(hipStreamSynchronize(p->stream));
            rc_check(fvm_batch_setup_device(p->batch, K, ns.data(), nt.data(), rows.data(), cols.data(), diags.data(),
                                            offs.data(), bnds.data()));
            rc_check(fvm_batch_solve_device(p->batch, bs.data(), xs.data(), r0.data(), rf.data(), cyc.data()));
            // the iteration's norm: the directions' initial norms summed in order
            double rnorm = r0[0];
            for (int k = 1; k < K; k++) rnorm += r0[k];
            for (int k = 0; k < K; k++) {
                rc_check(fvm_system_post_solve(sys[k], xs[k]));      // postSolve
            }
            free_systems();
            // updateSolution: e += delta (all directions)
            fvm::k_axpy<<<fvm::grid_for((long long)K * nct, 256), 256, 0, p->stream>>>((long long)K * nct, p->delta.p,
                                                                                     p->e.p);
            FVM_HIP(hipGetLastError());
            if (!p->have_initial) { p->initial_norm = rnorm; p->have_initial = true; }
            if (p->niters < 5) p->initial_norm = std::max(p->initial_norm, rnorm);   // setMax
            const double ratio = rnorm / p->initial_norm;
            p->norms.push_back(rnorm);
            if (done < cap) norms[done] = rnorm;
            p->niters++;
            done++;
            if (rnorm < abs_tol || ratio < rel_tol) { conv = 1; break; }
            fvm::k_update_TL<<<fvm::grid_for(nct, 256), 256, 0, p->stream>>>(nct, K, p->e.p, p->coef.p, p->TL.p);
            fvm::k_update_e0<<<fvm::grid_for((long long)K * nct, 256), 256, 0, p->stream>>>(nct, K, p->TL.p,
                                                                                          p->coef.p, p->e0.p);
            phonon_walls(p);
        }
        FVM_HIP(hipStreamSynchronize(p->stream));
    } catch (...) {
        free_systems();
        throw;
    }
    if (iterations) *iterations = done;
    if (converged) *converged = conv;
    return FVM_OK;
    FVM_API_END
}

int fvm_phonon_get(const fvm_phonon *p, double *TL, double *e, double *e0) {
    FVM_API_BEGIN
    FVM_CHECK(p, FVM_E_INVALID, "null handle");
    DevGuard d(p->device);
    FVM_HIP(hipStreamSynchronize(p->stream));
    const size_t tot = (size_t)p->K * p->nct;
    if (TL) FVM_HIP(hipMemcpy(TL, p->TL.p, sizeof(double) * p->nct, hipMemcpyDeviceToHost));
    if (e) FVM_HIP(hipMemcpy(e, p->e.p, sizeof(double) * tot, hipMemcpyDeviceToHost));
    if (e0) FVM_HIP(hipMemcpy(e0, p->e0.p, sizeof(double) * tot, hipMemcpyDeviceToHost));
    return FVM_OK;
    FVM_API_END
}

int fvm_phonon_set_field(fvm_phonon *p, const double *e) {
    FVM_API_BEGIN
    FVM_CHECK(p && e, FVM_E_INVALID, "null argument");
    DevGuard d(p->device);
    FVM_HIP(hipMemcpy(p->e.p, e, sizeof(double) * (size_t)p->K * p->nct, hipMemcpyHostToDevice));
    return FVM_OK;
    FVM_API_END
}

}  // extern "C"
