// BGK kinetic model loop on the device: KineticModel::advance
// (src/fvm/src/modules/esbgkbase/KineticModel.h:3360-3462) around the batched AMG.
//
// The distribution function is discretised on a velocity quadrature (Quadrature.h);
// each discrete velocity c_j is one scalar finite-volume system: collision
// V nu (f - feq) plus upwind convection with the face flux A . c_j, with the
// kinetic boundary faces (KineticModel.h:1505-1674). Per outer iteration the
// reference linearizes and solves every direction (AMG::solve), adds its initial
// norm to the iteration's norm, then postSolve / updateSolution; after all
// directions: the initial norm (the largest of the first five iterations' norms),
// then always the boundary values (callBoundaryConditions: diffuse walls, zero
// gradient, specular symmetry), the moments (ComputeMacroparameters), the
// collision frequency and the equilibrium (fgamma 0: the Maxwellian of the stored
// coefficients around the new velocity; 1: the BGK equilibrium whose coefficients
// a per-cell Newton solve fits to the cell's discrete moments; 2: that, then the
// ES-BGK equilibrium fitted to the temperature tensor, which the collision term then
// relaxes to), then the stop test.
//
// The directions are independent within an iteration: here they are assembled
// one after another on the device and solved together by one batched hierarchy
// (each system's result is the one its own AMG solve gives); the norm is summed in
// direction order on the host. The per-cell kernels (moments, collision frequency,
// Newton, equilibrium) and the per-face boundary kernels keep the reference's
// operation order per cell and per face.
#include "amg.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

struct fvm_kinetic {
    fvm_mesh *mesh = nullptr;          // borrowed
    int device = 0;
    hipStream_t stream = nullptr;
    int K = 0, nc = 0, nct = 0, ngroups = 0;
    fvm_kinetic_options opt{};
    std::vector<double> cx, cy, cz, w;
    std::vector<int> gptr, kind;       // kind per group: 0 diffuse wall, 1 zero gradient, 2 specular symmetry
    std::vector<double> wall_u, wall_T;
    fvm::DArr<double> d_q;             // cx, cy, cz, w (K each)
    fvm::DArr<double> f, feq, delta;   // K x nct, direction-major
    fvm::DArr<double> rho, vel, T, nu, coeff;   // vel: 3 x nct; coeff: 5 x nct (component-major)
    fvm::DArr<double> coeffg, T6, feqes;        // ES-BGK: 10 x nct, 6 x nct (Txx Tyy Tzz Txy Tyz Tzx), K x nct
    fvm::DArr<int> d_grp, d_refl;      // gptr + kind; reflection tables (K per group, -1 unused)
    fvm::DArr<double> d_wall;          // uw (3) and Tw per group
    fvm::DArr<int> d_agrp;             // the assembly's groups: gptr + kind (0 wall-like, 1 zero gradient)
    fvm::DArr<double> d_awall;         // the assembly's wall velocities (3 per group; zero for symmetry)
    fvm::DArr<double> d_F;             // assembly scratch: face fluxes
    fvm::DArr<uint8_t> d_inc;          // assembly scratch: incoming flags
    std::vector<fvm_system *> sys;     // one system per velocity, kept from iteration to iteration
    fvm_batch *batch = nullptr;
    fvm_amg_params prm{};
    int niters = 0;
    bool have_initial = false;
    double initial_norm = 0.0;
    std::vector<double> norms;
};

namespace fvm {
namespace {

constexpr double KPI = 3.141592653589793238462643383279502884;   // acos(-1) in double

__global__ void k_axpy_kin(long long n, const double *d, double *x) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] += d[i];
}

// initializeMaxwellian (KineticModel.h:1216-1240) from the rest state of
// InitializeMacroparameters (:349-390): density 1, velocity 0, temperature 1
__global__ void k_kin_init(int nct, int K, const double *q, double *f, double *rho, double *vel, double *T,
                           double *coeff) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)K * nct;
         i += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(i / nct), c = (int)(i % nct);
        const double d = 1.0, t = 1.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
        const double a = cx[j] - u0, b = cy[j] - u1, e = cz[j] - u2;
        f[i] = d / pow(KPI * t, 1.5) * exp(-(a * a + b * b + e * e) / t);
        if (j == 0) {
            rho[c] = d; T[c] = t;
            vel[c] = u0; vel[nct + c] = u1; vel[2 * (size_t)nct + c] = u2;
            coeff[c] = d / pow(KPI * t, 1.5);
            coeff[nct + c] = 1 / t;
            coeff[2 * (size_t)nct + c] = 0.0; coeff[3 * (size_t)nct + c] = 0.0; coeff[4 * (size_t)nct + c] = 0.0;
        }
    }
}

// ComputeMacroparameters (KineticModel.h:471-548) for every cell, ghosts included
__global__ void k_kin_moments(int nct, int K, const double *q, const double *f, double *rho, double *vel,
                              double *T) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K, *w = q + 3 * K;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        double d = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, t = 0.0;
        for (int j = 0; j < K; j++) {
            const double fj = f[(size_t)j * nct + c];
            d = d + w[j] * fj;
            v0 = v0 + (cx[j] * fj) * w[j];
            v1 = v1 + (cy[j] * fj) * w[j];
            v2 = v2 + (cz[j] * fj) * w[j];
            t = t + (cx[j] * cx[j] + cy[j] * cy[j] + cz[j] * cz[j]) * fj * w[j];
        }
        v0 = v0 / d;
        v1 = v1 / d;
        v2 = v2 / d;
        t = t - (v0 * v0 + v1 * v1 + v2 * v2) * d;
        t = t / (1.5 * d);
        rho[c] = d; T[c] = t;
        vel[c] = v0; vel[nct + c] = v1; vel[2 * (size_t)nct + c] = v2;
    }
}

// ComputeCollisionfrequency (KineticModel.h:645-677)
__global__ void k_kin_collision(int nct, const double *rho, const double *T, double T_init, double Tmuref,
                                double muref, double mu_w, double mu0, double prandtl, int fgamma, double *nu) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        const double visc = muref * pow(T[c] * T_init / Tmuref, mu_w);
        double v = rho[c] * T[c] / visc * mu0;
        if (fgamma == 2) v = prandtl * v;
        nu[c] = v;
    }
}

// inverseGauss (fvmbase/MatrixOperation.h:200-290) on an N x N matrix held by one
// thread: scaled partial pivoting, multipliers kept below the diagonal, then the
// unit right-hand sides eliminated and back-substituted
template <int N>
__device__ void inverse_gauss(double a[N][N], double x[N][N]) {
    int indx[N];
    double cs[N], b[N][N];
    for (int i = 0; i < N; i++) {
        indx[i] = i;
        for (int j = 0; j < N; j++) { b[i][j] = i == j ? 1.0 : 0.0; x[i][j] = 0.0; }
    }
    for (int i = 0; i < N; i++) {
        double c1 = 0;
        for (int j = 0; j < N; j++)
            if (fabs(a[i][j]) > c1) c1 = fabs(a[i][j]);
        cs[i] = c1;
    }
    for (int j = 0; j < N - 1; j++) {
        double pi1 = 0.0;
        int pk = j;
        for (int i = j; i < N; i++) {
            const double pi = fabs(a[indx[i]][j]) / cs[indx[i]];
            if (pi > pi1) { pi1 = pi; pk = i; }
        }
        const int itmp = indx[j];
        indx[j] = indx[pk];
        indx[pk] = itmp;
        for (int i = j + 1; i < N; i++) {
            const double pj = a[indx[i]][j] / a[indx[j]][j];
            a[indx[i]][j] = pj;
            for (int k = j + 1; k < N; k++) a[indx[i]][k] = a[indx[i]][k] - pj * a[indx[j]][k];
        }
    }
    for (int i = 0; i < N - 1; i++)
        for (int j = i + 1; j < N; j++)
            for (int k = 0; k < N; k++) b[indx[j]][k] = b[indx[j]][k] - a[indx[j]][i] * b[indx[i]][k];
    for (int col = 0; col < N; col++) {
        x[N - 1][col] = b[indx[N - 1]][col] / a[indx[N - 1]][N - 1];
        for (int j = N - 2; j >= 0; j = j - 1) {
            x[j][col] = b[indx[j]][col];
            for (int k = j + 1; k < N; k++) x[j][col] = x[j][col] - a[indx[j]][k] * x[k][col];
            x[j][col] = x[j][col] / a[indx[j]][j];
        }
    }
}

// NewtonsMethodBGK + setJacobianBGK (KineticModel.h:829-905, 957-990), one cell per thread
__global__ void k_kin_newton_bgk(int nct, int K, const double *q, const double *rho, const double *vel,
                                 const double *T, int ktrial, double tolx, double tolf, double *coeff) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K, *w = q + 3 * K;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        const double d = rho[c], u0 = vel[c], u1 = vel[nct + c], u2 = vel[2 * (size_t)nct + c], t = T[c];
        double xn[5];
        for (int r = 0; r < 5; r++) xn[r] = coeff[(size_t)r * nct + c];
        for (int trial = 0; trial < ktrial; trial++) {
            double fjac[5][5], fvec[5];
            for (int r = 0; r < 5; r++)
                for (int k = 0; k < 5; k++) fjac[r][k] = 0.0;
            fvec[0] = d;
            fvec[1] = d * u0;
            fvec[2] = d * u1;
            fvec[3] = d * u2;
            fvec[4] = 1.5 * d * t + d * (u0 * u0 + u1 * u1 + u2 * u2);
            for (int j = 0; j < K; j++) {
                const double a0 = cx[j] - u0, a1 = cy[j] - u1, a2 = cz[j] - u2;
                const double Cc = a0 * a0 + a1 * a1 + a2 * a2;
                const double E = xn[0] * exp(-xn[1] * Cc + xn[2] * a0 + xn[3] * a1 + xn[4] * a2) * w[j];
                const double m[5] = {1.0, cx[j], cy[j], cz[j], cx[j] * cx[j] + cy[j] * cy[j] + cz[j] * cz[j]};
                for (int r = 0; r < 5; r++) fvec[r] += -E * m[r];
                const double me[5] = {-E / xn[0], E * Cc, -E * a0, -E * a1, -E * a2};
                for (int r = 0; r < 5; r++)
                    for (int k = 0; k < 5; k++) fjac[r][k] += m[r] * me[k];
            }
            double errf = 0.;
            for (int r = 0; r < 5; r++) errf += fabs(fvec[r]);
            if (errf <= tolf) break;
            double inv[5][5];
            inverse_gauss<5>(fjac, inv);
            double errx = 0.;
            for (int r = 0; r < 5; r++) {
                double xr = 0.0;
                for (int k = 0; k < 5; k++) xr += inv[r][k] * -fvec[k];
                errx += fabs(xr);
                xn[r] += xr;
            }
            if (errx <= tolx) break;
        }
        for (int r = 0; r < 5; r++) coeff[(size_t)r * nct + c] = xn[r];
    }
}

// initializeMaxwellianEq / the BGK equilibrium (KineticModel.h:796-803, 944-951)
__global__ void k_kin_equilibrium(int nct, int K, const double *q, const double *coeff, const double *vel,
                                  double *feq) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)K * nct;
         i += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(i / nct), c = (int)(i % nct);
        const double a0 = cx[j] - vel[c], a1 = cy[j] - vel[nct + c], a2 = cz[j] - vel[2 * (size_t)nct + c];
        feq[i] = coeff[c] * exp(-coeff[nct + c] * (a0 * a0 + a1 * a1 + a2 * a2) + coeff[2 * (size_t)nct + c] * a0 +
                                coeff[3 * (size_t)nct + c] * a1 + coeff[4 * (size_t)nct + c] * a2);
    }
}

// the ES-BGK equilibrium's exponent (KineticModel.h:1151-1161, 1196-1200)
__device__ __forceinline__ double es_exponent(const double *xn, double cxj, double cyj, double czj, double c1,
                                              double c2, double c3) {
    return -xn[1] * (c1 * c1) + xn[2] * c1 - xn[3] * (c2 * c2) + xn[4] * c2 - xn[5] * (c3 * c3) + xn[6] * c3 +
           xn[7] * cxj * cyj + xn[8] * cyj * czj + xn[9] * czj * cxj;
}

// ComputeMacroparametersESBGK (KineticModel.h:565-631): the temperature tensor of
// (1 - 1/Pr) f + 1/Pr feq (feq the BGK equilibrium), per cell
__global__ void k_kin_tensor(int nct, int K, const double *q, const double *f, const double *feq, const double *rho,
                             const double *vel, double Pr, double *T6) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K, *w = q + 3 * K;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        const double u0 = vel[c], u1 = vel[nct + c], u2 = vel[2 * (size_t)nct + c];
        double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int j = 0; j < K; j++) {
            const double g = (1 - 1 / Pr) * f[(size_t)j * nct + c] + 1 / Pr * feq[(size_t)j * nct + c];
            const double d0 = cx[j] - u0, d1 = cy[j] - u1, d2 = cz[j] - u2;
            t[0] = t[0] + d0 * d0 * g * w[j];
            t[1] = t[1] + d1 * d1 * g * w[j];
            t[2] = t[2] + d2 * d2 * g * w[j];
            t[3] = t[3] + d0 * d1 * g * w[j];
            t[4] = t[4] + d1 * d2 * g * w[j];
            t[5] = t[5] + d2 * d0 * g * w[j];
        }
        for (int a = 0; a < 6; a++) T6[(size_t)a * nct + c] = t[a] / rho[c];
    }
}

// NewtonsMethodESBGK + setJacobianESBGK (KineticModel.h:1019-1108, 1166-1213), one cell per thread
__global__ void k_kin_newton_esbgk(int nct, int K, const double *q, const double *rho, const double *vel,
                                   const double *T6, int ktrial, double tolx, double tolf, double *coeffg) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K, *w = q + 3 * K;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        const double d = rho[c], u0 = vel[c], u1 = vel[nct + c], u2 = vel[2 * (size_t)nct + c];
        double Tt[6];
        for (int a = 0; a < 6; a++) Tt[a] = T6[(size_t)a * nct + c];
        double xn[10];
        for (int r = 0; r < 10; r++) xn[r] = coeffg[(size_t)r * nct + c];
        for (int trial = 0; trial < ktrial; trial++) {
            double fjac[10][10], fvec[10];
            for (int r = 0; r < 10; r++)
                for (int k = 0; k < 10; k++) fjac[r][k] = 0.0;
            fvec[0] = d;
            fvec[1] = d * u0;
            fvec[2] = d * u1;
            fvec[3] = d * u2;
            fvec[4] = d * (u0 * u0 + Tt[0]);
            fvec[5] = d * (u1 * u1 + Tt[1]);
            fvec[6] = d * (u2 * u2 + Tt[2]);
            fvec[7] = d * (u0 * u1 + Tt[3]);
            fvec[8] = d * (u1 * u2 + Tt[4]);
            fvec[9] = d * (u2 * u0 + Tt[5]);
            for (int j = 0; j < K; j++) {
                const double c1 = cx[j] - u0, c2 = cy[j] - u1, c3 = cz[j] - u2;
                const double E = xn[0] * exp(es_exponent(xn, cx[j], cy[j], cz[j], c1, c2, c3)) * w[j];
                const double m[10] = {1.0, cx[j], cy[j], cz[j], cx[j] * cx[j], cy[j] * cy[j], cz[j] * cz[j],
                                      cx[j] * cy[j], cy[j] * cz[j], cz[j] * cx[j]};
                for (int r = 0; r < 10; r++) fvec[r] += -E * m[r];
                const double me[10] = {-E / xn[0], E * (c1 * c1), -E * c1, E * (c2 * c2), -E * c2, E * (c3 * c3),
                                       -E * c3, -E * cx[j] * cy[j], -E * cy[j] * cz[j], -E * cz[j] * cx[j]};
                for (int r = 0; r < 10; r++)
                    for (int k = 0; k < 10; k++) fjac[r][k] += m[r] * me[k];
            }
            double errf = 0.;
            for (int r = 0; r < 10; r++) errf += fabs(fvec[r]);
            if (errf <= tolf) break;
            double inv[10][10];
            inverse_gauss<10>(fjac, inv);
            double errx = 0.;
            for (int r = 0; r < 10; r++) {
                double xr = 0.0;
                for (int k = 0; k < 10; k++) xr += inv[r][k] * -fvec[k];
                errx += fabs(xr);
                xn[r] += xr;
            }
            if (errx <= tolx) break;
        }
        for (int r = 0; r < 10; r++) coeffg[(size_t)r * nct + c] = xn[r];
    }
}

// the ES-BGK equilibrium (EquilibriumDistributionESBGK, KineticModel.h:1146-1162)
__global__ void k_kin_equilibrium_es(int nct, int K, const double *q, const double *coeffg, const double *vel,
                                     double *feqes) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)K * nct;
         i += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(i / nct), c = (int)(i % nct);
        double xn[10];
        for (int r = 0; r < 10; r++) xn[r] = coeffg[(size_t)r * nct + c];
        const double c1 = cx[j] - vel[c], c2 = cy[j] - vel[nct + c], c3 = cz[j] - vel[2 * (size_t)nct + c];
        feqes[i] = xn[0] * exp(es_exponent(xn, cx[j], cy[j], cz[j], c1, c2, c3));
    }
}

// InitializeMacroparameters' ES-BGK coefficients (KineticModel.h:392-402)
__global__ void k_kin_init_es(int nct, const double *coeff, double *coeffg) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        const double a0 = coeff[c], a1 = coeff[nct + c], a2 = coeff[2 * (size_t)nct + c],
                     a3 = coeff[3 * (size_t)nct + c], a4 = coeff[4 * (size_t)nct + c];
        const double g[10] = {a0, a1, a2, a1, a3, a1, a4, 0.0, 0.0, 0.0};
        for (int r = 0; r < 10; r++) coeffg[(size_t)r * nct + c] = g[r];
    }
}

__device__ __forceinline__ int face_group(int f, int ngroups, const int *gptr) {
    int g = 1;
    while (g + 1 < ngroups && f >= gptr[g + 1]) g++;
    return g;
}

// callBoundaryConditions (KineticModel.h:3271-3330), one boundary face per thread:
// diffuse wall (KineticBoundaryConditions.h:65-140), zero gradient (:385-406),
// specular symmetry with the group's reflection table (:287-330). Each face
// reads its interior cell and writes its own ghost cell only.
__global__ void k_kin_walls(int nct, int K, int first_bface, int nf, int ngroups, const int *grp,
                            const int *refl, const double *wall, const int *fcells, const double *fA,
                            const double *fAm, const double *q, double eps, double *f, double *rho, double *vel,
                            double *T) {
    const double *cx = q, *cy = q + K, *cz = q + 2 * K, *w = q + 3 * K;
    for (int fi = first_bface + blockIdx.x * blockDim.x + threadIdx.x; fi < nf; fi += gridDim.x * blockDim.x) {
        const int g = face_group(fi, ngroups, grp);
        const int kind = grp[ngroups + 1 + g];
        const int c0 = fcells[2 * fi], c1 = fcells[2 * fi + 1];
        const double e0 = fA[3 * fi] / fAm[fi], e1 = fA[3 * fi + 1] / fAm[fi], e2 = fA[3 * fi + 2] / fAm[fi];
        if (kind == 3) {
            // a group without a boundary-value rule ("CopyBC" in the reference's
            // Couette script): callBoundaryConditions leaves its ghosts alone
        } else if (kind == 1) {
            for (int j = 0; j < K; j++) f[(size_t)j * nct + c1] = f[(size_t)j * nct + c0];
        } else if (kind == 2) {
            const int *tab = refl + (size_t)g * K;
            for (int j = 0; j < K; j++) {
                const double cn = cx[j] * e0 + cy[j] * e1 + cz[j] * e2;
                f[(size_t)j * nct + c1] = cn < 0.0 ? f[(size_t)tab[j] * nct + c0] : f[(size_t)j * nct + c0];
            }
        } else {
            const double uw = wall[4 * g], vw = wall[4 * g + 1], ww = wall[4 * g + 2], Tw = wall[4 * g + 3];
            vel[c1] = uw; vel[nct + c1] = vw; vel[2 * (size_t)nct + c1] = ww;
            T[c1] = Tw;
            double Nmr = 0.0, Dmr = 0.0;
            const double un = uw * e0 + vw * e1 + ww * e2;
            for (int j = 0; j < K; j++) {
                const double cn = cx[j] * e0 + cy[j] * e1 + cz[j] * e2;
                const double a0 = cx[j] - uw, a1 = cy[j] - vw, a2 = cz[j] - ww;
                const double fwall = 1.0 / pow(KPI * Tw, 1.5) * exp(-(a0 * a0 + a1 * a1 + a2 * a2) / Tw);
                if (cn - un < eps) Dmr = Dmr - fwall * w[j] * (cn - un);
                else Nmr = Nmr + f[(size_t)j * nct + c0] * w[j] * (cn - un);
            }
            const double nwall = Nmr / Dmr;
            rho[c1] = nwall;
            for (int j = 0; j < K; j++) {
                const double cn = cx[j] * e0 + cy[j] * e1 + cz[j] * e2;
                if (cn - un < eps) {
                    const double a0 = cx[j] - uw, a1 = cy[j] - vw, a2 = cz[j] - ww;
                    f[(size_t)j * nct + c1] = nwall / pow(KPI * Tw, 1.5) * exp(-(a0 * a0 + a1 * a1 + a2 * a2) / Tw);
                } else {
                    f[(size_t)j * nct + c1] = f[(size_t)j * nct + c0];
                }
            }
        }
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
struct KDevGuard {
    int prev = -1;
    explicit KDevGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        FVM_HIP(hipSetDevice(d));
    }
    ~KDevGuard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

void krc(int rc) {
    if (rc != FVM_OK) throw fvm::Error(rc, fvm_last_error());
}

// boundary values, moments, collision frequency and equilibrium: the part of
// KineticModel::advance after the directions' updates (KineticModel.h:3440-3452)
void kinetic_update(fvm_kinetic *p) {
    int nc, ng, nf, nn, ngr;
    krc(fvm_mesh_info(p->mesh, &nc, &ng, &nf, &nn, &ngr));
    const double *fA, *fAm;
    const int *fcells;
    fvm::mesh_device_metrics(p->mesh, &fcells, &fA, &fAm);
    const int first_bface = p->gptr.size() > 1 ? p->gptr[1] : nf;
    hipStream_t s = p->stream;
    if (nf > first_bface)
        fvm::k_kin_walls<<<fvm::grid_for(nf - first_bface, 64), 64, 0, s>>>(
            p->nct, p->K, first_bface, nf, p->ngroups, p->d_grp.p, p->d_refl.p, p->d_wall.p, fcells, fA, fAm,
            p->d_q.p, p->opt.epsilon, p->f.p, p->rho.p, p->vel.p, p->T.p);
    fvm::k_kin_moments<<<fvm::grid_for(p->nct, 256), 256, 0, s>>>(p->nct, p->K, p->d_q.p, p->f.p, p->rho.p, p->vel.p,
                                                                  p->T.p);
    const fvm_kinetic_options &o = p->opt;
    const double R = 8314.0 / o.molecularWeight;
    const double mu0 = o.rho_init * R * o.T_init * o.nonDimLt / pow(2 * R * o.T_init, 0.5);
    fvm::k_kin_collision<<<fvm::grid_for(p->nct, 256), 256, 0, s>>>(p->nct, p->rho.p, p->T.p, o.T_init, o.Tmuref,
                                                                    o.muref, o.mu_w, mu0, o.Prandtl, o.fgamma,
                                                                    p->nu.p);
    if (o.fgamma >= 1)
        fvm::k_kin_newton_bgk<<<fvm::grid_for(p->nct, 64), 64, 0, s>>>(p->nct, p->K, p->d_q.p, p->rho.p, p->vel.p,
                                                                      p->T.p, o.ktrial, o.ToleranceX, o.ToleranceF,
                                                                      p->coeff.p);
    const long long tot = (long long)p->K * p->nct;
    fvm::k_kin_equilibrium<<<fvm::grid_for(tot, 256), 256, 0, s>>>(p->nct, p->K, p->d_q.p, p->coeff.p, p->vel.p,
                                                                   p->feq.p);
    if (o.fgamma == 2) {
        // EquilibriumDistributionESBGK (KineticModel.h:1112-1163): the tensor of the
        // new f and BGK equilibrium, the ten-coefficient Newton fit, the equilibrium
        fvm::k_kin_tensor<<<fvm::grid_for(p->nct, 256), 256, 0, s>>>(p->nct, p->K, p->d_q.p, p->f.p, p->feq.p,
                                                                     p->rho.p, p->vel.p, o.Prandtl, p->T6.p);
        fvm::k_kin_newton_esbgk<<<fvm::grid_for(p->nct, 64), 64, 0, s>>>(p->nct, p->K, p->d_q.p, p->rho.p,
                                                                        p->vel.p, p->T6.p, o.ktrial, o.ToleranceX,
                                                                        o.ToleranceF, p->coeffg.p);
        fvm::k_kin_equilibrium_es<<<fvm::grid_for(tot, 256), 256, 0, s>>>(p->nct, p->K, p->d_q.p, p->coeffg.p,
                                                                          p->vel.p, p->feqes.p);
    }
    FVM_HIP(hipGetLastError());
}
}  // namespace

extern "C" {

int fvm_quadrature_cartesian(int N1, int N2, int N3, double clim, double T2, double *cx, double *cy, double *cz,
                             double *w) {
    FVM_API_BEGIN
    FVM_CHECK(cx && cy && cz && w, FVM_E_INVALID, "null argument");
    FVM_CHECK(N1 >= 2 && N2 >= 2 && N3 >= 2, FVM_E_INVALID, "need at least two abscissae per axis");
    FVM_CHECK(std::isfinite(clim) && clim > 0 && std::isfinite(T2) && T2 > 0, FVM_E_INVALID,
              "clim and T2 must be positive");
    // Quadrature.h:83-181: abscissae by repeated addition of the spacing
    const int N[3] = {N1, N2, N3};
    std::vector<double> ab[3];
    double sp[3];
    for (int a = 0; a < 3; a++) {
        const double cmin = -clim * sqrt(0.5 * T2), cmax = clim * sqrt(0.5 * T2);
        const double d = (cmax - cmin) / (N[a] - 1.0);
        ab[a].resize(N[a]);
        ab[a][0] = cmin;
        for (int j = 1; j < N[a]; j++) ab[a][j] = ab[a][j - 1] + d;
        sp[a] = d;
    }
    int j = 0;
    for (int j1 = 0; j1 < N1; j1++)
        for (int j2 = 0; j2 < N2; j2++)
            for (int j3 = 0; j3 < N3; j3++, j++) {
                cx[j] = ab[0][j1]; cy[j] = ab[1][j2]; cz[j] = ab[2][j3];
                w[j] = sp[0] * sp[1] * sp[2];
            }
    return FVM_OK;
    FVM_API_END
}

int fvm_kinetic_default_options(fvm_kinetic_options *o) {
    if (!o) return FVM_E_INVALID;
    // KineticModelOptions (esbgkbase/KineticBC.h:61-113)
    o->rho_init = 9.28e-6; o->T_init = 273.15; o->nonDimLt = 1.0; o->molecularWeight = 40.0;
    o->Tmuref = 273.15; o->muref = 2.117e-5; o->mu_w = 0.81; o->Prandtl = 2.0 / 3.0;
    o->fgamma = 2; o->ktrial = 50; o->ToleranceX = 1e-8; o->ToleranceF = 1e-16; o->epsilon = 1e-50;
    return FVM_OK;
}

int fvm_kinetic_create(fvm_kinetic **out, fvm_mesh *mesh, int K, const double *cx, const double *cy,
                       const double *cz, const double *w, const fvm_kinetic_options *opt, const int32_t *bc_kind,
                       const double *wall_u, const double *wall_T) {
    FVM_API_BEGIN
    FVM_CHECK(out && mesh && cx && cy && cz && w && opt, FVM_E_INVALID, "null argument");
    FVM_CHECK(K >= 1, FVM_E_INVALID, "need at least one direction");
    FVM_CHECK(opt->fgamma >= 0 && opt->fgamma <= 2, FVM_E_INVALID,
              "fgamma must be 0 (Maxwellian), 1 (BGK) or 2 (ES-BGK)");
    FVM_CHECK(opt->fgamma != 2 || (std::isfinite(opt->Prandtl) && opt->Prandtl != 0), FVM_E_INVALID,
              "ES-BGK needs a finite non-zero Prandtl number");
    FVM_CHECK(opt->ktrial >= 0, FVM_E_INVALID, "ktrial must be >= 0");
    int nc, ng, nf, nn, ngr;
    krc(fvm_mesh_info(mesh, &nc, &ng, &nf, &nn, &ngr));
    FVM_CHECK(ngr >= 1 && ngr <= 64, FVM_E_INVALID, "bad face groups");
    FVM_CHECK(ngr <= 1 || (bc_kind && wall_u && wall_T), FVM_E_INVALID, "boundary groups need their conditions");
    for (int g = 1; g < ngr; g++) {
        FVM_CHECK(bc_kind[g] >= 0 && bc_kind[g] <= 3, FVM_E_INVALID,
                  "kinetic boundary kind must be 0 (diffuse wall), 1 (zero gradient), 2 (symmetry) or 3 (no "
                  "boundary-value rule)");
        FVM_CHECK(bc_kind[g] != 0 || (std::isfinite(wall_T[g]) && wall_T[g] > 0), FVM_E_INVALID,
                  "a diffuse wall needs a positive temperature");
    }
    for (int j = 0; j < K; j++)
        FVM_CHECK(std::isfinite(cx[j]) && std::isfinite(cy[j]) && std::isfinite(cz[j]) && std::isfinite(w[j]),
                  FVM_E_INVALID, "quadrature must be finite");
    const int dev = fvm::mesh_device(mesh);
    fvm_kinetic *p = new fvm_kinetic();
    try {
        p->mesh = mesh;
        p->K = K; p->nc = nc; p->nct = nc + ng; p->ngroups = ngr;
        p->opt = *opt;
        p->cx.assign(cx, cx + K); p->cy.assign(cy, cy + K); p->cz.assign(cz, cz + K); p->w.assign(w, w + K);
        p->gptr.resize(ngr + 1);
        krc(fvm_mesh_get(mesh, nullptr, nullptr, nullptr, p->gptr.data()));
        p->kind.assign(ngr, 0);
        p->wall_u.assign(3 * (size_t)ngr, 0.0);
        p->wall_T.assign(ngr, 1.0);
        for (int g = 1; g < ngr; g++) {
            p->kind[g] = bc_kind[g];
            for (int a = 0; a < 3; a++) p->wall_u[3 * g + a] = wall_u[3 * g + a];
            p->wall_T[g] = wall_T[g];
        }
        p->device = dev;
        KDevGuard d(dev);
        FVM_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamDefault));
        const size_t tot = (size_t)K * p->nct;
        p->f.alloc(std::max<size_t>(tot, 1)); p->feq.alloc(std::max<size_t>(tot, 1));
        p->delta.alloc(std::max<size_t>(tot, 1));
        p->rho.alloc(std::max(p->nct, 1)); p->T.alloc(std::max(p->nct, 1)); p->nu.alloc(std::max(p->nct, 1));
        p->vel.alloc(3 * (size_t)std::max(p->nct, 1)); p->coeff.alloc(5 * (size_t)std::max(p->nct, 1));
        if (opt->fgamma == 2) {
            p->coeffg.alloc(10 * (size_t)std::max(p->nct, 1));
            p->T6.alloc(6 * (size_t)std::max(p->nct, 1));
            p->feqes.alloc(std::max<size_t>(tot, 1));
        }
        std::vector<double> q(4 * (size_t)K);
        for (int j = 0; j < K; j++) { q[j] = cx[j]; q[K + j] = cy[j]; q[2 * K + j] = cz[j]; q[3 * K + j] = w[j]; }
        p->d_q.alloc(q.size());
        FVM_HIP(hipMemcpy(p->d_q.p, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice));
        // group table: gptr (ngr + 1) then kind (ngr); walls: uw, vw, ww, Tw per group
        std::vector<int> grp(2 * (size_t)ngr + 1, 0);
        for (int g = 0; g <= ngr; g++) grp[g] = p->gptr[g];
        for (int g = 0; g < ngr; g++) grp[ngr + 1 + g] = p->kind[g];
        std::vector<double> wall(4 * (size_t)ngr, 0.0);
        for (int g = 0; g < ngr; g++) {
            for (int a = 0; a < 3; a++) wall[4 * g + a] = p->wall_u[3 * g + a];
            wall[4 * g + 3] = p->wall_T[g];
        }
        p->d_grp.alloc(grp.size());
        p->d_wall.alloc(wall.size());
        FVM_HIP(hipMemcpy(p->d_grp.p, grp.data(), sizeof(int) * grp.size(), hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(p->d_wall.p, wall.data(), sizeof(double) * wall.size(), hipMemcpyHostToDevice));
        // specular reflection tables (KineticModel.h:290-318): from the normal of the
        // group's first face, the direction nearest to c - 2 (c.n) n, first minimum
        std::vector<int> refl((size_t)ngr * K, -1);
        bool any_sym = false;
        for (int g = 1; g < ngr; g++) any_sym = any_sym || p->kind[g] == 2;
        if (any_sym) {
            const double *fA, *fAm;
            const int *fcells;
            fvm::mesh_device_metrics(mesh, &fcells, &fA, &fAm);
            FVM_HIP(hipDeviceSynchronize());   // the metrics are computed on the mesh's stream
            for (int g = 1; g < ngr; g++) {
                if (p->kind[g] != 2 || p->gptr[g + 1] <= p->gptr[g]) continue;
                const int f0 = p->gptr[g];
                double A[3], Am;
                FVM_HIP(hipMemcpy(A, fA + 3 * (size_t)f0, sizeof(double) * 3, hipMemcpyDeviceToHost));
                FVM_HIP(hipMemcpy(&Am, fAm + f0, sizeof(double), hipMemcpyDeviceToHost));
                const double en[3] = {A[0] / Am, A[1] / Am, A[2] / Am};
                for (int j = 0; j < K; j++) {
                    const double cn = cx[j] * en[0] + cy[j] * en[1] + cz[j] * en[2];
                    const double ix = cx[j] - 2.0 * cn * en[0], iy = cy[j] - 2.0 * cn * en[1],
                                 iz = cz[j] - 2.0 * cn * en[2];
                    int arg = 0;
                    double best = 1e54;
                    for (int js = 0; js < K; js++) {
                        const double dd = pow(ix - cx[js], 2) + pow(iy - cy[js], 2) + pow(iz - cz[js], 2);
                        if (dd < best) { best = dd; arg = js; }
                    }
                    refl[(size_t)g * K + j] = arg;
                }
            }
        }
        // the assembly's group table: symmetry groups take the wall-like incoming test
        // with zero velocity; zero-gradient groups take every direction as incoming
        std::vector<int> agrp(2 * (size_t)ngr + 1, 0);
        for (int g = 0; g <= ngr; g++) agrp[g] = p->gptr[g];
        std::vector<double> awall(3 * (size_t)ngr, 0.0);
        for (int g = 1; g < ngr; g++) {
            agrp[ngr + 1 + g] = p->kind[g] == 1 ? 1 : 0;
            if (p->kind[g] == 0)
                for (int a = 0; a < 3; a++) awall[3 * g + a] = p->wall_u[3 * g + a];
        }
        p->d_agrp.alloc(agrp.size());
        p->d_awall.alloc(awall.size());
        FVM_HIP(hipMemcpy(p->d_agrp.p, agrp.data(), sizeof(int) * agrp.size(), hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(p->d_awall.p, awall.data(), sizeof(double) * awall.size(), hipMemcpyHostToDevice));
        p->d_F.alloc(std::max(nf, 1));
        p->d_inc.alloc(std::max(nf, 1));
        p->d_refl.alloc(std::max<size_t>(refl.size(), 1));
        FVM_HIP(hipMemcpy(p->d_refl.p, refl.data(), sizeof(int) * refl.size(), hipMemcpyHostToDevice));
        // KineticModel constructor (KineticModel.h:81-120): rest Maxwellian, then its
        // moments, collision frequency and the equilibrium of the initial coefficients
        fvm::k_kin_init<<<fvm::grid_for((long long)tot, 256), 256, 0, p->stream>>>(p->nct, K, p->d_q.p, p->f.p,
                                                                                   p->rho.p, p->vel.p, p->T.p,
                                                                                   p->coeff.p);
        fvm::k_kin_moments<<<fvm::grid_for(p->nct, 256), 256, 0, p->stream>>>(p->nct, K, p->d_q.p, p->f.p, p->rho.p,
                                                                              p->vel.p, p->T.p);
        {
            const fvm_kinetic_options &o = p->opt;
            const double R = 8314.0 / o.molecularWeight;
            const double mu0 = o.rho_init * R * o.T_init * o.nonDimLt / pow(2 * R * o.T_init, 0.5);
            fvm::k_kin_collision<<<fvm::grid_for(p->nct, 256), 256, 0, p->stream>>>(
                p->nct, p->rho.p, p->T.p, o.T_init, o.Tmuref, o.muref, o.mu_w, mu0, o.Prandtl, o.fgamma, p->nu.p);
        }
        fvm::k_kin_equilibrium<<<fvm::grid_for((long long)tot, 256), 256, 0, p->stream>>>(p->nct, K, p->d_q.p,
                                                                                          p->coeff.p, p->vel.p,
                                                                                          p->feq.p);
        if (opt->fgamma == 2) {
            fvm::k_kin_init_es<<<fvm::grid_for(p->nct, 256), 256, 0, p->stream>>>(p->nct, p->coeff.p, p->coeffg.p);
            fvm::k_kin_equilibrium_es<<<fvm::grid_for((long long)tot, 256), 256, 0, p->stream>>>(
                p->nct, K, p->d_q.p, p->coeffg.p, p->vel.p, p->feqes.p);
        }
        FVM_HIP(hipGetLastError());
        FVM_HIP(hipStreamSynchronize(p->stream));
        krc(fvm_amg_default_params(&p->prm));
        p->prm.verbosity = 0;
        krc(fvm_batch_create(&p->batch, dev));
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

int fvm_kinetic_destroy(fvm_kinetic *p) {
    FVM_API_BEGIN
    if (!p) return FVM_OK;
    {
        KDevGuard d(p->device);
        (void)hipStreamSynchronize(p->stream);
        if (p->batch) fvm_batch_destroy(p->batch);
        for (auto *S : p->sys)
            if (S) fvm_system_destroy(S);
        p->sys.clear();
        p->d_agrp.release(); p->d_awall.release(); p->d_F.release(); p->d_inc.release();
        p->coeffg.release(); p->T6.release(); p->feqes.release();
        p->f.release(); p->feq.release(); p->delta.release(); p->rho.release(); p->vel.release();
        p->T.release(); p->nu.release(); p->coeff.release(); p->d_q.release(); p->d_grp.release();
        p->d_refl.release(); p->d_wall.release();
        (void)hipStreamDestroy(p->stream);
    }
    delete p;
    return FVM_OK;
    FVM_API_END
}

int fvm_kinetic_set_params(fvm_kinetic *p, const fvm_amg_params *prm) {
    FVM_API_BEGIN
    FVM_CHECK(p && prm, FVM_E_INVALID, "null argument");
    p->prm = *prm;
    return FVM_OK;
    FVM_API_END
}

int fvm_kinetic_advance(fvm_kinetic *p, int niter, double rel_tol, double abs_tol, int *iterations, double *norms,
                        int cap, int *converged) {
    FVM_API_BEGIN
    FVM_CHECK(p && niter >= 0 && (cap == 0 || norms), FVM_E_INVALID, "bad argument");
    KDevGuard d(p->device);
    const int K = p->K, nct = p->nct;
    krc(fvm_batch_set_params(p->batch, &p->prm));
    if (p->sys.empty()) {
        // one system per velocity, with the mesh's pattern; re-assembled in place
        p->sys.assign(K, nullptr);
        for (int k = 0; k < K; k++) p->sys[k] = fvm::system_for_mesh(p->mesh);
    }
    hipStream_t ms = fvm::mesh_stream(p->mesh);
    int done = 0, conv = 0;
    std::vector<int32_t> ns(K), nt(K);
    std::vector<const int32_t *> rows(K), cols(K);
    std::vector<const double *> diags(K), offs(K), bs(K);
    std::vector<const uint8_t *> bnds(K);
    std::vector<double *> xs(K);
    std::vector<double> r0(K), rf(K);
    std::vector<int> cyc(K);
    for (int k = 0; k < K; k++) {
        int s_n, s_t;
        long long nz;
        krc(fvm_system_info(p->sys[k], &s_n, &s_t, &nz));
        ns[k] = s_n; nt[k] = s_t;
        const double *bb;
        krc(fvm_system_device(p->sys[k], &rows[k], &cols[k], &diags[k], &offs[k], &bb, &bnds[k]));
        bs[k] = bb;
        xs[k] = p->delta.p + (size_t)k * nct;
    }
    // phase timing to stderr when FVM_AMD_KINETIC_TIMING is set
    const bool timing = getenv("FVM_AMD_KINETIC_TIMING") != nullptr;
    auto now = [&]() {
        FVM_HIP(hipDeviceSynchronize());
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    for (int n = 0; n < niter; n++) {
        FVM_HIP(hipStreamSynchronize(p->stream));   // f, feq, nu of the previous iteration are final
        const double t_a = timing ? now() : 0.0;
        for (int k = 0; k < K; k++) {
            // initKineticModelLinearization + linearizeKineticModel + initSolve for velocity k
            const double c3[3] = {p->cx[k], p->cy[k], p->cz[k]};
            const double *feq_lin = p->opt.fgamma == 2 ? p->feqes.p : p->feq.p;
            fvm::kinetic_assemble_into(p->mesh, c3, p->nu.p, feq_lin + (size_t)k * nct, p->f.p + (size_t)k * nct,
                                       p->d_agrp.p, p->d_awall.p, p->opt.epsilon, p->sys[k], p->d_F.p, p->d_inc.p);
        }
        FVM_HIP(hipMemsetAsync(p->delta.p, 0, sizeof(double) * (size_t)K * nct, p->stream));
        FVM_HIP(hipStreamSynchronize(ms));
        FVM_HIP(hipStreamSynchronize(p->stream));
        const double t_b = timing ? now() : 0.0;
        krc(fvm_batch_setup_device(p->batch, K, ns.data(), nt.data(), rows.data(), cols.data(), diags.data(),
                                   offs.data(), bnds.data()));
        const double t_c = timing ? now() : 0.0;
        krc(fvm_batch_solve_device(p->batch, bs.data(), xs.data(), r0.data(), rf.data(), cyc.data()));
        const double t_d = timing ? now() : 0.0;
        double rnorm = r0[0];
        for (int k = 1; k < K; k++) rnorm += r0[k];
        // postSolve (every direction) and updateSolution
        for (int k = 0; k < K; k++) fvm::system_post_solve_async(p->sys[k], xs[k], p->stream);
        fvm::k_axpy_kin<<<fvm::grid_for((long long)K * nct, 256), 256, 0, p->stream>>>((long long)K * nct,
                                                                                     p->delta.p, p->f.p);
        FVM_HIP(hipGetLastError());
        if (!p->have_initial) { p->initial_norm = rnorm; p->have_initial = true; }
        if (p->niters < 5) p->initial_norm = std::max(p->initial_norm, rnorm);   // setMax
        const double ratio = rnorm / p->initial_norm;
        p->norms.push_back(rnorm);
        if (done < cap) norms[done] = rnorm;
        p->niters++;
        done++;
        kinetic_update(p);
        if (timing) {
            const double t_e = now();
            int cmax = 0;
            for (int k = 0; k < K; k++) cmax = std::max(cmax, cyc[k]);
            fprintf(stderr, "[kinetic] iteration %d: assembly %.1f ms, setup %.1f ms, solve %.1f ms (cycles <= %d), "
                            "post-solve + update %.1f ms\n", p->niters, 1e3 * (t_b - t_a), 1e3 * (t_c - t_b),
                    1e3 * (t_d - t_c), cmax, 1e3 * (t_e - t_d));
        }
        if (rnorm < abs_tol || ratio < rel_tol) { conv = 1; break; }
    }
    FVM_HIP(hipStreamSynchronize(p->stream));
    if (iterations) *iterations = done;
    if (converged) *converged = conv;
    return FVM_OK;
    FVM_API_END
}

int fvm_kinetic_get(const fvm_kinetic *p, double *f, double *feq, double *density, double *velocity,
                    double *temperature, double *nu) {
    FVM_API_BEGIN
    FVM_CHECK(p, FVM_E_INVALID, "null handle");
    KDevGuard d(p->device);
    FVM_HIP(hipStreamSynchronize(p->stream));
    const size_t tot = (size_t)p->K * p->nct;
    if (f) FVM_HIP(hipMemcpy(f, p->f.p, sizeof(double) * tot, hipMemcpyDeviceToHost));
    if (feq) FVM_HIP(hipMemcpy(feq, p->opt.fgamma == 2 ? p->feqes.p : p->feq.p, sizeof(double) * tot,
                               hipMemcpyDeviceToHost));
    if (density) FVM_HIP(hipMemcpy(density, p->rho.p, sizeof(double) * p->nct, hipMemcpyDeviceToHost));
    if (velocity) FVM_HIP(hipMemcpy(velocity, p->vel.p, sizeof(double) * 3 * p->nct, hipMemcpyDeviceToHost));
    if (temperature) FVM_HIP(hipMemcpy(temperature, p->T.p, sizeof(double) * p->nct, hipMemcpyDeviceToHost));
    if (nu) FVM_HIP(hipMemcpy(nu, p->nu.p, sizeof(double) * p->nct, hipMemcpyDeviceToHost));
    return FVM_OK;
    FVM_API_END
}

}  // extern "C"
