/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Sequential restatement of the c-PRIMED/fvm finite-volume assembly of the
 * thermal (diffusion) equation, used to check the HIP assembly kernels:
 *   mesh metrics         MeshMetricsCalculator_impl.h:59-236 (centroids),
 *                        :239-305 (areas), :385 (area magnitude), :393-462 (volumes)
 *   cellCells pattern    CRConnectivity.cpp:167-270 (multiply, first appearance)
 *   diffusion            DiffusionDiscretization.h:19-27, :153-230
 *   source               SourceDiscretization.h:40-60
 *   Dirichlet / Neumann  GenericBCS.h:77-157
 *   boundary elimination CRMatrix.h:899-944, :1064-1085 (LinearSystem::initSolve)
 * (paths relative to src/fvm/src/modules/fvmbase/). Loops run in the
 * reference's order so every FP64 value is bit-identical to it when built
 * without FMA contraction.
 *
 * Mesh input (unstructured): n_cells interior cells, n_ghost ghost cells (one
 * per boundary face, numbered n_cells + boundary-face rank), faces with 4 nodes
 * each (quad faces; area formula of MeshMetricsCalculator for numNodes == 4),
 * face_cells[2f] = (c0, c1), and face groups [gptr[g], gptr[g+1]) with group 0
 * the interior faces.
 *
 * Pinned by tests/test_fv_oracle.py: closed forms on uniform meshes (|A| = h^2,
 * V = h^3, interior coefficient k h, Dirichlet coefficient 2 k h, Neumann rows
 * eliminate to zero flux) and divergence-theorem identities on jittered meshes
 * (sum of volumes = domain volume, closed-surface sum of A = 0); by
 * tests/test_fv_closed_forms.py: the exact series-resistance solution of a
 * layered conductivity (harmonic face mean) and the affine closed form of the
 * coefficient k |A|^2 / (A . ds) on sheared parallelepiped meshes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fv_oracle.h"

static void v_sub(const double *a, const double *b, double *c) { c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2]; }
static double v_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void v_cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

void ofv_metrics(int n_faces, const double *nodes, const int *face_nodes, const int *face_cells,
                 int n_cells, int n_ghost, int n_groups, const int *gptr,
                 double *fA, double *fAm, double *fc, double *cc, double *vol)
{
    const int nct = n_cells + n_ghost;
    /* areas (quad): 0.5 * (x2 - x0) x (x3 - x1) */
    for (int f = 0; f < n_faces; f++) {
        const double *x0 = nodes + 3 * face_nodes[4 * f], *x1 = nodes + 3 * face_nodes[4 * f + 1];
        const double *x2 = nodes + 3 * face_nodes[4 * f + 2], *x3 = nodes + 3 * face_nodes[4 * f + 3];
        double d20[3], d31[3], c[3];
        v_sub(x2, x0, d20);
        v_sub(x3, x1, d31);
        v_cross(d20, d31, c);
        for (int k = 0; k < 3; k++) fA[3 * f + k] = 0.5 * c[k];
        fAm[f] = sqrt(v_dot(fA + 3 * f, fA + 3 * f));
    }
    /* face centroids: node average then the non-planar quad correction */
    const double twoThirds = 2. / 3.;
    for (int f = 0; f < n_faces; f++) {
        double p[3];
        const int *fn = face_nodes + 4 * f;
        for (int k = 0; k < 3; k++) p[k] = nodes[3 * fn[0] + k];
        for (int q = 1; q < 4; q++)
            for (int k = 0; k < 3; k++) p[k] += nodes[3 * fn[q] + k];
        for (int k = 0; k < 3; k++) p[k] /= 4.0;
        double en[3];
        for (int k = 0; k < 3; k++) en[k] = fA[3 * f + k] / fAm[f];
        double denom = 0.0, cfc[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < 4; q++) {
            const double *a = nodes + 3 * fn[q], *b = nodes + 3 * fn[(q + 1) % 4];
            double rc0[3], rc1[3], cr[3], tri[3], xm[3];
            v_sub(a, p, rc0);
            v_sub(b, p, rc1);
            v_cross(rc0, rc1, cr);
            for (int k = 0; k < 3; k++) tri[k] = 0.5 * cr[k];
            const double triP = v_dot(tri, en);
            for (int k = 0; k < 3; k++) xm[k] = 0.5 * (a[k] + b[k]);
            for (int k = 0; k < 3; k++) cfc[k] += (twoThirds * (xm[k] - p[k])) * triP;
            denom += triP;
        }
        for (int k = 0; k < 3; k++) cfc[k] /= denom;
        for (int k = 0; k < 3; k++) fc[3 * f + k] = p[k] + cfc[k];
    }
    /* cell centroids: area-weighted face centroids, all faces in order */
    double *w = (double *)calloc((size_t)nct, sizeof(double));
    memset(cc, 0, sizeof(double) * 3 * (size_t)nct);
    for (int f = 0; f < n_faces; f++)
        for (int s = 0; s < 2; s++) {
            const int c = face_cells[2 * f + s];
            for (int k = 0; k < 3; k++) cc[3 * c + k] += fc[3 * f + k] * fAm[f];
            w[c] += fAm[f];
        }
    for (int c = 0; c < n_cells; c++)
        for (int k = 0; k < 3; k++) cc[3 * c + k] /= w[c];
    for (int g = 1; g < n_groups; g++)
        for (int f = gptr[g]; f < gptr[g + 1]; f++) {
            const int c1 = face_cells[2 * f + 1];
            for (int k = 0; k < 3; k++) cc[3 * c1 + k] = fc[3 * f + k];
        }
    free(w);
    /* volumes */
    memset(vol, 0, sizeof(double) * (size_t)nct);
    const double dim = 3.0;
    for (int f = 0; f < n_faces; f++) {
        double d[3];
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        v_sub(fc + 3 * f, cc + 3 * c0, d);
        vol[c0] += v_dot(d, fA + 3 * f) / dim;
        v_sub(fc + 3 * f, cc + 3 * c1, d);
        vol[c1] -= v_dot(d, fA + 3 * f) / dim;
    }
    for (int g = 1; g < n_groups; g++)
        for (int f = gptr[g]; f < gptr[g + 1]; f++)
            vol[face_cells[2 * f + 1]] = vol[face_cells[2 * f]];
}

/* cellCells = cellFaces x faceCells, first appearance, implicit diagonal
 * (CRConnectivity::multiply); also the pair->CSR positions of every face. */
static void cell_cells(int n_faces, const int *face_cells, int nct, int *row, int **col_out,
                       int *pos01, int *pos10)
{
    int *fcnt = (int *)calloc((size_t)nct + 1, sizeof(int));
    for (int f = 0; f < n_faces; f++) { fcnt[face_cells[2 * f] + 1]++; fcnt[face_cells[2 * f + 1] + 1]++; }
    for (int c = 0; c < nct; c++) fcnt[c + 1] += fcnt[c];
    int *cf = (int *)malloc(sizeof(int) * (size_t)(fcnt[nct] > 0 ? fcnt[nct] : 1));
    int *fill = (int *)malloc(sizeof(int) * (size_t)(nct > 0 ? nct : 1));
    for (int c = 0; c < nct; c++) fill[c] = fcnt[c];
    for (int f = 0; f < n_faces; f++) {   /* ascending f: cellFaces rows are sorted */
        cf[fill[face_cells[2 * f]]++] = f;
        cf[fill[face_cells[2 * f + 1]]++] = f;
    }
    int *col = (int *)malloc(sizeof(int) * (size_t)(fcnt[nct] > 0 ? fcnt[nct] : 1));
    row[0] = 0;
    int k = 0;
    for (int c = 0; c < nct; c++) {
        const int b = k;
        for (int q = fcnt[c]; q < fcnt[c + 1]; q++) {
            const int f = cf[q];
            for (int s = 0; s < 2; s++) {
                const int o = face_cells[2 * f + s];
                if (o == c) continue;
                int seen = 0;
                for (int e = b; e < k; e++) if (col[e] == o) { seen = 1; break; }
                if (!seen) col[k++] = o;
            }
        }
        row[c + 1] = k;
    }
    for (int f = 0; f < n_faces; f++) {
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        for (int e = row[c0]; e < row[c0 + 1]; e++) if (col[e] == c1) { pos01[f] = e; break; }
        for (int e = row[c1]; e < row[c1 + 1]; e++) if (col[e] == c0) { pos10[f] = e; break; }
    }
    free(fcnt); free(cf); free(fill);
    *col_out = col;
}

int ofv_thermal_nnz(int n_faces, const int *face_cells, int n_cells, int n_ghost)
{
    const int nct = n_cells + n_ghost;
    int *row = (int *)malloc(sizeof(int) * ((size_t)nct + 1));
    int *p01 = (int *)malloc(sizeof(int) * (size_t)(n_faces > 0 ? n_faces : 1));
    int *p10 = (int *)malloc(sizeof(int) * (size_t)(n_faces > 0 ? n_faces : 1));
    int *col;
    cell_cells(n_faces, face_cells, nct, row, &col, p01, p10);
    const int nnz = row[nct];
    free(row); free(col); free(p01); free(p10);
    return nnz;
}

static double harmonic(double x0, double x1)
{
    const double sum = x0 + x1;
    if (x0 + x1 != 0.0) return 2.0 * x0 * x1 / sum;
    return sum;
}

int ofv_thermal(int n_faces, const double *nodes, const int *face_nodes, const int *face_cells,
                int n_cells, int n_ghost, int n_groups, const int *gptr, const int *bc_type,
                const double *bc_value, const double *kcell, const double *src, const double *x,
                int *row, int *col, double *diag, double *off, double *b, unsigned char *is_bnd)
{
    const int nct = n_cells + n_ghost;
    double *fA = (double *)malloc(sizeof(double) * 3 * (size_t)n_faces);
    double *fAm = (double *)malloc(sizeof(double) * (size_t)n_faces);
    double *fc = (double *)malloc(sizeof(double) * 3 * (size_t)n_faces);
    double *cc = (double *)malloc(sizeof(double) * 3 * (size_t)nct);
    double *vol = (double *)malloc(sizeof(double) * (size_t)nct);
    ofv_metrics(n_faces, nodes, face_nodes, face_cells, n_cells, n_ghost, n_groups, gptr, fA, fAm, fc, cc, vol);
    int *p01 = (int *)malloc(sizeof(int) * (size_t)n_faces);
    int *p10 = (int *)malloc(sizeof(int) * (size_t)n_faces);
    int *c2;
    cell_cells(n_faces, face_cells, nct, row, &c2, p01, p10);
    memcpy(col, c2, sizeof(int) * (size_t)row[nct]);
    free(c2);
    /* LinearSystem::initAssembly: zero matrix and b */
    memset(diag, 0, sizeof(double) * (size_t)nct);
    memset(off, 0, sizeof(double) * (size_t)row[nct]);
    memset(b, 0, sizeof(double) * (size_t)nct);
    memset(is_bnd, 0, (size_t)nct);
    /* DiffusionDiscretization::discretize over all face groups (gradient of a
     * uniform field: zero) */
    for (int f = 0; f < n_faces; f++) {
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        const double vol0 = vol[c0], vol1 = vol[c1];
        double ds[3];
        v_sub(cc + 3 * c1, cc + 3 * c0, ds);
        double kf;
        if (vol0 == 0.) kf = kcell[c1];
        else if (vol1 == 0.) kf = kcell[c0];
        else kf = harmonic(kcell[c0], kcell[c1]);
        const double diffMetric = fAm[f] * fAm[f] / v_dot(fA + 3 * f, ds);
        const double diffCoeff = kf * diffMetric;
        double sec[3], gradF[3];
        for (int k = 0; k < 3; k++) sec[k] = kf * (fA[3 * f + k] - ds[k] * diffMetric);
        for (int k = 0; k < 3; k++) gradF[k] = (0.0 * vol0 + 0.0 * vol1) / (vol0 + vol1);
        double dsec = 0.0;
        for (int k = 0; k < 3; k++) dsec += gradF[k] * sec[k];
        const double dFlux = diffCoeff * (x[c1] - x[c0]) + dsec;
        b[c0] += dFlux;
        b[c1] -= dFlux;
        off[p01[f]] += diffCoeff;
        off[p10[f]] += diffCoeff;
        diag[c0] -= diffCoeff;
        diag[c1] -= diffCoeff;
    }
    /* SourceDiscretization */
    for (int c = 0; c < n_cells; c++) b[c] += vol[c] * src[c];
    /* boundary conditions, groups in order */
    for (int g = 1; g < n_groups; g++)
        for (int f = gptr[g]; f < gptr[g + 1]; f++) {
            const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
            if (bc_type[g] == OFV_BC_DIRICHLET) {
                const double dX = bc_value[g] - x[c1];
                b[c0] += off[p01[f]] * dX;
                off[p01[f]] = 0.0;
                off[p10[f]] = 0.0;
                b[c1] = 0.0;
                diag[c1] = -1.0;
            } else {
                const double fluxB = -b[c1];
                b[c1] = bc_value[g] * fAm[f] - fluxB;
                is_bnd[c1] = 1;
            }
        }
    /* LinearSystem::initSolve -> eliminateBoundaryEquations (ghost rows, ascending) */
    for (int j = n_cells; j < nct; j++) {
        if (!is_bnd[j]) continue;
        const double ajj = diag[j];
        for (int nb = row[j]; nb < row[j + 1]; nb++) {
            const int i = col[nb];
            int pij = -1;
            for (int e = row[i]; e < row[i + 1]; e++) if (col[e] == j) { pij = e; break; }
            if (pij < 0) continue;
            const double aij = off[pij];
            for (int nb2 = row[j]; nb2 < row[j + 1]; nb2++) {
                const int k = col[nb2];
                const double ajk = off[nb2];
                if (i != k) {
                    for (int e = row[i]; e < row[i + 1]; e++)
                        if (col[e] == k) { off[e] -= aij * (ajk / ajj); break; }
                } else {
                    diag[i] -= aij * (ajk / ajj);
                }
            }
            b[i] -= aij * (b[j] / ajj);
            off[pij] = 0.0;
        }
    }
    free(fA); free(fAm); free(fc); free(cc); free(vol); free(p01); free(p10);
    return 0;
}

/*
 * Gray phonon-BTE style scalar system of one direction (discrete ordinate):
 *   collision   PhononCollisionDiscretization.h:54-68  (phononbase/):
 *               r[c] -= V (x - e0) / tau,  diag[c] -= V / tau       (self cells)
 *   convection  PhononConvectionDiscretization.h:150-175 (upwind, no grid flux):
 *               F = A . vg;  F > 0: r0 -= F x0, r1 += F x0, d0 -= F, a10 += F
 *                            else : r0 -= F x1, r1 += F x1, d1 += F, a01 -= F
 *   BCs         PhononModel.h:498-523 + GenericPhononBCS.h:67-131: per boundary
 *               face, vg . (A/|A|) < 0 -> Dirichlet with the ghost's own value,
 *               else extrapolation (ghost row marked boundary)
 *   initSolve   eliminate the marked ghost rows (CRMatrix.h:899-944)
 * The linearizer runs collision before convection (PhononModel.h:449-459).
 * Loops run in the reference's order (faces ascending, groups in order).
 */
int ofv_phonon(int n_faces, const double *nodes, const int *face_nodes, const int *face_cells,
               int n_cells, int n_ghost, int n_groups, const int *gptr, const double *vg, double tau,
               const double *e0, const double *x, int *row, int *col, double *diag, double *off,
               double *b, unsigned char *is_bnd)
{
    const int nct = n_cells + n_ghost;
    double *fA = (double *)malloc(sizeof(double) * 3 * (size_t)n_faces);
    double *fAm = (double *)malloc(sizeof(double) * (size_t)n_faces);
    double *fc = (double *)malloc(sizeof(double) * 3 * (size_t)n_faces);
    double *cc = (double *)malloc(sizeof(double) * 3 * (size_t)nct);
    double *vol = (double *)malloc(sizeof(double) * (size_t)nct);
    ofv_metrics(n_faces, nodes, face_nodes, face_cells, n_cells, n_ghost, n_groups, gptr, fA, fAm, fc, cc, vol);
    int *p01 = (int *)malloc(sizeof(int) * (size_t)n_faces);
    int *p10 = (int *)malloc(sizeof(int) * (size_t)n_faces);
    int *c2;
    cell_cells(n_faces, face_cells, nct, row, &c2, p01, p10);
    memcpy(col, c2, sizeof(int) * (size_t)row[nct]);
    free(c2);
    memset(diag, 0, sizeof(double) * (size_t)nct);
    memset(off, 0, sizeof(double) * (size_t)row[nct]);
    memset(b, 0, sizeof(double) * (size_t)nct);
    memset(is_bnd, 0, (size_t)nct);
    /* the boundary-condition code writes the field (Dirichlet ghosts) */
    double *xf = (double *)malloc(sizeof(double) * (size_t)nct);
    memcpy(xf, x, sizeof(double) * (size_t)nct);
    for (int c = 0; c < n_cells; c++) {
        b[c] -= vol[c] * (xf[c] - e0[c]) / tau;
        diag[c] -= vol[c] / tau;
    }
    for (int f = 0; f < n_faces; f++) {
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        const double F = fA[3 * f] * vg[0] + fA[3 * f + 1] * vg[1] + fA[3 * f + 2] * vg[2];
        double flux;
        if (F > 0.0) {
            flux = F * xf[c0];
            diag[c0] -= F;
            off[p10[f]] += F;
        } else {
            flux = F * xf[c1];
            diag[c1] += F;
            off[p01[f]] -= F;
        }
        b[c0] -= flux;
        b[c1] += flux;
    }
    for (int g = 1; g < n_groups; g++)
        for (int f = gptr[g]; f < gptr[g + 1]; f++) {
            const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
            double en[3];
            for (int k = 0; k < 3; k++) en[k] = fA[3 * f + k] / fAm[f];
            const double vn = vg[0] * en[0] + vg[1] * en[1] + vg[2] * en[2];
            if (vn < 0.0) {
                const double bval = xf[c1];
                const double dX = bval - xf[c1];
                b[c0] += off[p01[f]] * dX;
                off[p01[f]] = 0.0;
                xf[c1] = bval;
                off[p10[f]] = 0.0;
                b[c1] = 0.0;
                diag[c1] = -1.0;
            } else {
                const double dFdX1 = -diag[c1];
                const double dx01 = xf[c0] - xf[c1];
                diag[c0] += dFdX1;
                b[c0] += dFdX1 * dx01;
                off[p01[f]] = 0.0;
                diag[c1] = -1.0;
                off[p10[f]] = 1.0;
                b[c1] = dx01;
                is_bnd[c1] = 1;
            }
        }
    for (int j = n_cells; j < nct; j++) {
        if (!is_bnd[j]) continue;
        const double ajj = diag[j];
        for (int nb = row[j]; nb < row[j + 1]; nb++) {
            const int i = col[nb];
            int pij = -1;
            for (int e = row[i]; e < row[i + 1]; e++) if (col[e] == j) { pij = e; break; }
            if (pij < 0) continue;
            const double aij = off[pij];
            for (int nb2 = row[j]; nb2 < row[j + 1]; nb2++) {
                const int k = col[nb2];
                const double ajk = off[nb2];
                if (i != k) {
                    for (int e = row[i]; e < row[i + 1]; e++)
                        if (col[e] == k) { off[e] -= aij * (ajk / ajj); break; }
                } else {
                    diag[i] -= aij * (ajk / ajj);
                }
            }
            b[i] -= aij * (b[j] / ajj);
            off[pij] = 0.0;
        }
    }
    free(xf); free(fA); free(fAm); free(fc); free(cc); free(vol); free(p01); free(p10);
    return 0;
}
