/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Sequential restatement of the c-PRIMED/fvm finite-volume assembly of the
 * thermal (convection-diffusion) equation and of one phonon direction, used to
 * check the HIP assembly kernels (paths relative to src/fvm/src/modules/fvmbase/):
 *   mesh metrics         MeshMetricsCalculator_impl.h:59-118 (face centroids),
 *                        :126-236 (cell centroids), :239-305 (face areas for 2, 3,
 *                        4 and more nodes), :385 (|A|), :393-462 (volumes, / dim)
 *   cellCells pattern    CRConnectivity.cpp:167-270 (multiply, first appearance)
 *   gradient             GradientModel.h:128-276 (3-D least squares),
 *                        :278-420 (2-D), :472-520 (compute, ghosts copy),
 *                        GradientMatrix.h:55-76 (getGradient)
 *   diffusion            DiffusionDiscretization.h:19-27, :153-230 (with the
 *                        non-orthogonal secondary flux)
 *   convection           ConvectionDiscretization.h:53-200 (upwind, or the
 *                        central-difference branch; continuity residual)
 *   source               SourceDiscretization.h:40-60
 *   boundary conditions  GenericBCS.h:77-212 (Dirichlet, Neumann, extrapolation),
 *                        ThermalModel_impl.h:300-330 (outflow Dirichlet faces
 *                        extrapolate when a convecting flux is given)
 *   boundary elimination CRMatrix.h:899-944, :1064-1085 (LinearSystem::initSolve)
 * Loops run in the reference's order so every FP64 value is bit-identical to it
 * when built without FMA contraction.
 *
 * Mesh input (unstructured): dimension 2 or 3, n_cells interior cells, n_ghost
 * ghost cells (one per boundary face, numbered n_cells + boundary-face rank),
 * faces as a CSR of nodes (fnptr[n_faces + 1], face_nodes), face_cells[2f] =
 * (c0, c1), and face groups [gptr[g], gptr[g+1]) with group 0 the interior
 * faces. All boundary groups are walls (ghost centroid = face centroid, ghost
 * gradient = its cell's gradient); symmetry and immersed-boundary faces are not
 * modelled. Each pair of cells shares at most one face (conformal meshes).
 *
 * Pinned by tests/test_fv_oracle.py: closed forms on uniform meshes (|A| = h^2,
 * V = h^3, interior coefficient k h, Dirichlet coefficient 2 k h, Neumann rows
 * eliminate to zero flux) and divergence-theorem identities on jittered meshes
 * (sum of volumes = domain volume, closed-surface sum of A = 0); by
 * tests/test_fv_closed_forms.py: the exact series-resistance solution of a
 * layered conductivity (harmonic face mean), the affine closed form of the
 * coefficient k |A|^2 / (A . ds) on sheared parallelepiped meshes, the
 * exactness of the least-squares gradient on linear fields, a vanishing
 * residual of a linear field on sheared meshes (secondary flux) and the upwind
 * convection closed forms; by tests/test_fluent_golden.py: the reference's
 * THERMAL_MATRIX golden matrix and right-hand side of the 32 x 32 cavity.
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
/* mag(v) of Vector.h: sqrt of the dot product */
static double v_mag(const double *a) { return sqrt(v_dot(a, a)); }

typedef struct {
    int dim, nf, nc, ng, ngroups;
    const double *nodes;
    const int *fnptr, *fnodes, *fcells, *gptr;
} mesh_t;

void ofv_metrics(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                 const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                 double *fA, double *fAm, double *fc, double *cc, double *vol)
{
    const int nct = n_cells + n_ghost;
    /* face areas by node count */
    for (int f = 0; f < n_faces; f++) {
        const int *fn = face_nodes + fnptr[f];
        const int nn = fnptr[f + 1] - fnptr[f];
        double *A = fA + 3 * f;
        if (nn == 2) {
            double dr[3];
            v_sub(nodes + 3 * fn[1], nodes + 3 * fn[0], dr);
            A[0] = dr[1];
            A[1] = -dr[0];
            A[2] = 0.;
        } else if (nn == 3) {
            double d10[3], d20[3], c[3];
            v_sub(nodes + 3 * fn[1], nodes + 3 * fn[0], d10);
            v_sub(nodes + 3 * fn[2], nodes + 3 * fn[0], d20);
            v_cross(d10, d20, c);
            for (int k = 0; k < 3; k++) A[k] = 0.5 * c[k];
        } else if (nn == 4) {
            double d20[3], d31[3], c[3];
            v_sub(nodes + 3 * fn[2], nodes + 3 * fn[0], d20);
            v_sub(nodes + 3 * fn[3], nodes + 3 * fn[1], d31);
            v_cross(d20, d31, c);
            for (int k = 0; k < 3; k++) A[k] = 0.5 * c[k];
        } else {
            A[0] = A[1] = A[2] = 0.0;
            for (int q = 0; q < nn; q++) {
                const double *a = nodes + 3 * fn[q], *b = nodes + 3 * fn[(q + 1) % nn];
                double xm[3], dr[3];
                for (int k = 0; k < 3; k++) xm[k] = 0.5 * (b[k] + a[k]);
                v_sub(b, a, dr);
                A[0] += xm[1] * dr[2];
                A[1] += xm[2] * dr[0];
                A[2] += xm[0] * dr[1];
            }
        }
        fAm[f] = v_mag(A);
    }
    /* face centroids: node average, then the correction of faces with more than 3 nodes */
    const double twoThirds = 2. / 3.;
    for (int f = 0; f < n_faces; f++) {
        double p[3];
        const int *fn = face_nodes + fnptr[f];
        const int nn = fnptr[f + 1] - fnptr[f];
        for (int k = 0; k < 3; k++) p[k] = nodes[3 * fn[0] + k];
        for (int q = 1; q < nn; q++)
            for (int k = 0; k < 3; k++) p[k] += nodes[3 * fn[q] + k];
        for (int k = 0; k < 3; k++) p[k] /= (double)nn;
        if (nn > 3) {
            double en[3];
            for (int k = 0; k < 3; k++) en[k] = fA[3 * f + k] / fAm[f];
            double denom = 0.0, cfc[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < nn; q++) {
                const double *a = nodes + 3 * fn[q], *b = nodes + 3 * fn[(q + 1) % nn];
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
            for (int k = 0; k < 3; k++) p[k] += cfc[k];
        }
        for (int k = 0; k < 3; k++) fc[3 * f + k] = p[k];
    }
    /* cell centroids: area-weighted face centroids, all faces in order */
    double *w = (double *)calloc((size_t)nct + 1, sizeof(double));
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
    const double rdim = (double)dim;
    for (int f = 0; f < n_faces; f++) {
        double d[3];
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        v_sub(fc + 3 * f, cc + 3 * c0, d);
        vol[c0] += v_dot(d, fA + 3 * f) / rdim;
        v_sub(fc + 3 * f, cc + 3 * c1, d);
        vol[c1] -= v_dot(d, fA + 3 * f) / rdim;
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

/* metrics and connectivity of one mesh */
typedef struct {
    double *fA, *fAm, *fc, *cc, *vol;
    int *row, *col, *p01, *p10;
} geom_t;

static void geom_build(const mesh_t *M, geom_t *G)
{
    const int nct = M->nc + M->ng, nf = M->nf;
    G->fA = (double *)malloc(sizeof(double) * 3 * (size_t)(nf + 1));
    G->fAm = (double *)malloc(sizeof(double) * (size_t)(nf + 1));
    G->fc = (double *)malloc(sizeof(double) * 3 * (size_t)(nf + 1));
    G->cc = (double *)malloc(sizeof(double) * 3 * (size_t)(nct + 1));
    G->vol = (double *)malloc(sizeof(double) * (size_t)(nct + 1));
    ofv_metrics(M->dim, nf, M->nodes, M->fnptr, M->fnodes, M->fcells, M->nc, M->ng, M->ngroups, M->gptr,
                G->fA, G->fAm, G->fc, G->cc, G->vol);
    G->row = (int *)malloc(sizeof(int) * ((size_t)nct + 1));
    G->p01 = (int *)malloc(sizeof(int) * (size_t)(nf + 1));
    G->p10 = (int *)malloc(sizeof(int) * (size_t)(nf + 1));
    cell_cells(nf, M->fcells, nct, G->row, &G->col, G->p01, G->p10);
}

static void geom_free(geom_t *G)
{
    free(G->fA); free(G->fAm); free(G->fc); free(G->cc); free(G->vol);
    free(G->row); free(G->col); free(G->p01); free(G->p10);
}

/* Least-squares gradient coefficients, three per cellCells entry
 * (GradientModel::getLeastSquaresGradientMatrix3D / 2D): unit steps to the
 * neighbours, the inverse of the weighted normal matrix applied per self row
 * when its determinant exceeds 1e-6 (3-D) or 1e-26 (2-D), then a division by
 * the step length; rows with a smaller determinant take the Green-Gauss face
 * term 0.5 A / V instead. */
static void lsq_coeffs(const mesh_t *M, const geom_t *G, double *gc)
{
    const int nct = M->nc + M->ng, nf = M->nf;
    const int *fcl = M->fcells;
    memset(gc, 0, sizeof(double) * 3 * (size_t)G->row[nct]);
    unsigned char *degen = (unsigned char *)calloc((size_t)nct + 1, 1);
    for (int f = 0; f < nf; f++) {
        const int c0 = fcl[2 * f], c1 = fcl[2 * f + 1];
        double ds[3];
        v_sub(G->cc + 3 * c1, G->cc + 3 * c0, ds);
        const double dsMag = v_mag(ds);
        for (int k = 0; k < 3; k++) {
            gc[3 * G->p01[f] + k] = ds[k] / dsMag;
            gc[3 * G->p10[f] + k] = -ds[k] / dsMag;
        }
    }
    for (int nc = 0; nc < M->nc; nc++) {
        const int rb = G->row[nc], re = G->row[nc + 1];
        if (M->dim == 2) {
            double Ixx = 0, Iyy = 0, Ixy = 0;
            for (int inb = rb; inb < re; inb++) {
                const double *ds = gc + 3 * inb;
                Ixx += ds[0] * ds[0];
                Iyy += ds[1] * ds[1];
                Ixy += ds[0] * ds[1];
            }
            const double det = Ixx * Iyy - Ixy * Ixy;
            if (det > 1e-26) {
                const double Kxx = Iyy / det, Kxy = -Ixy / det, Kyy = Ixx / det;
                for (int inb = rb; inb < re; inb++) {
                    const double d0 = gc[3 * inb], d1 = gc[3 * inb + 1];
                    gc[3 * inb] = (Kxx * d0 + Kxy * d1);
                    gc[3 * inb + 1] = (Kxy * d0 + Kyy * d1);
                    gc[3 * inb + 2] = 0;
                }
            } else {
                degen[nc] = 1;
            }
        } else {
            double Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
            for (int inb = rb; inb < re; inb++) {
                const double *ds = gc + 3 * inb;
                Ixx += ds[0] * ds[0];
                Iyy += ds[1] * ds[1];
                Izz += ds[2] * ds[2];
                Ixy += ds[0] * ds[1];
                Ixz += ds[0] * ds[2];
                Iyz += ds[1] * ds[2];
            }
            const double det = Ixx * (Iyy * Izz - Iyz * Iyz) - Ixy * (Ixy * Izz - Iyz * Ixz)
                               + Ixz * (Ixy * Iyz - Iyy * Ixz);
            if (det > 1e-6) {
                const double Kxx = (Iyy * Izz - Iyz * Iyz) / det;
                const double Kxy = -(Ixy * Izz - Iyz * Ixz) / det;
                const double Kxz = (Ixy * Iyz - Iyy * Ixz) / det;
                const double Kyy = (Ixx * Izz - Ixz * Ixz) / det;
                const double Kyz = -(Ixx * Iyz - Ixy * Ixz) / det;
                const double Kzz = (Ixx * Iyy - Ixy * Ixy) / det;
                for (int inb = rb; inb < re; inb++) {
                    const double d0 = gc[3 * inb], d1 = gc[3 * inb + 1], d2 = gc[3 * inb + 2];
                    gc[3 * inb] = (Kxx * d0 + Kxy * d1 + Kxz * d2);
                    gc[3 * inb + 1] = (Kxy * d0 + Kyy * d1 + Kyz * d2);
                    gc[3 * inb + 2] = (Kxz * d0 + Kyz * d1 + Kzz * d2);
                }
            } else {
                degen[nc] = 1;
            }
        }
    }
    for (int f = 0; f < nf; f++) {
        const int c0 = fcl[2 * f], c1 = fcl[2 * f + 1];
        double ds[3];
        v_sub(G->cc + 3 * c1, G->cc + 3 * c0, ds);
        const double dsMag = v_mag(ds);
        for (int k = 0; k < 3; k++) {
            gc[3 * G->p01[f] + k] /= dsMag;
            gc[3 * G->p10[f] + k] /= dsMag;
        }
    }
    for (int f = 0; f < nf; f++) {
        const int c0 = fcl[2 * f], c1 = fcl[2 * f + 1];
        if (degen[c0])
            for (int k = 0; k < 3; k++) gc[3 * G->p01[f] + k] = 0.5 * G->fA[3 * f + k] / G->vol[c0];
        if (degen[c1])
            for (int k = 0; k < 3; k++) gc[3 * G->p10[f] + k] = -0.5 * G->fA[3 * f + k] / G->vol[c1];
    }
    free(degen);
}

/* grad[c] = sum over cellCells entries of coeff * (x[j] - x[c]) for self cells;
 * a ghost cell takes the gradient of its boundary face's cell. */
static void gradient(const mesh_t *M, const geom_t *G, const double *gc, const double *x, double *grad)
{
    for (int nr = 0; nr < M->nc; nr++) {
        double g[3] = {0.0, 0.0, 0.0};
        for (int nb = G->row[nr]; nb < G->row[nr + 1]; nb++) {
            const double v = x[G->col[nb]] - x[nr];
            for (int k = 0; k < 3; k++) g[k] += gc[3 * nb + k] * v;
        }
        for (int k = 0; k < 3; k++) grad[3 * nr + k] = g[k];
    }
    for (int g = 1; g < M->ngroups; g++)
        for (int f = M->gptr[g]; f < M->gptr[g + 1]; f++) {
            const int c0 = M->fcells[2 * f], c1 = M->fcells[2 * f + 1];
            for (int k = 0; k < 3; k++) grad[3 * c1 + k] = grad[3 * c0 + k];
        }
}

int ofv_gradient(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                 const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                 const double *x, double *grad)
{
    const mesh_t M = {dim, n_faces, n_cells, n_ghost, n_groups, nodes, fnptr, face_nodes, face_cells, gptr};
    geom_t G;
    geom_build(&M, &G);
    double *gc = (double *)malloc(sizeof(double) * 3 * (size_t)(G.row[n_cells + n_ghost] + 1));
    lsq_coeffs(&M, &G, gc);
    gradient(&M, &G, gc, x, grad);
    free(gc);
    geom_free(&G);
    return 0;
}

static double harmonic(double x0, double x1)
{
    const double sum = x0 + x1;
    if (x0 + x1 != 0.0) return 2.0 * x0 * x1 / sum;
    return sum;
}

/* LinearSystem::initSolve -> eliminateBoundaryEquations (ghost rows, ascending) */
static void eliminate(int n_cells, int nct, const int *row, const int *col, const unsigned char *is_bnd,
                      double *diag, double *off, double *b)
{
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
}

int ofv_thermal(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                const int *bc_type, const double *bc_value, const double *kcell, const double *src,
                const double *x, const double *conv_flux, const double *cont_res, int central,
                int *row, int *col, double *diag, double *off, double *b, unsigned char *is_bnd)
{
    const mesh_t M = {dim, n_faces, n_cells, n_ghost, n_groups, nodes, fnptr, face_nodes, face_cells, gptr};
    const int nct = n_cells + n_ghost;
    geom_t G;
    geom_build(&M, &G);
    const int *p01 = G.p01, *p10 = G.p10;
    memcpy(row, G.row, sizeof(int) * ((size_t)nct + 1));
    memcpy(col, G.col, sizeof(int) * (size_t)G.row[nct]);
    /* GradientModel::compute of the current field */
    double *gc = (double *)malloc(sizeof(double) * 3 * (size_t)(G.row[nct] + 1));
    double *grad = (double *)malloc(sizeof(double) * 3 * (size_t)(nct + 1));
    lsq_coeffs(&M, &G, gc);
    gradient(&M, &G, gc, x, grad);
    /* LinearSystem::initAssembly: zero matrix and b */
    memset(diag, 0, sizeof(double) * (size_t)nct);
    memset(off, 0, sizeof(double) * (size_t)row[nct]);
    memset(b, 0, sizeof(double) * (size_t)nct);
    memset(is_bnd, 0, (size_t)nct);
    /* the boundary conditions write the field (Dirichlet ghosts) */
    double *xf = (double *)malloc(sizeof(double) * (size_t)(nct + 1));
    memcpy(xf, x, sizeof(double) * (size_t)nct);
    /* DiffusionDiscretization::discretize over all face groups */
    for (int f = 0; f < n_faces; f++) {
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        const double vol0 = G.vol[c0], vol1 = G.vol[c1];
        double ds[3];
        v_sub(G.cc + 3 * c1, G.cc + 3 * c0, ds);
        double kf;
        if (vol0 == 0.) kf = kcell[c1];
        else if (vol1 == 0.) kf = kcell[c0];
        else kf = harmonic(kcell[c0], kcell[c1]);
        const double *A = G.fA + 3 * f;
        const double diffMetric = G.fAm[f] * G.fAm[f] / v_dot(A, ds);
        const double diffCoeff = kf * diffMetric;
        double sec[3], gradF[3];
        for (int k = 0; k < 3; k++) sec[k] = kf * (A[k] - ds[k] * diffMetric);
        for (int k = 0; k < 3; k++) gradF[k] = (grad[3 * c0 + k] * vol0 + grad[3 * c1 + k] * vol1) / (vol0 + vol1);
        double dsec = 0.0;
        for (int k = 0; k < 3; k++) dsec += gradF[k] * sec[k];
        const double dFlux = diffCoeff * (xf[c1] - xf[c0]) + dsec;
        b[c0] += dFlux;
        b[c1] -= dFlux;
        off[p01[f]] += diffCoeff;
        off[p10[f]] += diffCoeff;
        diag[c0] -= diffCoeff;
        diag[c1] -= diffCoeff;
    }
    /* ConvectionDiscretization::discretize (no grid flux, no immersed boundary) */
    if (conv_flux) {
        for (int f = 0; f < n_faces; f++) {
            const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
            const double F = conv_flux[f];
            if (central) {
                /* the reference's central branch evaluates 0.5 F (x[c0] + x[c0]) and
                 * linearizes as upwind */
                const double varFlux = 0.5 * F * (xf[c0] + xf[c0]);
                b[c0] -= varFlux;
                b[c1] += varFlux;
                if (F > 0.0) {
                    diag[c0] -= F;
                    off[p10[f]] += F;
                } else {
                    diag[c1] += F;
                    off[p01[f]] -= F;
                }
            } else {
                double varFlux;
                if (F > 0.0) {
                    varFlux = F * xf[c0];
                    diag[c0] -= F;
                    off[p10[f]] += F;
                } else {
                    varFlux = F * xf[c1];
                    diag[c1] += F;
                    off[p01[f]] -= F;
                }
                b[c0] -= varFlux;
                b[c1] += varFlux;
            }
        }
        for (int c = 0; c < n_cells; c++) diag[c] += cont_res ? cont_res[c] : 0.0;
    }
    /* SourceDiscretization */
    for (int c = 0; c < n_cells; c++) b[c] += G.vol[c] * src[c];
    /* boundary conditions, groups in order */
    for (int g = 1; g < n_groups; g++)
        for (int f = gptr[g]; f < gptr[g + 1]; f++) {
            const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
            if (bc_type[g] == OFV_BC_DIRICHLET && !(conv_flux && conv_flux[f] > 0.)) {
                const double dX = bc_value[g] - xf[c1];
                b[c0] += off[p01[f]] * dX;
                off[p01[f]] = 0.0;
                xf[c1] = bc_value[g];
                off[p10[f]] = 0.0;
                b[c1] = 0.0;
                diag[c1] = -1.0;
            } else if (bc_type[g] == OFV_BC_DIRICHLET) {
                /* outflow face of a specified-temperature group: extrapolation */
                const double dFdX1 = -diag[c1];
                const double dx01 = xf[c0] - xf[c1];
                diag[c0] += dFdX1;
                b[c0] += dFdX1 * dx01;
                off[p01[f]] = 0.0;
                diag[c1] = -1.0;
                off[p10[f]] = 1.0;
                b[c1] = dx01;
                is_bnd[c1] = 1;
            } else {
                const double fluxB = -b[c1];
                b[c1] = bc_value[g] * G.fAm[f] - fluxB;
                is_bnd[c1] = 1;
            }
        }
    eliminate(n_cells, nct, row, col, is_bnd, diag, off, b);
    free(xf); free(gc); free(grad);
    geom_free(&G);
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
/* One velocity direction of the ESBGK kinetic model: KineticModel::linearizeKineticModel
 * (esbgkbase/KineticModel.h:1505-1674). Discretizations in the reference's order:
 * the collision term first (CollisionTermDiscretization.h:46-62:
 * r -= V nu (f - feq), diag -= V nu, self cells), then upwind convection with the
 * face flux A.c (ConvectionDiscretization_Kmodel.h:171-197, no grid flux, no central
 * difference), then the boundary faces group by group (KineticModel.h:1596-1661):
 * bc_kind[g] = 1 (ZeroGradBC): every direction takes the Dirichlet form with the
 * ghost's own value; bc_kind[g] = 0 (WallBC, SymmetryBC and the other wall-like
 * groups): Dirichlet when c.en - uw.en < eps (incoming), else extrapolation
 * (GenericKineticBCS.h:76-116 and :139-173), uw = wall_u[3g..3g+2]. Then the
 * elimination of LinearSystem::initSolve. */
int ofv_kinetic(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                const double *c, const double *nue, const double *feq, const double *x, const int *bc_kind,
                const double *wall_u, double eps, int *row, int *col, double *diag, double *off, double *b,
                unsigned char *is_bnd)
{
    const mesh_t M = {dim, n_faces, n_cells, n_ghost, n_groups, nodes, fnptr, face_nodes, face_cells, gptr};
    const int nct = n_cells + n_ghost;
    geom_t G;
    geom_build(&M, &G);
    const int *p01 = G.p01, *p10 = G.p10;
    memcpy(row, G.row, sizeof(int) * ((size_t)nct + 1));
    memcpy(col, G.col, sizeof(int) * (size_t)G.row[nct]);
    memset(diag, 0, sizeof(double) * (size_t)nct);
    memset(off, 0, sizeof(double) * (size_t)row[nct]);
    memset(b, 0, sizeof(double) * (size_t)nct);
    memset(is_bnd, 0, (size_t)nct);
    double *xf = (double *)malloc(sizeof(double) * (size_t)(nct + 1));
    memcpy(xf, x, sizeof(double) * (size_t)nct);
    for (int cc = 0; cc < n_cells; cc++) {
        b[cc] -= G.vol[cc] * nue[cc] * (xf[cc] - feq[cc]);
        diag[cc] -= G.vol[cc] * nue[cc];
    }
    for (int f = 0; f < n_faces; f++) {
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        const double *A = G.fA + 3 * f;
        const double F = A[0] * c[0] + A[1] * c[1] + A[2] * c[2];
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
            int incoming = 1;
            if (bc_kind[g] != 1) {
                double en[3];
                for (int k = 0; k < 3; k++) en[k] = G.fA[3 * f + k] / G.fAm[f];
                const double cn = c[0] * en[0] + c[1] * en[1] + c[2] * en[2];
                const double *uw = wall_u + 3 * g;
                const double un = uw[0] * en[0] + uw[1] * en[1] + uw[2] * en[2];
                incoming = cn - un < eps;
            }
            if (incoming) {
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
    eliminate(n_cells, nct, row, col, is_bnd, diag, off, b);
    free(xf);
    geom_free(&G);
    return 0;
}

int ofv_phonon(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
               const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
               const double *vg, double tau, const double *e0, const double *x, int *row, int *col,
               double *diag, double *off, double *b, unsigned char *is_bnd)
{
    const mesh_t M = {dim, n_faces, n_cells, n_ghost, n_groups, nodes, fnptr, face_nodes, face_cells, gptr};
    const int nct = n_cells + n_ghost;
    geom_t G;
    geom_build(&M, &G);
    const int *p01 = G.p01, *p10 = G.p10;
    memcpy(row, G.row, sizeof(int) * ((size_t)nct + 1));
    memcpy(col, G.col, sizeof(int) * (size_t)G.row[nct]);
    memset(diag, 0, sizeof(double) * (size_t)nct);
    memset(off, 0, sizeof(double) * (size_t)row[nct]);
    memset(b, 0, sizeof(double) * (size_t)nct);
    memset(is_bnd, 0, (size_t)nct);
    /* the boundary-condition code writes the field (Dirichlet ghosts) */
    double *xf = (double *)malloc(sizeof(double) * (size_t)(nct + 1));
    memcpy(xf, x, sizeof(double) * (size_t)nct);
    for (int c = 0; c < n_cells; c++) {
        b[c] -= G.vol[c] * (xf[c] - e0[c]) / tau;
        diag[c] -= G.vol[c] / tau;
    }
    for (int f = 0; f < n_faces; f++) {
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        const double *A = G.fA + 3 * f;
        const double F = A[0] * vg[0] + A[1] * vg[1] + A[2] * vg[2];
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
            for (int k = 0; k < 3; k++) en[k] = G.fA[3 * f + k] / G.fAm[f];
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
    eliminate(n_cells, nct, row, col, is_bnd, diag, off, b);
    free(xf);
    geom_free(&G);
    return 0;
}
