/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the c-PRIMED/fvm algebraic multigrid
 * (pairwise agglomeration, Galerkin coarse operators, V/W/F cycles with
 * Gauss-Seidel or Jacobi smoothing) used to check the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker or the timed CPU baseline; the product library
 * never calls it.
 *
 * Every floating-point expression keeps the operand order of the reference
 * loops so that results are bit-identical to the reference algorithm compiled
 * without FMA contraction (build with -ffp-contract=off).
 *
 * Citation prefix: src/fvm/src/modules/fvmbase/ (M:) of the reference.
 *
 * Pinned by (tests/test_oracle_kat.py and tests/test_oracle_props.py):
 *   - KAT-226 (MatrixMarket226.dat): coarse sizes 108/48/20/8/3, r0 6981.57,
 *     40 cycles, final 5.32223e-05 (src/fvm/test/testLinearSolver.out:5-11)
 *   - KAT-cav32 (THERMAL_MATRIX golden matrix): r0 63200, 56 cycles,
 *     5.75812e-05 at relTol 1e-9 (AMG_MERGING_THERMAL/proc1/GOLDEN/convergence.dat)
 *   - residual / GS / Jacobi against dense numpy algebra (closed forms)
 *   - coarsening against hand-derived maps on chains and stars, invariants
 *   - Galerkin against dense P^T A P
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amg_oracle.h"

/* ------------------------------------------------------------------ */
/* matrix helpers                                                     */
/* ------------------------------------------------------------------ */

static void om_mat_free(om_mat *m)
{
    free(m->row); free(m->col); free(m->diag); free(m->off); free(m->is_bnd);
    memset(m, 0, sizeof(*m));
}

static int om_mat_alloc(om_mat *m, int n_self, int n_total, int nnz)
{
    memset(m, 0, sizeof(*m));
    m->n_self = n_self;
    m->n_total = n_total;
    m->nnz = nnz;
    m->row = (int *)calloc((size_t)n_total + 1, sizeof(int));
    m->col = (int *)calloc(nnz > 0 ? (size_t)nnz : 1, sizeof(int));
    m->diag = (double *)calloc(n_total > 0 ? (size_t)n_total : 1, sizeof(double));
    m->off = (double *)calloc(nnz > 0 ? (size_t)nnz : 1, sizeof(double));
    m->is_bnd = (unsigned char *)calloc(n_total > 0 ? (size_t)n_total : 1, 1);
    if (!m->row || !m->col || !m->diag || !m->off || !m->is_bnd) {
        om_mat_free(m);
        return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* MatrixMarket ingest: M:../importers/MMReader.cpp:80-184             */
/* two passes, off-diagonals appended in file order, b[i] = -rhs[i]    */
/* ------------------------------------------------------------------ */

static int mm_header(FILE *fp, int *nr, int *nc, int *nz, int *sym)
{
    char a[64], b[64], c[64], d[64], e[64];
    if (fscanf(fp, "%63s", a) != 1) return -1;
    if (strcmp(a, "%%") == 0) {
        if (fscanf(fp, "%63s", b) != 1 || strcmp(b, "MatrixMarket") != 0) return -1;
    } else if (strcmp(a, "%%MatrixMarket") != 0) return -1;
    if (fscanf(fp, "%63s %63s %63s %63s", b, c, d, e) != 4) return -1;
    if (strcmp(b, "matrix") || strcmp(c, "coordinate") || strcmp(d, "real")) return -1;
    if (strcmp(e, "symmetric") == 0) *sym = 1;
    else if (strcmp(e, "general") == 0) *sym = 0;
    else return -1;
    if (fscanf(fp, "%d %d %d", nr, nc, nz) != 3) return -1;
    return (*nr == *nc) ? 0 : -1;
}

int om_read_mm(const char *mat_path, const char *rhs_path, om_mat *m,
               double **b_out, int *is_symmetric)
{
    FILE *fp = fopen(mat_path, "r");
    if (!fp) return -1;
    int nr, nc, nz, sym;
    if (mm_header(fp, &nr, &nc, &nz, &sym)) { fclose(fp); return -2; }
    int *cnt = (int *)calloc((size_t)nr + 1, sizeof(int));
    int total = 0;
    for (int k = 0; k < nz; k++) {
        int i, j; double v;
        if (fscanf(fp, "%d %d %lf", &i, &j, &v) != 3) { free(cnt); fclose(fp); return -3; }
        i -= 1; j -= 1;
        if (i != j) {
            cnt[i]++; total++;
            if (sym) { cnt[j]++; total++; }
        }
    }
    if (om_mat_alloc(m, nr, nr, total)) { free(cnt); fclose(fp); return -4; }
    m->row[0] = 0;
    for (int i = 0; i < nr; i++) m->row[i + 1] = m->row[i] + cnt[i];
    int *fill = cnt;
    for (int i = 0; i < nr; i++) fill[i] = m->row[i];
    rewind(fp);
    mm_header(fp, &nr, &nc, &nz, &sym);
    for (int k = 0; k < nz; k++) {
        int i, j; double v;
        if (fscanf(fp, "%d %d %lf", &i, &j, &v) != 3) break;
        i -= 1; j -= 1;
        if (i != j) {
            int p = fill[i]++;
            m->col[p] = j; m->off[p] = v;
            if (sym) { p = fill[j]++; m->col[p] = i; m->off[p] = v; }
        } else {
            m->diag[i] = v;
        }
    }
    fclose(fp);
    free(cnt);
    double *b = (double *)calloc((size_t)nr, sizeof(double));
    FILE *fb = fopen(rhs_path, "r");
    if (!fb) { free(b); om_mat_free(m); return -5; }
    for (int i = 0; i < nr; i++) {
        double r = 0;
        if (fscanf(fb, "%lf", &r) != 1) { r = 0; }
        b[i] = -r;
    }
    fclose(fb);
    *b_out = b;
    *is_symmetric = sym;
    return 0;
}

/* ------------------------------------------------------------------ */
/* kernels: M:CRMatrix.h                                              */
/* ------------------------------------------------------------------ */

/* r_i = b_i + a_ii x_i + sum_j a_ij x_j over self rows; CRMatrix.h:407-426 */
void om_residual(const om_mat *A, const double *x, const double *b, double *r)
{
    for (int i = 0; i < A->n_self; i++) {
        double s = b[i] + A->diag[i] * x[i];
        for (int p = A->row[i]; p < A->row[i + 1]; p++) s += A->off[p] * x[A->col[p]];
        r[i] = s;
    }
}

/* sequential 1-norm over self rows; Array.h:302-311, NumType.h:108 */
double om_one_norm(const double *r, int n)
{
    double s = 0;
    for (int i = 0; i < n; i++) s += fabs(r[i]);
    return s;
}

/* CRMatrix.h:303-321 */
void om_forward_gs(const om_mat *A, double *x, const double *b)
{
    for (int i = 0; i < A->n_self; i++) {
        double s = b[i];
        for (int p = A->row[i]; p < A->row[i + 1]; p++) s += A->off[p] * x[A->col[p]];
        x[i] = -s / A->diag[i];
    }
}

/* CRMatrix.h:328-346 */
void om_reverse_gs(const om_mat *A, double *x, const double *b)
{
    for (int i = A->n_self - 1; i >= 0; i--) {
        double s = b[i];
        for (int p = A->row[i]; p < A->row[i + 1]; p++) s += A->off[p] * x[A->col[p]];
        x[i] = -s / A->diag[i];
    }
}

/* CRMatrix.h:353-374 then copy of self rows, MultiFieldMatrix.cpp:168-217 */
void om_jacobi(const om_mat *A, double *x, const double *b, double *tmp)
{
    for (int i = 0; i < A->n_self; i++) {
        double s = b[i];
        for (int p = A->row[i]; p < A->row[i + 1]; p++) s += A->off[p] * x[A->col[p]];
        tmp[i] = -s / A->diag[i];
    }
    memcpy(x, tmp, sizeof(double) * (size_t)A->n_self);
}

/* ------------------------------------------------------------------ */
/* setup: CRMatrix.h:468-758, MultiFieldMatrix.cpp:627-667            */
/* ------------------------------------------------------------------ */

/* greedy pairwise agglomeration, CRMatrix.h:468-586 */
int om_coarsen(const om_mat *A, int group_size, double theta, int *ci)
{
    const int n = A->n_self;
    for (int i = 0; i < A->n_total; i++) ci[i] = -1;
    int *count = (int *)calloc(n > 0 ? (size_t)n : 1, sizeof(int));
    int nc = 0;
    for (int nr = 0; nr < n; nr++) {
        if (ci[nr] != -1 || A->is_bnd[nr]) continue;
        int current = nr;
        int col_g = -1, col_u = -1;
        int ng;
        ci[current] = nc;
        for (ng = 1; ng < group_size; ng++) {
            double max_u = 0, max_g = 0;
            col_g = -1; col_u = -1;
            for (int p = A->row[current]; p < A->row[current + 1]; p++) {
                const int c = A->col[p];
                if (c < n && !A->is_bnd[c]) {
                    const double d0 = fabs(A->diag[nr]);   /* seed row, not current */
                    const double d1 = fabs(A->diag[c]);
                    const double cm = fabs(A->off[p]);
                    const double w = fabs(cm / (d0 < d1 ? d1 : d0));   /* std::max(d0,d1) */
                    if (ci[c] == -1) {
                        if (col_u == -1 || w > max_u) { col_u = c; max_u = w; }
                    } else if (ci[c] != ci[nr]) {
                        if (col_g == -1 || w > max_g) { col_g = c; max_g = w; }
                    }
                }
            }
            if (col_u != -1 && (col_g == -1 || max_u > theta * max_g)) {
                ci[col_u] = ci[current];
                count[ci[current]]++;
                current = col_u;
            } else {
                break;
            }
        }
        if (ng > 1 || col_g == -1 || count[ci[col_g]] > group_size + 2) {
            count[ci[nr]]++;
            nc++;
        } else {
            ci[nr] = ci[col_g];
            count[ci[col_g]]++;
        }
    }
    free(count);
    return nc;
}

/* transpose of the fine->coarse map, ascending fine index;
 * MultiFieldMatrix.cpp:627-667 */
void om_coarse_to_fine(const int *ci, int n_total_fine, int nc, int *c2f_row, int *c2f_col)
{
    for (int I = 0; I <= nc; I++) c2f_row[I] = 0;
    for (int i = 0; i < n_total_fine; i++)
        if (ci[i] >= 0) c2f_row[ci[i] + 1]++;
    for (int I = 0; I < nc; I++) c2f_row[I + 1] += c2f_row[I];
    int *fill = (int *)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    for (int I = 0; I < nc; I++) fill[I] = c2f_row[I];
    for (int i = 0; i < n_total_fine; i++)
        if (ci[i] >= 0) c2f_col[fill[ci[i]]++] = i;
    free(fill);
}

/* coarse connectivity (CRMatrix.h:597-691) and Galerkin values
 * (CRMatrix.h:699-758) */
int om_coarse_matrix(const om_mat *A, const int *ci, int nc, om_mat *C)
{
    int *c2f_row = (int *)malloc(sizeof(int) * ((size_t)nc + 1));
    int *c2f_col = (int *)malloc(sizeof(int) * (size_t)(A->n_total > 0 ? A->n_total : 1));
    om_coarse_to_fine(ci, A->n_total, nc, c2f_row, c2f_col);
    unsigned char *mark = (unsigned char *)calloc(nc > 0 ? (size_t)nc : 1, 1);
    int *rowcnt = (int *)calloc((size_t)nc + 1, sizeof(int));
    long long total = 0;
    for (int I = 0; I < nc; I++) {
        int k = 0;
        for (int q = c2f_row[I]; q < c2f_row[I + 1]; q++) {
            const int i = c2f_col[q];
            for (int p = A->row[i]; p < A->row[i + 1]; p++) {
                const int J = ci[A->col[p]];
                if (J >= 0 && J != I && !mark[J]) { mark[J] = 1; k++; }
            }
        }
        for (int q = c2f_row[I]; q < c2f_row[I + 1]; q++) {
            const int i = c2f_col[q];
            for (int p = A->row[i]; p < A->row[i + 1]; p++) {
                const int J = ci[A->col[p]];
                if (J >= 0) mark[J] = 0;
            }
        }
        rowcnt[I] = k;
        total += k;
    }
    if (om_mat_alloc(C, nc, nc, (int)total)) return -1;
    C->row[0] = 0;
    for (int I = 0; I < nc; I++) C->row[I + 1] = C->row[I] + rowcnt[I];
    int *pos = (int *)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    for (int I = 0; I < nc; I++) {
        int k = C->row[I];
        for (int q = c2f_row[I]; q < c2f_row[I + 1]; q++) {
            const int i = c2f_col[q];
            for (int p = A->row[i]; p < A->row[i + 1]; p++) {
                const int J = ci[A->col[p]];
                if (J >= 0 && J != I && !mark[J]) { mark[J] = 1; C->col[k++] = J; }
            }
        }
        for (int p = C->row[I]; p < C->row[I + 1]; p++) mark[C->col[p]] = 0;
        /* Galerkin row I: D_I and A_IJ accumulated in fine-row then CSR order */
        for (int p = C->row[I]; p < C->row[I + 1]; p++) pos[C->col[p]] = p;
        double d = 0;
        for (int q = c2f_row[I]; q < c2f_row[I + 1]; q++) {
            const int i = c2f_col[q];
            d += A->diag[i];
            for (int p = A->row[i]; p < A->row[i + 1]; p++) {
                const int J = ci[A->col[p]];
                if (J < 0) continue;
                if (J != I) C->off[pos[J]] += A->off[p];
                else d += A->off[p];
            }
        }
        C->diag[I] = d;
    }
    free(pos); free(mark); free(rowcnt); free(c2f_row); free(c2f_col);
    return 0;
}

/* ------------------------------------------------------------------ */
/* AMG object: M:AMG.cpp                                              */
/* ------------------------------------------------------------------ */

struct om_amg {
    om_params prm;
    int nlev;                 /* number of levels including the finest */
    om_mat lev[OM_MAX_LEVELS];
    int *ci[OM_MAX_LEVELS];   /* fine->coarse map of level l (length n_total of l) */
    double *b[OM_MAX_LEVELS], *x[OM_MAX_LEVELS], *r[OM_MAX_LEVELS];
    int owns_finest;
    long long total_iterations;
};

void om_default_params(om_params *p)
{
    /* AMG.cpp:15-36, LinearSolver.h:19-24 */
    p->max_coarse_levels = 30;
    p->n_pre_sweeps = 0;
    p->n_post_sweeps = 1;
    p->coarse_group_size = 2;
    p->weight_ratio_threshold = 0.65;
    p->cycle_type = OM_V_CYCLE;
    p->smoother_type = OM_GAUSS_SEIDEL;
    p->scale_corrections = 1;
    p->n_max_iterations = 100;
    p->verbosity = 0;
    p->relative_tolerance = 1e-8;
    p->absolute_tolerance = 1e-50;
    p->parallel_stop_rule = 1;
    p->is_symmetric = 0;
}

om_amg *om_amg_create(void)
{
    om_amg *a = (om_amg *)calloc(1, sizeof(om_amg));
    om_default_params(&a->prm);
    return a;
}

void om_amg_cleanup(om_amg *a)
{
    for (int l = 0; l < a->nlev; l++) {
        if (l > 0 || a->owns_finest) om_mat_free(&a->lev[l]);
        free(a->ci[l]); free(a->b[l]); free(a->x[l]); free(a->r[l]);
        a->ci[l] = NULL; a->b[l] = a->x[l] = a->r[l] = NULL;
    }
    a->nlev = 0;
}

void om_amg_destroy(om_amg *a)
{
    if (!a) return;
    om_amg_cleanup(a);
    free(a);
}

void om_amg_set_params(om_amg *a, const om_params *p) { a->prm = *p; }
void om_amg_get_params(const om_amg *a, om_params *p) { *p = a->prm; }

/* hierarchy construction, AMG.cpp:149-210 (stop rule C12 selectable) */
int om_amg_setup(om_amg *a, int n_self, int n_total, const int *row, const int *col,
                 const double *diag, const double *off, const unsigned char *is_bnd)
{
    om_amg_cleanup(a);
    om_mat *F = &a->lev[0];
    const int nnz = row[n_total];
    if (om_mat_alloc(F, n_self, n_total, nnz)) return -1;
    memcpy(F->row, row, sizeof(int) * ((size_t)n_total + 1));
    memcpy(F->col, col, sizeof(int) * (size_t)nnz);
    memcpy(F->diag, diag, sizeof(double) * (size_t)n_total);
    memcpy(F->off, off, sizeof(double) * (size_t)nnz);
    if (is_bnd) memcpy(F->is_bnd, is_bnd, (size_t)n_total);
    a->owns_finest = 1;
    a->nlev = 1;
    for (int k = 0; k < a->prm.max_coarse_levels && a->nlev < OM_MAX_LEVELS; k++) {
        om_mat *A = &a->lev[a->nlev - 1];
        int *ci = (int *)malloc(sizeof(int) * (size_t)(A->n_total > 0 ? A->n_total : 1));
        const int nc = om_coarsen(A, a->prm.coarse_group_size, a->prm.weight_ratio_threshold, ci);
        if (nc == A->n_total) { free(ci); break; }          /* size unchanged: stop */
        if (!a->prm.parallel_stop_rule && nc <= 3) { free(ci); break; }
        om_mat C;
        if (om_coarse_matrix(A, ci, nc, &C)) { free(ci); return -1; }
        a->ci[a->nlev - 1] = ci;
        a->lev[a->nlev] = C;
        a->nlev++;
        if (a->prm.verbosity > 1)
            printf("Created coarse level %d of size %d\n", k, nc);
        if (a->prm.parallel_stop_rule && nc <= 3) break;
    }
    for (int l = 0; l < a->nlev; l++) {
        size_t n = (size_t)(a->lev[l].n_total > 0 ? a->lev[l].n_total : 1);
        a->b[l] = (double *)calloc(n, sizeof(double));
        a->x[l] = (double *)calloc(n, sizeof(double));
        a->r[l] = (double *)calloc(n, sizeof(double));
    }
    return 0;
}

static void do_sweeps(om_amg *a, int nsweeps, int l)
{
    const om_mat *A = &a->lev[l];
    for (int s = 0; s < nsweeps; s++) {
        if (a->prm.smoother_type == OM_GAUSS_SEIDEL) {
            om_forward_gs(A, a->x[l], a->b[l]);
            om_reverse_gs(A, a->x[l], a->b[l]);
        } else {
            om_jacobi(A, a->x[l], a->b[l], a->r[l]);
            om_jacobi(A, a->x[l], a->b[l], a->r[l]);
        }
    }
}

/* AMG.cpp:70-147 */
static void cycle(om_amg *a, int type, int l)
{
    do_sweeps(a, a->prm.n_pre_sweeps, l);
    if (l < a->nlev - 1) {
        const om_mat *A = &a->lev[l];
        const om_mat *C = &a->lev[l + 1];
        om_residual(A, a->x[l], a->b[l], a->r[l]);
        memset(a->b[l + 1], 0, sizeof(double) * (size_t)C->n_total);
        memset(a->x[l + 1], 0, sizeof(double) * (size_t)C->n_total);
        const int *ci = a->ci[l];
        for (int i = 0; i < A->n_self; i++)            /* Array.h:427-436 */
            if (ci[i] >= 0) a->b[l + 1][ci[i]] += a->r[l][i];
        cycle(a, type, l + 1);
        if (type == OM_W_CYCLE) cycle(a, OM_W_CYCLE, l + 1);
        else if (type == OM_F_CYCLE) cycle(a, OM_V_CYCLE, l + 1);
        if (a->prm.is_symmetric && a->prm.scale_corrections) {
            /* AMG.cpp:118-135: scale = -(x.b)/(x^T A x), limited to [1,1] */
            const double *xc = a->x[l + 1], *bc = a->b[l + 1];
            double xb = 0;
            for (int I = 0; I < C->n_self; I++) xb += xc[I] * bc[I];
            double mxb = -xb;
            double xax = 0;
            for (int I = 0; I < C->n_self; I++) {
                double sn = C->diag[I] * xc[I];
                for (int p = C->row[I]; p < C->row[I + 1]; p++) sn += C->off[p] * xc[C->col[p]];
                xax += sn * xc[I];
            }
            double scale = mxb;
            if (xax != 0) scale /= xax;
            if (scale < 1.0) scale = 1.0;
            else if (scale > 1.0) scale = 1.0;
            for (int i = 0; i < A->n_self; i++)       /* Array.h:451-455 */
                if (ci[i] >= 0) a->x[l][i] += xc[ci[i]] * scale;
        } else {
            for (int i = 0; i < A->n_self; i++)       /* Array.h:462-465 */
                if (ci[i] >= 0) a->x[l][i] += a->x[l + 1][ci[i]];
        }
    }
    do_sweeps(a, a->prm.n_post_sweeps, l);
}

/* AMG.cpp:219-282; x_inout is delta (the correction); returns 0 */
int om_amg_solve(om_amg *a, const double *b, double *x_inout, double *rnorm0,
                 double *rnorm_final, int *cycles, double *history, int hist_cap)
{
    if (a->nlev < 1) return -1;
    const om_mat *F = &a->lev[0];
    memcpy(a->b[0], b, sizeof(double) * (size_t)F->n_total);
    memcpy(a->x[0], x_inout, sizeof(double) * (size_t)F->n_total);
    om_residual(F, a->x[0], a->b[0], a->r[0]);
    const double r0 = om_one_norm(a->r[0], F->n_self);
    if (a->prm.verbosity > 0) printf("0: %g\n", r0);
    if (history && hist_cap > 0) history[0] = r0;
    double rn = r0;
    int it = 0;
    if (!(r0 < a->prm.absolute_tolerance)) {
        for (int i = 1; i < a->prm.n_max_iterations; i++) {
            a->total_iterations++;
            cycle(a, a->prm.cycle_type, 0);
            om_residual(F, a->x[0], a->b[0], a->r[0]);
            rn = om_one_norm(a->r[0], F->n_self);
            double ratio = rn;
            if (r0 != 0) ratio /= r0;                 /* NumType.h:114 */
            it = i;
            if (history && i < hist_cap) history[i] = rn;
            if (a->prm.verbosity > 0) printf("%d: %g\n", i, rn);
            if (rn < a->prm.absolute_tolerance || ratio < a->prm.relative_tolerance) break;
        }
    }
    memcpy(x_inout, a->x[0], sizeof(double) * (size_t)F->n_total);
    if (rnorm0) *rnorm0 = r0;
    if (rnorm_final) *rnorm_final = rn;
    if (cycles) *cycles = it;
    return 0;
}

/* AMG.cpp:285-298: one cycle from the given x */
int om_amg_smooth(om_amg *a, const double *b, double *x_inout)
{
    if (a->nlev < 1) return -1;
    const om_mat *F = &a->lev[0];
    memcpy(a->b[0], b, sizeof(double) * (size_t)F->n_total);
    memcpy(a->x[0], x_inout, sizeof(double) * (size_t)F->n_total);
    cycle(a, a->prm.cycle_type, 0);
    memcpy(x_inout, a->x[0], sizeof(double) * (size_t)F->n_total);
    return 0;
}

long long om_amg_total_iterations(const om_amg *a) { return a->total_iterations; }
int om_amg_num_levels(const om_amg *a) { return a->nlev; }

int om_amg_level_info(const om_amg *a, int l, int *n_self, int *n_total, int *nnz)
{
    if (l < 0 || l >= a->nlev) return -1;
    *n_self = a->lev[l].n_self;
    *n_total = a->lev[l].n_total;
    *nnz = a->lev[l].nnz;
    return 0;
}

int om_amg_level_get(const om_amg *a, int l, int *row, int *col, double *diag,
                     double *off, int *ci)
{
    if (l < 0 || l >= a->nlev) return -1;
    const om_mat *A = &a->lev[l];
    if (row) memcpy(row, A->row, sizeof(int) * ((size_t)A->n_total + 1));
    if (col) memcpy(col, A->col, sizeof(int) * (size_t)A->nnz);
    if (diag) memcpy(diag, A->diag, sizeof(double) * (size_t)A->n_total);
    if (off) memcpy(off, A->off, sizeof(double) * (size_t)A->nnz);
    if (ci) {
        if (l < a->nlev - 1) memcpy(ci, a->ci[l], sizeof(int) * (size_t)A->n_total);
        else for (int i = 0; i < A->n_total; i++) ci[i] = -1;
    }
    return 0;
}

/* flat entry points for ctypes (test use) */
int om_mm_load(const char *mat_path, const char *rhs_path, int *n_out, int *nnz_out,
               int *sym_out, int *row, int *col, double *diag, double *off, double *b)
{
    om_mat m; double *bb = NULL; int sym = 0;
    int rc = om_read_mm(mat_path, rhs_path, &m, &bb, &sym);
    if (rc) return rc;
    *n_out = m.n_self; *nnz_out = m.nnz; *sym_out = sym;
    if (row) memcpy(row, m.row, sizeof(int) * ((size_t)m.n_total + 1));
    if (col) memcpy(col, m.col, sizeof(int) * (size_t)m.nnz);
    if (diag) memcpy(diag, m.diag, sizeof(double) * (size_t)m.n_total);
    if (off) memcpy(off, m.off, sizeof(double) * (size_t)m.nnz);
    if (b) memcpy(b, bb, sizeof(double) * (size_t)m.n_total);
    om_mat_free(&m);
    free(bb);
    return 0;
}

void om_residual_flat(int n_self, const int *row, const int *col, const double *diag,
                      const double *off, const double *x, const double *b, double *r)
{
    om_mat A = {0};
    A.n_self = n_self; A.row = (int *)row; A.col = (int *)col;
    A.diag = (double *)diag; A.off = (double *)off;
    om_residual(&A, x, b, r);
}

void om_gs_flat(int n_self, const int *row, const int *col, const double *diag,
                const double *off, double *x, const double *b, int reverse)
{
    om_mat A = {0};
    A.n_self = n_self; A.row = (int *)row; A.col = (int *)col;
    A.diag = (double *)diag; A.off = (double *)off;
    if (reverse) om_reverse_gs(&A, x, b); else om_forward_gs(&A, x, b);
}

int om_coarsen_flat(int n_self, int n_total, const int *row, const int *col,
                    const double *diag, const double *off, const unsigned char *is_bnd,
                    int group_size, double theta, int *ci)
{
    om_mat A = {0};
    A.n_self = n_self; A.n_total = n_total; A.row = (int *)row; A.col = (int *)col;
    A.diag = (double *)diag; A.off = (double *)off;
    unsigned char *zb = NULL;
    if (!is_bnd) { zb = (unsigned char *)calloc((size_t)(n_total > 0 ? n_total : 1), 1); A.is_bnd = zb; }
    else A.is_bnd = (unsigned char *)is_bnd;
    int nc = om_coarsen(&A, group_size, theta, ci);
    free(zb);
    return nc;
}
