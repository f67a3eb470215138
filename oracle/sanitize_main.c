/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Driver for the sanitizer build of the oracle (oracle/build.sh sanitize, run by
 * tests/test_oracle_sanitize.py): exercises every entry point of amg_oracle.c
 * and fv_oracle.c under AddressSanitizer and UndefinedBehaviorSanitizer.
 *   sanitize_main MATRIX RHS MESHFILE
 * MATRIX / RHS: a MatrixMarket system (the KAT-226 pair); the AMG runs V, W and F
 * cycles with both smoothers and checks the KAT's 40 cycles.
 * MESHFILE: int32 header (dim, n_nodes, n_faces, n_cells, n_ghost, n_groups,
 * n_face_nodes) then doubles nodes[3 n_nodes], int32 fnptr[n_faces + 1],
 * face_nodes, face_cells[2 n_faces], gptr[n_groups + 1]; it is assembled
 * (diffusion with a non-uniform field, convection, phonon direction) and solved.
 * Exit status 0 when every run completed; the sanitizers abort on a fault.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amg_oracle.h"
#include "fv_oracle.h"

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) { fprintf(stderr, "out of memory\n"); exit(3); }
    return p;
}

static int read_exact(FILE *f, void *p, size_t size, size_t n) {
    return fread(p, size, n, f) == n ? 0 : -1;
}

static int solve_system(int n_self, int n_total, const int *row, const int *col, const double *diag,
                        const double *off, const unsigned char *isb, const double *b, int cycle, int smoother,
                        int *cycles_out) {
    om_amg *a = om_amg_create();
    om_params p;
    om_default_params(&p);
    p.cycle_type = cycle;
    p.smoother_type = smoother;
    p.n_max_iterations = 60;
    om_amg_set_params(a, &p);
    if (om_amg_setup(a, n_self, n_total, row, col, diag, off, isb)) { om_amg_destroy(a); return -1; }
    double *x = (double *)calloc((size_t)n_total + 1, sizeof(double));
    double hist[128], r0, rf;
    int cyc = 0;
    om_amg_solve(a, b, x, &r0, &rf, &cyc, hist, 128);
    om_amg_smooth(a, b, x);
    for (int l = 0; l < om_amg_num_levels(a); l++) {
        int ns, nt, nz;
        om_amg_level_info(a, l, &ns, &nt, &nz);
        int *r = (int *)xmalloc(sizeof(int) * ((size_t)nt + 1)), *c = (int *)xmalloc(sizeof(int) * (size_t)nz);
        double *d = (double *)xmalloc(sizeof(double) * (size_t)nt), *o = (double *)xmalloc(sizeof(double) * (size_t)nz);
        int *ci = (int *)xmalloc(sizeof(int) * (size_t)nt);
        om_amg_level_get(a, l, r, c, d, o, ci);
        free(r); free(c); free(d); free(o); free(ci);
    }
    om_amg_cleanup(a);
    om_amg_destroy(a);
    free(x);
    if (cycles_out) *cycles_out = cyc;
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 4) { fprintf(stderr, "usage: %s MATRIX RHS MESHFILE\n", argv[0]); return 2; }
    /* KAT-226 through the MatrixMarket reader */
    int n, nnz, sym;
    if (om_mm_load(argv[1], argv[2], &n, &nnz, &sym, NULL, NULL, NULL, NULL, NULL)) return 4;
    int *row = (int *)xmalloc(sizeof(int) * ((size_t)n + 1)), *col = (int *)xmalloc(sizeof(int) * (size_t)nnz);
    double *diag = (double *)xmalloc(sizeof(double) * (size_t)n), *off = (double *)xmalloc(sizeof(double) * (size_t)nnz);
    double *b = (double *)xmalloc(sizeof(double) * (size_t)n);
    if (om_mm_load(argv[1], argv[2], &n, &nnz, &sym, row, col, diag, off, b)) return 4;
    int cycles = 0;
    for (int cyc = 0; cyc < 3; cyc++)
        for (int sm = 0; sm < 2; sm++) {
            int c = 0;
            if (solve_system(n, n, row, col, diag, off, NULL, b, cyc, sm, &c)) return 5;
            if (cyc == 0 && sm == 0) cycles = c;
        }
    if (cycles != 40) { fprintf(stderr, "KAT-226: %d cycles, expected 40\n", cycles); return 6; }
    double *r = (double *)xmalloc(sizeof(double) * (size_t)n), *x = (double *)calloc((size_t)n, sizeof(double));
    om_residual_flat(n, row, col, diag, off, x, b, r);
    om_gs_flat(n, row, col, diag, off, x, b, 0);
    om_gs_flat(n, row, col, diag, off, x, b, 1);
    int *ci = (int *)xmalloc(sizeof(int) * (size_t)n);
    om_coarsen_flat(n, n, row, col, diag, off, NULL, 2, 0.65, ci);
    free(row); free(col); free(diag); free(off); free(b); free(r); free(x); free(ci);

    /* mesh: metrics, gradient, assembly (diffusion + convection, phonon), solve */
    FILE *f = fopen(argv[3], "rb");
    if (!f) return 7;
    int hdr[7];
    if (read_exact(f, hdr, sizeof(int), 7)) return 7;
    const int dim = hdr[0], nn = hdr[1], nf = hdr[2], nc = hdr[3], ng = hdr[4], ngr = hdr[5], nfn = hdr[6];
    const int nct = nc + ng;
    double *nodes = (double *)xmalloc(sizeof(double) * 3 * (size_t)nn);
    int *fnptr = (int *)xmalloc(sizeof(int) * ((size_t)nf + 1)), *fnodes = (int *)xmalloc(sizeof(int) * (size_t)nfn);
    int *fcells = (int *)xmalloc(sizeof(int) * 2 * (size_t)nf), *gptr = (int *)xmalloc(sizeof(int) * ((size_t)ngr + 1));
    if (read_exact(f, nodes, sizeof(double), 3 * (size_t)nn) || read_exact(f, fnptr, sizeof(int), (size_t)nf + 1) ||
        read_exact(f, fnodes, sizeof(int), (size_t)nfn) || read_exact(f, fcells, sizeof(int), 2 * (size_t)nf) ||
        read_exact(f, gptr, sizeof(int), (size_t)ngr + 1))
        return 7;
    fclose(f);
    double *fA = (double *)xmalloc(sizeof(double) * 3 * (size_t)nf), *fAm = (double *)xmalloc(sizeof(double) * (size_t)nf);
    double *fc = (double *)xmalloc(sizeof(double) * 3 * (size_t)nf), *cc = (double *)xmalloc(sizeof(double) * 3 * (size_t)nct);
    double *vol = (double *)xmalloc(sizeof(double) * (size_t)nct);
    ofv_metrics(dim, nf, nodes, fnptr, fnodes, fcells, nc, ng, ngr, gptr, fA, fAm, fc, cc, vol);
    double *k = (double *)xmalloc(sizeof(double) * (size_t)nct), *src = (double *)xmalloc(sizeof(double) * (size_t)nct);
    double *xf = (double *)xmalloc(sizeof(double) * (size_t)nct), *grad = (double *)xmalloc(sizeof(double) * 3 * (size_t)nct);
    for (int c = 0; c < nct; c++) { k[c] = 1.0 + 0.5 * (c % 7); src[c] = 0.1 * (c % 5); xf[c] = 0.01 * (c % 11); }
    ofv_gradient(dim, nf, nodes, fnptr, fnodes, fcells, nc, ng, ngr, gptr, xf, grad);
    double *flux = (double *)xmalloc(sizeof(double) * (size_t)nf), *cres = (double *)xmalloc(sizeof(double) * (size_t)nc);
    for (int q = 0; q < nf; q++) flux[q] = 0.3 * fA[3 * q] - 0.2 * fA[3 * q + 1] + 0.1 * fA[3 * q + 2];
    for (int c = 0; c < nc; c++) cres[c] = 0.0;
    int *bt = (int *)xmalloc(sizeof(int) * (size_t)ngr);
    double *bv = (double *)xmalloc(sizeof(double) * (size_t)ngr);
    for (int g = 0; g < ngr; g++) { bt[g] = g % 2 == 1 ? OFV_BC_DIRICHLET : OFV_BC_NEUMANN; bv[g] = 1.0 * g; }
    const int mnnz = ofv_thermal_nnz(nf, fcells, nc, ng);
    int *srow = (int *)xmalloc(sizeof(int) * ((size_t)nct + 1)), *scol = (int *)xmalloc(sizeof(int) * (size_t)mnnz);
    double *sdiag = (double *)xmalloc(sizeof(double) * (size_t)nct), *soff = (double *)xmalloc(sizeof(double) * (size_t)mnnz);
    double *sb = (double *)xmalloc(sizeof(double) * (size_t)nct);
    unsigned char *isb = (unsigned char *)xmalloc((size_t)nct);
    for (int central = 0; central < 2; central++) {
        ofv_thermal(dim, nf, nodes, fnptr, fnodes, fcells, nc, ng, ngr, gptr, bt, bv, k, src, xf,
                    central ? flux : NULL, central ? cres : NULL, central, srow, scol, sdiag, soff, sb, isb);
        if (solve_system(nc, nct, srow, scol, sdiag, soff, isb, sb, 0, 0, NULL)) return 8;
    }
    const double vg[3] = {0.6, -0.3, 0.74};
    ofv_phonon(dim, nf, nodes, fnptr, fnodes, fcells, nc, ng, ngr, gptr, vg, 0.7, src, xf, srow, scol, sdiag, soff,
               sb, isb);
    if (solve_system(nc, nct, srow, scol, sdiag, soff, isb, sb, 1, 0, NULL)) return 9;
    free(nodes); free(fnptr); free(fnodes); free(fcells); free(gptr); free(fA); free(fAm); free(fc); free(cc);
    free(vol); free(k); free(src); free(xf); free(grad); free(flux); free(cres); free(bt); free(bv);
    free(srow); free(scol); free(sdiag); free(soff); free(sb); free(isb);
    printf("sanitized oracle runs: ok\n");
    return 0;
}
