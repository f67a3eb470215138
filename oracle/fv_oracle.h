/* CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see fv_oracle.c). */
#ifndef FVM_AMD_FV_ORACLE_H
#define FVM_AMD_FV_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif
enum { OFV_BC_DIRICHLET = 0, OFV_BC_NEUMANN = 1 };
/* mesh arguments: dimension (2 or 3), faces as a node CSR (fnptr, face_nodes),
 * face_cells[2 f], interior and ghost cell counts, face groups (group 0 interior) */
void ofv_metrics(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                 const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                 double *fA, double *fAm, double *fc, double *cc, double *vol);
int ofv_thermal_nnz(int n_faces, const int *face_cells, int n_cells, int n_ghost);
int ofv_gradient(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                 const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                 const double *x, double *grad);
/* conv_flux (n_faces) and cont_res (n_cells) may be NULL: no convection term */
int ofv_thermal(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                const int *bc_type, const double *bc_value, const double *kcell, const double *src,
                const double *x, const double *conv_flux, const double *cont_res, int central,
                int *row, int *col, double *diag, double *off, double *b, unsigned char *is_bnd);
int ofv_kinetic(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
                const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
                const double *c, const double *nue, const double *feq, const double *x, const int *bc_kind,
                const double *wall_u, double eps, int *row, int *col, double *diag, double *off, double *b,
                unsigned char *is_bnd);
int ofv_phonon(int dim, int n_faces, const double *nodes, const int *fnptr, const int *face_nodes,
               const int *face_cells, int n_cells, int n_ghost, int n_groups, const int *gptr,
               const double *vg, double tau, const double *e0, const double *x, int *row, int *col,
               double *diag, double *off, double *b, unsigned char *is_bnd);
#ifdef __cplusplus
}
#endif
#endif
