// Finite-volume assembly of the thermal (diffusion) equation on the GPU.
//
// Reference semantics (src/fvm/src/modules/fvmbase/):
//   MeshMetricsCalculator_impl.h:59-236 (face / cell centroids), :239-305
//     (quad face areas), :385 (|A|), :393-462 (cell volumes)
//   CRConnectivity.cpp:167-270 (cellCells = cellFaces x faceCells, first appearance)
//   CRMatrix.h:117-160 + CRConnectivity.cpp:729-792 (face -> CSR positions)
//   DiffusionDiscretization.h:19-27, :153-230; SourceDiscretization.h:40-60
//   GenericBCS.h:77-157 (Dirichlet / Neumann); CRMatrix.h:899-944, :1064-1085
//     (LinearSystem::initSolve eliminates boundary rows)
//
// The reference runs sequential face loops that scatter into cells. Here every
// per-face quantity is computed once by a face-parallel kernel and every cell
// gathers its faces in ascending face order, which is the order in which the
// face loop touched that cell; with no FP contraction each cell's sums are
// therefore bit-identical to the reference's.
#include "amg.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace fvm {
namespace {

constexpr int BS = 256;
constexpr int MAXG = 64;   // face groups

struct Groups {
    int n;
    int ptr[MAXG + 1];
    int type[MAXG];
    double value[MAXG];
};

__device__ __forceinline__ int group_of(const Groups &G, int f) {
    int g = 0;
    while (g + 1 < G.n && f >= G.ptr[g + 1]) g++;
    return g;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
// uniform in [0,1): 53 random bits
__device__ __forceinline__ double u01(unsigned long long seed, unsigned long long idx, unsigned stream) {
    const unsigned long long h = mix64(mix64(seed ^ (0xD1B54A32D192ED03ULL * (stream + 1))) ^ idx);
    return (double)(h >> 11) * 0x1.0p-53;
}

// ------------------------------------------------------------------ jittered hex mesh
__device__ __forceinline__ int nid(int N, int i, int j, int k) { return i + (N + 1) * (j + (N + 1) * k); }
__device__ __forceinline__ int cid(int N, int i, int j, int k) { return i + N * (j + N * k); }

__global__ void k_hex_nodes(int N, double jitter, unsigned long long seed, double *xyz) {
    const long long nn = (long long)(N + 1) * (N + 1) * (N + 1);
    const double h = 1.0 / N;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < nn;
         id += (long long)gridDim.x * blockDim.x) {
        int idx[3];
        idx[0] = (int)(id % (N + 1));
        idx[1] = (int)((id / (N + 1)) % (N + 1));
        idx[2] = (int)(id / ((long long)(N + 1) * (N + 1)));
        for (int a = 0; a < 3; a++) {
            double v = (double)idx[a] * h;
            if (idx[a] > 0 && idx[a] < N)       // walls stay planar: tangential jitter only
                v += (2.0 * u01(seed, (unsigned long long)id, a) - 1.0) * jitter * h;
            xyz[3 * id + a] = v;
        }
    }
}

// Faces: interior x-, y-, z-faces (lexicographic, x fastest), then boundary
// groups x0, x1, y0, y1, z0, z1 (N^2 faces each). Face nodes are ordered so
// that 0.5 (x2-x0) x (x3-x1) points from c0 to c1 (outward on the boundary).
__global__ void k_hex_faces(int N, int *fnodes, int *fcells) {
    const long long Fx = (long long)(N - 1) * N * N;
    const long long Fint = 3 * Fx;
    const long long nf = Fint + 6LL * N * N;
    const int n = N * N * N;
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < nf;
         f += (long long)gridDim.x * blockDim.x) {
        int c0, c1, q[4];
        if (f < Fint) {
            const int dir = (int)(f / Fx);
            const long long r = f - dir * Fx;
            if (dir == 0) {
                const int i = (int)(r % (N - 1)), j = (int)((r / (N - 1)) % N), k = (int)(r / ((long long)(N - 1) * N));
                c0 = cid(N, i, j, k); c1 = cid(N, i + 1, j, k);
                q[0] = nid(N, i + 1, j, k); q[1] = nid(N, i + 1, j + 1, k);
                q[2] = nid(N, i + 1, j + 1, k + 1); q[3] = nid(N, i + 1, j, k + 1);
            } else if (dir == 1) {
                const int i = (int)(r % N), j = (int)((r / N) % (N - 1)), k = (int)(r / ((long long)N * (N - 1)));
                c0 = cid(N, i, j, k); c1 = cid(N, i, j + 1, k);
                q[0] = nid(N, i, j + 1, k); q[1] = nid(N, i, j + 1, k + 1);
                q[2] = nid(N, i + 1, j + 1, k + 1); q[3] = nid(N, i + 1, j + 1, k);
            } else {
                const int i = (int)(r % N), j = (int)((r / N) % N), k = (int)(r / ((long long)N * N));
                c0 = cid(N, i, j, k); c1 = cid(N, i, j, k + 1);
                q[0] = nid(N, i, j, k + 1); q[1] = nid(N, i + 1, j, k + 1);
                q[2] = nid(N, i + 1, j + 1, k + 1); q[3] = nid(N, i, j + 1, k + 1);
            }
        } else {
            const long long bf = f - Fint;
            const int grp = (int)(bf / ((long long)N * N));
            const int b = (int)(bf % ((long long)N * N));
            const int a0 = b % N, a1 = b / N;
            c1 = n + (int)bf;
            switch (grp) {
            case 0:  // x = 0, outward -x
                c0 = cid(N, 0, a0, a1);
                q[0] = nid(N, 0, a0, a1); q[1] = nid(N, 0, a0, a1 + 1);
                q[2] = nid(N, 0, a0 + 1, a1 + 1); q[3] = nid(N, 0, a0 + 1, a1);
                break;
            case 1:  // x = 1, +x
                c0 = cid(N, N - 1, a0, a1);
                q[0] = nid(N, N, a0, a1); q[1] = nid(N, N, a0 + 1, a1);
                q[2] = nid(N, N, a0 + 1, a1 + 1); q[3] = nid(N, N, a0, a1 + 1);
                break;
            case 2:  // y = 0, -y
                c0 = cid(N, a0, 0, a1);
                q[0] = nid(N, a0, 0, a1); q[1] = nid(N, a0 + 1, 0, a1);
                q[2] = nid(N, a0 + 1, 0, a1 + 1); q[3] = nid(N, a0, 0, a1 + 1);
                break;
            case 3:  // y = 1, +y
                c0 = cid(N, a0, N - 1, a1);
                q[0] = nid(N, a0, N, a1); q[1] = nid(N, a0, N, a1 + 1);
                q[2] = nid(N, a0 + 1, N, a1 + 1); q[3] = nid(N, a0 + 1, N, a1);
                break;
            case 4:  // z = 0, -z
                c0 = cid(N, a0, a1, 0);
                q[0] = nid(N, a0, a1, 0); q[1] = nid(N, a0, a1 + 1, 0);
                q[2] = nid(N, a0 + 1, a1 + 1, 0); q[3] = nid(N, a0 + 1, a1, 0);
                break;
            default:  // z = 1, +z
                c0 = cid(N, a0, a1, N - 1);
                q[0] = nid(N, a0, a1, N); q[1] = nid(N, a0 + 1, a1, N);
                q[2] = nid(N, a0 + 1, a1 + 1, N); q[3] = nid(N, a0, a1 + 1, N);
                break;
            }
        }
        fcells[2 * f] = c0;
        fcells[2 * f + 1] = c1;
        for (int t = 0; t < 4; t++) fnodes[4 * f + t] = q[t];
    }
}

// ------------------------------------------------------------------ connectivity
__global__ void k_cf_count(int nf, const int *fcells, int *cnt) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        atomicAdd(&cnt[fcells[2 * f]], 1);
        atomicAdd(&cnt[fcells[2 * f + 1]], 1);
    }
}
__global__ void k_cf_fill(int nf, const int *fcells, const int *row, int *fill, int *col) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        const int a = fcells[2 * f], b = fcells[2 * f + 1];
        col[row[a] + atomicAdd(&fill[a], 1)] = f;
        col[row[b] + atomicAdd(&fill[b], 1)] = f;
    }
}
__global__ void k_sort_segments(int nseg, const int *ptr, int *v) {
    for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < nseg; I += gridDim.x * blockDim.x) {
        const int b = ptr[I], e = ptr[I + 1];
        for (int a = b + 1; a < e; a++) {
            const int key = v[a];
            int j = a - 1;
            while (j >= b && v[j] > key) { v[j + 1] = v[j]; j--; }
            v[j + 1] = key;
        }
    }
}
// cellCells row lengths (pass 0) / entries and face positions (pass 1)
__global__ void k_cc_build(int nct, const int *cf_row, const int *cf_col, const int *fcells,
                           const int *cc_row, int *cnt, int *cc_col, int *pos01, int *pos10, int pass) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        const int base = pass ? cc_row[c] : 0;
        int k = 0;
        // pass 0 has no storage: count distinct neighbours by rescanning
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            for (int s = 0; s < 2; s++) {
                const int o = fcells[2 * f + s];
                if (o == c) continue;
                int pos = -1;
                if (pass) {
                    for (int e = 0; e < k; e++)
                        if (cc_col[base + e] == o) { pos = e; break; }
                } else {
                    // seen before iff an earlier face of c (or an earlier slot of f) has o
                    for (int q2 = cf_row[c]; q2 <= q && pos < 0; q2++) {
                        const int f2 = cf_col[q2];
                        for (int s2 = 0; s2 < 2; s2++) {
                            if (q2 == q && s2 >= s) break;
                            if (fcells[2 * f2 + s2] == o) { pos = 0; break; }
                        }
                    }
                }
                if (pos < 0) {
                    if (pass) cc_col[base + k] = o;
                    pos = k++;
                }
                if (pass) {
                    if (s == 1) pos01[f] = base + pos;    // c is c0, neighbour c1
                    else pos10[f] = base + pos;          // c is c1, neighbour c0
                }
            }
        }
        if (!pass) cnt[c] = k;
    }
}

// ------------------------------------------------------------------ metrics
// Face area vectors by node count (MeshMetricsCalculator_impl.h:239-305: 2 nodes
// in 2-D, triangles, quads, general polygons), |A|, and face centroids (node
// average, then the correction of faces with more than 3 nodes, :59-118).
__global__ void k_face_metrics(int nf, const double *xyz, const int *fnptr, const int *fnodes, double *fA,
                               double *fAm, double *fc) {
    const double twoThirds = 2. / 3.;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        const int *q = fnodes + fnptr[f];
        const int nn = fnptr[f + 1] - fnptr[f];
        double A[3];
        if (nn == 2) {
            const double *x0 = xyz + 3 * q[0], *x1 = xyz + 3 * q[1];
            A[0] = x1[1] - x0[1];
            A[1] = -(x1[0] - x0[0]);
            A[2] = 0.;
        } else if (nn == 3) {
            const double *x0 = xyz + 3 * q[0], *x1 = xyz + 3 * q[1], *x2 = xyz + 3 * q[2];
            double d10[3], d20[3];
            for (int a = 0; a < 3; a++) { d10[a] = x1[a] - x0[a]; d20[a] = x2[a] - x0[a]; }
            A[0] = 0.5 * (d10[1] * d20[2] - d10[2] * d20[1]);
            A[1] = 0.5 * (d10[2] * d20[0] - d10[0] * d20[2]);
            A[2] = 0.5 * (d10[0] * d20[1] - d10[1] * d20[0]);
        } else if (nn == 4) {
            const double *x0 = xyz + 3 * q[0], *x1 = xyz + 3 * q[1], *x2 = xyz + 3 * q[2], *x3 = xyz + 3 * q[3];
            double d20[3], d31[3];
            for (int a = 0; a < 3; a++) { d20[a] = x2[a] - x0[a]; d31[a] = x3[a] - x1[a]; }
            A[0] = 0.5 * (d20[1] * d31[2] - d20[2] * d31[1]);
            A[1] = 0.5 * (d20[2] * d31[0] - d20[0] * d31[2]);
            A[2] = 0.5 * (d20[0] * d31[1] - d20[1] * d31[0]);
        } else {
            A[0] = A[1] = A[2] = 0.0;
            for (int t = 0; t < nn; t++) {
                const double *u = xyz + 3 * q[t], *v = xyz + 3 * q[(t + 1) % nn];
                double xm[3], dr[3];
                for (int a = 0; a < 3; a++) { xm[a] = 0.5 * (v[a] + u[a]); dr[a] = v[a] - u[a]; }
                A[0] += xm[1] * dr[2];
                A[1] += xm[2] * dr[0];
                A[2] += xm[0] * dr[1];
            }
        }
        const double Am = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        double p[3];
        for (int a = 0; a < 3; a++) p[a] = xyz[3 * q[0] + a];
        for (int t = 1; t < nn; t++)
            for (int a = 0; a < 3; a++) p[a] += xyz[3 * q[t] + a];
        for (int a = 0; a < 3; a++) p[a] /= (double)nn;
        if (nn > 3) {
            double en[3];
            for (int a = 0; a < 3; a++) en[a] = A[a] / Am;
            double denom = 0.0, cfc[3] = {0.0, 0.0, 0.0};
            for (int t = 0; t < nn; t++) {
                const double *u = xyz + 3 * q[t], *v = xyz + 3 * q[(t + 1) % nn];
                double r0[3], r1[3], tri[3], xm[3];
                for (int a = 0; a < 3; a++) { r0[a] = u[a] - p[a]; r1[a] = v[a] - p[a]; }
                tri[0] = 0.5 * (r0[1] * r1[2] - r0[2] * r1[1]);
                tri[1] = 0.5 * (r0[2] * r1[0] - r0[0] * r1[2]);
                tri[2] = 0.5 * (r0[0] * r1[1] - r0[1] * r1[0]);
                const double triP = tri[0] * en[0] + tri[1] * en[1] + tri[2] * en[2];
                for (int a = 0; a < 3; a++) xm[a] = 0.5 * (u[a] + v[a]);
                for (int a = 0; a < 3; a++) cfc[a] += (twoThirds * (xm[a] - p[a])) * triP;
                denom += triP;
            }
            for (int a = 0; a < 3; a++) cfc[a] /= denom;
            for (int a = 0; a < 3; a++) p[a] += cfc[a];
        }
        for (int a = 0; a < 3; a++) {
            fA[3 * f + a] = A[a];
            fc[3 * f + a] = p[a];
        }
        fAm[f] = Am;
    }
}

__global__ void k_iota_stride(int n, int stride, int *v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) v[i] = stride * i;
}

__global__ void k_cell_centroids(int n_cells, int nct, const int *cf_row, const int *cf_col,
                                 const double *fAm, const double *fc, double *cc) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        if (c >= n_cells) {   // ghost: centroid of its boundary face
            const int f = cf_col[cf_row[c]];
            for (int a = 0; a < 3; a++) cc[3 * c + a] = fc[3 * f + a];
            continue;
        }
        double s[3] = {0.0, 0.0, 0.0}, w = 0.0;
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            for (int a = 0; a < 3; a++) s[a] += fc[3 * f + a] * fAm[f];
            w += fAm[f];
        }
        for (int a = 0; a < 3; a++) cc[3 * c + a] = s[a] / w;
    }
}

__global__ void k_cell_volumes(int n_cells, double rdim, const int *cf_row, const int *cf_col, const int *fcells,
                               const double *fA, const double *fc, const double *cc, double *vol) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cells; c += gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            double d[3];
            for (int a = 0; a < 3; a++) d[a] = fc[3 * f + a] - cc[3 * c + a];
            const double t = (d[0] * fA[3 * f] + d[1] * fA[3 * f + 1] + d[2] * fA[3 * f + 2]) / rdim;
            if (fcells[2 * f] == c) v += t;
            else v -= t;
        }
        vol[c] = v;
    }
}

__global__ void k_ghost_volumes(int n_cells, int nct, const int *cf_row, const int *cf_col,
                                const int *fcells, double *vol) {
    for (int g = n_cells + blockIdx.x * blockDim.x + threadIdx.x; g < nct; g += gridDim.x * blockDim.x)
        vol[g] = vol[fcells[2 * cf_col[cf_row[g]]]];
}

// ------------------------------------------------------------------ thermal assembly
__device__ __forceinline__ double harmonic(double x0, double x1) {
    const double sum = x0 + x1;
    if (x0 + x1 != 0.0) return 2.0 * x0 * x1 / sum;
    return sum;
}

// ------------------------------------------------------------------ least-squares gradient
// GradientModel::getLeastSquaresGradientMatrix3D / 2D (GradientModel.h:128-420):
// per cellCells entry of a self cell, K s^ / |ds| with s^ = ds / |ds| the unit step to
// the neighbour and K the inverse of the cell's normal matrix sum s^ s^T (summed in
// CSR order, used when its determinant exceeds 1e-6 in 3-D / 1e-26 in 2-D);
// Green-Gauss face terms 0.5 A / V (sign of the face side) for degenerate cells;
// ghost rows keep -s^ / |ds|. Each row's entries are its faces in ascending order
// (conformal meshes: one face per neighbour), so the sums follow the CSR order.
__device__ __forceinline__ void unit_step(const double *cc, int c0, int c1, bool first, double *u,
                                          double &dsMag) {
    double ds[3];
    for (int a = 0; a < 3; a++) ds[a] = cc[3 * c1 + a] - cc[3 * c0 + a];
    dsMag = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
    for (int a = 0; a < 3; a++) u[a] = first ? ds[a] / dsMag : -ds[a] / dsMag;
}

__global__ void k_lsq_coeffs(int n_cells, int nct, int dim, const int *cf_row, const int *cf_col,
                             const int *fcells, const int *pos01, const int *pos10, const double *cc,
                             const double *fA, const double *vol, double *gc) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        const int qb = cf_row[c], qe = cf_row[c + 1];
        bool degen = false;
        double K[6] = {0, 0, 0, 0, 0, 0};   // xx xy xz yy yz zz
        if (c < n_cells) {
            double Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
            for (int q = qb; q < qe; q++) {
                const int f = cf_col[q];
                double u[3], m;
                unit_step(cc, fcells[2 * f], fcells[2 * f + 1], fcells[2 * f] == c, u, m);
                Ixx += u[0] * u[0];
                Iyy += u[1] * u[1];
                Izz += u[2] * u[2];
                Ixy += u[0] * u[1];
                Ixz += u[0] * u[2];
                Iyz += u[1] * u[2];
            }
            if (dim == 2) {
                const double det = Ixx * Iyy - Ixy * Ixy;
                if (det > 1e-26) { K[0] = Iyy / det; K[1] = -Ixy / det; K[3] = Ixx / det; }
                else degen = true;
            } else {
                const double det = Ixx * (Iyy * Izz - Iyz * Iyz) - Ixy * (Ixy * Izz - Iyz * Ixz)
                                   + Ixz * (Ixy * Iyz - Iyy * Ixz);
                if (det > 1e-6) {
                    K[0] = (Iyy * Izz - Iyz * Iyz) / det;
                    K[1] = -(Ixy * Izz - Iyz * Ixz) / det;
                    K[2] = (Ixy * Iyz - Iyy * Ixz) / det;
                    K[3] = (Ixx * Izz - Ixz * Ixz) / det;
                    K[4] = -(Ixx * Iyz - Ixy * Ixz) / det;
                    K[5] = (Ixx * Iyy - Ixy * Ixy) / det;
                } else {
                    degen = true;
                }
            }
        }
        for (int q = qb; q < qe; q++) {
            const int f = cf_col[q];
            const bool first = fcells[2 * f] == c;
            double u[3], m, w[3];
            unit_step(cc, fcells[2 * f], fcells[2 * f + 1], first, u, m);
            if (c >= n_cells) {
                for (int a = 0; a < 3; a++) w[a] = u[a] / m;
            } else if (degen) {
                const double sgn = first ? 0.5 : -0.5;
                for (int a = 0; a < 3; a++) w[a] = sgn * fA[3 * f + a] / vol[c];
            } else if (dim == 2) {
                w[0] = (K[0] * u[0] + K[1] * u[1]) / m;
                w[1] = (K[1] * u[0] + K[3] * u[1]) / m;
                w[2] = 0.0 / m;
            } else {
                w[0] = (K[0] * u[0] + K[1] * u[1] + K[2] * u[2]) / m;
                w[1] = (K[1] * u[0] + K[3] * u[1] + K[4] * u[2]) / m;
                w[2] = (K[2] * u[0] + K[4] * u[1] + K[5] * u[2]) / m;
            }
            const int p = first ? pos01[f] : pos10[f];
            for (int a = 0; a < 3; a++) gc[3 * (size_t)p + a] = w[a];
        }
    }
}

// GradientMatrix::getGradient (GradientMatrix.h:55-76) of the self cells, entries in
// CSR order; a ghost cell then takes its face's cell's gradient (GradientModel.h:500-515)
__global__ void k_gradient(int n_cells, const int *row, const int *col, const double *gc, const double *x,
                           double *grad) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cells; c += gridDim.x * blockDim.x) {
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        const double xc = x[c];
        for (int nb = row[c]; nb < row[c + 1]; nb++) {
            const double v = x[col[nb]] - xc;
            g0 += gc[3 * (size_t)nb] * v;
            g1 += gc[3 * (size_t)nb + 1] * v;
            g2 += gc[3 * (size_t)nb + 2] * v;
        }
        grad[3 * c] = g0;
        grad[3 * c + 1] = g1;
        grad[3 * c + 2] = g2;
    }
}

__global__ void k_grad_ghost(int n_cells, int nct, const int *cf_row, const int *cf_col, const int *fcells,
                             double *grad) {
    for (int g = n_cells + blockIdx.x * blockDim.x + threadIdx.x; g < nct; g += gridDim.x * blockDim.x) {
        const int c0 = fcells[2 * cf_col[cf_row[g]]];
        for (int a = 0; a < 3; a++) grad[3 * g + a] = grad[3 * c0 + a];
    }
}

// per face: diffusion coefficient and flux with the secondary (non-orthogonal)
// term from the volume-weighted face gradient (DiffusionDiscretization.h:172-212)
__global__ void k_face_diffusion(int nf, const int *fcells, const double *fA, const double *fAm,
                                 const double *cc, const double *vol, const double *kc,
                                 const double *x, const double *grad, double *dcoef, double *dflux) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        const int c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
        const double vol0 = vol[c0], vol1 = vol[c1];
        double ds[3];
        for (int a = 0; a < 3; a++) ds[a] = cc[3 * c1 + a] - cc[3 * c0 + a];
        double kf;
        if (vol0 == 0.) kf = kc[c1];
        else if (vol1 == 0.) kf = kc[c0];
        else kf = harmonic(kc[c0], kc[c1]);
        const double *A = fA + 3 * f;
        const double diffMetric = fAm[f] * fAm[f] / (A[0] * ds[0] + A[1] * ds[1] + A[2] * ds[2]);
        const double diffCoeff = kf * diffMetric;
        double sec[3], gradF[3];
        for (int a = 0; a < 3; a++) sec[a] = kf * (A[a] - ds[a] * diffMetric);
        for (int a = 0; a < 3; a++) gradF[a] = (grad[3 * c0 + a] * vol0 + grad[3 * c1 + a] * vol1) / (vol0 + vol1);
        double dsec = 0.0;
        for (int a = 0; a < 3; a++) dsec += gradF[a] * sec[a];
        dcoef[f] = diffCoeff;
        dflux[f] = diffCoeff * (x[c1] - x[c0]) + dsec;
    }
}

// per cell: diffusion sums in face order, then the convection terms of its faces in
// face order (ConvectionDiscretization.h:160-185 upwind; :131-147 central, whose
// face value is 0.5 (x0 + x0) with the upwind linearization), the continuity
// residual, then the source (interior cells) -- the order of the reference's
// discretization list (ThermalModel_impl.h:238-265)
__global__ void k_cell_assemble(int n_cells, int nct, const int *cf_row, const int *cf_col,
                                const int *fcells, const int *pos01, const int *pos10,
                                const int *row, const double *dcoef, const double *dflux,
                                const double *F, const double *cont_res, int central, const double *x,
                                const double *vol, const double *src, double *diag, double *off,
                                double *b) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        for (int p = row[c]; p < row[c + 1]; p++) off[p] = 0.0;
        double d = 0.0, r = 0.0;
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            const bool first = (fcells[2 * f] == c);
            off[first ? pos01[f] : pos10[f]] += dcoef[f];
            d -= dcoef[f];
            if (first) r += dflux[f];
            else r -= dflux[f];
        }
        if (F) {
            for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
                const int f = cf_col[q];
                const int c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
                const double Ff = F[f];
                double flux;
                if (central) flux = 0.5 * Ff * (x[c0] + x[c0]);
                else flux = Ff > 0.0 ? Ff * x[c0] : Ff * x[c1];
                if (Ff > 0.0) {
                    if (c == c0) d -= Ff;
                    else off[pos10[f]] += Ff;
                } else {
                    if (c == c1) d += Ff;
                    else off[pos01[f]] -= Ff;
                }
                if (c == c0) r -= flux;
                else r += flux;
            }
            if (c < n_cells) d += cont_res ? cont_res[c] : 0.0;
        }
        if (c < n_cells) r += vol[c] * src[c];
        diag[c] = d;
        b[c] = r;
    }
}

// A specified-temperature face with an outflow convecting flux extrapolates
// (ThermalModel_impl.h:306-318).
__device__ __forceinline__ bool bc_dirichlet(const Groups &G, int g, const double *F, int f) {
    return G.type[g] == FVM_BC_DIRICHLET && !(F && F[f] > 0.);
}

// GenericBCS, interior side, faces in order: Dirichlet adds coeff01 (Tb - x_ghost)
// to b[c0] and drops the coefficient (GenericBCS.h:77-100); extrapolation folds
// the ghost's diagonal into the cell (:170-185). Reads the ghost diagonal before
// k_bc_ghost resets it.
__global__ void k_bc_interior(int n_cells, const int *cf_row, const int *cf_col, const int *fcells,
                             const int *pos01, Groups G, const double *F, const double *x, double *diag,
                             double *off, double *b) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cells; c += gridDim.x * blockDim.x) {
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            if (f < G.ptr[1]) continue;
            const int g = group_of(G, f);
            const int c1 = fcells[2 * f + 1];
            if (bc_dirichlet(G, g, F, f)) {
                const double dX = G.value[g] - x[c1];
                b[c] += off[pos01[f]] * dX;
                off[pos01[f]] = 0.0;
            } else if (G.type[g] == FVM_BC_DIRICHLET) {
                const double dFdX1 = -diag[c1];
                const double dx01 = x[c] - x[c1];
                diag[c] += dFdX1;
                b[c] += dFdX1 * dx01;
                off[pos01[f]] = 0.0;
            }
        }
    }
}

// GenericBCS, ghost side (GenericBCS.h:100-110 Dirichlet, :129-150 Neumann,
// :187-192 extrapolation)
__global__ void k_bc_ghost(int n_cells, int nct, const int *cf_row, const int *cf_col, const int *fcells,
                           const int *pos10, const double *fAm, Groups G, const double *F, const double *x,
                           double *diag, double *off, double *b, uint8_t *isb) {
    for (int gc = n_cells + blockIdx.x * blockDim.x + threadIdx.x; gc < nct; gc += gridDim.x * blockDim.x) {
        const int f = cf_col[cf_row[gc]];
        const int g = group_of(G, f);
        isb[gc] = 0;
        if (bc_dirichlet(G, g, F, f)) {
            off[pos10[f]] = 0.0;
            b[gc] = 0.0;
            diag[gc] = -1.0;
        } else if (G.type[g] == FVM_BC_DIRICHLET) {
            diag[gc] = -1.0;
            off[pos10[f]] = 1.0;
            b[gc] = x[fcells[2 * f]] - x[gc];
            isb[gc] = 1;
        } else {
            const double fluxB = -b[gc];
            b[gc] = G.value[g] * fAm[f] - fluxB;
            isb[gc] = 1;
        }
    }
}

// LinearSystem::initSolve: eliminate the marked ghost rows (one neighbour each)
// into their interior cell, ghosts in ascending order (CRMatrix.h:899-944).
// LinearSystem::postSolve -> CRMatrix::solveBoundary (LinearSystem.cpp:250-252,
// CRMatrix.h:433-454): every boundary row nr >= n_self takes
// x[nr] = -(b[nr] + sum off x[col]) / diag[nr], the sum in CSR order.
// flag: some boundary row reads another boundary row (then rows go in ascending order).
__global__ void k_boundary_dep(int n_self, int n_total, const int *row, const int *col, const uint8_t *isb,
                               int *flag) {
    for (int nr = n_self + blockIdx.x * blockDim.x + threadIdx.x; nr < n_total; nr += gridDim.x * blockDim.x) {
        if (!isb[nr]) continue;
        for (int nb = row[nr]; nb < row[nr + 1]; nb++) {
            const int j = col[nb];
            if (j >= n_self && j != nr && isb[j]) { atomicOr(flag, 1); break; }
        }
    }
}

template <bool SEQ>
__global__ void k_solve_boundary(int n_self, int n_total, const int *row, const int *col, const double *diag,
                                 const double *off, const uint8_t *isb, const double *b, double *x) {
    const int start = SEQ ? n_self : n_self + blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = SEQ ? 1 : gridDim.x * blockDim.x;
    for (int nr = start; nr < n_total; nr += stride) {
        if (!isb[nr]) continue;
        double sum = b[nr];
        for (int nb = row[nr]; nb < row[nr + 1]; nb++) sum += off[nb] * x[col[nb]];
        x[nr] = -sum / diag[nr];
    }
}

__global__ void k_eliminate(int n_cells, const int *cf_row, const int *cf_col, const int *fcells,
                            const int *pos01, const int *pos10, const uint8_t *isb, double *diag,
                            double *off, double *b) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cells; c += gridDim.x * blockDim.x) {
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            const int j = fcells[2 * f + 1];
            if (j < n_cells || fcells[2 * f] != c || !isb[j]) continue;
            const double ajj = diag[j];
            const double aij = off[pos01[f]];
            const double ajk = off[pos10[f]];
            diag[c] -= aij * (ajk / ajj);
            b[c] -= aij * (b[j] / ajj);
            off[pos01[f]] = 0.0;
        }
    }
}


// ------------------------------------------------------------------ phonon direction
// Upwind convection of a constant direction vg plus relaxation towards e0
// (PhononConvectionDiscretization.h:150-175, PhononCollisionDiscretization.h:54-68,
// collision linearized first as in PhononModel.h:449-459).
__global__ void k_face_conv(int nf, const double *fA, double v0, double v1, double v2, double *F) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x)
        F[f] = fA[3 * f] * v0 + fA[3 * f + 1] * v1 + fA[3 * f + 2] * v2;
}

// per cell: relaxation, then the upwind terms of its faces in face order (the
// order in which the reference's face loop reaches this cell's entries)
__global__ void k_cell_phonon(int n_cells, int nct, const int *cf_row, const int *cf_col, const int *fcells,
                              const int *pos01, const int *pos10, const int *row, const double *F,
                              const double *vol, const double *e0, const double *x, double tau, double *diag,
                              double *off, double *b) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        for (int p = row[c]; p < row[c + 1]; p++) off[p] = 0.0;
        double d = 0.0, r = 0.0;
        if (c < n_cells) {
            r -= vol[c] * (x[c] - e0[c]) / tau;
            d -= vol[c] / tau;
        }
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            const int c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
            const double Ff = F[f];
            double flux;
            if (Ff > 0.0) {
                flux = Ff * x[c0];
                if (c == c0) d -= Ff;
                else off[pos10[f]] += Ff;
            } else {
                flux = Ff * x[c1];
                if (c == c1) d += Ff;
                else off[pos01[f]] -= Ff;
            }
            if (c == c0) r -= flux;
            else r += flux;
        }
        diag[c] = d;
        b[c] = r;
    }
}

__device__ __forceinline__ bool phonon_inflow(const double *fA, const double *fAm, int f, double v0, double v1,
                                              double v2) {
    const double e0 = fA[3 * f] / fAm[f], e1 = fA[3 * f + 1] / fAm[f], e2 = fA[3 * f + 2] / fAm[f];
    return v0 * e0 + v1 * e1 + v2 * e2 < 0.0;
}

// boundary faces, interior side (GenericPhononBCS.h:67-88 Dirichlet with the
// ghost's own value, :106-124 extrapolation), faces in order per cell; reads the
// ghost diagonal before k_bc_phonon_ghost resets it
__global__ void k_bc_phonon_interior(int n_cells, const int *cf_row, const int *cf_col, const int *fcells,
                                     const int *pos01, int first_bface, const double *fA, const double *fAm,
                                     double v0, double v1, double v2, const double *x, double *diag,
                                     double *off, double *b) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cells; c += gridDim.x * blockDim.x) {
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            if (f < first_bface) continue;
            const int c1 = fcells[2 * f + 1];
            if (phonon_inflow(fA, fAm, f, v0, v1, v2)) {
                const double bval = x[c1];
                const double dX = bval - x[c1];
                b[c] += off[pos01[f]] * dX;
                off[pos01[f]] = 0.0;
            } else {
                const double dFdX1 = -diag[c1];
                const double dx01 = x[c] - x[c1];
                diag[c] += dFdX1;
                b[c] += dFdX1 * dx01;
                off[pos01[f]] = 0.0;
            }
        }
    }
}

__global__ void k_bc_phonon_ghost(int n_cells, int nct, const int *cf_row, const int *cf_col, const int *fcells,
                                  const int *pos10, const double *fA, const double *fAm, double v0, double v1,
                                  double v2, const double *x, double *diag, double *off, double *b,
                                  uint8_t *isb) {
    for (int gc = n_cells + blockIdx.x * blockDim.x + threadIdx.x; gc < nct; gc += gridDim.x * blockDim.x) {
        const int f = cf_col[cf_row[gc]];
        diag[gc] = -1.0;
        if (phonon_inflow(fA, fAm, f, v0, v1, v2)) {
            off[pos10[f]] = 0.0;
            b[gc] = 0.0;
            isb[gc] = 0;
        } else {
            off[pos10[f]] = 1.0;
            b[gc] = x[fcells[2 * f]] - x[gc];
            isb[gc] = 1;
        }
    }
}

// ------------------------------------------------------------------ kinetic (ESBGK) directions
// Collision term first (CollisionTermDiscretization.h:46-62: r -= V nu (f - feq),
// diag -= V nu on self cells), then upwind convection with F = A . c over the
// cell's faces in order (ConvectionDiscretization_Kmodel.h:171-197), as the
// reference's Linearizer runs the two discretizations in that order.
__global__ void k_cell_kinetic(int n_cells, int nct, const int *cf_row, const int *cf_col, const int *fcells,
                               const int *pos01, const int *pos10, const int *row, const double *F,
                               const double *vol, const double *nue, const double *feq, const double *x,
                               double *diag, double *off, double *b) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nct; c += gridDim.x * blockDim.x) {
        for (int p = row[c]; p < row[c + 1]; p++) off[p] = 0.0;
        double d = 0.0, r = 0.0;
        if (c < n_cells) {
            r -= vol[c] * nue[c] * (x[c] - feq[c]);
            d -= vol[c] * nue[c];
        }
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            const int c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
            const double Ff = F[f];
            double flux;
            if (Ff > 0.0) {
                flux = Ff * x[c0];
                if (c == c0) d -= Ff;
                else off[pos10[f]] += Ff;
            } else {
                flux = Ff * x[c1];
                if (c == c1) d += Ff;
                else off[pos01[f]] -= Ff;
            }
            if (c == c0) r -= flux;
            else r += flux;
        }
        diag[c] = d;
        b[c] = r;
    }
}

// incoming[f - first_bface] for every boundary face (KineticModel.h:1622-1661):
// a zero-gradient group (kind 1) takes every direction as incoming; the wall-like
// groups (kind 0) when c.en - uw.en < eps, en = A / |A|, uw the group's velocity.
// grp = [gptr[0..ng], kind[0..ng-1]] as ints; wall_u: 3 per group.
__global__ void k_kinetic_inflow(int first_bface, int nf, int ngroups, const int *grp, const double *wall_u,
                                 const double *fA, const double *fAm, double c0, double c1, double c2, double eps,
                                 uint8_t *incoming) {
    for (int f = first_bface + blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        int g = 1;
        while (g + 1 < ngroups && f >= grp[g + 1]) g++;
        const int kind = grp[ngroups + 1 + g];
        uint8_t in = 1;
        if (kind != 1) {
            const double e0 = fA[3 * f] / fAm[f], e1 = fA[3 * f + 1] / fAm[f], e2 = fA[3 * f + 2] / fAm[f];
            const double cn = c0 * e0 + c1 * e1 + c2 * e2;
            const double *uw = wall_u + 3 * g;
            const double un = uw[0] * e0 + uw[1] * e1 + uw[2] * e2;
            in = (cn - un < eps) ? 1 : 0;
        }
        incoming[f - first_bface] = in;
    }
}

// boundary faces, interior side: Dirichlet with the ghost's own value for incoming
// directions (GenericKineticBCS.h:76-116), extrapolation otherwise (:139-173);
// faces in order per cell; reads the ghost diagonal before k_bc_kinetic_ghost resets it
__global__ void k_bc_kinetic_interior(int n_cells, const int *cf_row, const int *cf_col, const int *fcells,
                                      const int *pos01, int first_bface, const uint8_t *incoming, const double *x,
                                      double *diag, double *off, double *b) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cells; c += gridDim.x * blockDim.x) {
        for (int q = cf_row[c]; q < cf_row[c + 1]; q++) {
            const int f = cf_col[q];
            if (f < first_bface) continue;
            const int c1 = fcells[2 * f + 1];
            if (incoming[f - first_bface]) {
                const double bval = x[c1];
                const double dX = bval - x[c1];
                b[c] += off[pos01[f]] * dX;
                off[pos01[f]] = 0.0;
            } else {
                const double dFdX1 = -diag[c1];
                const double dx01 = x[c] - x[c1];
                diag[c] += dFdX1;
                b[c] += dFdX1 * dx01;
                off[pos01[f]] = 0.0;
            }
        }
    }
}

__global__ void k_bc_kinetic_ghost(int n_cells, int nct, const int *cf_row, const int *cf_col, const int *fcells,
                                   const int *pos10, int first_bface, const uint8_t *incoming, const double *x,
                                   double *diag, double *off, double *b, uint8_t *isb) {
    for (int gc = n_cells + blockIdx.x * blockDim.x + threadIdx.x; gc < nct; gc += gridDim.x * blockDim.x) {
        const int f = cf_col[cf_row[gc]];
        diag[gc] = -1.0;
        if (incoming[f - first_bface]) {
            off[pos10[f]] = 0.0;
            b[gc] = 0.0;
            isb[gc] = 0;
        } else {
            off[pos10[f]] = 1.0;
            b[gc] = x[fcells[2 * f]] - x[gc];
            isb[gc] = 1;
        }
    }
}

// ------------------------------------------------------------------ fields
__global__ void k_lognormal(int N, int n_cells, int block, double sigma, unsigned long long seed,
                            double *k) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_cells; c += gridDim.x * blockDim.x) {
        unsigned long long key = (unsigned long long)c;
        if (block > 1 && N > 0) {
            const int i = c % N, j = (c / N) % N, kk = c / (N * N);
            const int nb = (N + block - 1) / block;
            key = (unsigned long long)(i / block) + (unsigned long long)nb * ((j / block) + (unsigned long long)nb * (kk / block));
        }
        const double u1 = 1.0 - u01(seed, key, 10);   // (0,1]
        const double u2 = u01(seed, key, 11);
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        k[c] = exp(sigma * z);
    }
}

__global__ void k_uniform(int n, double lo, double hi, unsigned long long seed, double *v) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        v[c] = lo + (hi - lo) * u01(seed, (unsigned long long)c, 12);
}

__global__ void k_ghost_copy(int n_cells, int nct, const int *cf_row, const int *cf_col,
                             const int *fcells, double *v) {
    for (int g = n_cells + blockIdx.x * blockDim.x + threadIdx.x; g < nct; g += gridDim.x * blockDim.x)
        v[g] = v[fcells[2 * cf_col[cf_row[g]]]];
}

}  // namespace
}  // namespace fvm

using fvm::DArr;

struct fvm_mesh {
    int device = 0;
    hipStream_t stream = nullptr;
    int hexN = 0;
    int dim = 3;
    int n_cells = 0, n_ghost = 0, n_faces = 0, n_nodes = 0, n_groups = 0;
    std::vector<int> gptr;
    DArr<double> nodes;
    DArr<int> fnptr, fnodes, fcells;   // face nodes as a CSR (fnptr[n_faces + 1])
    long long n_face_nodes = 0;
    // derived
    bool conn = false, metrics = false, lsq = false;
    DArr<int> cf_row, cf_col, cc_row, cc_col, pos01, pos10;
    long long nnz = 0;
    DArr<double> fA, fAm, fc, cc, vol;
    DArr<double> gcoef;                // least-squares gradient coefficients, 3 per cellCells entry
};

struct fvm_system {
    int device = 0;
    int n_self = 0, n_total = 0;
    long long nnz = 0;
    DArr<int> row, col;
    DArr<double> diag, off, b;
    DArr<uint8_t> isb;
};

namespace fvm {

namespace {
struct Dev {
    int prev = -1;
    explicit Dev(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        FVM_HIP(hipSetDevice(d));
    }
    ~Dev() { if (prev >= 0) (void)hipSetDevice(prev); }
};

void mesh_connectivity(fvm_mesh *m) {
    if (m->conn) return;
    hipStream_t s = m->stream;
    const int nct = m->n_cells + m->n_ghost, nf = m->n_faces;
    Scanner sc;
    DArr<int> cnt((size_t)nct + 1), fill((size_t)nct + 1);
    FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    FVM_HIP(hipMemsetAsync(fill.p, 0, fill.bytes(), s));
    k_cf_count<<<grid_for(nf, BS, 4096), BS, 0, s>>>(nf, m->fcells.p, cnt.p);
    m->cf_row.alloc((size_t)nct + 1);
    sc.exclusive(cnt.p, m->cf_row.p, nct + 1, s);
    m->cf_col.alloc(2 * (size_t)std::max(nf, 1));
    k_cf_fill<<<grid_for(nf, BS, 4096), BS, 0, s>>>(nf, m->fcells.p, m->cf_row.p, fill.p, m->cf_col.p);
    k_sort_segments<<<grid_for(nct, BS, 4096), BS, 0, s>>>(nct, m->cf_row.p, m->cf_col.p);
    FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    k_cc_build<<<grid_for(nct, BS, 4096), BS, 0, s>>>(nct, m->cf_row.p, m->cf_col.p, m->fcells.p, nullptr,
                                                      cnt.p, nullptr, nullptr, nullptr, 0);
    m->cc_row.alloc((size_t)nct + 1);
    sc.exclusive(cnt.p, m->cc_row.p, nct + 1, s);
    m->nnz = read_scalar(m->cc_row.p + nct, s);
    m->cc_col.alloc(std::max<long long>(m->nnz, 1));
    m->pos01.alloc(std::max(nf, 1));
    m->pos10.alloc(std::max(nf, 1));
    k_cc_build<<<grid_for(nct, BS, 4096), BS, 0, s>>>(nct, m->cf_row.p, m->cf_col.p, m->fcells.p,
                                                      m->cc_row.p, nullptr, m->cc_col.p, m->pos01.p,
                                                      m->pos10.p, 1);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(s));
    m->conn = true;
}

void mesh_metrics(fvm_mesh *m) {
    mesh_connectivity(m);
    if (m->metrics) return;
    hipStream_t s = m->stream;
    const int nct = m->n_cells + m->n_ghost, nf = m->n_faces;
    m->fA.alloc(3 * (size_t)nf); m->fAm.alloc(nf); m->fc.alloc(3 * (size_t)nf);
    m->cc.alloc(3 * (size_t)nct); m->vol.alloc(nct);
    k_face_metrics<<<grid_for(nf, BS), BS, 0, s>>>(nf, m->nodes.p, m->fnptr.p, m->fnodes.p, m->fA.p, m->fAm.p,
                                                   m->fc.p);
    k_cell_centroids<<<grid_for(nct, BS), BS, 0, s>>>(m->n_cells, nct, m->cf_row.p, m->cf_col.p, m->fAm.p,
                                                      m->fc.p, m->cc.p);
    k_cell_volumes<<<grid_for(m->n_cells, BS), BS, 0, s>>>(m->n_cells, (double)m->dim, m->cf_row.p, m->cf_col.p,
                                                           m->fcells.p, m->fA.p, m->fc.p, m->cc.p, m->vol.p);
    if (m->n_ghost)
        k_ghost_volumes<<<grid_for(m->n_ghost, BS), BS, 0, s>>>(m->n_cells, nct, m->cf_row.p, m->cf_col.p,
                                                                m->fcells.p, m->vol.p);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(s));
    m->metrics = true;
}

// least-squares gradient coefficients of the mesh (computed once)
void mesh_gradient_coeffs(fvm_mesh *m) {
    mesh_metrics(m);
    if (m->lsq) return;
    FVM_CHECK(m->nnz == 2LL * m->n_faces, FVM_E_INVALID,
              "the gradient needs a conformal mesh: two cells share at most one face");
    hipStream_t s = m->stream;
    const int nct = m->n_cells + m->n_ghost;
    m->gcoef.alloc(3 * (size_t)std::max<long long>(m->nnz, 1));
    k_lsq_coeffs<<<grid_for(nct, BS), BS, 0, s>>>(m->n_cells, nct, m->dim, m->cf_row.p, m->cf_col.p, m->fcells.p,
                                                  m->pos01.p, m->pos10.p, m->cc.p, m->fA.p, m->vol.p,
                                                  m->gcoef.p);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(s));
    m->lsq = true;
}

// GradientModel::compute of a cell field into grad (3 per cell, ghosts included)
void mesh_gradient(fvm_mesh *m, const double *x, double *grad, hipStream_t s) {
    const int nct = m->n_cells + m->n_ghost;
    if (m->n_cells)
        k_gradient<<<grid_for(m->n_cells, BS), BS, 0, s>>>(m->n_cells, m->cc_row.p, m->cc_col.p, m->gcoef.p, x,
                                                           grad);
    if (m->n_ghost)
        k_grad_ghost<<<grid_for(m->n_ghost, BS), BS, 0, s>>>(m->n_cells, nct, m->cf_row.p, m->cf_col.p,
                                                             m->fcells.p, grad);
}

Groups make_groups(const fvm_mesh *m, const int32_t *type, const double *value) {
    Groups G{};
    G.n = m->n_groups;
    for (int g = 0; g <= m->n_groups; g++) G.ptr[g] = m->gptr[g];
    for (int g = 0; g < m->n_groups; g++) {
        G.type[g] = g == 0 ? -1 : type[g];
        G.value[g] = g == 0 ? 0.0 : value[g];
        if (g > 0)
            FVM_CHECK(type[g] == FVM_BC_DIRICHLET || type[g] == FVM_BC_NEUMANN, FVM_E_INVALID, "bad bc type");
    }
    return G;
}

}  // namespace
}  // namespace fvm

namespace fvm {
namespace {
__global__ void k_fill_u8(int n, uint8_t v, uint8_t *p) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// a system with the mesh's cellCells pattern
void system_alloc(fvm_mesh *m, fvm_system *S) {
    const int nct = m->n_cells + m->n_ghost;
    S->device = m->device;
    S->n_self = m->n_cells;
    S->n_total = nct;
    S->nnz = m->nnz;
    S->row.alloc((size_t)nct + 1);
    S->col.alloc(std::max<long long>(m->nnz, 1));
    S->off.alloc(std::max<long long>(m->nnz, 1));
    S->diag.alloc(std::max(nct, 1)); S->b.alloc(std::max(nct, 1)); S->isb.alloc(std::max(nct, 1));
    FVM_HIP(hipMemcpyAsync(S->row.p, m->cc_row.p, S->row.bytes(), hipMemcpyDeviceToDevice, m->stream));
    if (m->nnz)
        FVM_HIP(hipMemcpyAsync(S->col.p, m->cc_col.p, sizeof(int) * m->nnz, hipMemcpyDeviceToDevice, m->stream));
}

// ThermalModel::linearize + LinearSystem::initSolve into S (pattern already set):
// gradient of x, diffusion with secondary flux, convection (d_flux may be null),
// source, boundary conditions, elimination. Scratch: dcoef, dflux (n_faces),
// grad (3 (n_cells + n_ghost)). Queued on the mesh's stream; no allocation and
// no host synchronization.
void assemble_thermal(fvm_mesh *m, fvm_system *S, const double *d_k, const double *d_source, const double *d_x,
                      const double *d_flux, const double *d_cont, int central, const Groups &G, double *dcoef,
                      double *dflux, double *grad) {
    hipStream_t s = m->stream;
    const int nc = m->n_cells, nct = m->n_cells + m->n_ghost, nf = m->n_faces;
    mesh_gradient(m, d_x, grad, s);
    if (nf)
        k_face_diffusion<<<grid_for(nf, 256), 256, 0, s>>>(nf, m->fcells.p, m->fA.p, m->fAm.p, m->cc.p, m->vol.p,
                                                           d_k, d_x, grad, dcoef, dflux);
    if (nct)
        k_cell_assemble<<<grid_for(nct, 256), 256, 0, s>>>(nc, nct, m->cf_row.p, m->cf_col.p, m->fcells.p,
                                                           m->pos01.p, m->pos10.p, m->cc_row.p, dcoef, dflux,
                                                           d_flux, d_cont, central, d_x, m->vol.p, d_source,
                                                           S->diag.p, S->off.p, S->b.p);
    if (nc)
        k_bc_interior<<<grid_for(nc, 256), 256, 0, s>>>(nc, m->cf_row.p, m->cf_col.p, m->fcells.p, m->pos01.p, G,
                                                        d_flux, d_x, S->diag.p, S->off.p, S->b.p);
    if (nct) k_fill_u8<<<grid_for(nct, 256), 256, 0, s>>>(nct, 0, S->isb.p);
    if (m->n_ghost)
        k_bc_ghost<<<grid_for(m->n_ghost, 256), 256, 0, s>>>(nc, nct, m->cf_row.p, m->cf_col.p, m->fcells.p,
                                                             m->pos10.p, m->fAm.p, G, d_flux, d_x, S->diag.p,
                                                             S->off.p, S->b.p, S->isb.p);
    if (nc)
        k_eliminate<<<grid_for(nc, 256), 256, 0, s>>>(nc, m->cf_row.p, m->cf_col.p, m->fcells.p, m->pos01.p,
                                                      m->pos10.p, S->isb.p, S->diag.p, S->off.p, S->b.p);
    FVM_HIP(hipGetLastError());
}
}  // namespace

// Scratch of the allocation-free re-assembly (fvm_convdiff_reassemble_noalloc):
// per-face diffusion coefficients and fluxes (2 n_faces) and the cell gradients
// (3 (n_cells + n_ghost)). One copy per device, held until the process ends.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)1 << 22;   // 32 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// device address of this device's scratch copy (looked up once per device, with
// that device current); the symbol's host-side name is not a device address
double *scratch_on(int device) {
    FVM_CHECK(device >= 0 && device < MAX_DEVICES, FVM_E_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        FVM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        FVM_CHECK(ptr != nullptr, FVM_E_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[device] = (double *)ptr;
    }
    return g_scratch_ptr[device];
}
}  // namespace

int mesh_device(const fvm_mesh *m) { return m->device; }

void mesh_device_metrics(fvm_mesh *m, const int **fcells, const double **fA, const double **fAm) {
    Dev d(m->device);
    mesh_metrics(m);
    *fcells = m->fcells.p;
    *fA = m->fA.p;
    *fAm = m->fAm.p;
}

hipStream_t mesh_stream(fvm_mesh *m) { return m->stream; }

// A system with the mesh's cell-cell pattern (row and column arrays copied on the
// mesh's stream), values unset; the caller owns it (fvm_system_destroy).
fvm_system *system_for_mesh(fvm_mesh *m) {
    Dev d(m->device);
    mesh_metrics(m);
    const int nct = m->n_cells + m->n_ghost;
    fvm_system *S = new fvm_system();
    try {
        S->device = m->device;
        S->n_self = m->n_cells;
        S->n_total = nct;
        S->nnz = m->nnz;
        S->row.alloc((size_t)nct + 1);
        S->col.alloc(std::max<long long>(m->nnz, 1));
        S->off.alloc(std::max<long long>(m->nnz, 1));
        S->diag.alloc(std::max(nct, 1)); S->b.alloc(std::max(nct, 1)); S->isb.alloc(std::max(nct, 1));
        FVM_HIP(hipMemcpyAsync(S->row.p, m->cc_row.p, S->row.bytes(), hipMemcpyDeviceToDevice, m->stream));
        if (m->nnz)
            FVM_HIP(hipMemcpyAsync(S->col.p, m->cc_col.p, sizeof(int) * m->nnz, hipMemcpyDeviceToDevice, m->stream));
    } catch (...) {
        delete S;
        throw;
    }
    return S;
}

// One kinetic velocity's system into S (from system_for_mesh on this mesh), all on
// the mesh's stream with no synchronisation: grp = [gptr (n_groups + 1), kind
// (n_groups)] and wall (3 per group) on the device; F (n_faces) and inc (n_faces)
// are scratch that the next call on the same stream may reuse.
void kinetic_assemble_into(fvm_mesh *m, const double *c, const double *d_nue, const double *d_feq,
                           const double *d_x, const int *grp, const double *wall, double eps, fvm_system *S,
                           double *F, uint8_t *inc) {
    hipStream_t s = m->stream;
    const int nc = m->n_cells, nct = m->n_cells + m->n_ghost, nf = m->n_faces;
    k_face_conv<<<grid_for(nf, 256), 256, 0, s>>>(nf, m->fA.p, c[0], c[1], c[2], F);
    k_cell_kinetic<<<grid_for(nct, 256), 256, 0, s>>>(nc, nct, m->cf_row.p, m->cf_col.p, m->fcells.p, m->pos01.p,
                                                      m->pos10.p, m->cc_row.p, F, m->vol.p, d_nue, d_feq, d_x,
                                                      S->diag.p, S->off.p, S->b.p);
    FVM_HIP(hipMemsetAsync(S->isb.p, 0, nct, s));
    const int first_bface = m->gptr.size() > 1 ? m->gptr[1] : nf;
    if (nf > first_bface) {
        k_kinetic_inflow<<<grid_for(nf - first_bface, 256), 256, 0, s>>>(first_bface, nf, m->n_groups, grp, wall,
                                                                         m->fA.p, m->fAm.p, c[0], c[1], c[2], eps,
                                                                         inc);
        k_bc_kinetic_interior<<<grid_for(nc, 256), 256, 0, s>>>(nc, m->cf_row.p, m->cf_col.p, m->fcells.p,
                                                                m->pos01.p, first_bface, inc, d_x, S->diag.p,
                                                                S->off.p, S->b.p);
        if (m->n_ghost)
            k_bc_kinetic_ghost<<<grid_for(m->n_ghost, 256), 256, 0, s>>>(nc, nct, m->cf_row.p, m->cf_col.p,
                                                                         m->fcells.p, m->pos10.p, first_bface, inc,
                                                                         d_x, S->diag.p, S->off.p, S->b.p, S->isb.p);
        k_eliminate<<<grid_for(nc, 256), 256, 0, s>>>(nc, m->cf_row.p, m->cf_col.p, m->fcells.p, m->pos01.p,
                                                      m->pos10.p, S->isb.p, S->diag.p, S->off.p, S->b.p);
    }
    FVM_HIP(hipGetLastError());
}

// postSolve of a system assembled on a mesh, on stream s without synchronising:
// a mesh's ghost row holds only its face's interior cell, so no boundary row reads
// another one and the rows are independent (CRMatrix.h:443-453 in any order)
void system_post_solve_async(const fvm_system *S, double *d_delta, hipStream_t s) {
    const int ng = S->n_total - S->n_self;
    if (ng <= 0) return;
    k_solve_boundary<false><<<grid_for(ng, 256), 256, 0, s>>>(S->n_self, S->n_total, S->row.p, S->col.p, S->diag.p,
                                                              S->off.p, S->isb.p, S->b.p, d_delta);
    FVM_HIP(hipGetLastError());
}
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

extern "C" {

int fvm_mesh_create_hex(fvm_mesh **out, int device, int n, double jitter, uint64_t seed) {
    FVM_API_BEGIN
    FVM_CHECK(out && n >= 2 && n <= 1024 && jitter >= 0.0 && jitter < 0.5, FVM_E_INVALID,
              "need 2 <= n <= 1024 and 0 <= jitter < 0.5");
    fvm::Dev d(device);
    fvm_mesh *m = new fvm_mesh();
    try {
        m->device = device;
        FVM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamDefault));
        m->hexN = n;
        m->n_cells = n * n * n;
        m->n_ghost = 6 * n * n;
        const long long Fint = 3LL * (n - 1) * n * n;
        m->n_faces = (int)(Fint + m->n_ghost);
        m->n_nodes = (n + 1) * (n + 1) * (n + 1);
        m->n_groups = 7;
        m->gptr.resize(8);
        m->gptr[0] = 0;
        m->gptr[1] = (int)Fint;
        for (int g = 0; g < 6; g++) m->gptr[g + 2] = (int)(Fint + (long long)(g + 1) * n * n);
        m->nodes.alloc(3 * (size_t)m->n_nodes);
        m->n_face_nodes = 4LL * m->n_faces;
        m->fnptr.alloc((size_t)m->n_faces + 1);
        m->fnodes.alloc(4 * (size_t)m->n_faces);
        m->fcells.alloc(2 * (size_t)m->n_faces);
        fvm::k_iota_stride<<<fvm::grid_for(m->n_faces + 1, 256), 256, 0, m->stream>>>(m->n_faces, 4, m->fnptr.p);
        fvm::k_hex_nodes<<<fvm::grid_for(m->n_nodes, 256), 256, 0, m->stream>>>(n, jitter, seed, m->nodes.p);
        fvm::k_hex_faces<<<fvm::grid_for(m->n_faces, 256), 256, 0, m->stream>>>(n, m->fnodes.p, m->fcells.p);
        FVM_HIP(hipGetLastError());
        FVM_HIP(hipStreamSynchronize(m->stream));
    } catch (...) {
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
        throw;
    }
    *out = m;
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_create_general(fvm_mesh **out, int device, int dim, int n_nodes, const double *nodes, int n_faces,
                            const int32_t *face_nptr, const int32_t *face_nodes, const int32_t *face_cells,
                            int n_cells, int n_ghost, int n_groups, const int32_t *group_ptr) {
    FVM_API_BEGIN
    FVM_CHECK(out && nodes && face_nptr && face_nodes && face_cells && group_ptr, FVM_E_INVALID, "null argument");
    FVM_CHECK(dim == 2 || dim == 3, FVM_E_INVALID, "dim must be 2 or 3");
    FVM_CHECK(n_nodes >= 0 && n_faces >= 0 && n_cells >= 0 && n_ghost >= 0, FVM_E_INVALID, "negative count");
    FVM_CHECK(n_groups >= 1 && n_groups <= 64 && group_ptr[0] == 0 && group_ptr[n_groups] == n_faces,
              FVM_E_INVALID, "bad face groups");
    for (int g = 0; g < n_groups; g++)
        FVM_CHECK(group_ptr[g] <= group_ptr[g + 1], FVM_E_INVALID, "face groups must be ascending");
    FVM_CHECK(group_ptr[1] == n_faces - n_ghost, FVM_E_INVALID, "one ghost cell per boundary face");
    FVM_CHECK(face_nptr[0] == 0, FVM_E_INVALID, "face_nptr[0] must be 0");
    const int nct = n_cells + n_ghost;
    for (int f = 0; f < n_faces; f++) {
        const int c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
        FVM_CHECK(c0 >= 0 && c0 < n_cells && c1 >= 0 && c1 < nct && c0 != c1, FVM_E_INVALID, "bad face cells");
        if (f >= group_ptr[1]) FVM_CHECK(c1 == n_cells + (f - group_ptr[1]), FVM_E_INVALID, "ghost numbering");
        const int nn = face_nptr[f + 1] - face_nptr[f];
        FVM_CHECK(dim == 2 ? nn == 2 : (nn >= 3 && nn <= 64), FVM_E_INVALID,
                  "faces need 2 nodes in 2-D and 3..64 nodes in 3-D");
        for (int t = face_nptr[f]; t < face_nptr[f + 1]; t++)
            FVM_CHECK(face_nodes[t] >= 0 && face_nodes[t] < n_nodes, FVM_E_INVALID, "bad face node");
    }
    fvm::Dev d(device);
    fvm_mesh *m = new fvm_mesh();
    try {
        m->device = device;
        FVM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamDefault));
        m->dim = dim;
        m->n_cells = n_cells; m->n_ghost = n_ghost; m->n_faces = n_faces; m->n_nodes = n_nodes;
        m->n_groups = n_groups;
        m->gptr.assign(group_ptr, group_ptr + n_groups + 1);
        m->n_face_nodes = face_nptr[n_faces];
        m->nodes.alloc(3 * (size_t)std::max(n_nodes, 1));
        m->fnptr.alloc((size_t)n_faces + 1);
        m->fnodes.alloc((size_t)std::max<long long>(m->n_face_nodes, 1));
        m->fcells.alloc(2 * (size_t)std::max(n_faces, 1));
        FVM_HIP(hipMemcpy(m->nodes.p, nodes, sizeof(double) * 3 * (size_t)n_nodes, hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(m->fnptr.p, face_nptr, sizeof(int) * ((size_t)n_faces + 1), hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(m->fnodes.p, face_nodes, sizeof(int) * (size_t)m->n_face_nodes, hipMemcpyHostToDevice));
        FVM_HIP(hipMemcpy(m->fcells.p, face_cells, sizeof(int) * 2 * (size_t)n_faces, hipMemcpyHostToDevice));
    } catch (...) {
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
        throw;
    }
    *out = m;
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_create(fvm_mesh **out, int device, int n_nodes, const double *nodes, int n_faces,
                    const int32_t *face_nodes, const int32_t *face_cells, int n_cells, int n_ghost,
                    int n_groups, const int32_t *group_ptr) {
    if (n_faces < 0) return FVM_E_INVALID;
    std::vector<int32_t> nptr((size_t)n_faces + 1);
    for (int f = 0; f <= n_faces; f++) nptr[f] = 4 * f;
    return fvm_mesh_create_general(out, device, 3, n_nodes, nodes, n_faces, nptr.data(), face_nodes, face_cells,
                                   n_cells, n_ghost, n_groups, group_ptr);
}

int fvm_mesh_destroy(fvm_mesh *m) {
    FVM_API_BEGIN
    if (!m) return FVM_OK;
    {
        fvm::Dev d(m->device);
        (void)hipStreamSynchronize(m->stream);
        m->nodes.release(); m->fnptr.release(); m->fnodes.release(); m->fcells.release(); m->gcoef.release();
        m->cf_row.release(); m->cf_col.release(); m->cc_row.release(); m->cc_col.release();
        m->pos01.release(); m->pos10.release();
        m->fA.release(); m->fAm.release(); m->fc.release(); m->cc.release(); m->vol.release();
        (void)hipStreamDestroy(m->stream);
    }
    delete m;
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_info(const fvm_mesh *m, int *n_cells, int *n_ghost, int *n_faces, int *n_nodes, int *n_groups) {
    if (!m) return FVM_E_INVALID;
    if (n_cells) *n_cells = m->n_cells;
    if (n_ghost) *n_ghost = m->n_ghost;
    if (n_faces) *n_faces = m->n_faces;
    if (n_nodes) *n_nodes = m->n_nodes;
    if (n_groups) *n_groups = m->n_groups;
    return FVM_OK;
}

int fvm_mesh_topology(const fvm_mesh *m, int *dim, long long *n_face_nodes, int32_t *face_nptr) {
    FVM_API_BEGIN
    FVM_CHECK(m, FVM_E_INVALID, "null mesh");
    if (dim) *dim = m->dim;
    if (n_face_nodes) *n_face_nodes = m->n_face_nodes;
    if (face_nptr) {
        fvm::Dev d(m->device);
        FVM_HIP(hipMemcpy(face_nptr, m->fnptr.p, sizeof(int) * ((size_t)m->n_faces + 1), hipMemcpyDeviceToHost));
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_gradient(fvm_mesh *m, const double *d_x, double *d_grad) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_x && d_grad, FVM_E_INVALID, "null argument");
    fvm::Dev d(m->device);
    fvm::mesh_gradient_coeffs(m);
    fvm::mesh_gradient(m, d_x, d_grad, m->stream);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_get(const fvm_mesh *m, double *nodes, int32_t *face_nodes, int32_t *face_cells, int32_t *group_ptr) {
    FVM_API_BEGIN
    FVM_CHECK(m, FVM_E_INVALID, "null mesh");
    fvm::Dev d(m->device);
    if (nodes) FVM_HIP(hipMemcpy(nodes, m->nodes.p, m->nodes.bytes(), hipMemcpyDeviceToHost));
    if (face_nodes && m->n_face_nodes)
        FVM_HIP(hipMemcpy(face_nodes, m->fnodes.p, sizeof(int) * (size_t)m->n_face_nodes, hipMemcpyDeviceToHost));
    if (face_cells && m->n_faces)
        FVM_HIP(hipMemcpy(face_cells, m->fcells.p, sizeof(int) * 2 * (size_t)m->n_faces, hipMemcpyDeviceToHost));
    if (group_ptr) for (int g = 0; g <= m->n_groups; g++) group_ptr[g] = m->gptr[g];
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_metrics(fvm_mesh *m, double *fA, double *fAm, double *fc, double *cc, double *vol) {
    FVM_API_BEGIN
    FVM_CHECK(m, FVM_E_INVALID, "null mesh");
    fvm::Dev d(m->device);
    fvm::mesh_metrics(m);
    if (fA) FVM_HIP(hipMemcpy(fA, m->fA.p, m->fA.bytes(), hipMemcpyDeviceToHost));
    if (fAm) FVM_HIP(hipMemcpy(fAm, m->fAm.p, m->fAm.bytes(), hipMemcpyDeviceToHost));
    if (fc) FVM_HIP(hipMemcpy(fc, m->fc.p, m->fc.bytes(), hipMemcpyDeviceToHost));
    if (cc) FVM_HIP(hipMemcpy(cc, m->cc.p, m->cc.bytes(), hipMemcpyDeviceToHost));
    if (vol) FVM_HIP(hipMemcpy(vol, m->vol.p, m->vol.bytes(), hipMemcpyDeviceToHost));
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_lognormal_field(fvm_mesh *m, double sigma, int block, uint64_t seed, double *d_out) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_out && block >= 1, FVM_E_INVALID, "bad argument");
    FVM_CHECK(block == 1 || m->hexN > 0, FVM_E_INVALID, "block fields need a hex mesh");
    fvm::Dev d(m->device);
    fvm::mesh_connectivity(m);
    const int nct = m->n_cells + m->n_ghost;
    fvm::k_lognormal<<<fvm::grid_for(m->n_cells, 256), 256, 0, m->stream>>>(m->hexN, m->n_cells, block, sigma,
                                                                            seed, d_out);
    if (m->n_ghost)
        fvm::k_ghost_copy<<<fvm::grid_for(m->n_ghost, 256), 256, 0, m->stream>>>(m->n_cells, nct, m->cf_row.p,
                                                                                 m->cf_col.p, m->fcells.p, d_out);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_mesh_uniform_field(fvm_mesh *m, double lo, double hi, uint64_t seed, double *d_out) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_out, FVM_E_INVALID, "bad argument");
    fvm::Dev d(m->device);
    const int nct = m->n_cells + m->n_ghost;
    fvm::k_uniform<<<fvm::grid_for(nct, 256), 256, 0, m->stream>>>(nct, lo, hi, seed, d_out);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_convdiff_assemble(fvm_mesh *m, const double *d_k, const double *d_source, const double *d_x,
                          const double *d_flux, const double *d_cont_res, int central, const int32_t *bc_type,
                          const double *bc_value, fvm_system **out) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_k && d_source && d_x && bc_type && bc_value && out, FVM_E_INVALID, "null argument");
    FVM_CHECK(central == 0 || central == 1, FVM_E_INVALID, "central must be 0 or 1");
    fvm::Dev d(m->device);
    fvm::mesh_gradient_coeffs(m);
    const fvm::Groups G = fvm::make_groups(m, bc_type, bc_value);
    const int nct = m->n_cells + m->n_ghost, nf = m->n_faces;
    fvm_system *S = new fvm_system();
    try {
        fvm::system_alloc(m, S);
        DArr<double> dcoef(std::max(nf, 1)), dflux(std::max(nf, 1)), grad(3 * (size_t)std::max(nct, 1));
        fvm::assemble_thermal(m, S, d_k, d_source, d_x, d_flux, d_cont_res, central, G, dcoef.p, dflux.p, grad.p);
        FVM_HIP(hipStreamSynchronize(m->stream));
    } catch (...) {
        delete S;
        throw;
    }
    *out = S;
    return FVM_OK;
    FVM_API_END
}

int fvm_thermal_assemble(fvm_mesh *m, const double *d_k, const double *d_source, const double *d_x,
                         const int32_t *bc_type, const double *bc_value, fvm_system **out) {
    return fvm_convdiff_assemble(m, d_k, d_source, d_x, nullptr, nullptr, 0, bc_type, bc_value, out);
}

int fvm_convdiff_reassemble_noalloc(fvm_mesh *m, const double *d_k, const double *d_source, const double *d_x,
                                    const double *d_flux, const double *d_cont_res, int central,
                                    const int32_t *bc_type, const double *bc_value, fvm_system *S) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_k && d_source && d_x && bc_type && bc_value && S, FVM_E_INVALID, "null argument");
    FVM_CHECK(central == 0 || central == 1, FVM_E_INVALID, "central must be 0 or 1");
    FVM_CHECK(m->lsq, FVM_E_INVALID,
              "the mesh has no metrics or gradient coefficients yet: assemble it once with fvm_convdiff_assemble");
    const int nct = m->n_cells + m->n_ghost, nf = m->n_faces;
    FVM_CHECK(S->device == m->device && S->n_self == m->n_cells && S->n_total == nct && S->nnz == m->nnz,
              FVM_E_INVALID, "the system was not assembled from this mesh");
    const size_t need = 2 * (size_t)nf + 3 * (size_t)nct;
    FVM_CHECK(need <= fvm::NOALLOC_SCRATCH_DOUBLES, FVM_E_INVALID,
              "mesh too large for the allocation-free re-assembly: 2 n_faces + 3 (n_cells + n_ghost) = " +
                  std::to_string(need) + " doubles, at most " + std::to_string(fvm::NOALLOC_SCRATCH_DOUBLES));
    fvm::Dev d(m->device);
    double *scr = fvm::scratch_on(m->device);
    const fvm::Groups G = fvm::make_groups(m, bc_type, bc_value);
    fvm::assemble_thermal(m, S, d_k, d_source, d_x, d_flux, d_cont_res, central, G, scr, scr + nf,
                          scr + 2 * (size_t)nf);
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_phonon_assemble(fvm_mesh *m, const double *vg, double tau, const double *d_e0, const double *d_x,
                        fvm_system **out) {
    FVM_API_BEGIN
    FVM_CHECK(m && vg && d_e0 && d_x && out, FVM_E_INVALID, "null argument");
    FVM_CHECK(tau != 0.0 && std::isfinite(tau), FVM_E_INVALID, "tau must be finite and non-zero");
    fvm::Dev d(m->device);
    fvm::mesh_metrics(m);
    hipStream_t s = m->stream;
    const int nc = m->n_cells, nct = m->n_cells + m->n_ghost, nf = m->n_faces;
    fvm_system *S = new fvm_system();
    try {
        S->device = m->device;
        S->n_self = nc;
        S->n_total = nct;
        S->nnz = m->nnz;
        S->row.alloc((size_t)nct + 1);
        S->col.alloc(std::max<long long>(m->nnz, 1));
        S->off.alloc(std::max<long long>(m->nnz, 1));
        S->diag.alloc(nct); S->b.alloc(nct); S->isb.alloc(nct);
        FVM_HIP(hipMemcpyAsync(S->row.p, m->cc_row.p, S->row.bytes(), hipMemcpyDeviceToDevice, s));
        FVM_HIP(hipMemcpyAsync(S->col.p, m->cc_col.p, sizeof(int) * m->nnz, hipMemcpyDeviceToDevice, s));
        DArr<double> F(std::max(nf, 1));
        fvm::k_face_conv<<<fvm::grid_for(nf, 256), 256, 0, s>>>(nf, m->fA.p, vg[0], vg[1], vg[2], F.p);
        fvm::k_cell_phonon<<<fvm::grid_for(nct, 256), 256, 0, s>>>(
            nc, nct, m->cf_row.p, m->cf_col.p, m->fcells.p, m->pos01.p, m->pos10.p, m->cc_row.p, F.p, m->vol.p,
            d_e0, d_x, tau, S->diag.p, S->off.p, S->b.p);
        const int first_bface = m->gptr.size() > 1 ? m->gptr[1] : nf;
        fvm::k_bc_phonon_interior<<<fvm::grid_for(nc, 256), 256, 0, s>>>(
            nc, m->cf_row.p, m->cf_col.p, m->fcells.p, m->pos01.p, first_bface, m->fA.p, m->fAm.p, vg[0], vg[1],
            vg[2], d_x, S->diag.p, S->off.p, S->b.p);
        FVM_HIP(hipMemsetAsync(S->isb.p, 0, nct, s));
        if (m->n_ghost)
            fvm::k_bc_phonon_ghost<<<fvm::grid_for(m->n_ghost, 256), 256, 0, s>>>(
                nc, nct, m->cf_row.p, m->cf_col.p, m->fcells.p, m->pos10.p, m->fA.p, m->fAm.p, vg[0], vg[1],
                vg[2], d_x, S->diag.p, S->off.p, S->b.p, S->isb.p);
        fvm::k_eliminate<<<fvm::grid_for(nc, 256), 256, 0, s>>>(nc, m->cf_row.p, m->cf_col.p, m->fcells.p,
                                                                m->pos01.p, m->pos10.p, S->isb.p, S->diag.p,
                                                                S->off.p, S->b.p);
        FVM_HIP(hipGetLastError());
        FVM_HIP(hipStreamSynchronize(s));
    } catch (...) {
        delete S;
        throw;
    }
    *out = S;
    return FVM_OK;
    FVM_API_END
}

int fvm_kinetic_assemble(fvm_mesh *m, const double *c, const double *d_nue, const double *d_feq, const double *d_x,
                         const int32_t *bc_kind, const double *wall_u, double eps, fvm_system **out) {
    FVM_API_BEGIN
    FVM_CHECK(m && c && d_nue && d_feq && d_x && out, FVM_E_INVALID, "null argument");
    FVM_CHECK(std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2]), FVM_E_INVALID,
              "the direction's velocity must be finite");
    const int ng = m->n_groups;
    FVM_CHECK(ng <= 1 || (bc_kind && wall_u), FVM_E_INVALID, "boundary groups need a kind and a wall velocity");
    for (int g = 1; g < ng; g++)
        FVM_CHECK(bc_kind[g] == 0 || bc_kind[g] == 1, FVM_E_INVALID, "kinetic boundary kind must be 0 or 1");
    fvm::Dev d(m->device);
    fvm_system *S = fvm::system_for_mesh(m);
    try {
        std::vector<int> grp((size_t)2 * ng + 1, 0);
        for (int g = 0; g <= ng; g++) grp[g] = m->gptr[g];
        for (int g = 1; g < ng; g++) grp[ng + 1 + g] = bc_kind[g];
        DArr<int> d_grp(grp.size());
        DArr<double> d_wall(3 * (size_t)std::max(ng, 1));
        DArr<double> F(std::max(m->n_faces, 1));
        DArr<uint8_t> inc(std::max(m->n_faces, 1));
        FVM_HIP(hipMemcpyAsync(d_grp.p, grp.data(), sizeof(int) * grp.size(), hipMemcpyHostToDevice, m->stream));
        if (ng > 1)
            FVM_HIP(hipMemcpyAsync(d_wall.p, wall_u, sizeof(double) * 3 * ng, hipMemcpyHostToDevice, m->stream));
        fvm::kinetic_assemble_into(m, c, d_nue, d_feq, d_x, d_grp.p, d_wall.p, eps, S, F.p, inc.p);
        FVM_HIP(hipStreamSynchronize(m->stream));   // the uploads and scratch above are freed on return
    } catch (...) {
        delete S;
        throw;
    }
    *out = S;
    return FVM_OK;
    FVM_API_END
}

int fvm_system_info(const fvm_system *S, int *n_self, int *n_total, long long *nnz) {
    if (!S) return FVM_E_INVALID;
    if (n_self) *n_self = S->n_self;
    if (n_total) *n_total = S->n_total;
    if (nnz) *nnz = S->nnz;
    return FVM_OK;
}

int fvm_system_device(const fvm_system *S, const int32_t **row, const int32_t **col, const double **diag,
                      const double **off, const double **b, const uint8_t **is_boundary) {
    if (!S) return FVM_E_INVALID;
    if (row) *row = S->row.p;
    if (col) *col = S->col.p;
    if (diag) *diag = S->diag.p;
    if (off) *off = S->off.p;
    if (b) *b = S->b.p;
    if (is_boundary) *is_boundary = S->isb.p;
    return FVM_OK;
}

int fvm_system_post_solve(const fvm_system *S, double *d_delta) {
    FVM_API_BEGIN
    FVM_CHECK(S, FVM_E_INVALID, "null system");
    FVM_CHECK(S->n_total == 0 || d_delta, FVM_E_INVALID, "null vector");
    const int ng = S->n_total - S->n_self;
    if (ng <= 0) return FVM_OK;
    fvm::Dev d(S->device);
    DArr<int> flag(1);
    FVM_HIP(hipMemset(flag.p, 0, sizeof(int)));
    const unsigned g = fvm::grid_for(ng, 256);
    fvm::k_boundary_dep<<<g, 256>>>(S->n_self, S->n_total, S->row.p, S->col.p, S->isb.p, flag.p);
    int dep = 0;
    FVM_HIP(hipMemcpy(&dep, flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (dep)   // a boundary row reads another one: the reference's ascending row order
        fvm::k_solve_boundary<true><<<1, 1>>>(S->n_self, S->n_total, S->row.p, S->col.p, S->diag.p, S->off.p,
                                              S->isb.p, S->b.p, d_delta);
    else
        fvm::k_solve_boundary<false><<<g, 256>>>(S->n_self, S->n_total, S->row.p, S->col.p, S->diag.p, S->off.p,
                                                 S->isb.p, S->b.p, d_delta);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipDeviceSynchronize());
    return FVM_OK;
    FVM_API_END
}

int fvm_system_get(const fvm_system *S, int32_t *row, int32_t *col, double *diag, double *off, double *b,
                   uint8_t *is_boundary) {
    FVM_API_BEGIN
    FVM_CHECK(S, FVM_E_INVALID, "null system");
    fvm::Dev d(S->device);
    if (row) FVM_HIP(hipMemcpy(row, S->row.p, S->row.bytes(), hipMemcpyDeviceToHost));
    if (col && S->nnz) FVM_HIP(hipMemcpy(col, S->col.p, sizeof(int) * S->nnz, hipMemcpyDeviceToHost));
    if (diag) FVM_HIP(hipMemcpy(diag, S->diag.p, S->diag.bytes(), hipMemcpyDeviceToHost));
    if (off && S->nnz) FVM_HIP(hipMemcpy(off, S->off.p, sizeof(double) * S->nnz, hipMemcpyDeviceToHost));
    if (b) FVM_HIP(hipMemcpy(b, S->b.p, S->b.bytes(), hipMemcpyDeviceToHost));
    if (is_boundary) FVM_HIP(hipMemcpy(is_boundary, S->isb.p, S->isb.bytes(), hipMemcpyDeviceToHost));
    return FVM_OK;
    FVM_API_END
}

int fvm_system_destroy(fvm_system *S) {
    FVM_API_BEGIN
    if (!S) return FVM_OK;
    {
        fvm::Dev d(S->device);
        S->row.release(); S->col.release(); S->diag.release(); S->off.release(); S->b.release(); S->isb.release();
    }
    delete S;
    return FVM_OK;
    FVM_API_END
}

int fvm_amg_setup_system(fvm_amg *h, const fvm_system *S) {
    FVM_API_BEGIN
    FVM_CHECK(h && S, FVM_E_INVALID, "null argument");
    FVM_CHECK(h->device == S->device, FVM_E_INVALID, "system and solver on different devices");
    return fvm_amg_setup_device(h, S->n_self, S->n_total, S->row.p, S->col.p, S->diag.p, S->off.p, S->isb.p);
    FVM_API_END
}

}  // extern "C"
