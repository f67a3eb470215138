// Learned correction for the AMG preconditioner: a 3-layer message-passing
// network over the graph of a CRMatrix (no counterpart in the reference).
//
// Graph of a system A (self rows, symmetric pattern): edge weight of entry p of
// row i c_p = -off_p / d_i, node feature l_i = (log|d_i| - mean) / std.
// For a residual r (delta form, A e = -r):
//   u = -r / d, s = rms(u), v = u / s, h0 = [v, C v, l, 0 x 13]      (16 features)
//   layer k: a = C h ; h' = relu(Ws_k h + Wn_k a + b_k)                 (64 features)
//   e = s (wo . h3 + bo)
// Weights are held as FP32 masters and used as bf16 copies: every dense layer
// runs on the matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation);
// activations are stored in bf16. Training (residual loss ||r + A e||^2 / ||r||^2,
// A symmetric) uses FP32 backward kernels with deterministic reductions and Adam.
#include "amg.hpp"

#include <cmath>

namespace fvm {
namespace {

constexpr int MH = 64;       // hidden features
constexpr int MF0 = 16;      // input features (3 used)
constexpr int MLAYERS = 3;
constexpr int BSZ = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr int fin_of(int k) { return k == 0 ? MF0 : MH; }
// packed parameter offsets: per layer Ws (64 x fin), Wn (64 x fin), b (64); then wo (64), bo (1)
constexpr int layer_off(int k) { return k == 0 ? 0 : (k == 1 ? 2 * MH * MF0 + MH : 2 * MH * MF0 + MH + (k - 1) * (2 * MH * MH + MH)); }
constexpr int WO_OFF = layer_off(MLAYERS);
constexpr int NPARAM = WO_OFF + MH + 1;

template <int B>
__device__ double bsum(double v) {
    __shared__ double sh[B / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < B / 64; k++) t += sh[k];
    __syncthreads();
    return t;
}

// ------------------------------------------------------------------ graph preparation
__global__ void k_mp_weights(int n, const int *row, const int *col, const double *off, const double *diag,
                             float *c) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double d = diag[i];
        for (int p = row[i]; p < row[i + 1]; p++) c[p] = (float)(-off[p] / d);
    }
}

// ct[q] for entry q = (j, i) of row j: the weight of entry (i, j) of row i; flag if absent
__global__ void k_mp_weights_t(int n, const int *row, const int *col, const float *c, float *ct, int *flag) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        for (int q = row[j]; q < row[j + 1]; q++) {
            const int i = col[q];
            float w = 0.f;
            bool found = false;
            for (int p = row[i]; p < row[i + 1]; p++)
                if (col[p] == j) { w = c[p]; found = true; break; }
            if (!found) atomicOr(flag, 1);
            ct[q] = w;
        }
}

__global__ void __launch_bounds__(BSZ) k_mp_logd_partial(int n, const double *diag, double *part) {
    double s = 0, s2 = 0;
    for (int i = blockIdx.x * BSZ + threadIdx.x; i < n; i += gridDim.x * BSZ) {
        const double l = log(fabs(diag[i]));
        s += l;
        s2 += l * l;
    }
    s = bsum<BSZ>(s);
    s2 = bsum<BSZ>(s2);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s; part[2 * blockIdx.x + 1] = s2; }
}

__global__ void k_mp_logd(int n, int nb, const double *diag, const double *part, float *l) {
    __shared__ double mean, sd;
    if (threadIdx.x == 0) {
        double s = 0, s2 = 0;
        for (int b = 0; b < nb; b++) { s += part[2 * b]; s2 += part[2 * b + 1]; }
        mean = s / n;
        double var = s2 / n - mean * mean;
        sd = sqrt(var > 0 ? var : 0.0) + 1e-12;
    }
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        l[i] = (float)((log(fabs(diag[i])) - mean) / sd);
}

// ------------------------------------------------------------------ forward
__global__ void __launch_bounds__(BSZ) k_mp_u(int n, const double *r, const double *diag, float *u, double *part) {
    double s = 0;
    for (int i = blockIdx.x * BSZ + threadIdx.x; i < n; i += gridDim.x * BSZ) {
        const double ui = -r[i] / diag[i];
        u[i] = (float)ui;
        s += ui * ui;
    }
    s = bsum<BSZ>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_mp_scale(int n, int nb, const double *part, double *s_out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int b = 0; b < nb; b++) s += part[b];
        *s_out = sqrt(s / (n > 0 ? n : 1)) + 1e-300;
    }
}

__global__ void k_mp_sum(int nb, const double *part, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int b = 0; b < nb; b++) s += part[b];
        *out = s;
    }
}

__global__ void k_mp_v(int n, const double *r, const double *diag, const double *s, float *v) {
    const double inv = 1.0 / *s;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[i] = (float)((-r[i] / diag[i]) * inv);
}

__global__ void k_mp_h0(int n, const int *row, const int *col, const float *c, const float *v, const float *l,
                        uint16_t *h0) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float a = 0.f;
        for (int p = row[i]; p < row[i + 1]; p++) a += c[p] * v[col[p]];
        uint16_t *o = h0 + (size_t)i * MF0;
        o[0] = f2bf(v[i]);
        o[1] = f2bf(a);
        o[2] = f2bf(l[i]);
        for (int f = 3; f < MF0; f++) o[f] = 0;
    }
}

// One layer for 64 nodes per workgroup (16 per wavefront):
//   a_i = sum_p c_p h_col(p)   (FP32 sums, stored bf16 for the backward pass)
//   h'_i = relu([h_i | a_i] W^T + b), W row o = [Ws(o,:) | Wn(o,:)] (bf16)
// The 16 x 2FIN tile [h | a] of a wavefront is staged in LDS; the products run as
// v_mfma_f32_16x16x32_bf16 over K = 2FIN in steps of 32, four 16-wide output blocks.
// Operand layout of that instruction (lane l): A row l%16, k = 8(l/16)..+8;
// B column l%16, same k; D rows 4(l/16)..+4, column l%16.
template <int FIN>
__global__ void __launch_bounds__(BSZ) k_mp_layer(int n, const int *row, const int *col, const float *c,
                                                  const uint16_t *H, const uint16_t *W, const float *bias,
                                                  uint16_t *Aout, uint16_t *Hn) {
    constexpr int K2 = 2 * FIN;
    constexpr int LDR = K2 + 8;   // padded LDS row (elements), keeps 16-byte alignment
    constexpr int FQ = FIN / 4;   // features per lane in the gather (4 lanes per node)
    __shared__ __attribute__((aligned(16))) uint16_t X[4][16 * LDR];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int node0 = (blockIdx.x * 4 + wave) * 16;
    uint16_t *Xw = X[wave];
    {
        const int nl = lane >> 2, q = lane & 3;
        const int i = node0 + nl;
        float acc[FQ];
#pragma unroll
        for (int j = 0; j < FQ; j++) acc[j] = 0.f;
        if (i < n) {
            const uint16_t *hi = H + (size_t)i * FIN + q * FQ;
#pragma unroll
            for (int j = 0; j < FQ; j++) Xw[nl * LDR + q * FQ + j] = hi[j];
            for (int p = row[i]; p < row[i + 1]; p++) {
                const float cp = c[p];
                const uint16_t *hj = H + (size_t)col[p] * FIN + q * FQ;
#pragma unroll
                for (int j = 0; j < FQ; j++) acc[j] += cp * bf2f(hj[j]);
            }
#pragma unroll
            for (int j = 0; j < FQ; j++) {
                const uint16_t a = f2bf(acc[j]);
                Xw[nl * LDR + FIN + q * FQ + j] = a;
                if (Aout) Aout[(size_t)i * FIN + q * FQ + j] = a;   // kept for the backward pass
            }
        } else {
#pragma unroll
            for (int j = 0; j < FQ; j++) {
                Xw[nl * LDR + q * FQ + j] = 0;
                Xw[nl * LDR + FIN + q * FQ + j] = 0;
            }
        }
    }
    __syncthreads();
    f32x4 d[4];
#pragma unroll
    for (int ob = 0; ob < 4; ob++) d[ob] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
    for (int kk = 0; kk < K2; kk += 32) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(Xw + r16 * LDR + kk + kq);
#pragma unroll
        for (int ob = 0; ob < 4; ob++) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(W + (size_t)(ob * 16 + r16) * K2 + kk + kq);
            d[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, d[ob], 0, 0, 0);
        }
    }
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        const int o = ob * 16 + r16;
        const float bo = bias[o];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int node = node0 + (lane >> 4) * 4 + t;
            if (node < n) {
                const float z = d[ob][t] + bo;
                Hn[(size_t)node * MH + o] = f2bf(z > 0.f ? z : 0.f);
            }
        }
    }
}

// e = s (wo . h3 + bo); accumulate: e += that
__global__ void k_mp_readout(int n, const uint16_t *h3, const float *wo, const float *bo, const double *s,
                             double *e, int accumulate, int *bad) {
    __shared__ float w[MH];
    if (threadIdx.x < MH) w[threadIdx.x] = wo[threadIdx.x];
    __syncthreads();
    const double sc = *s;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float y = bo[0];
        const uint16_t *hi = h3 + (size_t)i * MH;
        for (int o = 0; o < MH; o++) y += w[o] * bf2f(hi[o]);
        const double ei = sc * (double)y;
        if (!isfinite(ei)) atomicOr(bad, 1);
        e[i] = accumulate ? e[i] + ei : ei;
    }
}

// ------------------------------------------------------------------ loss and backward
// y = add + alpha A x (add may be null)
__global__ void k_mp_spmv(int n, const int *row, const int *col, const double *off, const double *diag,
                          const double *x, const double *add, double alpha, double *y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = diag[i] * x[i];
        for (int p = row[i]; p < row[i + 1]; p++) s += off[p] * x[col[p]];
        y[i] = add ? add[i] + alpha * s : alpha * s;
    }
}

__global__ void k_mp_sub(int n, const double *a, const double *b, double *y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a[i] - b[i];
}

__global__ void __launch_bounds__(BSZ) k_mp_dot_partial(int n, const double *a, const double *b, double *part) {
    double s = 0;
    for (int i = blockIdx.x * BSZ + threadIdx.x; i < n; i += gridDim.x * BSZ) s += a[i] * b[i];
    s = bsum<BSZ>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// dy_i = s * 2 (A t)_i / ||r||^2
__global__ void k_mp_dy(int n, const double *At, const double *s, const double *rr, float *dy) {
    const double f = 2.0 * (*s) / (*rr);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dy[i] = (float)(f * At[i]);
}

// dH3[i][o] = dy_i wo[o]; partial sums of dy_i h3[i][o] and dy_i per block of 256 nodes
__global__ void __launch_bounds__(BSZ) k_mp_out_bwd(int n, const float *dy, const uint16_t *h3, const float *wo,
                                                    float *dH, float *part) {
    __shared__ float acc[4][MH + 1];
    const int o = threadIdx.x & 63, g = threadIdx.x >> 6;
    float s = 0.f, sb = 0.f;
    const float w = wo[o];
    for (int k = g; k < BSZ; k += 4) {
        const int i = blockIdx.x * BSZ + k;
        if (i >= n) break;
        const float d = dy[i];
        dH[(size_t)i * MH + o] = d * w;
        s += d * bf2f(h3[(size_t)i * MH + o]);
        sb += d;
    }
    acc[g][o] = s;
    if (o == 0) acc[g][MH] = sb;
    __syncthreads();
    if (g == 0) {
        part[(size_t)blockIdx.x * (MH + 1) + o] = acc[0][o] + acc[1][o] + acc[2][o] + acc[3][o];
        if (o == 0) part[(size_t)blockIdx.x * (MH + 1) + MH] = acc[0][MH] + acc[1][MH] + acc[2][MH] + acc[3][MH];
    }
}

// Weight gradients of layer k over a block of 64 nodes (delta = dH' * [h' > 0]):
// part[blk][o][f] = sum_i delta[i][o] x_i[f] with x_i = [h_i | a_i], part[blk][64*2FIN + o] = sum_i delta[i][o]
template <int FIN>
__global__ void __launch_bounds__(BSZ) k_mp_wgrad(int n, const float *dH, const uint16_t *Hn, const uint16_t *H,
                                                  const uint16_t *A, float *part) {
    constexpr int K2 = 2 * FIN;
    constexpr int NB = 64;
    __shared__ float D[NB][MH + 1];
    __shared__ float Xs[NB][K2 + 1];
    const int i0 = blockIdx.x * NB;
    for (int t = threadIdx.x; t < NB * MH; t += BSZ) {
        const int k = t / MH, o = t % MH, i = i0 + k;
        float d = 0.f;
        if (i < n) d = bf2f(Hn[(size_t)i * MH + o]) > 0.f ? dH[(size_t)i * MH + o] : 0.f;
        D[k][o] = d;
    }
    for (int t = threadIdx.x; t < NB * K2; t += BSZ) {
        const int k = t / K2, f = t % K2, i = i0 + k;
        float x = 0.f;
        if (i < n) x = f < FIN ? bf2f(H[(size_t)i * FIN + f]) : bf2f(A[(size_t)i * FIN + f - FIN]);
        Xs[k][f] = x;
    }
    __syncthreads();
    float *pb = part + (size_t)blockIdx.x * (MH * K2 + MH);
    for (int t = threadIdx.x; t < MH * K2; t += BSZ) {
        const int o = t / K2, f = t % K2;
        float s = 0.f;
        for (int k = 0; k < NB; k++) s += D[k][o] * Xs[k][f];
        pb[t] = s;
    }
    for (int o = threadIdx.x; o < MH; o += BSZ) {
        float s = 0.f;
        for (int k = 0; k < NB; k++) s += D[k][o];
        pb[MH * K2 + o] = s;
    }
}

// G[i][f] = sum_o delta[i][o] Wn[o][f], P[i][f] = sum_o delta[i][o] Ws[o][f]
template <int FIN>
__global__ void k_mp_bwd_gp(int n, const float *dH, const uint16_t *Hn, const uint16_t *W, float *G, float *P) {
    constexpr int K2 = 2 * FIN;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < (long long)n * FIN;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t / FIN), f = (int)(t % FIN);
        float g = 0.f, pv = 0.f;
        for (int o = 0; o < MH; o++) {
            const float d = bf2f(Hn[(size_t)i * MH + o]) > 0.f ? dH[(size_t)i * MH + o] : 0.f;
            pv += d * bf2f(W[(size_t)o * K2 + f]);
            g += d * bf2f(W[(size_t)o * K2 + FIN + f]);
        }
        G[t] = g;
        P[t] = pv;
    }
}

// dH_k[i][f] = P[i][f] + sum over entries q of row i of ct[q] G[col q][f]
template <int FIN>
__global__ void k_mp_bwd_agg(int n, const int *row, const int *col, const float *ct, const float *G,
                             const float *P, float *dH) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < (long long)n * FIN;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t / FIN), f = (int)(t % FIN);
        float s = P[t];
        for (int q = row[i]; q < row[i + 1]; q++) s += ct[q] * G[(size_t)col[q] * FIN + f];
        dH[t] = s;
    }
}

// dst[off + j] (+)= sum over blocks of part[blk * stride + j], blocks in order
__global__ void k_mp_reduce(int nblk, int m, int stride, const float *part, float *dst) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int b = 0; b < nblk; b++) s += part[(size_t)b * stride + j];
        dst[j] += s;
    }
}

// layer gradient block [o][2FIN] + b[o] -> packed Ws (64 x FIN), Wn (64 x FIN), b (64)
template <int FIN>
__global__ void k_mp_scatter_wgrad(const float *g, float *grad_layer) {
    constexpr int K2 = 2 * FIN;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < MH * K2 + MH; t += gridDim.x * blockDim.x) {
        if (t < MH * K2) {
            const int o = t / K2, f = t % K2;
            if (f < FIN) grad_layer[o * FIN + f] += g[t];
            else grad_layer[MH * FIN + o * FIN + (f - FIN)] += g[t];
        } else {
            grad_layer[2 * MH * FIN + (t - MH * K2)] += g[t];
        }
    }
}

__global__ void k_mp_adam(int np, float *theta, float *grad, float *m, float *v, float lr, float b1, float b2,
                          float eps, float inv_count, float c1, float c2) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < np; j += gridDim.x * blockDim.x) {
        const float g = grad[j] * inv_count;
        const float mj = b1 * m[j] + (1.f - b1) * g;
        const float vj = b2 * v[j] + (1.f - b2) * g * g;
        m[j] = mj;
        v[j] = vj;
        theta[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
        grad[j] = 0.f;
    }
}

// bf16 copies of the layer matrices: row o = [Ws(o,:) | Wn(o,:)]
__global__ void k_mp_quantize(const float *theta, uint16_t *w0, uint16_t *w1, uint16_t *w2) {
    for (int k = 0; k < MLAYERS; k++) {
        const int fin = fin_of(k), off = layer_off(k);
        uint16_t *w = k == 0 ? w0 : (k == 1 ? w1 : w2);
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < MH * 2 * fin; t += gridDim.x * blockDim.x) {
            const int o = t / (2 * fin), f = t % (2 * fin);
            w[t] = f2bf(f < fin ? theta[off + o * fin + f] : theta[off + MH * fin + o * fin + (f - fin)]);
        }
    }
}

__global__ void k_mp_gather(int n, const int *perm, const double *src, double *dst) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) dst[p] = src[perm[p]];
}
__global__ void k_mp_scatter_sub(int n, const int *perm, const double *a, const double *b, double *dst) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        dst[perm ? perm[p] : p] = a[p] - b[p];
}
__global__ void k_mp_add_gathered(int n, const int *perm, const double *e, double *z) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        z[p] += e[perm ? perm[p] : p];
}

}  // namespace
}  // namespace fvm

using fvm::DArr;

extern "C" int fvm_mpnn_set_graph(fvm_mpnn *m, int n, const int32_t *d_row, const int32_t *d_col,
                                  const double *d_off, const double *d_diag);
extern "C" const char *fvm_last_error(void);

struct fvm_mpnn {
    int device = 0;
    hipStream_t stream = nullptr;
    fvm::Pool pool;
    DArr<float> theta, grad, adam_m, adam_v;
    DArr<uint16_t> wbf[fvm::MLAYERS];
    int adam_t = 0;
    int grad_count = 0;
    // graph (borrowed CSR)
    int n = 0;
    const int *row = nullptr, *col = nullptr;
    const double *off = nullptr, *diag = nullptr;
    DArr<float> c, ct, l;
    // activations and work
    DArr<uint16_t> H[fvm::MLAYERS + 1], A[fvm::MLAYERS];
    DArr<float> v, dy, dH, dH2, G, P, part;
    DArr<double> s, rr, e, t, At, dpart, tmp_a, tmp_b;
    DArr<int> flag;
};

namespace fvm {
namespace {

struct MpScope {
    int prev = -1;
    PoolScope ps;
    explicit MpScope(fvm_mpnn *m) : ps(&m->pool) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        FVM_HIP(hipSetDevice(m->device));
    }
    ~MpScope() { if (prev >= 0) (void)hipSetDevice(prev); }
};

void mp_alloc_acts(fvm_mpnn *m) {
    const size_t n = std::max(m->n, 1);
    const size_t nb = (n + BSZ - 1) / BSZ;
    m->H[0].ensure(n * MF0);
    for (int k = 1; k <= MLAYERS; k++) m->H[k].ensure(n * MH);
    m->A[0].ensure(n * MF0);
    for (int k = 1; k < MLAYERS; k++) m->A[k].ensure(n * MH);
    m->v.ensure(n);
    m->s.ensure(4);
    m->rr.ensure(4);
    m->dpart.ensure(2 * std::max<size_t>(nb, 1) + 8);
    m->flag.ensure(1);
}

// forward pass of the current graph on residual r (device, n); result in e_out or added to it;
// keep: store the aggregates for a backward pass
void mp_forward(fvm_mpnn *m, const double *r, double *e_out, int accumulate, bool keep = false) {
    hipStream_t st = m->stream;
    const int n = m->n;
    mp_alloc_acts(m);
    const unsigned nb = grid_for(n, BSZ, 1024);
    k_mp_u<<<nb, BSZ, 0, st>>>(n, r, m->diag, m->v.p, m->dpart.p);
    k_mp_scale<<<1, 64, 0, st>>>(n, (int)nb, m->dpart.p, m->s.p);
    k_mp_v<<<grid_for(n, BSZ), BSZ, 0, st>>>(n, r, m->diag, m->s.p, m->v.p);
    k_mp_h0<<<grid_for(n, BSZ), BSZ, 0, st>>>(n, m->row, m->col, m->c.p, m->v.p, m->l.p, m->H[0].p);
    const unsigned gl = (unsigned)((n + 63) / 64);
    for (int k = 0; k < MLAYERS; k++) {
        const float *bias = m->theta.p + layer_off(k) + 2 * MH * fin_of(k);
        uint16_t *aout = keep ? m->A[k].p : nullptr;
        if (k == 0)
            k_mp_layer<MF0><<<gl, BSZ, 0, st>>>(n, m->row, m->col, m->c.p, m->H[0].p, m->wbf[0].p, bias, aout,
                                               m->H[1].p);
        else
            k_mp_layer<MH><<<gl, BSZ, 0, st>>>(n, m->row, m->col, m->c.p, m->H[k].p, m->wbf[k].p, bias, aout,
                                              m->H[k + 1].p);
    }
    FVM_HIP(hipMemsetAsync(m->flag.p, 0, sizeof(int), st));
    k_mp_readout<<<grid_for(n, BSZ), BSZ, 0, st>>>(n, m->H[MLAYERS].p, m->theta.p + WO_OFF,
                                                   m->theta.p + WO_OFF + MH, m->s.p, e_out, accumulate, m->flag.p);
    FVM_HIP(hipGetLastError());
}

// forward + backward on one residual; gradient added to m->grad; returns the loss:
// et null: ||r + A e||^2 / ||r||^2; else (e - et)^T (-A) (e - et) / et^T (-A) et
double mp_loss_grad(fvm_mpnn *m, const double *r, const double *et) {
    hipStream_t st = m->stream;
    const int n = m->n;
    mp_alloc_acts(m);
    const size_t nn = std::max(n, 1);
    m->e.ensure(nn); m->t.ensure(nn); m->At.ensure(nn);
    m->dy.ensure(nn); m->dH.ensure(nn * MH); m->dH2.ensure(nn * MH);
    m->G.ensure(nn * MH); m->P.ensure(nn * MH);
    mp_forward(m, r, m->e.p, 0, true);
    const unsigned gs = grid_for(n, BSZ);
    const unsigned nb = grid_for(n, BSZ, 1024);
    // rr[0] = denominator, rr[1] = numerator; At = d(numerator)/de / 2
    if (!et) {
        // t = r + A e ; At = A t (A symmetric)
        k_mp_spmv<<<gs, BSZ, 0, st>>>(n, m->row, m->col, m->off, m->diag, m->e.p, r, 1.0, m->t.p);
        k_mp_spmv<<<gs, BSZ, 0, st>>>(n, m->row, m->col, m->off, m->diag, m->t.p, nullptr, 1.0, m->At.p);
        k_mp_dot_partial<<<nb, BSZ, 0, st>>>(n, r, r, m->dpart.p);
        k_mp_dot_partial<<<nb, BSZ, 0, st>>>(n, m->t.p, m->t.p, m->dpart.p + nb);
    } else {
        // t = e - et ; At = -A t ; den = et^T (-A et) (the -A et product goes to e, no longer needed)
        k_mp_sub<<<gs, BSZ, 0, st>>>(n, m->e.p, et, m->t.p);
        k_mp_spmv<<<gs, BSZ, 0, st>>>(n, m->row, m->col, m->off, m->diag, m->t.p, nullptr, -1.0, m->At.p);
        k_mp_spmv<<<gs, BSZ, 0, st>>>(n, m->row, m->col, m->off, m->diag, et, nullptr, -1.0, m->e.p);
        k_mp_dot_partial<<<nb, BSZ, 0, st>>>(n, et, m->e.p, m->dpart.p);
        k_mp_dot_partial<<<nb, BSZ, 0, st>>>(n, m->t.p, m->At.p, m->dpart.p + nb);
    }
    k_mp_sum<<<1, 64, 0, st>>>((int)nb, m->dpart.p, m->rr.p);
    k_mp_sum<<<1, 64, 0, st>>>((int)nb, m->dpart.p + nb, m->rr.p + 1);
    k_mp_dy<<<gs, BSZ, 0, st>>>(n, m->At.p, m->s.p, m->rr.p, m->dy.p);
    // readout backward
    const unsigned nbo = (unsigned)((nn + BSZ - 1) / BSZ);
    m->part.ensure((size_t)nbo * (MH + 1));
    k_mp_out_bwd<<<nbo, BSZ, 0, st>>>(n, m->dy.p, m->H[MLAYERS].p, m->theta.p + WO_OFF, m->dH.p, m->part.p);
    k_mp_reduce<<<1, 128, 0, st>>>((int)nbo, MH + 1, MH + 1, m->part.p, m->grad.p + WO_OFF);
    // layers
    float *dHk = m->dH.p, *dHo = m->dH2.p;
    const unsigned nbw = (unsigned)((nn + 63) / 64);
    for (int k = MLAYERS - 1; k >= 0; k--) {
        const int fin = fin_of(k);
        const size_t blk = (size_t)MH * 2 * fin + MH;
        m->part.ensure((size_t)nbw * blk);
        m->tmp_a.ensure((blk + 1) / 2 + 1);
        float *gsum = (float *)m->tmp_a.p;
        FVM_HIP(hipMemsetAsync(gsum, 0, blk * sizeof(float), st));
        if (fin == MF0)
            k_mp_wgrad<MF0><<<nbw, BSZ, 0, st>>>(n, dHk, m->H[k + 1].p, m->H[k].p, m->A[k].p, m->part.p);
        else
            k_mp_wgrad<MH><<<nbw, BSZ, 0, st>>>(n, dHk, m->H[k + 1].p, m->H[k].p, m->A[k].p, m->part.p);
        k_mp_reduce<<<grid_for((long long)blk, BSZ), BSZ, 0, st>>>((int)nbw, (int)blk, (int)blk, m->part.p, gsum);
        float *gl = m->grad.p + layer_off(k);
        if (fin == MF0) k_mp_scatter_wgrad<MF0><<<grid_for((long long)blk, BSZ), BSZ, 0, st>>>(gsum, gl);
        else k_mp_scatter_wgrad<MH><<<grid_for((long long)blk, BSZ), BSZ, 0, st>>>(gsum, gl);
        if (k > 0) {   // gradient of the layer input (fin = 64 here)
            const unsigned gg = grid_for((long long)n * MH, BSZ);
            k_mp_bwd_gp<MH><<<gg, BSZ, 0, st>>>(n, dHk, m->H[k + 1].p, m->wbf[k].p, m->G.p, m->P.p);
            k_mp_bwd_agg<MH><<<gg, BSZ, 0, st>>>(n, m->row, m->col, m->ct.p, m->G.p, m->P.p, dHo);
            std::swap(dHk, dHo);
        }
    }
    FVM_HIP(hipGetLastError());
    m->grad_count++;
    double h_rr[2];
    FVM_HIP(hipMemcpyAsync(h_rr, m->rr.p, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    FVM_HIP(hipStreamSynchronize(st));
    return h_rr[1] / h_rr[0];
}

void mp_quantize(fvm_mpnn *m) {
    k_mp_quantize<<<64, BSZ, 0, m->stream>>>(m->theta.p, m->wbf[0].p, m->wbf[1].p, m->wbf[2].p);
    FVM_HIP(hipGetLastError());
}

}  // namespace

// Preconditioner correction inside FGMRES: z += NN(A z - v) on the finest level.
// v, z in the AMG handle's layout (perm: coloured position -> row, or null).
void mpnn_correct(fvm_mpnn *m, hipStream_t st, const int *perm, const double *Az, const double *v, double *z,
                  int n, int *bad_out) {
    MpScope g(m);
    FVM_CHECK(n == m->n, FVM_E_INVALID, "learned correction: graph and system sizes differ");
    const hipStream_t keep = m->stream;
    m->stream = st;
    m->tmp_b.ensure((size_t)std::max(n, 1) * 2);
    double *rn = m->tmp_b.p, *en = m->tmp_b.p + std::max(n, 1);
    k_mp_scatter_sub<<<grid_for(n, BSZ), BSZ, 0, st>>>(n, perm, Az, v, rn);
    mp_forward(m, rn, en, 0);
    k_mp_add_gathered<<<grid_for(n, BSZ), BSZ, 0, st>>>(n, perm, en, z);
    FVM_HIP(hipGetLastError());
    m->stream = keep;
    if (bad_out) *bad_out = read_scalar(m->flag.p, st);
}

void mpnn_bind(fvm_mpnn *m, int n, const int *row, const int *col, const double *off, const double *diag) {
    if (m->n == n && m->row == row && m->col == col && m->off == off && m->diag == diag) return;
    const int rc = fvm_mpnn_set_graph(m, n, row, col, off, diag);
    if (rc) throw Error(rc, std::string("learned correction: ") + fvm_last_error());
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

int fvm_mpnn_num_params(void) { return fvm::NPARAM; }

int fvm_mpnn_create(fvm_mpnn **out, int device, uint64_t seed) {
    FVM_API_BEGIN
    FVM_CHECK(out, FVM_E_INVALID, "null output handle");
    int ndev = 0;
    FVM_HIP(hipGetDeviceCount(&ndev));
    FVM_CHECK(device >= 0 && device < ndev, FVM_E_INVALID, "device ordinal out of range");
    fvm_mpnn *m = new fvm_mpnn();
    m->device = device;
    try {
        fvm::MpScope g(m);
        FVM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamDefault));
        const int np = fvm::NPARAM;
        m->theta.alloc(np); m->grad.alloc(np); m->adam_m.alloc(np); m->adam_v.alloc(np);
        for (int k = 0; k < fvm::MLAYERS; k++) m->wbf[k].alloc((size_t)fvm::MH * 2 * fvm::fin_of(k));
        // He-style initialization from a seeded generator (host side, once)
        std::vector<float> th(np, 0.f);
        uint64_t x = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
        auto gauss = [&]() {
            auto u01 = [&]() {
                x ^= x << 13; x ^= x >> 7; x ^= x << 17;
                return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            };
            const double a = u01(), b = u01();
            return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
        };
        for (int k = 0; k < fvm::MLAYERS; k++) {
            const int fin = fvm::fin_of(k), off = fvm::layer_off(k);
            for (int j = 0; j < 2 * fvm::MH * fin; j++) th[off + j] = (float)(gauss() * std::sqrt(1.0 / fin));
        }
        for (int j = 0; j < fvm::MH; j++) th[fvm::WO_OFF + j] = (float)(gauss() * std::sqrt(1.0 / fvm::MH) * 0.1);
        FVM_HIP(hipMemcpyAsync(m->theta.p, th.data(), sizeof(float) * np, hipMemcpyHostToDevice, m->stream));
        FVM_HIP(hipMemsetAsync(m->grad.p, 0, sizeof(float) * np, m->stream));
        FVM_HIP(hipMemsetAsync(m->adam_m.p, 0, sizeof(float) * np, m->stream));
        FVM_HIP(hipMemsetAsync(m->adam_v.p, 0, sizeof(float) * np, m->stream));
        fvm::mp_quantize(m);
        FVM_HIP(hipStreamSynchronize(m->stream));
    } catch (...) {
        delete m;
        throw;
    }
    *out = m;
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_destroy(fvm_mpnn *m) {
    FVM_API_BEGIN
    if (!m) return FVM_OK;
    {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(m->device);
        hipStream_t st = m->stream;
        delete m;
        if (st) (void)hipStreamDestroy(st);
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_set_params(fvm_mpnn *m, const float *theta) {
    FVM_API_BEGIN
    FVM_CHECK(m && theta, FVM_E_INVALID, "null argument");
    fvm::MpScope g(m);
    FVM_HIP(hipMemcpyAsync(m->theta.p, theta, sizeof(float) * fvm::NPARAM, hipMemcpyHostToDevice, m->stream));
    fvm::mp_quantize(m);
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_get_params(const fvm_mpnn *m, float *theta) {
    FVM_API_BEGIN
    FVM_CHECK(m && theta, FVM_E_INVALID, "null argument");
    int prev = -1;
    (void)hipGetDevice(&prev);
    FVM_HIP(hipSetDevice(m->device));
    FVM_HIP(hipMemcpyAsync(theta, m->theta.p, sizeof(float) * fvm::NPARAM, hipMemcpyDeviceToHost, m->stream));
    FVM_HIP(hipStreamSynchronize(m->stream));
    if (prev >= 0) (void)hipSetDevice(prev);
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_set_graph(fvm_mpnn *m, int n, const int32_t *d_row, const int32_t *d_col, const double *d_off,
                       const double *d_diag) {
    FVM_API_BEGIN
    FVM_CHECK(m && n >= 0 && d_row && (n == 0 || d_diag), FVM_E_INVALID, "null argument");
    fvm::MpScope g(m);
    hipStream_t st = m->stream;
    const long long nnz = n > 0 ? fvm::read_scalar(d_row + n, st) : 0;
    FVM_CHECK(nnz == 0 || (d_col && d_off), FVM_E_INVALID, "null col/off arrays");
    m->n = n; m->row = d_row; m->col = d_col; m->off = d_off; m->diag = d_diag;
    m->c.alloc(std::max<long long>(nnz, 1));
    m->ct.alloc(std::max<long long>(nnz, 1));
    m->l.alloc(std::max(n, 1));
    fvm::mp_alloc_acts(m);
    if (n > 0) {
        FVM_HIP(hipMemsetAsync(m->flag.p, 0, sizeof(int), st));
        fvm::k_mp_weights<<<fvm::grid_for(n, fvm::BSZ), fvm::BSZ, 0, st>>>(n, d_row, d_col, d_off, d_diag, m->c.p);
        fvm::k_mp_weights_t<<<fvm::grid_for(n, fvm::BSZ), fvm::BSZ, 0, st>>>(n, d_row, d_col, m->c.p, m->ct.p,
                                                                             m->flag.p);
        const unsigned nb = fvm::grid_for(n, fvm::BSZ, 1024);
        fvm::k_mp_logd_partial<<<nb, fvm::BSZ, 0, st>>>(n, d_diag, m->dpart.p);
        fvm::k_mp_logd<<<fvm::grid_for(n, fvm::BSZ), fvm::BSZ, 0, st>>>(n, (int)nb, d_diag, m->dpart.p, m->l.p);
        FVM_HIP(hipGetLastError());
        FVM_CHECK(fvm::read_scalar(m->flag.p, st) == 0, FVM_E_INVALID,
                  "learned correction needs a structurally symmetric pattern");
    }
    FVM_HIP(hipStreamSynchronize(st));
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_forward(fvm_mpnn *m, const double *d_r, double *d_e, int accumulate) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_r && d_e, FVM_E_INVALID, "null argument");
    fvm::MpScope g(m);
    if (m->n > 0) fvm::mp_forward(m, d_r, d_e, accumulate);
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_loss_grad(fvm_mpnn *m, const double *d_r, double *loss) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_r, FVM_E_INVALID, "null argument");
    FVM_CHECK(m->n > 0, FVM_E_NOT_SETUP, "no graph");
    fvm::MpScope g(m);
    const double L = fvm::mp_loss_grad(m, d_r, nullptr);
    FVM_HIP(hipStreamSynchronize(m->stream));
    if (loss) *loss = L;
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_loss_grad_pair(fvm_mpnn *m, const double *d_r, const double *d_e, double *loss) {
    FVM_API_BEGIN
    FVM_CHECK(m && d_r && d_e, FVM_E_INVALID, "null argument");
    FVM_CHECK(m->n > 0, FVM_E_NOT_SETUP, "no graph");
    fvm::MpScope g(m);
    const double L = fvm::mp_loss_grad(m, d_r, d_e);
    FVM_HIP(hipStreamSynchronize(m->stream));
    if (loss) *loss = L;
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_get_grad(fvm_mpnn *m, float *grad, int reset) {
    FVM_API_BEGIN
    FVM_CHECK(m, FVM_E_INVALID, "null handle");
    fvm::MpScope g(m);
    if (grad)
        FVM_HIP(hipMemcpyAsync(grad, m->grad.p, sizeof(float) * fvm::NPARAM, hipMemcpyDeviceToHost, m->stream));
    if (reset) {
        FVM_HIP(hipMemsetAsync(m->grad.p, 0, sizeof(float) * fvm::NPARAM, m->stream));
        m->grad_count = 0;
    }
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

int fvm_mpnn_adam_step(fvm_mpnn *m, float lr) {
    FVM_API_BEGIN
    FVM_CHECK(m, FVM_E_INVALID, "null handle");
    FVM_CHECK(m->grad_count > 0, FVM_E_INVALID, "no accumulated gradient");
    fvm::MpScope g(m);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m->adam_t++;
    const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
    fvm::k_mp_adam<<<fvm::grid_for(fvm::NPARAM, fvm::BSZ), fvm::BSZ, 0, m->stream>>>(
        fvm::NPARAM, m->theta.p, m->grad.p, m->adam_m.p, m->adam_v.p, lr, b1, b2, eps, 1.f / m->grad_count, c1, c2);
    fvm::mp_quantize(m);
    m->grad_count = 0;
    FVM_HIP(hipStreamSynchronize(m->stream));
    return FVM_OK;
    FVM_API_END
}

}  // extern "C"
