// AMG cycle and solve loop on the GPU.
//
// Reference semantics (src/fvm/src/modules/fvmbase/):
//   AMG::solve   AMG.cpp:219-282  (residual, 1-norm, stop rule, returns r0)
//   AMG::cycle   AMG.cpp:70-147   (V/W/F recursion, restriction, correction)
//   doSweeps     AMG.cpp:43-68    (symmetric GS = forward + reverse, or 2 Jacobi)
//   residual     CRMatrix.h:407-426, GS CRMatrix.h:303-346, Jacobi :353-374
//   inject       Array.h:427-436, correct Array.h:438-467
//
// Every row keeps the reference's accumulation order (b + a_ii x_i, then the
// off-diagonals in CSR order) and the library is built without FP contraction,
// so the exact layout reproduces the reference bit for bit; the coloured
// layout changes only the Gauss-Seidel row order.
//
// Two provably-equal shortcuts of the reference are taken: a coarse level
// entered with delta = 0 and no pre-sweeps restricts b itself (its residual is
// b + A*0), and the finest residual of the convergence check is reused as the
// residual of the next cycle's restriction (delta has not changed in between).
#include "amg.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace fvm {

hipEvent_t timing_begin(fvm_amg *h) {
    if (!h->ktime) return nullptr;
    hipEvent_t e;
    if (h->ev_pool.empty()) {
        FVM_HIP(hipEventCreate(&e));
    } else {
        e = h->ev_pool.back();
        h->ev_pool.pop_back();
    }
    FVM_HIP(hipEventRecord(e, h->stream));
    return e;
}

void timing_end(fvm_amg *h, int kind, hipEvent_t start, double bytes) {
    if (!h->ktime || !start) return;
    hipEvent_t e;
    if (h->ev_pool.empty()) {
        FVM_HIP(hipEventCreate(&e));
    } else {
        e = h->ev_pool.back();
        h->ev_pool.pop_back();
    }
    FVM_HIP(hipEventRecord(e, h->stream));
    h->ev_open.push_back({kind, start, e, bytes});
}

// after the stream is synchronized: accumulate and recycle the events
void timing_harvest(fvm_amg *h) {
    for (auto &it : h->ev_open) {
        float ms = 0.f;
        FVM_HIP(hipEventElapsedTime(&ms, it.start, it.stop));
        h->k_launches[it.kind]++;
        h->k_ms[it.kind] += ms;
        h->k_bytes[it.kind] += it.bytes;
        h->ev_pool.push_back(it.start);
        h->ev_pool.push_back(it.stop);
    }
    h->ev_open.clear();
}

namespace {

constexpr int BS = 256;
constexpr int NORM_BLOCKS = 1024;

__device__ __forceinline__ double ldv(const double *p, size_t i) { return p[i]; }
__device__ __forceinline__ float ldv(const float *p, size_t i) { return p[i]; }
__device__ __forceinline__ float ldv(const uint16_t *p, size_t i) {
    return __uint_as_float(((unsigned)p[i]) << 16);
}
// value of a stored matrix entry (bf16 entries are the high half of an FP32)
__device__ __forceinline__ double cvt_val(double v) { return v; }
__device__ __forceinline__ float cvt_val(float v) { return v; }
__device__ __forceinline__ float cvt_val(uint16_t v) { return __uint_as_float(((unsigned)v) << 16); }

// ------------------------------------------------------------------ CSR (reference order)
template <class V, class D, class X>
__global__ void k_csr_residual(int n, const int *__restrict__ row, const int *__restrict__ col,
                               const V *__restrict__ off, const D *__restrict__ diag,
                               const X *__restrict__ x, const X *__restrict__ b, X *__restrict__ r) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        X s = b[i] + (X)diag[i] * x[i];
        for (int p = row[i]; p < row[i + 1]; p++) s += (X)ldv(off, p) * x[col[p]];
        r[i] = s;
    }
}

template <class V, class D, class X>
__global__ void k_csr_gs_level(const int *__restrict__ ptr, int lvl, const int *__restrict__ rows,
                               const int *__restrict__ row, const int *__restrict__ col,
                               const V *__restrict__ off, const D *__restrict__ diag, X *x,
                               const X *__restrict__ b, const uint8_t *__restrict__ mask) {
    const int beg = ptr[lvl], end = ptr[lvl + 1];
    for (int k = beg + blockIdx.x * blockDim.x + threadIdx.x; k < end; k += gridDim.x * blockDim.x) {
        const int i = rows[k];
        if (mask && !mask[i]) continue;   // row of a system that has stopped cycling
        X s = b[i];
        for (int p = row[i]; p < row[i + 1]; p++) s += (X)ldv(off, p) * x[col[p]];
        x[i] = -s / (X)diag[i];
    }
}

// A whole reference-order sweep of a level in one workgroup: the dependency
// levels one after another, separated by workgroup barriers instead of launches
// (the rows of one dependency level never read each other, as in k_csr_gs_level).
// The level offsets are staged through LDS in chunks.
constexpr int GSWG_CHUNK = 2048;
template <class V, class D, class X>
__global__ void __launch_bounds__(1024)
k_csr_gs_sweep_wg(int nlev, const int *__restrict__ ptr, const int *__restrict__ rows, const int *__restrict__ row,
                  const int *__restrict__ col, const V *__restrict__ off, const D *__restrict__ diag, X *x,
                  const X *__restrict__ b, const uint8_t *__restrict__ mask) {
    __shared__ int s_ptr[GSWG_CHUNK + 1];
    for (int c0 = 0; c0 < nlev; c0 += GSWG_CHUNK) {
        const int cn = min(GSWG_CHUNK, nlev - c0);
        for (int k = threadIdx.x; k <= cn; k += blockDim.x) s_ptr[k] = ptr[c0 + k];
        __syncthreads();
        for (int lv = 0; lv < cn; lv++) {
            const int beg = s_ptr[lv], end = s_ptr[lv + 1];
            for (int k = beg + threadIdx.x; k < end; k += blockDim.x) {
                const int i = rows[k];
                if (mask && !mask[i]) continue;   // row of a system that has stopped cycling
                X s = b[i];
                for (int p = row[i]; p < row[i + 1]; p++) s += (X)ldv(off, p) * x[col[p]];
                x[i] = -s / (X)diag[i];
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------- reference order as dataflow
// A reference-order Gauss-Seidel sweep in one launch: thread k updates row
// rows[k] of the dependency-level schedule (rows of earlier levels first) as soon
// as every neighbour it reads the new value of has published its own. The new
// values go to a separate array whose words are first set to a "pending" NaN
// that the arithmetic never produces (a NaN result is stored as the canonical
// quiet NaN), so one word carries both the value and its readiness: relaxed
// device-scope atomic loads and stores of that word need no fence. The
// neighbours whose old values a row reads are only read (from the input array).
// Blocks take schedule chunks in the order they start (a ticket counter), so a
// row waits only on rows of blocks already running or done: every wait ends.
// A wave that sees no block finish for two seconds gives up and raises an error
// flag (a broken schedule would otherwise spin forever); every wave that sees
// the flag stops at once, and blocks that start after it skip their rows. The
// pending words left behind are NaN, so the result is visibly invalid as well.
template <class X> struct FlowWord;
template <> struct FlowWord<double> {
    using U = unsigned long long;
    static constexpr U pend = 0x7FF4A5A5A5A5A5A5ull;
    static constexpr U qnan = 0x7FF8000000000000ull;
    static __device__ __forceinline__ U bits(double v) { return (U)__double_as_longlong(v); }
    static __device__ __forceinline__ double val(U u) { return __longlong_as_double((long long)u); }
};
template <> struct FlowWord<float> {
    using U = unsigned int;
    static constexpr U pend = 0x7FA5A5A5u;
    static constexpr U qnan = 0x7FC00000u;
    static __device__ __forceinline__ U bits(float v) { return __float_as_uint(v); }
    static __device__ __forceinline__ float val(U u) { return __uint_as_float(u); }
};
template <class X>
__device__ __forceinline__ typename FlowWord<X>::U flow_load(const X *p) {
    return __hip_atomic_load((const typename FlowWord<X>::U *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class X>
__device__ __forceinline__ void flow_store(X *p, X v) {
    using W = FlowWord<X>;
    const typename W::U u = v != v ? W::qnan : W::bits(v);
    __hip_atomic_store((typename W::U *)p, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// pending words for the sweep's output; ctl[0] (tickets) and ctl[1] (finished
// blocks) to zero; ctl[2] (error) is kept until the host reads it
template <class X>
__global__ void k_flow_prep(int n, X *xn, int *ctl) {
    if (blockIdx.x == 0 && threadIdx.x < 2) ctl[threadIdx.x] = 0;
    using W = FlowWord<X>;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        ((typename W::U *)xn)[i] = W::pend;
}

constexpr int FLOW_BS = 256;
constexpr int FLOW_RC = 16;   // entries of a row kept in registers while it waits
// Waiting rows poll the one new value that usually comes last (the nearest row in
// sweep order), then all of them; a wavefront with nothing to do naps between
// polls. Measured on the 200^3 hierarchy (ms per reference-order V-cycle): 96.2
// with up to 32 naps, 97.7 with 8, 103.4 without; polling every new value each
// time: 107-110 (the polls' own traffic slows the rows that are ready).
constexpr int FLOW_NAPS = 32;
template <class V, class D, class X>
__global__ void __launch_bounds__(FLOW_BS)
k_csr_gs_flow(int n, int dir, const int *__restrict__ rows, const int *__restrict__ row, const int *__restrict__ col,
              const V *__restrict__ off, const D *__restrict__ diag, const X *__restrict__ xo, X *xn,
              const X *__restrict__ b, const uint8_t *__restrict__ mask, int *ctl, long long give_up,
              int spin_limit) {
    using W = FlowWord<X>;
    __shared__ int s_chunk, s_abort;
    if (threadIdx.x == 0) {
        s_chunk = atomicAdd(&ctl[0], 1);
        s_abort = __hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int k = s_abort ? n : s_chunk * FLOW_BS + threadIdx.x;
    int i = 0, rb = 0, re = 0;
    bool todo = false;
    X bi = 0, di = 1;
    int c[FLOW_RC];
    X v[FLOW_RC], xv[FLOW_RC];
    unsigned lowm = 0;   // entries (of the first FLOW_RC) that read a new value
    int key = -1;        // the new value that usually comes last: polled alone until it is there
    if (k < n) {
        i = rows[k];
        if (mask && !mask[i]) {
            flow_store(xn + i, xo[i]);   // row of a system that has stopped cycling
        } else {
            todo = true;
            rb = row[i];
            re = row[i + 1];
            bi = b[i];
            di = (X)diag[i];
#pragma unroll
            for (int u = 0; u < FLOW_RC; u++) {
                c[u] = 0;
                v[u] = 0;
                xv[u] = 0;
                if (rb + u < re) {
                    c[u] = col[rb + u];
                    v[u] = (X)ldv(off, rb + u);
                }
            }
#pragma unroll
            for (int u = 0; u < FLOW_RC; u++)
                if (rb + u < re) {
                    if (dir ? c[u] > i : c[u] < i) {
                        lowm |= 1u << u;
                        // nearest in sweep order: the largest lower (forward) / smallest higher (backward) row
                        if (key < 0 || (dir ? c[u] < key : c[u] > key)) key = c[u];
                    } else {
                        xv[u] = xo[c[u]];
                    }
                }
        }
    }
    int spins = 0, seen = -1, nap = 0, waits = 0;
    long long t_seen = 0;
    while (__ballot(todo)) {
        bool moved = false;
        if (todo && key >= 0 && flow_load(xn + key) != W::pend) key = -1;
        if (todo && key < 0) {
            // the new values this row reads: all of them published?
            bool ready = true;
#pragma unroll
            for (int u = 0; u < FLOW_RC; u++)
                if (lowm >> u & 1u) {
                    const typename W::U w = flow_load(xn + c[u]);
                    ready = ready && w != W::pend;
                    xv[u] = W::val(w);
                }
            for (int p = rb + FLOW_RC; p < re && ready; p++) {
                const int j = col[p];
                if (dir ? j > i : j < i) ready = flow_load(xn + j) != W::pend;
            }
            if (ready) {
                // the reference's arithmetic: b, then the entries in CSR order, over -diag
                X sacc = bi;
#pragma unroll
                for (int u = 0; u < FLOW_RC; u++)
                    if (rb + u < re) sacc += v[u] * xv[u];
                for (int p = rb + FLOW_RC; p < re; p++) {
                    const int j = col[p];
                    const X xj = (dir ? j > i : j < i) ? W::val(flow_load(xn + j)) : xo[j];
                    sacc += (X)ldv(off, p) * xj;
                }
                flow_store(xn + i, -sacc / di);
                todo = false;
                moved = true;
            }
        }
        if (__ballot(todo)) {
            // back off while nothing in the wavefront moves (up to FLOW_NAPS naps of 128 cycles)
            nap = __ballot(moved) ? 0 : (nap < FLOW_NAPS ? nap + 1 : FLOW_NAPS);
            if (spin_limit > 0 && ++waits >= spin_limit) {   // a caller's bound on the waits (0: none)
                if (__lane_id() == 0) atomicOr(&ctl[2], 1);
                break;
            }
            for (int z = 0; z < nap; z++) __builtin_amdgcn_s_sleep(2);
            if (++spins == 4096) {   // every few milliseconds: has any block finished since?
                spins = 0;
                // one lane's view decides for the whole wavefront
                const int pr = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(&ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                const int ab = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                const long long now = (long long)wall_clock64();
                if (ab) break;   // another wavefront gave up
                if (pr != seen) {
                    seen = pr;
                    t_seen = now;
                } else if (now - t_seen > give_up) {
                    if (__lane_id() == 0) atomicOr(&ctl[2], 1);
                    break;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&ctl[1], 1);
}

// Launch plan of a reference-order sweep from its dependency-level offsets hp
// (nlev levels): each maximal run of consecutive levels of at most W rows is one
// single-workgroup launch; every wider level is a launch of its own.
void sweep_runs(const std::vector<int> &hp, int nlev, int W, std::vector<Level::SweepRun> &out) {
    out.clear();
    for (int k = 0; k < nlev; k++) {
        const int w = hp[k + 1] - hp[k];
        if (w > W) {
            out.push_back({k, 1, w, false});
        } else if (!out.empty() && out.back().wg) {
            out.back().nlev++;
            out.back().width = std::max(out.back().width, w);
        } else {
            out.push_back({k, 1, w, true});
        }
    }
}

template <class V, class D, class X>
__global__ void k_csr_jacobi(int n, const int *__restrict__ row, const int *__restrict__ col,
                             const V *__restrict__ off, const D *__restrict__ diag,
                             const X *__restrict__ xo, const X *__restrict__ b, X *__restrict__ xn,
                             const uint8_t *__restrict__ mask) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (mask && !mask[i]) { xn[i] = xo[i]; continue; }
        X s = b[i];
        for (int p = row[i]; p < row[i + 1]; p++) s += (X)ldv(off, p) * xo[col[p]];
        xn[i] = -s / (X)diag[i];
    }
}

// ------------------------------------------------------------------ SELL-64 (coloured order)

// The hardware hands workgroups to the 8 XCDs of the chip round-robin (block b
// to XCD b mod 8), and each XCD has its own L2. Renumbering the blocks so that
// XCD j processes the j-th contiguous eighth of the rows keeps the gathers of a
// row's neighbours (nearby rows, also in the coloured order) inside one L2.
// A bijection of [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block() {
    const int G = gridDim.x, b = blockIdx.x;
    const int j = b & 7, q = G >> 3, r = G & 7;
    return j * q + (j < r ? j : r) + (b >> 3);
}

// acc + sum over the SELL-64 entries of one row, in storage order (the row's CSR
// order). Loads of a group of U entries are issued before their products are
// added, so a lane keeps U gathers in flight; the additions stay in order.
// NT: matrix entries read with non-temporal loads (the finest levels stream
// their matrices once per sweep; coarse levels keep theirs in cache from cycle
// to cycle).
template <int U, bool NT, class V, class X>
__device__ __forceinline__ X sell_row(X acc, const int *__restrict__ scol, const V *__restrict__ sval, int o,
                                      int w, int lane, const X *x) {
    for (int k = 0; k < w; k += U) {
        int c[U];
        V v[U];   // raw storage values: a conversion here would wait for each load in turn
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) {
                const int e = o + (k + u) * 64 + lane;
                c[u] = NT ? __builtin_nontemporal_load(scol + e) : scol[e];
                v[u] = NT ? __builtin_nontemporal_load(sval + e) : sval[e];
            }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) acc += (X)cvt_val(v[u]) * xv[u];
    }
    return acc;
}
constexpr int SELL_U = 8;

// Launch a SELL-64 row kernel with the load depth U of the level (all loads of a
// row chunk of U entries are issued before the first is used: a wide coarse row
// costs one round trip to memory per chunk, so coarse levels take U = 16 or 32).
// The summation order does not depend on U.
// Non-temporal matrix loads on the nt_levels finest levels (SNT): they stream
// far more than the caches hold and would evict the coarse levels, which are
// re-read every cycle.
// FP64 values take at most U = 16 (32 doubles of values and gathers per lane would
// cost occupancy on the large levels).
#define FVM_SELL_U3(u_, CALL)                                                          \
    do {                                                                               \
        if ((u_) >= 32 && sizeof(*val) < 8) { constexpr int SU = 32; CALL; }           \
        else if ((u_) >= 16) { constexpr int SU = 16; CALL; }                          \
        else { constexpr int SU = 8; CALL; }                                           \
    } while (0)
#define FVM_SELL_U(u_, l_, CALL)                                                        \
    do {                                                                                \
        if ((l_) < R.h->nt_levels) { constexpr bool SNT = true; FVM_SELL_U3(u_, CALL); }  \
        else { constexpr bool SNT = false; FVM_SELL_U3(u_, CALL); }                      \
    } while (0)

template <int U, bool NT, class V, class D, class X>
__global__ void __launch_bounds__(512)
k_sell_residual(int n, const int *__restrict__ soff, const int *__restrict__ scol,
                const V *__restrict__ sval, const D *__restrict__ diag, const X *__restrict__ x,
                const X *__restrict__ b, X *__restrict__ r, int xcd) {
    const int p = (xcd ? xcd_block() : (int)blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = p >> 6, lane = p & 63;
    const int o = soff[s];
    const int w = (soff[s + 1] - o) >> 6;
    const X acc = b[p] + (X)diag[p] * x[p];
    r[p] = sell_row<U, NT>(acc, scol, sval, o, w, lane, x);
}

template <int U, bool NT, class V, class D, class X>
__global__ void __launch_bounds__(512)
k_sell_gs(int beg, int end, const int *__restrict__ soff, const int *__restrict__ scol,
          const V *__restrict__ sval, const D *__restrict__ diag, X *x, const X *__restrict__ b,
          const uint8_t *__restrict__ mask, int xcd) {
    // threads start at the slice holding beg: each wavefront covers exactly one slice
    const int p = (beg & ~63) + (xcd ? xcd_block() : (int)blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < beg || p >= end) return;
    if (mask && !mask[p]) return;   // row of a system that has stopped cycling
    const int s = p >> 6, lane = p & 63;
    const int o = soff[s];
    const int w = (soff[s + 1] - o) >> 6;
    const X acc = sell_row<U, NT>(b[p], scol, sval, o, w, lane, (const X *)x);
    x[p] = -acc / (X)diag[p];
}

template <int U, bool NT, class V, class D, class X>
__global__ void __launch_bounds__(512)
k_sell_jacobi(int n, const int *__restrict__ soff, const int *__restrict__ scol,
              const V *__restrict__ sval, const D *__restrict__ diag, const X *__restrict__ xo,
              const X *__restrict__ b, X *__restrict__ xn, const uint8_t *__restrict__ mask, int xcd) {
    const int p = (xcd ? xcd_block() : (int)blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (mask && !mask[p]) { xn[p] = xo[p]; return; }
    const int s = p >> 6, lane = p & 63;
    const int o = soff[s];
    const int w = (soff[s + 1] - o) >> 6;
    const X acc = sell_row<U, NT>(b[p], scol, sval, o, w, lane, xo);
    xn[p] = -acc / (X)diag[p];
}

// ------------------------------------------------------------------ transfer operators
// b_c[I] = sum of r over the fine rows of I in ascending order (inject, Array.h:427-436)
template <class X>
__global__ void k_restrict(int nc, const int *__restrict__ c2f_row, const int *__restrict__ c2f_col,
                           const X *__restrict__ r, X *__restrict__ bc) {
    for (int I = xcd_block() * blockDim.x + threadIdx.x; I < nc; I += gridDim.x * blockDim.x) {
        X s = 0;
        for (int q = c2f_row[I]; q < c2f_row[I + 1]; q++) s += r[c2f_col[q]];
        bc[I] = s;
    }
}

// x_i += X_c[ci_i] (correct, Array.h:438-467); set != 0: x was zero, write X_c[ci_i]
template <class X>
__global__ void k_prolong(int n, const int *__restrict__ ci, const X *__restrict__ xc, X *x, int set,
                          const double *scale, const uint8_t *__restrict__ mask) {
    for (int i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (mask && !mask[i]) continue;
        const int I = ci[i];
        if (scale) {
            const X sc = (X)*scale;
            if (I >= 0) x[i] = set ? (X)0 + xc[I] * sc : x[i] + xc[I] * sc;
            else if (set) x[i] = 0;
        } else {
            if (I >= 0) x[i] = set ? (X)0 + xc[I] : x[i] + xc[I];
            else if (set) x[i] = 0;
        }
    }
}

template <class X>
__global__ void k_fill(long long n, X *x, X v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

// ------------------------------------------------------------------ reductions
// Deterministic: fixed row->thread assignment and fixed reduction trees.
template <int B>
__device__ double block_sum(double v) {
    __shared__ double sh[B / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < B / 64; k++) t += sh[k];
    __syncthreads();
    return t;
}

template <class X>
__global__ void __launch_bounds__(BS) k_norm1_partial(int n, const X *__restrict__ r, double *part) {
    double s = 0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) s += fabs((double)r[i]);
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <class X>
__global__ void __launch_bounds__(BS) k_dot_partial(int n, const X *__restrict__ a,
                                                    const X *__restrict__ b, double *part) {
    double s = 0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) s += (double)a[i] * (double)b[i];
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x^T A x over the reference-order CSR (quadProduct, CRMatrix.h:273-296)
template <class V, class D, class X>
__global__ void __launch_bounds__(BS) k_quad_partial(int n, const int *row, const int *col, const V *off,
                                                     const D *diag, const X *x, double *part) {
    double s = 0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        X sn = (X)diag[i] * x[i];
        for (int p = row[i]; p < row[i + 1]; p++) sn += (X)ldv(off, p) * x[col[p]];
        s += (double)(sn * x[i]);
    }
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <class V, class D, class X>
__global__ void __launch_bounds__(BS) k_quad_partial_sell(int n, const int *soff, const int *scol,
                                                          const V *sval, const D *diag, const X *x,
                                                          double *part) {
    double s = 0;
    for (int p = blockIdx.x * BS + threadIdx.x; p < n; p += gridDim.x * BS) {
        const int sl = p >> 6, lane = p & 63;
        const int o = soff[sl];
        const int w = (soff[sl + 1] - o) >> 6;
        X sn = (X)diag[p] * x[p];
        for (int k = 0; k < w; k++) {
            const int e = o + k * 64 + lane;
            sn += (X)ldv(sval, e) * x[scol[e]];
        }
        s += (double)(sn * x[p]);
    }
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(BS) k_final_sum(int m, const double *part, double *out) {
    double s = 0;
    for (int i = threadIdx.x; i < m; i += BS) s += part[i];
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) *out = s;
}

// scale = -(x.b) / (x^T A x) with safeDivide, then limit(1,1): AMG.cpp:124-133,
// MultiFieldReduction.cpp:107-118, NumType.h:159-167
__global__ void k_scale(const double *xb, const double *xax, double *scale) {
    double s = -(*xb);
    if (*xax != 0) s /= *xax;
    if (s < 1.0) s = 1.0;
    else if (s > 1.0) s = 1.0;
    *scale = s;
}

// ------------------------------------------------------------------ precision / layout moves
__global__ void k_d2f(int n, const double *a, float *b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)a[i];
}
__global__ void k_d2f_neg(int n, const double *a, float *b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)(-a[i]);
}
__global__ void k_f2d(int n, const float *e, double *x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = (double)e[i];
}
__global__ void k_add_f2d(int n, const float *e, double *x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += (double)e[i];
}
__global__ void k_gather(int n, const int *perm, const double *src, double *dst) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        dst[p] = src[perm ? perm[p] : p];
}
__global__ void k_scatter(int n, const int *perm, const double *src, double *dst) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        dst[perm ? perm[p] : p] = src[p];
}

// Row i of A^{-1} b (dense FP64 inverse in reference order) by one wavefront: lane l
// sums j = l, l + 64, ... in order, then a butterfly over the lanes leaves the same
// total in every lane. Used by the per-level path and the tail alike, so both give
// the same bits. The caller's loop over rows must be uniform across the wavefront.
template <class X>
__device__ __forceinline__ double dense_row(const double *__restrict__ inv, int n, int i, const X *b,
                                           const int *iperm, int lane) {
    double s = 0;
    for (int j = lane; j < n; j += 64) s += inv[(size_t)i * n + j] * (double)b[iperm ? iperm[j] : j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// x = -A^{-1} b on the direct level, one wavefront per row
template <class X>
__global__ void k_dense_apply(int n, const double *inv, const int *perm, const int *iperm, const X *b,
                              X *x) {
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; p < n; p += nw) {
        const double s = dense_row(inv, n, perm ? perm[p] : p, b, iperm, lane);
        if (lane == 0) x[p] = (X)(-s);
    }
}

// ------------------------------------------------------------------ single-workgroup tail
// The small levels of a V-cycle in one launch: every step (restriction, a
// colour of a sweep, prolongation) is a loop of the workgroup over the rows of
// one level, steps separated by workgroup barriers. Same arithmetic per row as
// the per-level kernels; it only removes the launches.
struct TailLv {
    int n, ncolor;
    const int *soff, *scol, *cptr, *perm, *iperm;
    const void *val, *diag;
    const int *pci;                 // fine -> coarse (level + 1), coloured indices
    const int *c2f_row, *c2f_col;   // rows of level + 1 -> rows of this level
    void *b, *x, *t;
    int lds;                        // bytes of its sweep data staged in LDS (upper bound)
};

__host__ __device__ constexpr int al16(long long b) { return (int)((b + 15) & ~15LL); }

template <class V, class D, class X>
__device__ __forceinline__ void tail_rows(const TailLv &L, int beg, int end, const X *xin, X *xout,
                                          const X *b) {
    const V *val = (const V *)L.val;
    const D *dg = (const D *)L.diag;
    for (int p = beg + threadIdx.x; p < end; p += blockDim.x) {
        const int s = p >> 6, lane = p & 63;
        const int o = L.soff[s];
        const int w = (L.soff[s + 1] - o) >> 6;
        // wide coarse rows: deeper load chunks (same summation order), at most 16 deep
        // and 8 for FP64 values (deeper chunks spill registers of this 1024-thread kernel)
        X acc;
        if (sizeof(V) < 8 && w > 8) acc = sell_row<16, false>(b[p], L.scol, val, o, w, lane, xin);
        else acc = sell_row<8, false>(b[p], L.scol, val, o, w, lane, xin);
        xout[p] = -acc / (X)dg[p];
    }
}

template <class V, class D, class X>
__device__ void tail_sweeps(const TailLv &L, int ns, int jacobi) {
    X *x = (X *)L.x;
    const X *b = (const X *)L.b;
    for (int k = 0; k < ns; k++) {
        if (!jacobi) {
            for (int c = 0; c < L.ncolor; c++) {
                tail_rows<V, D, X>(L, L.cptr[c], L.cptr[c + 1], x, x, b);
                __syncthreads();
            }
            for (int c = L.ncolor - 2; c >= 0; c--) {   // the last colour is not repeated
                tail_rows<V, D, X>(L, L.cptr[c], L.cptr[c + 1], x, x, b);
                __syncthreads();
            }
        } else {
            X *t = (X *)L.t;
            tail_rows<V, D, X>(L, 0, L.n, x, t, b);
            __syncthreads();
            tail_rows<V, D, X>(L, 0, L.n, t, x, b);
            __syncthreads();
        }
    }
}

// The sweeps of a level whose matrix and vectors fit in the workgroup's LDS run
// from there: its SELL columns and values, slice offsets, diagonal, b and x are
// copied in, swept with the same row code (same arithmetic, same order), and x
// is copied back. A colour step then waits on LDS instead of the L2.
template <class V, class D, class X>
__device__ void tail_sweeps_lds(const TailLv &L, int ns, int jacobi, unsigned char *sm, int cap) {
    if (jacobi || ns <= 0 || L.lds == 0 || L.lds > cap) {
        tail_sweeps<V, D, X>(L, ns, jacobi);
        return;
    }
    const int n = L.n, nsl = (n + 63) >> 6;
    const int E = L.soff[nsl];
    int o = 0;
    int *scol = (int *)(sm + o); o += al16(4LL * E);
    V *val = (V *)(sm + o); o += al16((long long)sizeof(V) * E);
    int *soff = (int *)(sm + o); o += al16(4LL * (nsl + 1));
    D *dg = (D *)(sm + o); o += al16((long long)sizeof(D) * n);
    X *b = (X *)(sm + o); o += al16((long long)sizeof(X) * n);
    X *x = (X *)(sm + o);
    const V *gv = (const V *)L.val;
    // eight entries per thread in flight at a time: the copy is a few round trips
    // to memory instead of one per entry
    constexpr int SB = 8;
    for (int i0 = threadIdx.x; i0 < E; i0 += SB * blockDim.x) {
        int c[SB];
        V v[SB];
#pragma unroll
        for (int u = 0; u < SB; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < E) {
                c[u] = L.scol[i];
                v[u] = gv[i];
            }
        }
#pragma unroll
        for (int u = 0; u < SB; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < E) {
                scol[i] = c[u];
                val[i] = v[u];
            }
        }
    }
    for (int i = threadIdx.x; i <= nsl; i += blockDim.x) soff[i] = L.soff[i];
    const D *gd = (const D *)L.diag;
    const X *gb = (const X *)L.b;
    X *gx = (X *)L.x;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        dg[i] = gd[i];
        b[i] = gb[i];
        x[i] = gx[i];
    }
    __syncthreads();
    TailLv Ls = L;
    Ls.scol = scol; Ls.val = val; Ls.soff = soff; Ls.diag = dg; Ls.b = b; Ls.x = x;
    tail_sweeps<V, D, X>(Ls, ns, jacobi);
    for (int i = threadIdx.x; i < n; i += blockDim.x) gx[i] = x[i];
    __syncthreads();
}

template <class V, class D, class X>
__device__ void tail_fill0(const TailLv &L) {
    X *x = (X *)L.x;
    for (int p = threadIdx.x; p < L.n; p += blockDim.x) x[p] = (X)0;
}

template <class V, class D, class X>
__global__ void __launch_bounds__(1024)
k_tail(const TailLv *lv, int l0, int last, int npre, int npost, int jacobi, const double *inv, int direct,
       int lds_cap) {
    extern __shared__ __align__(16) unsigned char s_tail[];   // lds_cap bytes
    // descend: level l0 holds its b, its x is zero
    for (int l = l0; l < last; l++) {
        const TailLv &L = lv[l];
        const X *rr = (const X *)L.b;
        if (npre > 0) {
            tail_fill0<V, D, X>(L);
            __syncthreads();
            tail_sweeps_lds<V, D, X>(L, npre, jacobi, s_tail, lds_cap);
            // residual r = b + A x into t
            const V *val = (const V *)L.val;
            const D *dg = (const D *)L.diag;
            const X *x = (const X *)L.x;
            const X *b = (const X *)L.b;
            X *r = (X *)L.t;
            for (int p = threadIdx.x; p < L.n; p += blockDim.x) {
                const int s = p >> 6, lane = p & 63;
                const int o = L.soff[s];
                const int w = (L.soff[s + 1] - o) >> 6;
                const X acc = b[p] + (X)dg[p] * x[p];
                r[p] = sell_row<SELL_U, false>(acc, L.scol, val, o, w, lane, x);
            }
            __syncthreads();
            rr = r;
        }
        const TailLv &C = lv[l + 1];
        X *bc = (X *)C.b;
        for (int I = threadIdx.x; I < C.n; I += blockDim.x) {
            X sum = 0;
            for (int q = L.c2f_row[I]; q < L.c2f_row[I + 1]; q++) sum += rr[L.c2f_col[q]];
            bc[I] = sum;
        }
        __syncthreads();
    }
    // coarsest level of the tail
    {
        const TailLv &L = lv[last];
        if (direct) {
            const X *b = (const X *)L.b;
            X *x = (X *)L.x;
            const int n = L.n;
            const int lane = threadIdx.x & 63;
            for (int p = threadIdx.x >> 6; p < n; p += blockDim.x >> 6) {
                const double sum = dense_row(inv, n, L.perm[p], b, L.iperm, lane);
                if (lane == 0) x[p] = (X)(-sum);
            }
        } else {
            tail_fill0<V, D, X>(L);
            __syncthreads();
            tail_sweeps_lds<V, D, X>(L, npre + npost, jacobi, s_tail, lds_cap);
        }
        __syncthreads();
    }
    // ascend
    for (int l = last - 1; l >= l0; l--) {
        const TailLv &L = lv[l];
        const X *xc = (const X *)lv[l + 1].x;
        X *x = (X *)L.x;
        for (int i = threadIdx.x; i < L.n; i += blockDim.x) {
            const int I = L.pci[i];
            if (npre > 0) {
                if (I >= 0) x[i] = x[i] + xc[I];
            } else {
                x[i] = I >= 0 ? (X)0 + xc[I] : (X)0;
            }
        }
        __syncthreads();
        tail_sweeps_lds<V, D, X>(L, npost, jacobi, s_tail, lds_cap);
    }
}

}  // namespace

// ====================================================================== host side

namespace {

enum VT { V64 = 0, V32 = 1, V16 = 2 };

struct Run {
    fvm_amg *h;
    bool fast;
    bool mixed;
    int last;       // last level of the recursion
    bool r0_valid;  // level-0 residual holds b + A x for the current x
    const uint8_t *mask0 = nullptr;   // batched solve: finest rows still cycling
    bool allow_graph = true;          // false while capturing or inside a captured batch iteration
};

inline int value_type(const Run &R, int l) {
    if (!R.mixed) return V64;
    return l == 0 ? V32 : V16;
}

// Algorithmic bytes of one pass over `rows` rows holding `nnz` off-diagonal
// entries: per row the diagonal, `vecs` vectors of xsize bytes (read or written
// once each) and a 4-byte row offset; per entry a 4-byte column and its value.
inline double pass_bytes(int vt, int xsize, int vecs, long long rows, long long nnz) {
    const int vsize = vt == V64 ? 8 : (vt == V32 ? 4 : 2);
    const int dsize = vt == V64 ? 8 : 4;
    return (double)rows * (dsize + vecs * xsize + 4) + (double)nnz * (4 + vsize);
}

template <class X> X *vb(Level &L);
template <class X> X *vx(Level &L);
template <class X> X *vr(Level &L);
template <class X> DArr<X> &vx_arr(Level &L);
template <class X> DArr<X> &vt_arr(Level &L);
template <> double *vb<double>(Level &L) { return L.b.p; }
template <> double *vx<double>(Level &L) { return L.x.p; }
template <> double *vr<double>(Level &L) { return L.r.p; }
template <> DArr<double> &vx_arr<double>(Level &L) { return L.x; }
template <> DArr<double> &vt_arr<double>(Level &L) { return L.t; }
template <> float *vb<float>(Level &L) { return L.b32.p; }
template <> float *vx<float>(Level &L) { return L.x32.p; }
template <> float *vr<float>(Level &L) { return L.r32.p; }
template <> DArr<float> &vx_arr<float>(Level &L) { return L.x32; }
template <> DArr<float> &vt_arr<float>(Level &L) { return L.t32; }

// Dispatch a kernel over the (value, diagonal) types of level l.
#define FVM_DISPATCH_VALS(R, l, L, X, FAST_CALL, CSR_CALL)                                        \
    do {                                                                                          \
        const int vt_ = value_type(R, l);                                                         \
        if (R.fast) {                                                                             \
            if (vt_ == V64) { auto *val = L.sell_val.p; auto *dg = L.pdiag.p; FAST_CALL; }        \
            else if (vt_ == V32) { auto *val = L.sell_val32.p; auto *dg = L.pdiag32.p; FAST_CALL; } \
            else { auto *val = L.sell_bf16.p; auto *dg = L.pdiag32.p; FAST_CALL; }                \
        } else {                                                                                  \
            if (vt_ == V64) { auto *val = L.off.p; auto *dg = L.diag.p; CSR_CALL; }               \
            else if (vt_ == V32) { auto *val = L.off32.p; auto *dg = L.diag32.p; CSR_CALL; }      \
            else { auto *val = L.off16.p; auto *dg = L.diag32.p; CSR_CALL; }                      \
        }                                                                                         \
    } while (0)

template <class X>
void residual(Run &R, int l, const X *x, const X *b, X *r) {
    Level &L = R.h->lev[l];
    hipStream_t s = R.h->stream;
    const int sb = R.h->spmv_block;
    const int n = L.n;
    if (n == 0) return;
    hipEvent_t t0 = (l == 0 && value_type(R, l) == V64) ? timing_begin(R.h) : nullptr;
    FVM_DISPATCH_VALS(R, l, L, X,
        FVM_SELL_U(L.sell_u, l, (k_sell_residual<SU, SNT><<<grid_for(n, sb), sb, 0, s>>>(n, L.sell_off.p, L.sell_col.p, val,
                                                                               dg, x, b, r, R.h->spmv_xcd))),
        (k_csr_residual<<<grid_for(n, BS), BS, 0, s>>>(n, L.row.p, L.col.p, val, dg, x, b, r)));
    if (t0) timing_end(R.h, KT_RESID64, t0, pass_bytes(V64, sizeof(X), 3, n, L.nnz));   // x, b read; r written
}

template <class X>
void gs_sweep(Run &R, int l, X *x, const X *b) {
    Level &L = R.h->lev[l];
    hipStream_t s = R.h->stream;
    const int sb = R.h->sell_block;
    if (L.n == 0) return;
    if (R.fast) {
        for (int dir = 0; dir < 2; dir++) {
            hipEvent_t t0 = (l == 0) ? timing_begin(R.h) : nullptr;
            long long rows = 0, nnz = 0;
            // the backward pass starts with the colour the forward pass ended with:
            // its rows read only rows of other colours, unchanged since, so that
            // launch would rewrite the same values and is skipped
            for (int k = dir ? 1 : 0; k < L.ncolor; k++) {
                const int c = dir ? L.ncolor - 1 - k : k;
                const int beg = L.color_ptr[c], end = L.color_ptr[c + 1];
                if (end <= beg) continue;
                rows += end - beg;
                nnz += L.color_nnz[c];
                FVM_DISPATCH_VALS(R, l, L, X,
                    FVM_SELL_U(L.sell_u, l,
                        (k_sell_gs<SU, SNT><<<grid_for(end - (beg & ~63), sb), sb, 0, s>>>(
                            beg, end, L.sell_off.p, L.sell_col.p, val, dg, x, b, l == 0 ? R.mask0 : nullptr,
                            R.h->sell_xcd))),
                    (void)0);
            }
            // b read, x read and written (the neighbours' x are counted once per updated row)
            if (t0) timing_end(R.h, KT_SWEEP, t0, pass_bytes(value_type(R, l), sizeof(X), 3, rows, nnz));
        }
    } else if (R.h->exact_flow) {
        fvm_amg *h = R.h;
        FVM_CHECK(h->flow_ctl.p, FVM_E_INVALID, "reference-order sweep before its control words exist");
        h->flow_used = true;
        DArr<X> &xa = vx_arr<X>(L);
        DArr<X> &ta = vt_arr<X>(L);
        FVM_CHECK(x == xa.p, FVM_E_INVALID, "reference-order sweep on a vector other than the level's own");
        const int n = L.n;
        for (int dir = 0; dir < 2; dir++) {
            const int *rows = dir ? L.bwd_rows.p : L.fwd_rows.p;
            k_flow_prep<X><<<grid_for(n, BS, 4096), BS, 0, s>>>(n, ta.p, h->flow_ctl.p);
            FVM_DISPATCH_VALS(R, l, L, X, (void)0,
                (k_csr_gs_flow<<<(unsigned)((n + FLOW_BS - 1) / FLOW_BS), FLOW_BS, 0, s>>>(
                    n, dir, rows, L.row.p, L.col.p, val, dg, xa.p, ta.p, b, l == 0 ? R.mask0 : nullptr,
                    h->flow_ctl.p, h->flow_ticks, h->flow_spin_limit)));
            std::swap(xa, ta);   // the new values become x (two swaps per sweep: the arrays end where they began)
        }
    } else {
        if (L.runs_width != R.h->exact_wg_width) {
            sweep_runs(L.h_fwd_ptr, L.fwd_nlev, R.h->exact_wg_width, L.fwd_runs);
            sweep_runs(L.h_bwd_ptr, L.bwd_nlev, R.h->exact_wg_width, L.bwd_runs);
            L.runs_width = R.h->exact_wg_width;
        }
        for (int dir = 0; dir < 2; dir++) {
            const int *ptr = dir ? L.bwd_ptr.p : L.fwd_ptr.p;
            const int *rows = dir ? L.bwd_rows.p : L.fwd_rows.p;
            for (const Level::SweepRun &run : dir ? L.bwd_runs : L.fwd_runs) {
                if (run.wg) {
                    // narrow dependency levels: the run in one workgroup, as wide as its widest level
                    const int wb = std::min(1024, std::max(64, (run.width + 63) & ~63));
                    FVM_DISPATCH_VALS(R, l, L, X, (void)0,
                        (k_csr_gs_sweep_wg<<<1, wb, 0, s>>>(run.nlev, ptr + run.lvl0, rows, L.row.p, L.col.p, val,
                                                            dg, x, b, l == 0 ? R.mask0 : nullptr)));
                } else {
                    FVM_DISPATCH_VALS(R, l, L, X, (void)0,
                        (k_csr_gs_level<<<grid_for(run.width, BS), BS, 0, s>>>(ptr, run.lvl0, rows, L.row.p, L.col.p,
                                                                               val, dg, x, b,
                                                                               l == 0 ? R.mask0 : nullptr)));
                }
            }
        }
    }
}

template <class X>
void jacobi_pass(Run &R, int l) {
    Level &L = R.h->lev[l];
    hipStream_t s = R.h->stream;
    const int sb = R.h->sell_block;
    const int n = L.n;
    if (n == 0) return;
    DArr<X> &xa = vx_arr<X>(L);
    DArr<X> &ta = vt_arr<X>(L);
    const X *b = vb<X>(L);
    FVM_DISPATCH_VALS(R, l, L, X,
        FVM_SELL_U(L.sell_u, l, (k_sell_jacobi<SU, SNT><<<grid_for(n, sb), sb, 0, s>>>(n, L.sell_off.p, L.sell_col.p, val, dg,
                                                                             xa.p, b, ta.p, l == 0 ? R.mask0 : nullptr,
                                                                             R.h->sell_xcd))),
        (k_csr_jacobi<<<grid_for(n, BS), BS, 0, s>>>(n, L.row.p, L.col.p, val, dg, xa.p, b, ta.p,
                                                      l == 0 ? R.mask0 : nullptr)));
    std::swap(xa, ta);   // xnew becomes x (MultiFieldMatrix.cpp:203-215 copies it back)
}

template <class X>
void sweeps(Run &R, int l, int ns) {
    Level &L = R.h->lev[l];
    for (int k = 0; k < ns; k++) {
        if (R.h->prm.smootherType == FVM_GAUSS_SEIDEL) {
            gs_sweep<X>(R, l, vx<X>(L), vb<X>(L));
        } else {
            jacobi_pass<X>(R, l);
            jacobi_pass<X>(R, l);
        }
    }
}

template <class X>
void fill0(Run &R, int l) {
    Level &L = R.h->lev[l];
    if (L.n) k_fill<X><<<grid_for(L.n, BS, 4096), BS, 0, R.h->stream>>>(L.n, vx<X>(L), (X)0);
}

double *scratch_scalar(fvm_amg *h, int k) { return h->ws.scal.p + k; }

void reduce_final(fvm_amg *h, int nparts, double *out) {
    k_final_sum<<<1, BS, 0, h->stream>>>(nparts, h->ws.partials.p, out);
}

template <class X>
void symmetric_scale(Run &R, int l, double *scale_out) {
    // coarse level l+1: scale = -(x.b)/(x^T A x) limited to [1,1]
    fvm_amg *h = R.h;
    Level &C = h->lev[l + 1];
    hipStream_t s = h->stream;
    const int n = C.n;
    const unsigned g = grid_for(n, BS, NORM_BLOCKS);
    k_dot_partial<X><<<g, BS, 0, s>>>(n, vx<X>(C), vb<X>(C), h->ws.partials.p);
    reduce_final(h, g, scratch_scalar(h, 8));
    FVM_DISPATCH_VALS(R, l + 1, C, X,
        (k_quad_partial_sell<<<g, BS, 0, s>>>(n, C.sell_off.p, C.sell_col.p, val, dg, vx<X>(C), h->ws.partials.p)),
        (k_quad_partial<<<g, BS, 0, s>>>(n, C.row.p, C.col.p, val, dg, vx<X>(C), h->ws.partials.p)));
    reduce_final(h, g, scratch_scalar(h, 9));
    k_scale<<<1, 1, 0, s>>>(scratch_scalar(h, 8), scratch_scalar(h, 9), scale_out);
}

// Descriptors of every level for the tail kernel (coloured layout), built once per precision.
template <class X>
void build_tail_desc(Run &R) {
    fvm_amg *h = R.h;
    const int prec = R.mixed ? 1 : 0;
    if (h->tail_prec == prec && h->tail_desc.p) return;
    std::vector<TailLv> d(h->lev.size());
    for (size_t l = 0; l < h->lev.size(); l++) {
        Level &L = h->lev[l];
        TailLv &t = d[l];
        t.n = L.n;
        t.ncolor = L.ncolor;
        t.soff = L.sell_off.p; t.scol = L.sell_col.p; t.cptr = L.d_color_ptr.p;
        t.perm = L.perm.p; t.iperm = L.iperm.p;
        const int vt = value_type(R, (int)l);
        t.val = vt == V64 ? (const void *)L.sell_val.p : vt == V32 ? (const void *)L.sell_val32.p
                                                                      : (const void *)L.sell_bf16.p;
        t.diag = vt == V64 ? (const void *)L.pdiag.p : (const void *)L.pdiag32.p;
        t.pci = L.pci.p; t.c2f_row = L.pc2f_row.p; t.c2f_col = L.pc2f_col.p;
        t.b = vb<X>(L); t.x = vx<X>(L); t.t = vt_arr<X>(L).p;
        const long long E = (long long)L.sell_col.n, ns = ((long long)L.n + 63) / 64;
        const int vs = vt == V64 ? 8 : vt == V32 ? 4 : 2, ds = vt == V64 ? 8 : 4, xs = (int)sizeof(X);
        const long long need = (long long)al16(4 * E) + al16(vs * E) + al16(4 * (ns + 1)) + al16((long long)ds * L.n) +
                               2LL * al16((long long)xs * L.n);
        t.lds = (L.n > 0 && need <= (1 << 30)) ? (int)need : 0;
    }
    auto &lb = h->tail_lds_bytes[prec];
    lb.assign(d.size(), 0);
    for (size_t l = 0; l < d.size(); l++) lb[l] = d[l].lds;
    // the device's LDS per workgroup, granted to the tail kernel once (outside any capture)
    if (h->tail_lds_lim < 0) {
        int dev = 0, lim = 0;
        FVM_HIP(hipGetDevice(&dev));
        FVM_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        h->tail_lds_lim = std::max(lim, 0);
        if (lim > 0) {
            FVM_HIP(hipFuncSetAttribute((const void *)k_tail<uint16_t, float, float>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lim));
            FVM_HIP(hipFuncSetAttribute((const void *)k_tail<double, double, double>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lim));
        }
    }
    h->tail_desc.alloc(sizeof(TailLv) * d.size());
    FVM_HIP(hipMemcpyAsync(h->tail_desc.p, d.data(), sizeof(TailLv) * d.size(), hipMemcpyHostToDevice, h->stream));
    FVM_HIP(hipStreamSynchronize(h->stream));
    h->tail_prec = prec;
}

template <class X>
void cycle(Run &R, int type, int l, bool x_zero);

// Levels >= 1 of a coloured V-cycle replayed from a HIP graph: the launch
// sequence below the finest level is the same every cycle (fixed buffers), so it
// is captured once per precision and replayed. Same kernels in the same order:
// results are bit-identical to the launched cycle.
template <class X>
void graph_cycle(Run &R, int type, int l, bool x_zero) {
    fvm_amg *h = R.h;
    hipStream_t s = h->stream;
    const int idx = R.mixed ? 1 : 0;
    if (!h->cg_exec[idx]) {
        if (h->tail_level >= 0) build_tail_desc<X>(R);   // allocates: outside the capture
        hipGraph_t g = nullptr;
        FVM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        Run Rc = R;
        Rc.allow_graph = false;
        try {
            cycle<X>(Rc, type, l, x_zero);
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        FVM_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&h->cg_exec[idx], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) h->cg_exec[idx] = nullptr;
        FVM_HIP(e);
    }
    FVM_HIP(hipGraphLaunch(h->cg_exec[idx], s));
}

template <class X>
void cycle(Run &R, int type, int l, bool x_zero) {
    fvm_amg *h = R.h;
    Level &L = h->lev[l];
    hipStream_t s = h->stream;
    if (R.allow_graph && h->cycle_graph && R.fast && l == 1 && type == FVM_V_CYCLE && x_zero && l <= R.last &&
        !(h->prm.isSymmetric && h->prm.scaleCorrections)) {
        graph_cycle<X>(R, type, l, x_zero);
        return;
    }
    if (R.fast && l == h->tail_level && type == FVM_V_CYCLE && x_zero && l <= R.last &&
        !(h->prm.isSymmetric && h->prm.scaleCorrections)) {
        build_tail_desc<X>(R);
        const TailLv *d = (const TailLv *)h->tail_desc.p;
        const int direct = (R.mixed && R.last == h->direct_level) ? 1 : 0;
        const int jac = h->prm.smootherType == FVM_JACOBI ? 1 : 0;
        // LDS for the largest tail level within the limit set up with the descriptors
        int cap = 0;
        if (h->tail_lds)
            for (int k = l; k <= R.last; k++) {
                const int need = h->tail_lds_bytes[R.mixed ? 1 : 0][k];
                if (need > cap && need <= h->tail_lds_lim) cap = need;
            }
        auto launch = [&](auto kern) {
            kern<<<1, 1024, cap, s>>>(d, l, R.last, h->prm.nPreSweeps, h->prm.nPostSweeps, jac, h->dense_inv.p,
                                      direct, cap);
        };
        if (R.mixed) launch(k_tail<uint16_t, float, float>);
        else launch(k_tail<double, double, double>);
        FVM_HIP(hipGetLastError());
        return;
    }
    if (R.mixed && l == h->direct_level) {
        const bool perm = R.fast;
        k_dense_apply<X><<<grid_for((long long)L.n * 64, 256), 256, 0, s>>>(L.n, h->dense_inv.p, perm ? L.perm.p : nullptr,
                                                          perm ? L.iperm.p : nullptr, vb<X>(L), vx<X>(L));
        return;
    }
    if (h->prm.nPreSweeps > 0) {
        if (x_zero) fill0<X>(R, l);
        sweeps<X>(R, l, h->prm.nPreSweeps);
        x_zero = false;
        if (l == 0) R.r0_valid = false;   // delta changed: the stored residual is stale
    }
    if (l < R.last) {
        Level &C = h->lev[l + 1];
        const X *rr;
        if (x_zero) {
            rr = vb<X>(L);
        } else if (l == 0 && R.r0_valid) {
            rr = vr<X>(L);
        } else {
            residual<X>(R, l, vx<X>(L), vb<X>(L), vr<X>(L));
            rr = vr<X>(L);
        }
        if (l == 0) R.r0_valid = false;
        const int *c2r = R.fast ? L.pc2f_row.p : L.c2f_row.p;
        const int *c2c = R.fast ? L.pc2f_col.p : L.c2f_col.p;
        if (C.n) k_restrict<X><<<grid_for(C.n, BS), BS, 0, s>>>(C.n, c2r, c2c, rr, vb<X>(C));
        cycle<X>(R, type, l + 1, true);
        if (type == FVM_W_CYCLE) cycle<X>(R, FVM_W_CYCLE, l + 1, false);
        else if (type == FVM_F_CYCLE) cycle<X>(R, FVM_V_CYCLE, l + 1, false);
        const double *scale = nullptr;
        if (h->prm.isSymmetric && h->prm.scaleCorrections) {
            symmetric_scale<X>(R, l, scratch_scalar(h, 10));
            scale = scratch_scalar(h, 10);
        }
        const int *ci = R.fast ? L.pci.p : L.ci.p;
        if (L.n)
            k_prolong<X><<<grid_for(L.n, BS), BS, 0, s>>>(L.n, ci, vx<X>(C), vx<X>(L), x_zero ? 1 : 0, scale,
                                                          l == 0 ? R.mask0 : nullptr);
        x_zero = false;
    }
    if (h->prm.nPostSweeps > 0) {
        if (x_zero) fill0<X>(R, l);
        sweeps<X>(R, l, h->prm.nPostSweeps);
    } else if (x_zero) {
        fill0<X>(R, l);
    }
    FVM_HIP(hipGetLastError());
}

Run make_run(fvm_amg *h) {
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "solve/smooth called before setup");
    Run R;
    R.h = h;
    R.fast = (h->prm.gsOrdering == FVM_GS_MULTICOLOR);
    R.mixed = (h->prm.precision == FVM_PREC_MIXED);
    R.r0_valid = false;
    if (R.fast) prepare_fast(h);
    else if (h->prm.smootherType == FVM_GAUSS_SEIDEL)
        for (auto &L : h->lev) prepare_exact_schedule(h, L);
    if (!R.fast && h->exact_flow && !h->flow_ctl.p) {   // before any graph capture of the cycle
        h->flow_ctl.alloc(4);
        FVM_HIP(hipMemsetAsync(h->flow_ctl.p, 0, h->flow_ctl.bytes(), h->stream));
        int khz = 0;
        FVM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device));
        h->flow_ticks = 2000LL * (khz > 0 ? khz : 100000);
    }
    if (R.mixed) {
        prepare_mixed(h, R.fast);
        if (h->direct_level < 0 && !h->direct_tried) prepare_direct(h);
    }
    R.last = (int)h->lev.size() - 1;
    if (R.mixed && h->direct_level >= 0) R.last = h->direct_level;
    h->tail_level = -1;
    if (R.fast) {
        for (int l = 1; l <= R.last; l++)
            if (h->lev[l].n <= h->tail_max_rows) { h->tail_level = l; break; }
    }
    return R;
}

// 1-norm of the finest FP64 residual after r = b + A x; returns the norm (host).
double finest_residual_norm(Run &R) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    Run R64 = R;
    R64.mixed = false;
    residual<double>(R64, 0, F.x.p, F.b.p, F.r.p);
    const unsigned g = grid_for(F.n, BS, NORM_BLOCKS);
    k_norm1_partial<double><<<g, BS, 0, s>>>(F.n, F.r.p, h->ws.partials.p);
    reduce_final(h, g, scratch_scalar(h, 0));
    FVM_HIP(hipGetLastError());
    R.r0_valid = true;
    return read_scalar(scratch_scalar(h, 0), s);
}

// one cycle on the finest level; in mixed mode an FP32 cycle on the FP64 residual
void one_cycle(Run &R) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    if (!R.mixed) {
        cycle<double>(R, h->prm.cycleType, 0, false);
        return;
    }
    if (!R.r0_valid) {
        Run R64 = R;
        R64.mixed = false;
        residual<double>(R64, 0, F.x.p, F.b.p, F.r.p);
    }
    k_d2f<<<grid_for(F.n, BS), BS, 0, s>>>(F.n, F.r.p, F.b32.p);
    R.r0_valid = false;
    cycle<float>(R, h->prm.cycleType, 0, true);
    k_add_f2d<<<grid_for(F.n, BS), BS, 0, s>>>(F.n, F.x32.p, F.x.p);
    FVM_HIP(hipGetLastError());
}

void load_finest(Run &R, const double *d_b, const double *d_x) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    const int *perm = R.fast ? F.perm.p : nullptr;
    if (F.n == 0) return;
    k_gather<<<grid_for(F.n, BS), BS, 0, h->stream>>>(F.n, perm, d_b, F.b.p);
    k_gather<<<grid_for(F.n, BS), BS, 0, h->stream>>>(F.n, perm, d_x, F.x.p);
}

void store_finest(Run &R, double *d_x) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    const int *perm = R.fast ? F.perm.p : nullptr;
    if (F.n == 0) return;
    k_scatter<<<grid_for(F.n, BS), BS, 0, h->stream>>>(F.n, perm, F.x.p, d_x);
}

}  // namespace

// After a solve's last synchronisation: a dataflow sweep that gave up (see
// k_csr_gs_flow) makes the call fail.
void flow_check(fvm_amg *h) {
    if (!h->flow_used) return;
    h->flow_used = false;
    const int e = read_scalar(h->flow_ctl.p + 2, h->stream);
    if (e) {
        FVM_HIP(hipMemsetAsync(h->flow_ctl.p + 2, 0, sizeof(int), h->stream));
        throw Error(FVM_E_NUMERIC, "a reference-order Gauss-Seidel sweep stopped waiting for a row that was never "
                                   "updated (dependency schedule inconsistent with the matrix)");
    }
}

void solve_device(fvm_amg *h, const double *d_b, double *d_x, double *rnorm0, double *rnorm_final,
                  int *cycles) {
    Run R = make_run(h);
    load_finest(R, d_b, d_x);
    const double r0 = finest_residual_norm(R);
    h->hist.assign(1, r0);
    if (h->prm.verbosity > 0) print_norm(h, "0", r0);
    double rn = r0;
    int it = 0;
    if (!(r0 < h->prm.absoluteTolerance)) {
        for (int i = 1; i < h->prm.nMaxIterations; i++) {
            h->total_iterations++;
            one_cycle(R);
            rn = finest_residual_norm(R);
            if (!std::isfinite(rn)) flow_check(h);   // a sweep that gave up is the more precise cause
            FVM_CHECK(std::isfinite(rn), FVM_E_NUMERIC,
                      "the cycle produced a non-finite residual (cycle " + std::to_string(i) +
                          "): the system holds non-finite or zero-diagonal entries, or the cycle diverged");
            double ratio = rn;
            if (r0 != 0) ratio /= r0;
            it = i;
            h->hist.push_back(rn);
            const bool stop = rn < h->prm.absoluteTolerance || ratio < h->prm.relativeTolerance;
            // the parallel build prints only the line where the loop ends (AMG.cpp:262-268)
            if (h->prm.verbosity > 0 && (h->print_style == 0 || stop || i == h->prm.nMaxIterations - 1))
                print_norm(h, std::to_string(i), rn);
            if (stop) break;
        }
    }
    store_finest(R, d_x);
    FVM_HIP(hipStreamSynchronize(h->stream));
    flow_check(h);
    timing_harvest(h);
    if (rnorm0) *rnorm0 = r0;
    if (rnorm_final) *rnorm_final = rn;
    if (cycles) *cycles = it;
}

void smooth_device(fvm_amg *h, const double *d_b, double *d_x) {
    Run R = make_run(h);
    load_finest(R, d_b, d_x);
    one_cycle(R);
    store_finest(R, d_x);
    FVM_HIP(hipStreamSynchronize(h->stream));
    flow_check(h);
}

// x = -A^{-1} b on the direct level with the cycle's own kernel, rows in the
// level's reference order (no colour permutation); device vectors of n entries.
void direct_apply(fvm_amg *h, const double *d_b, double *d_x) {
    FVM_CHECK(h->direct_level >= 0, FVM_E_INVALID, "the hierarchy has no direct level");
    const int n = h->lev[h->direct_level].n;
    k_dense_apply<double><<<grid_for((long long)n * 64, 256), 256, 0, h->stream>>>(n, h->dense_inv.p, nullptr,
                                                                                   nullptr, d_b, d_x);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(h->stream));
}


// ===========================================================================
// Autotuner: launch shapes that cannot change a result (every row keeps its
// arithmetic; the single-workgroup tail is bit-identical to per-level launches)
// are timed on the handle's own hierarchy and parameters, and the fastest kept.
//   sell_block     threads per block of the SELL-64 kernels (row blocks of
//                  sell_block/64 slices)
//   tail_max_rows  levels at or below this size run in the single-workgroup tail
//   cycle_graph    levels >= 1 of the coloured V-cycle replayed from a captured graph
//   nt_levels      finest levels whose matrices are read with non-temporal loads
// Each candidate: one warm-up cycle, then the minimum over `reps` cycles from
// x = 0 on a fixed right-hand side, timed with HIP events on the handle's stream.
// ===========================================================================
namespace {
// Reads m doubles (a buffer larger than the last-level cache), so that the launch
// timed after it finds its matrix and vectors in HBM, as it does inside a solve
// where Gram-Schmidt and the cycle stream gigabytes between two products.
__global__ void k_tune_evict(long long m, const double *__restrict__ a, double *__restrict__ sink) {
    double v = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
        v += a[i];
    if (v == 1.0) sink[0] = v;   // never true for the zero buffer: keeps the loads
}

__global__ void k_tune_rhs(int n, double *b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        z = (z ^ (z >> 31)) * 0xBF58476D1CE4E5B9ull;
        z ^= z >> 29;
        b[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
}
}  // namespace

std::string autotune(fvm_amg *h, int reps) {
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "autotune called before setup");
    hipStream_t s = h->stream;
    Level &F = h->lev[0];
    const int n = F.n;
    const int keep_block = h->sell_block;
    const long long keep_tail = h->tail_max_rows;
    const bool keep_kt = h->ktime;
    h->ktime = false;
    std::vector<int> blocks = {128, 256, 512};
    const int keep_sx = h->sell_xcd, keep_px = h->spmv_xcd, keep_cg = h->cycle_graph;
    std::vector<long long> tails = {0, 1000, 2500, 10000, 40000};
    hipEvent_t e0, e1;
    FVM_HIP(hipEventCreate(&e0));
    FVM_HIP(hipEventCreate(&e1));
    std::string rep = "{\"candidates\": [";
    double best = 1e300;
    int best_b = keep_block, best_x = keep_sx, best_g = keep_cg;
    long long best_t = keep_tail;
    bool first = true;
    try {
        for (int cg = 0; cg < 2; cg++)
        for (int xm = 0; xm < 2; xm++)
        for (int bsz : blocks)
            for (long long tl : tails) {
                h->cycle_graph = cg;
                h->sell_block = bsz;
                h->tail_max_rows = tl;
                h->sell_xcd = xm;
                invalidate_cycle_graphs(h);
                Run R = make_run(h);
                // the tail and the captured cycle only exist in the coloured layout
                if (!R.fast && (tl != tails[0] || cg)) continue;
                double tmin = 1e300;
                for (int r = 0; r <= reps; r++) {
                    if (n) {
                        k_tune_rhs<<<grid_for(n, BS), BS, 0, s>>>(n, F.b.p);
                        k_fill<double><<<grid_for(n, BS, 4096), BS, 0, s>>>(n, F.x.p, 0.0);
                    }
                    R.r0_valid = false;
                    FVM_HIP(hipEventRecord(e0, s));
                    one_cycle(R);
                    FVM_HIP(hipEventRecord(e1, s));
                    FVM_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    FVM_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0) tmin = std::min(tmin, (double)ms);
                }
                char buf[240];
                snprintf(buf, sizeof buf,
                         "%s{\"sell_block\": %d, \"tail_max_rows\": %lld, \"sell_xcd\": %d, \"cycle_graph\": %d, "
                         "\"cycle_ms\": %.4f}",
                         first ? "" : ", ", bsz, tl, xm, cg, tmin);
                rep += buf;
                first = false;
                if (tmin < best) { best = tmin; best_b = bsz; best_t = tl; best_x = xm; best_g = cg; }
            }
    } catch (...) {
        h->sell_block = keep_block;
        h->tail_max_rows = keep_tail;
        h->sell_xcd = keep_sx;
        h->cycle_graph = keep_cg;
        invalidate_cycle_graphs(h);
        h->ktime = keep_kt;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    h->sell_block = best_b;
    h->tail_max_rows = best_t;
    h->sell_xcd = best_x;
    h->cycle_graph = best_g;
    invalidate_cycle_graphs(h);
    // levels with non-temporal matrix loads, with the shapes chosen above
    std::string nrep;
    int best_nt = h->nt_levels;
    {
        const int keep_nt = h->nt_levels;
        try {
            double bnt = 1e300;
            for (int nt = 0; nt <= 3; nt++) {
                h->nt_levels = nt;
                invalidate_cycle_graphs(h);
                Run R = make_run(h);
                double tmin = 1e300;
                for (int r = 0; r <= reps; r++) {
                    if (n) {
                        k_tune_rhs<<<grid_for(n, BS), BS, 0, s>>>(n, F.b.p);
                        k_fill<double><<<grid_for(n, BS, 4096), BS, 0, s>>>(n, F.x.p, 0.0);
                    }
                    R.r0_valid = false;
                    FVM_HIP(hipEventRecord(e0, s));
                    one_cycle(R);
                    FVM_HIP(hipEventRecord(e1, s));
                    FVM_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    FVM_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0) tmin = std::min(tmin, (double)ms);
                }
                char b3[120];
                snprintf(b3, sizeof b3, "%s{\"nt_levels\": %d, \"cycle_ms\": %.4f}", nrep.empty() ? "" : ", ", nt, tmin);
                nrep += b3;
                if (tmin < bnt) { bnt = tmin; best_nt = nt; }
            }
        } catch (...) {
            h->nt_levels = keep_nt;
            invalidate_cycle_graphs(h);
            h->ktime = keep_kt;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        h->nt_levels = best_nt;
        invalidate_cycle_graphs(h);
    }
    // residual / SpMV kernels of the finest level on their own
    const int keep_spmv = h->spmv_block;
    int best_s = keep_spmv, best_px = keep_px;
    double best_sms = 1e300;
    std::string srep;
    if (n) {
        try {
            Run R = make_run(h);
            Run R64 = R;
            R64.mixed = false;
            // timed with cold caches: 512 MiB read before every launch
            const long long m_ev = 64LL << 20;
            DArr<double> ev((size_t)m_ev), sink(1);
            FVM_HIP(hipMemsetAsync(ev.p, 0, ev.bytes(), s));
            for (int xm = 0; xm < 2; xm++)
            for (int bsz : blocks) {
                h->spmv_block = bsz;
                h->spmv_xcd = xm;
                double tmin = 1e300;
                for (int r = 0; r <= reps + 2; r++) {
                    k_tune_evict<<<2048, BS, 0, s>>>(m_ev, ev.p, sink.p);
                    FVM_HIP(hipEventRecord(e0, s));
                    residual<double>(R64, 0, F.x.p, F.b.p, F.r.p);
                    FVM_HIP(hipEventRecord(e1, s));
                    FVM_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    FVM_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0) tmin = std::min(tmin, (double)ms);
                }
                char b2[160];
                snprintf(b2, sizeof b2, "%s{\"spmv_block\": %d, \"spmv_xcd\": %d, \"residual_ms\": %.4f}",
                         srep.empty() ? "" : ", ", bsz, xm, tmin);
                srep += b2;
                if (tmin < best_sms) { best_sms = tmin; best_s = bsz; best_px = xm; }
            }
        } catch (...) {
            h->spmv_block = keep_spmv;
            h->spmv_xcd = keep_px;
            h->ktime = keep_kt;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->spmv_block = best_s;
    h->spmv_xcd = best_px;
    h->ktime = keep_kt;
    std::string tail_rep = "], \"nt_candidates\": [" + nrep + "], \"spmv_candidates\": [" + srep + "], \"chosen\": ";
    char buf[400];
    snprintf(buf, sizeof buf,
             "{\"sell_block\": %d, \"tail_max_rows\": %lld, \"sell_xcd\": %d, \"cycle_graph\": %d, "
             "\"nt_levels\": %d, \"cycle_ms\": %.4f, \"spmv_block\": %d, \"spmv_xcd\": %d, \"residual_ms\": %.4f}}",
             best_b, best_t, best_x, best_g, best_nt, best, best_s, best_px, best_sms == 1e300 ? 0.0 : best_sms);
    rep += tail_rep;
    rep += buf;
    h->tune_report = rep;
    return rep;
}


// ===========================================================================
// Batched solve of independent systems on one block-diagonal hierarchy.
// Each system follows AMG::solve (AMG.cpp:219-282) on its own: its own initial
// norm, stop test and cycle count. A system that has met its test stops
// changing (its finest rows are masked out of the smoother and the
// correction), so its result is the one its own unbatched solve returns.
// Per-system 1-norms use the same blocks and summation tree as the unbatched
// norm of a system of that size.
// ===========================================================================
namespace {

__device__ __forceinline__ int batch_sys(const int *off, int K, int i) {
    int lo = 0, hi = K;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// finest vectors in the hierarchy's layout (perm: coloured position -> row; null: identity)
__global__ void k_batch_gather(int total, int K, const int *off, const int *perm, const double *const *bp,
                               double *const *xp, double *b, double *x) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const int i = perm ? perm[p] : p;
        const int k = batch_sys(off, K, i);
        const int li = i - off[k];
        b[p] = bp[k][li];
        x[p] = xp[k][li];
    }
}

__global__ void k_batch_scatter(int total, int K, const int *off, const int *perm, const double *x,
                                double *const *xp) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const int i = perm ? perm[p] : p;
        const int k = batch_sys(off, K, i);
        xp[k][i - off[k]] = x[p];
    }
}

__global__ void k_batch_syskey(int total, int K, const int *off, const int *perm, int *key, int *idx) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        key[p] = batch_sys(off, K, perm[p]);
        idx[p] = p;
    }
}

// block j of system k covers local rows j*BS + t, stride g_k*BS: the unbatched
// k_norm1_partial on that system's rows with g_k = its block count
// sysperm (coloured layout): positions of system k's rows, in its own coloured order
__global__ void __launch_bounds__(BS) k_seg_norm1_partial(int K, const int *off, const int *blk_off,
                                                          const int *sysperm, const double *__restrict__ r,
                                                          double *part) {
    const int k = batch_sys(blk_off, K, blockIdx.x);
    const int j = blockIdx.x - blk_off[k];
    const int gk = blk_off[k + 1] - blk_off[k];
    const int n = off[k + 1] - off[k];
    const int o = off[k];
    double s = 0;
    for (int i = j * BS + threadIdx.x; i < n; i += gk * BS) s += fabs(r[sysperm ? sysperm[o + i] : o + i]);
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(BS) k_seg_final(const int *blk_off, const double *part, double *out) {
    const int k = blockIdx.x;
    const int m = blk_off[k + 1] - blk_off[k];
    const double *pk = part + blk_off[k];
    double s = 0;
    for (int i = threadIdx.x; i < m; i += BS) s += pk[i];
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) out[k] = s;
}

__global__ void __launch_bounds__(BS) k_batch_init(int K, const int *off, const double *norm, double abs_tol,
                                                   double *r0, double *rn, uint8_t *act, int *cyc, int *ctr) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int c = 0;
    for (int k = threadIdx.x; k < K; k += BS) {
        r0[k] = norm[k];
        rn[k] = norm[k];
        cyc[k] = 0;
        const bool a = !(norm[k] < abs_tol);
        act[k] = a ? 1 : 0;
        c += a;
    }
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) { ctr[0] = 0; ctr[1] = cnt; }
}

// stop test of AMG.cpp:262-272 for every system still cycling
__global__ void __launch_bounds__(BS) k_batch_check(int K, const double *norm, double rel_tol, double abs_tol,
                                                    const double *r0, double *rn, uint8_t *act, int *cyc,
                                                    int *ctr) {
    __shared__ int cnt, it;
    if (threadIdx.x == 0) { cnt = 0; it = ctr[0] + 1; }
    __syncthreads();
    int c = 0;
    for (int k = threadIdx.x; k < K; k += BS) {
        if (!act[k]) continue;
        const double v = norm[k];
        double ratio = v;
        if (r0[k] != 0) ratio /= r0[k];
        rn[k] = v;
        cyc[k] = it;
        if (v < abs_tol || ratio < rel_tol) act[k] = 0;
        else c++;
    }
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) { ctr[0] = it; ctr[1] = cnt; }
}

__global__ void k_batch_mask(int total, int K, const int *off, const int *perm, const uint8_t *act,
                             uint8_t *mask) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x)
        mask[p] = act[batch_sys(off, K, perm ? perm[p] : p)];
}

// finest residual and per-system norms into B.norm
void batch_norms(Run &R, BatchInfo &B) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    Run R64 = R;
    R64.mixed = false;
    residual<double>(R64, 0, F.x.p, F.b.p, F.r.p);
    if (B.nblk > 0)
        k_seg_norm1_partial<<<B.nblk, BS, 0, s>>>(B.K, B.d_off0.p, B.blk_off.p, R.fast ? B.sysperm.p : nullptr,
                                                   F.r.p, B.partials.p);
    k_seg_final<<<B.K, BS, 0, s>>>(B.blk_off.p, B.partials.p, B.norm.p);
    R.r0_valid = true;
}

void batch_iteration(Run &R, BatchInfo &B) {
    fvm_amg *h = R.h;
    hipStream_t s = h->stream;
    one_cycle(R);
    batch_norms(R, B);
    k_batch_check<<<1, BS, 0, s>>>(B.K, B.norm.p, h->prm.relativeTolerance, h->prm.absoluteTolerance, B.r0.p,
                                   B.rn.p, B.act.p, B.cyc.p, B.counters.p);
    const int total = h->lev[0].n;
    if (total)
        k_batch_mask<<<grid_for(total, BS), BS, 0, s>>>(total, B.K, B.d_off0.p, R.fast ? h->lev[0].perm.p : nullptr,
                                                        B.act.p, B.mask0.p);
}

}  // namespace

void batch_solve(fvm_amg *h, BatchInfo &B, const double *const *d_b, double *const *d_x, double *rnorm0,
                 double *rnorm_final, int *cycles) {
    FVM_CHECK(h->ready, FVM_E_NOT_SETUP, "batched solve called before setup");
    Run R = make_run(h);
    R.allow_graph = false;   // the batched iteration is captured as a whole
    FVM_CHECK(!R.mixed, FVM_E_INVALID, "batched solve supports the FP64 cycle only");
    FVM_CHECK(!(h->prm.isSymmetric && h->prm.scaleCorrections), FVM_E_INVALID,
              "batched solve does not support scaled corrections");
    hipStream_t s = h->stream;
    const int K = B.K;
    const int total = h->lev[0].n;
    const std::vector<int> &off0 = B.off[0];
    if (B.blk_off.n != (size_t)K + 1) {
        std::vector<int> blk(K + 1, 0);
        for (int k = 0; k < K; k++) {
            const int n = off0[k + 1] - off0[k];
            blk[k + 1] = blk[k] + (n > 0 ? (int)grid_for(n, BS, NORM_BLOCKS) : 0);
        }
        B.nblk = blk[K];
        B.blk_off.alloc(K + 1);
        FVM_HIP(hipMemcpyAsync(B.blk_off.p, blk.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice, s));
        B.partials.alloc(std::max(B.nblk, 1));
        B.norm.alloc(K); B.r0.alloc(K); B.rn.alloc(K);
        B.cyc.alloc(K); B.counters.alloc(4);
        B.act.alloc(K); B.mask0.alloc(std::max(total, 1));
        B.d_bptr.alloc(K); B.d_xptr.alloc(K);
        if (!B.h_active) FVM_HIP(hipHostMalloc((void **)&B.h_active, sizeof(int) * 4, hipHostMallocDefault));
    }
    FVM_HIP(hipMemcpyAsync(B.d_bptr.p, d_b, sizeof(double *) * K, hipMemcpyHostToDevice, s));
    FVM_HIP(hipMemcpyAsync(B.d_xptr.p, d_x, sizeof(double *) * K, hipMemcpyHostToDevice, s));
    Level &F = h->lev[0];
    const int *perm0 = R.fast ? F.perm.p : nullptr;
    if (R.fast && B.sysperm.n != (size_t)std::max(total, 1)) {
        // positions of the coloured layout grouped by system (stable: each system's own order)
        B.sysperm.alloc(std::max(total, 1));
        if (total) {
            DArr<int> key(total), key_out(total), idx(total);
            k_batch_syskey<<<grid_for(total, BS), BS, 0, s>>>(total, K, B.d_off0.p, perm0, key.p, idx.p);
            int bits = 1;
            while ((1 << bits) < K) bits++;
            size_t bytes = 0;
            FVM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key.p, key_out.p, idx.p, B.sysperm.p, total,
                                                       0, bits, s));
            DArr<unsigned char> tmp(bytes + 16);
            FVM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, key.p, key_out.p, idx.p, B.sysperm.p, total, 0,
                                                       bits, s));
            FVM_HIP(hipStreamSynchronize(s));
        }
    }
    if (total)
        k_batch_gather<<<grid_for(total, BS), BS, 0, s>>>(total, K, B.d_off0.p, perm0, B.d_bptr.p, B.d_xptr.p,
                                                          F.b.p, F.x.p);
    R.mask0 = B.mask0.p;
    batch_norms(R, B);
    k_batch_init<<<1, BS, 0, s>>>(K, B.d_off0.p, B.norm.p, h->prm.absoluteTolerance, B.r0.p, B.rn.p, B.act.p,
                                  B.cyc.p, B.counters.p);
    if (total) k_batch_mask<<<grid_for(total, BS), BS, 0, s>>>(total, K, B.d_off0.p, perm0, B.act.p, B.mask0.p);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipMemcpyAsync(B.h_active, B.counters.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    FVM_HIP(hipStreamSynchronize(s));
    int active = *B.h_active;
    // One iteration (cycle, residual, norms, stop test, masks) is the same launch
    // sequence every time: captured once into a graph and replayed. The first
    // iteration runs uncaptured, so that one-time preparations (which allocate)
    // happen outside the capture.
    const bool saved_ktime = h->ktime;
    for (int i = 1; i < h->prm.nMaxIterations && active > 0; i++) {
        if (B.use_graph && i > 1) {
            if (!B.graph_ok) {
                h->ktime = false;
                hipGraph_t g = nullptr;
                FVM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                Run Rc = R;
                try {
                    batch_iteration(Rc, B);
                } catch (...) {
                    (void)hipStreamEndCapture(s, &g);
                    if (g) (void)hipGraphDestroy(g);
                    h->ktime = saved_ktime;
                    throw;
                }
                FVM_HIP(hipStreamEndCapture(s, &g));
                h->ktime = saved_ktime;
                if (B.graph) { (void)hipGraphExecDestroy(B.graph); B.graph = nullptr; }
                hipError_t e = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                FVM_HIP(e);
                B.graph_ok = true;
            }
            FVM_HIP(hipGraphLaunch(B.graph, s));
        } else {
            batch_iteration(R, B);
        }
        h->total_iterations++;
        FVM_HIP(hipMemcpyAsync(B.h_active, B.counters.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        FVM_HIP(hipStreamSynchronize(s));
        active = *B.h_active;
    }
    if (total)
        k_batch_scatter<<<grid_for(total, BS), BS, 0, s>>>(total, K, B.d_off0.p, perm0, F.x.p, B.d_xptr.p);
    FVM_HIP(hipGetLastError());
    if (rnorm0) FVM_HIP(hipMemcpyAsync(rnorm0, B.r0.p, sizeof(double) * K, hipMemcpyDeviceToHost, s));
    if (rnorm_final) FVM_HIP(hipMemcpyAsync(rnorm_final, B.rn.p, sizeof(double) * K, hipMemcpyDeviceToHost, s));
    if (cycles) FVM_HIP(hipMemcpyAsync(cycles, B.cyc.p, sizeof(int) * K, hipMemcpyDeviceToHost, s));
    FVM_HIP(hipStreamSynchronize(s));
    flow_check(h);
}


// ===========================================================================
// FP64 flexible GMRES(m), right-preconditioned by one AMG cycle (FP64 or mixed).
// Works on finest-level vectors in the active layout of the AMG handle.
// Textbook algorithm: Saad, "Iterative Methods for Sparse Linear Systems",
// 2nd ed., Alg. 9.6 (FGMRES), with classical Gram-Schmidt applied twice.
// ===========================================================================
namespace {

template <class V, class D>
__global__ void __launch_bounds__(512)
k_sell_spmv(int n, const int *__restrict__ soff, const int *__restrict__ scol, const V *__restrict__ sval,
            const D *__restrict__ diag, const double *__restrict__ x, double *__restrict__ y, int xcd) {
    const int p = (xcd ? xcd_block() : (int)blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = p >> 6, lane = p & 63;
    const int o = soff[s];
    const int w = (soff[s + 1] - o) >> 6;
    const double acc = (double)diag[p] * x[p];
    y[p] = sell_row<SELL_U, true>(acc, scol, sval, o, w, lane, x);
}

__global__ void k_csr_spmv(int n, const int *__restrict__ row, const int *__restrict__ col,
                           const double *__restrict__ off, const double *__restrict__ diag,
                           const double *__restrict__ x, double *__restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = diag[i] * x[i];
        for (int p = row[i]; p < row[i + 1]; p++) s += off[p] * x[col[p]];
        y[i] = s;
    }
}

// partial dots of up to 8 basis vectors V_j (stride n) with w; nslot >= 0: also
// the partial sums of w.w into slot nslot
__global__ void __launch_bounds__(BS) k_multidot8(int n, const double *__restrict__ V, int j0, int k,
                                                  const double *__restrict__ w, double *part, int nblk,
                                                  int nslot) {
    double acc[8];
    double ww = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = 0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double wi = w[i];
        ww += wi * wi;
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (q < k) acc[q] += V[(size_t)(j0 + q) * n + i] * wi;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (q < k) {
            const double t = block_sum<BS>(acc[q]);
            if (threadIdx.x == 0) part[(size_t)(j0 + q) * nblk + blockIdx.x] = t;
        }
    }
    if (nslot >= 0) {
        const double t = block_sum<BS>(ww);
        if (threadIdx.x == 0) part[(size_t)nslot * nblk + blockIdx.x] = t;
    }
}

// First classical Gram-Schmidt update and the dots of the second pass in one sweep
// over the basis: w' = w - sum_j c_j V_j, then the partial sums of V_j . w' into
// slots j (the second pass reads the updated w element by element, so its dots
// need no sweep of their own). Up to KMAX basis vectors.
template <int KMAX>
__global__ void __launch_bounds__(BS) k_axpy_dots(int n, const double *__restrict__ V, int k,
                                                  const double *__restrict__ c, double *w, double *part,
                                                  int nblk) {
    double cj[KMAX], acc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        cj[j] = j < k ? c[j] : 0.0;
        acc[j] = 0.0;
    }
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        double vj[KMAX];
        double s = 0;
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
            vj[j] = j < k ? V[(size_t)j * n + i] : 0.0;
            if (j < k) s += cj[j] * vj[j];
        }
        const double v = w[i] - s;
        w[i] = v;
#pragma unroll
        for (int j = 0; j < KMAX; j++)
            if (j < k) acc[j] += vj[j] * v;
    }
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j < k) {
            const double t = block_sum<BS>(acc[j]);
            if (threadIdx.x == 0) part[(size_t)j * nblk + blockIdx.x] = t;
        }
    }
}

// out[b] = sum of the partials of slot b (b < k), out[k] = those of slot nslot
__global__ void __launch_bounds__(BS) k_multi_final(int nblk, const double *part, double *out, int k, int nslot) {
    const int j = (int)blockIdx.x < k ? (int)blockIdx.x : nslot;
    double s = 0;
    for (int i = threadIdx.x; i < nblk; i += BS) s += part[(size_t)j * nblk + i];
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// w += sign * sum_j c_j V_j, and the partial sums of the new w.w into slot nslot
__global__ void __launch_bounds__(BS) k_multi_axpy_norm(int n, const double *__restrict__ V, int k,
                                                        const double *__restrict__ c, double sign, double *w,
                                                        double *part, int nblk, int nslot) {
    double ww = 0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        double s = 0;
#pragma unroll 4
        for (int j = 0; j < k; j++) s += c[j] * V[(size_t)j * n + i];
        const double v = w[i] + sign * s;
        w[i] = v;
        ww += v * v;
    }
    const double t = block_sum<BS>(ww);
    if (threadIdx.x == 0) part[(size_t)nslot * nblk + blockIdx.x] = t;
}

// w += sign * sum_j c_j V_j
__global__ void k_multi_axpy(int n, const double *__restrict__ V, int k, const double *__restrict__ c,
                             double sign, double *w) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0;
        for (int j = 0; j < k; j++) s += c[j] * V[(size_t)j * n + i];
        w[i] += sign * s;
    }
}

__global__ void k_scale_into(int n, const double *a, double alpha, double *b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = alpha * a[i];
}

__global__ void k_neg_into(int n, const double *a, double *b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = -a[i];
}

__global__ void __launch_bounds__(BS) k_norm2sq_partial(int n, const double *__restrict__ r, double *part) {
    double s = 0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) s += r[i] * r[i];
    s = block_sum<BS>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

}  // namespace

// blocks of the Krylov basis kernels (dots, updates): enough resident wavefronts
// per CU to keep HBM busy with their independent streams
constexpr unsigned KRYLOV_BLOCKS = 2048;

struct KrylovWork {
    int n = 0, m = 0;
    DArr<double> V, Z, w, b, x, r, part, dev_h;
};

namespace {

void spmv64(Run &R, const double *x, double *y) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    const int n = F.n;
    if (n == 0) return;
    struct T {
        fvm_amg *h; hipEvent_t e;
        long long n, nnz;
        ~T() { if (e) timing_end(h, KT_RESID64, e, pass_bytes(V64, 8, 2, n, nnz)); }   // x read, y written
    } timer{h, timing_begin(h), n, F.nnz};
    if (R.fast)
        k_sell_spmv<<<grid_for(n, h->spmv_block), h->spmv_block, 0, h->stream>>>(n, F.sell_off.p, F.sell_col.p, F.sell_val.p,
                                                           F.pdiag.p, x, y, h->spmv_xcd);
    else
        k_csr_spmv<<<grid_for(n, BS), BS, 0, h->stream>>>(n, F.row.p, F.col.p, F.off.p, F.diag.p, x, y);
}

// z = M(v) ~= A^{-1} v : one cycle from zero on A z + (-v) = 0
void precond(Run &R, const double *v, double *z) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    const int n = F.n;
    R.r0_valid = false;
    if (!R.mixed) {
        k_neg_into<<<grid_for(n, BS), BS, 0, s>>>(n, v, F.b.p);
        cycle<double>(R, h->prm.cycleType, 0, true);
        FVM_HIP(hipMemcpyAsync(z, F.x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    } else {
        k_d2f_neg<<<grid_for(n, BS), BS, 0, s>>>(n, v, F.b32.p);
        cycle<float>(R, h->prm.cycleType, 0, true);
        k_f2d<<<grid_for(n, BS), BS, 0, s>>>(n, F.x32.p, z);
    }
    FVM_HIP(hipGetLastError());
}

double reduce_to_host(fvm_amg *h, unsigned g, int slot) {
    k_final_sum<<<1, BS, 0, h->stream>>>(g, h->ws.partials.p, scratch_scalar(h, slot));
    return read_scalar(scratch_scalar(h, slot), h->stream);
}

double norm1_of(Run &R, const double *r, int n) {
    fvm_amg *h = R.h;
    const unsigned g = grid_for(n, BS, NORM_BLOCKS);
    k_norm1_partial<double><<<g, BS, 0, h->stream>>>(n, r, h->ws.partials.p);
    return reduce_to_host(h, g, 1);
}

double norm2_of(Run &R, const double *r, int n) {
    fvm_amg *h = R.h;
    const unsigned g = grid_for(n, BS, NORM_BLOCKS);
    k_norm2sq_partial<<<g, BS, 0, h->stream>>>(n, r, h->ws.partials.p);
    return std::sqrt(reduce_to_host(h, g, 2));
}

// out[0..k) = V_{0..k}^T w, on the device
void multidot_dev(Run &R, KrylovWork &W, int k, const double *w, double *out) {
    fvm_amg *h = R.h;
    hipStream_t s = h->stream;
    const int n = W.n;
    const unsigned g = grid_for(n, BS, KRYLOV_BLOCKS);
    for (int j0 = 0; j0 < k; j0 += 8)
        k_multidot8<<<g, BS, 0, s>>>(n, W.V.p, j0, std::min(8, k - j0), w, W.part.p, (int)g, -1);
    k_multi_final<<<k, BS, 0, s>>>((int)g, W.part.p, out, k, 0);
    FVM_HIP(hipGetLastError());
}

// w -= sum_j c_j V_j (c on the device), and out[0..k) = V_j . (the new w)
void axpy_dots_dev(Run &R, KrylovWork &W, int k, const double *c, double *w, double *out) {
    fvm_amg *h = R.h;
    hipStream_t s = h->stream;
    const unsigned g = grid_for(W.n, BS, KRYLOV_BLOCKS);
    if (k <= 8) k_axpy_dots<8><<<g, BS, 0, s>>>(W.n, W.V.p, k, c, w, W.part.p, (int)g);
    else if (k <= 16) k_axpy_dots<16><<<g, BS, 0, s>>>(W.n, W.V.p, k, c, w, W.part.p, (int)g);
    else k_axpy_dots<32><<<g, BS, 0, s>>>(W.n, W.V.p, k, c, w, W.part.p, (int)g);
    k_multi_final<<<k, BS, 0, s>>>((int)g, W.part.p, out, k, 0);
    FVM_HIP(hipGetLastError());
}

// w -= sum_j c_j V_j over the first k basis vectors (c on the device); ww != null:
// the new w.w into *ww (device)
void axpy_dev(Run &R, KrylovWork &W, int k, const double *c, double *w, double *ww) {
    fvm_amg *h = R.h;
    hipStream_t s = h->stream;
    const unsigned g = grid_for(W.n, BS, KRYLOV_BLOCKS);
    const int nslot = W.m + 1;
    k_multi_axpy_norm<<<g, BS, 0, s>>>(W.n, W.V.p, k, c, -1.0, w, W.part.p, (int)g, nslot);
    if (ww) k_multi_final<<<1, BS, 0, s>>>((int)g, W.part.p, ww, 0, nslot);
    FVM_HIP(hipGetLastError());
}

void multi_axpy(Run &R, KrylovWork &W, const double *basis, int k, const std::vector<double> &c, double sign,
                double *w) {
    fvm_amg *h = R.h;
    hipStream_t s = h->stream;
    FVM_HIP(hipMemcpyAsync(W.dev_h.p, c.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    k_multi_axpy<<<grid_for(W.n, BS), BS, 0, s>>>(W.n, basis, k, W.dev_h.p, sign, w);
    FVM_HIP(hipGetLastError());
}

}  // namespace

void fgmres_solve(fvm_amg *h, KrylovWork &W, const fvm_fgmres_params &p, const double *d_b, double *d_x,
                  double *rnorm0, double *rnorm_final, int *iterations, std::vector<double> &hist,
                  const LearnedHook &nn) {
    Run R = make_run(h);
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    const int n = F.n;
    const int m = std::max(1, p.restart);
    if (W.n != n || W.m != m) {
        PoolScope no_pool(nullptr);   // the Krylov basis may outlive the preconditioner's pool
        W.n = n; W.m = m;
        W.V.alloc((size_t)(m + 1) * std::max(n, 1));
        W.Z.alloc((size_t)m * std::max(n, 1));
        W.w.alloc(std::max(n, 1)); W.b.alloc(std::max(n, 1)); W.x.alloc(std::max(n, 1));
        W.r.alloc(std::max(n, 1));
        W.part.alloc((size_t)(m + 2) * KRYLOV_BLOCKS);
        W.dev_h.alloc(2 * (size_t)(m + 1) + 1);
    }
    const int *perm = R.fast ? F.perm.p : nullptr;
    k_gather<<<grid_for(n, BS), BS, 0, s>>>(n, perm, d_b, W.b.p);
    k_gather<<<grid_for(n, BS), BS, 0, s>>>(n, perm, d_x, W.x.p);
    auto residual_into_r = [&]() {
        // r = b + A x, with the residual kernel of the FP64 finest level
        Run R64 = R;
        R64.mixed = false;
        residual<double>(R64, 0, W.x.p, W.b.p, W.r.p);
    };
    residual_into_r();
    const double r0 = norm1_of(R, W.r.p, n);
    FVM_CHECK(std::isfinite(r0), FVM_E_NUMERIC, "FGMRES: non-finite initial residual");
    hist.assign(1, r0);
    double rn = r0;
    int total = 0;
    // learned correction of the preconditioner: on until the residual stalls
    bool nn_on = nn.m != nullptr && n > 0;
    if (nn_on) mpnn_bind(nn.m, F.n, F.row.p, F.col.p, F.off.p, F.diag.p);
    if (nn.used) *nn.used = 0;
    if (nn.fallback_at) *nn.fallback_at = -1;
    std::vector<double> est_hist;
    if (!(r0 < p.absoluteTolerance) && n > 0) {
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), hc, hbuf, y;
        for (;;) {
            const double beta = norm2_of(R, W.r.p, n);
            FVM_CHECK(std::isfinite(beta), FVM_E_NUMERIC, "FGMRES: non-finite residual at a restart");
            if (!(beta > 0)) break;
            // The stop rule is on the 1-norm of the true residual, the Arnoldi
            // estimate is a 2-norm: aim the estimate at the reduction the 1-norm
            // still needs (rn -> tol * r0), with a margin for the ratio of the two
            // norms drifting, so that a restart rarely ends just short of it.
            const double goal2 = beta * (p.relativeTolerance * r0 / rn) * 0.5;
            k_scale_into<<<grid_for(n, BS), BS, 0, s>>>(n, W.r.p, -1.0 / beta, W.V.p);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int jend = 0;
            for (int j = 0; j < m; j++) {
                double *vj = W.V.p + (size_t)j * n;
                double *zj = W.Z.p + (size_t)j * n;
                precond(R, vj, zj);
                if (nn_on) {
                    // z += NN(A z - v): the network's estimate of the remaining error
                    spmv64(R, zj, W.w.p);
                    int bad = 0;
                    mpnn_correct(nn.m, s, perm, W.w.p, vj, zj, n, &bad);
                    if (bad) {   // non-finite output: plain preconditioner from here on
                        nn_on = false;
                        if (nn.fallback_at) *nn.fallback_at = total;
                        precond(R, vj, zj);
                    } else if (nn.used) {
                        (*nn.used)++;
                    }
                }
                spmv64(R, zj, W.w.p);
                // classical Gram-Schmidt applied twice ("twice is enough": Giraud,
                // Langou, Rozloznik, Comput. Math. Appl. 50 (2005)); coefficients stay
                // on the device, the norm of w comes from the last pass, and the
                // host reads them once per iteration
                double *h1 = W.dev_h.p, *h2 = W.dev_h.p + (m + 1), *hw = W.dev_h.p + 2 * (m + 1);
                multidot_dev(R, W, j + 1, W.w.p, h1);
                if (j + 1 <= 32) {
                    axpy_dots_dev(R, W, j + 1, h1, W.w.p, h2);   // first update + second dots
                } else {
                    axpy_dev(R, W, j + 1, h1, W.w.p, nullptr);
                    multidot_dev(R, W, j + 1, W.w.p, h2);
                }
                axpy_dev(R, W, j + 1, h2, W.w.p, hw);
                hbuf.resize(2 * (size_t)(m + 1) + 1);
                FVM_HIP(hipMemcpyAsync(hbuf.data(), W.dev_h.p, sizeof(double) * hbuf.size(), hipMemcpyDeviceToHost, s));
                FVM_HIP(hipStreamSynchronize(s));
                hc.assign(j + 1, 0.0);
                for (int q = 0; q <= j; q++) hc[q] = hbuf[q] + hbuf[(size_t)(m + 1) + q];
                const double ww1 = hbuf[2 * (size_t)(m + 1)];
                const double hn = std::sqrt(ww1);
                for (int q = 0; q <= j; q++) H[(size_t)q * m + j] = hc[q];
                H[(size_t)(j + 1) * m + j] = hn;
                if (hn > 0)
                    k_scale_into<<<grid_for(n, BS), BS, 0, s>>>(n, W.w.p, 1.0 / hn, W.V.p + (size_t)(j + 1) * n);
                for (int q = 0; q < j; q++) {
                    const double a = H[(size_t)q * m + j], b = H[(size_t)(q + 1) * m + j];
                    H[(size_t)q * m + j] = cs[q] * a + sn[q] * b;
                    H[(size_t)(q + 1) * m + j] = -sn[q] * a + cs[q] * b;
                }
                const double a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
                const double den = std::hypot(a, b);
                cs[j] = den > 0 ? a / den : 1.0;
                sn[j] = den > 0 ? b / den : 0.0;
                H[(size_t)j * m + j] = den;
                H[(size_t)(j + 1) * m + j] = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                total++;
                jend = j + 1;
                const double est = std::fabs(g[j + 1]);
                if (p.verbosity > 1) printf("fgmres %d: est2 %g\n", total, est);
                est_hist.push_back(est);
                const int win = 5;
                if (nn_on && (int)est_hist.size() > win &&
                    est > nn.stall_ratio * est_hist[est_hist.size() - 1 - win]) {
                    nn_on = false;   // stalled: plain preconditioner for the rest of the solve
                    if (nn.fallback_at) *nn.fallback_at = total;
                }
                if (est <= goal2 || est < p.absoluteTolerance || !(hn > 0) || total >= p.nMaxIterations)
                    break;
            }
            y.assign(jend, 0.0);
            for (int q = jend - 1; q >= 0; q--) {
                double t = g[q];
                for (int c = q + 1; c < jend; c++) t -= H[(size_t)q * m + c] * y[c];
                y[q] = t / H[(size_t)q * m + q];
            }
            multi_axpy(R, W, W.Z.p, jend, y, 1.0, W.x.p);
            residual_into_r();
            rn = norm1_of(R, W.r.p, n);
            FVM_CHECK(std::isfinite(rn), FVM_E_NUMERIC,
                      "FGMRES produced a non-finite residual (iteration " + std::to_string(total) +
                          "): the system holds non-finite or zero-diagonal entries, or the preconditioner diverged");
            hist.push_back(rn);
            if (p.verbosity > 0) printf("%d: %g\n", total, rn);
            double ratio = rn;
            if (r0 != 0) ratio /= r0;
            if (rn < p.absoluteTolerance || ratio < p.relativeTolerance) break;
            if (total >= p.nMaxIterations) break;
        }
    }
    k_scatter<<<grid_for(n, BS), BS, 0, s>>>(n, perm, W.x.p, d_x);
    FVM_HIP(hipStreamSynchronize(s));
    flow_check(h);
    timing_harvest(h);
    if (rnorm0) *rnorm0 = r0;
    if (rnorm_final) *rnorm_final = rn;
    if (iterations) *iterations = total;
}

// ===========================================================================
// BiCGStab around AMG::smooth (BCGStab.cpp:27-171): the reference's own Krylov
// loop, in FP64 on the device, with the handle's cycle as the preconditioner.
// The reference's operations in its order: rho = r.rt; p = (p - omega v) beta + r;
// p^ = smooth(b = p, x = 0); v = A p^; alpha = rho / (rt.v); x -= alpha p^,
// r -= alpha v; ||r||_1 < absTol ends the loop; s^ = smooth(b = r, x = 0);
// t = A s^; omega = (t.r) / (t.t); x -= omega s^, r -= omega t; the 1-norm of r,
// relative to r0, is the stop rule; at the end the true residual is recomputed.
// ===========================================================================
namespace {

__global__ void k_bicg_p(int n, double *p, const double *v, const double *r, double omega, double beta) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = p[i] - omega * v[i];
        t *= beta;
        p[i] = t + r[i];
    }
}

// x -= a * xs, r -= a * rs
__global__ void k_bicg_update(int n, double a, const double *xs, const double *rs, double *x, double *r) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] -= a * xs[i];
        r[i] -= a * rs[i];
    }
}

double dot_of(Run &R, const double *a, const double *b, int n) {
    fvm_amg *h = R.h;
    const unsigned g = grid_for(n, BS, NORM_BLOCKS);
    k_dot_partial<double><<<g, BS, 0, h->stream>>>(n, a, b, h->ws.partials.p);
    return reduce_to_host(h, g, 3);
}

// y = the handle's cycle on A y + b = 0 from y = 0 (AMG::smooth with a zero delta)
void smooth_zero(Run &R, const double *b, double *y) {
    fvm_amg *h = R.h;
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    FVM_HIP(hipMemcpyAsync(F.b.p, b, sizeof(double) * F.n, hipMemcpyDeviceToDevice, s));
    FVM_HIP(hipMemsetAsync(F.x.p, 0, sizeof(double) * F.n, s));
    R.r0_valid = false;
    one_cycle(R);
    FVM_HIP(hipMemcpyAsync(y, F.x.p, sizeof(double) * F.n, hipMemcpyDeviceToDevice, s));
}

}  // namespace

void bcgstab_solve(fvm_amg *h, const fvm_linsolver_params &p, const double *d_b, double *d_x, double *rnorm0,
                   double *rnorm_final, int *iterations, std::vector<double> &hist) {
    Run R = make_run(h);
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    const int n = F.n;
    const size_t n1 = std::max(n, 1);
    const int *perm = R.fast ? F.perm.p : nullptr;
    DArr<double> b(n1), x(n1), r(n1), rt(n1), pv(n1), ph(n1), v(n1), t(n1);
    const unsigned gb = grid_for(n, BS);
    if (n > 0) {
        k_gather<<<gb, BS, 0, s>>>(n, perm, d_b, b.p);
        k_gather<<<gb, BS, 0, s>>>(n, perm, d_x, x.p);
    }
    Run R64 = R;
    R64.mixed = false;
    residual<double>(R64, 0, x.p, b.p, r.p);
    const double r0 = norm1_of(R, r.p, n);
    FVM_CHECK(std::isfinite(r0), FVM_E_NUMERIC, "BCGStab: non-finite initial residual");
    hist.assign(1, r0);
    if (p.verbosity > 0) print_norm(h, "0", r0);
    FVM_HIP(hipMemcpyAsync(rt.p, r.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    double rho = 0, alpha = 0, omega = 0;
    bool have_p = false;
    int it = 0;
    for (int i = 0; i < p.nMaxIterations && n > 0; i++) {
        it = i + 1;
        const double rho_prev = rho;
        rho = dot_of(R, r.p, rt.p, n);
        if (!have_p) {
            FVM_HIP(hipMemcpyAsync(pv.p, r.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            have_p = true;
        } else {
            const double beta = (rho / rho_prev) * (alpha / omega);
            k_bicg_p<<<gb, BS, 0, s>>>(n, pv.p, v.p, r.p, omega, beta);
        }
        smooth_zero(R, pv.p, ph.p);
        spmv64(R, ph.p, v.p);
        alpha = rho / dot_of(R, rt.p, v.p, n);
        k_bicg_update<<<gb, BS, 0, s>>>(n, alpha, ph.p, v.p, x.p, r.p);
        double rn = norm1_of(R, r.p, n);
        FVM_CHECK(std::isfinite(rn), FVM_E_NUMERIC,
                  "BCGStab produced a non-finite residual (iteration " + std::to_string(i + 1) + ")");
        if (rn < p.absoluteTolerance) break;
        smooth_zero(R, r.p, ph.p);   // s^ takes p^'s storage, as in the reference
        spmv64(R, ph.p, t.p);
        const double tdotr = dot_of(R, t.p, r.p, n);
        const double tdott = dot_of(R, t.p, t.p, n);
        omega = tdotr / tdott;
        k_bicg_update<<<gb, BS, 0, s>>>(n, omega, ph.p, t.p, x.p, r.p);
        rn = norm1_of(R, r.p, n);
        FVM_CHECK(std::isfinite(rn), FVM_E_NUMERIC,
                  "BCGStab produced a non-finite residual (iteration " + std::to_string(i + 1) + ")");
        hist.push_back(rn);
        if (p.verbosity > 0) print_norm(h, std::to_string(i + 1), rn);
        if (rn < p.absoluteTolerance || rn / r0 < p.relativeTolerance) break;
    }
    residual<double>(R64, 0, x.p, b.p, r.p);
    const double rf = norm1_of(R, r.p, n);
    if (p.verbosity > 0) print_norm(h, "n", rf);
    if (n > 0) k_scatter<<<gb, BS, 0, s>>>(n, perm, x.p, d_x);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(s));
    flow_check(h);
    if (rnorm0) *rnorm0 = r0;
    if (rnorm_final) *rnorm_final = rf;
    if (iterations) *iterations = it;
}

// ===========================================================================
// Preconditioned CG around AMG::smooth (CG.cpp:25-137), FP64 on the device, in
// the reference's order: z = smooth(b = r, x = 0); rho = r.z; p = z, or
// p = p beta + z with beta = rho / rho_prev; q = A p; alpha = rho / (p.q);
// x -= alpha p, r -= alpha q; the 1-norm of r is checked after every iteration.
// ===========================================================================
namespace {

// p = p * beta + z
__global__ void k_cg_p(int n, double *p, const double *z, double beta) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = p[i] * beta;
        p[i] = t + z[i];
    }
}

}  // namespace

void cg_solve(fvm_amg *h, const fvm_linsolver_params &p, const double *d_b, double *d_x, double *rnorm0,
              double *rnorm_final, int *iterations, std::vector<double> &hist) {
    Run R = make_run(h);
    Level &F = h->lev[0];
    hipStream_t s = h->stream;
    const int n = F.n;
    const size_t n1 = std::max(n, 1);
    const int *perm = R.fast ? F.perm.p : nullptr;
    DArr<double> b(n1), x(n1), r(n1), z(n1), pv(n1), q(n1);
    const unsigned gb = grid_for(n, BS);
    if (n > 0) {
        k_gather<<<gb, BS, 0, s>>>(n, perm, d_b, b.p);
        k_gather<<<gb, BS, 0, s>>>(n, perm, d_x, x.p);
    }
    Run R64 = R;
    R64.mixed = false;
    residual<double>(R64, 0, x.p, b.p, r.p);
    const double r0 = norm1_of(R, r.p, n);
    FVM_CHECK(std::isfinite(r0), FVM_E_NUMERIC, "CG: non-finite initial residual");
    hist.assign(1, r0);
    if (p.verbosity > 0) print_norm(h, "0", r0);
    double rho = 0;
    bool have_p = false;
    int it = 0;
    for (int i = 0; i < p.nMaxIterations && n > 0; i++) {
        it = i + 1;
        smooth_zero(R, r.p, z.p);
        const double rho_prev = rho;
        rho = dot_of(R, r.p, z.p, n);
        if (!have_p) {
            FVM_HIP(hipMemcpyAsync(pv.p, z.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            have_p = true;
        } else {
            k_cg_p<<<gb, BS, 0, s>>>(n, pv.p, z.p, rho / rho_prev);
        }
        spmv64(R, pv.p, q.p);
        const double alpha = rho / dot_of(R, pv.p, q.p, n);
        k_bicg_update<<<gb, BS, 0, s>>>(n, alpha, pv.p, q.p, x.p, r.p);
        const double rn = norm1_of(R, r.p, n);
        FVM_CHECK(std::isfinite(rn), FVM_E_NUMERIC,
                  "CG produced a non-finite residual (iteration " + std::to_string(i + 1) + ")");
        hist.push_back(rn);
        if (p.verbosity > 0) print_norm(h, std::to_string(i + 1), rn);
        if (rn < p.absoluteTolerance || rn / r0 < p.relativeTolerance) break;
    }
    residual<double>(R64, 0, x.p, b.p, r.p);
    const double rf = norm1_of(R, r.p, n);
    if (p.verbosity > 0) print_norm(h, "n", rf);
    if (n > 0) k_scatter<<<gb, BS, 0, s>>>(n, perm, x.p, d_x);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(s));
    flow_check(h);
    if (rnorm0) *rnorm0 = r0;
    if (rnorm_final) *rnorm_final = rf;
    if (iterations) *iterations = it;
}

}  // namespace fvm

namespace fvm {
KrylovWork *new_krylov_work() { return new KrylovWork(); }
void delete_krylov_work(KrylovWork *w) { delete w; }
}  // namespace fvm
