// AMG hierarchy construction on the GPU.
//
// Reference semantics (src/fvm/src/modules/fvmbase/):
//   coarsening            CRMatrix.h:468-586   (greedy pairwise agglomeration)
//   coarse-to-fine map    MultiFieldMatrix.cpp:627-667
//   coarse connectivity   CRMatrix.h:597-691   (first-appearance column order)
//   Galerkin operator     CRMatrix.h:699-758   (fixed summation order)
//   level loop/stop rule  AMG.cpp:149-210
//
// The greedy agglomeration is sequential in the reference (row nr's decision
// reads the state left by rows < nr). Here it runs in rounds: a row decides
// as soon as every row that could still change its inputs has decided, which
// yields exactly the sequential result (see k_coarsen_round).
#include "amg.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <climits>
#include <cmath>
#include <cstdlib>

namespace fvm {

namespace {

constexpr int BS = 256;
constexpr int MEMW = 5;  // members per group: at most groupSize + 3 for groupSize 2

__device__ __forceinline__ bool isb(const uint8_t *bnd, int i) { return bnd && bnd[i]; }

// ------------------------------------------------------------------ matrix signature
// Order-independent 64-bit signature of an array: the sum over its elements of a
// mix of (element bits, salted index) with the splitmix64 finaliser (a bijection,
// so changing one element changes the sum). Used to recognise a setup of the
// same matrix; not a defence against crafted collisions.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
template <class T>
__global__ void k_signature(long long m, const T *a, unsigned long long salt, unsigned long long *out) {
    __shared__ unsigned long long s_acc[BS / 64];
    unsigned long long acc = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x) {
        unsigned long long v = 0;
        if (sizeof(T) == 8) v = *(const unsigned long long *)(a + i);
        else if (sizeof(T) == 4) v = *(const unsigned int *)(a + i);
        else v = *(const unsigned char *)(a + i);
        acc += mix64(v ^ mix64((unsigned long long)i + salt));
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) s_acc[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < BS / 64; w++) t += s_acc[w];
        atomicAdd(out, t);
    }
}

// ------------------------------------------------------------------ input checks
__global__ void k_validate(int n_self, int n_total, long long nnz, const int *row, const int *col,
                           const double *diag, int *flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_total; i += gridDim.x * blockDim.x) {
        const int b = row[i], e = row[i + 1];
        if (b < 0 || e < b || (long long)e > nnz) { atomicOr(flag, 1); continue; }
        if (i >= n_self) continue;
        for (int p = b; p < e; p++) {
            const int c = col[p];
            if (c < 0 || c >= n_total || c == i) atomicOr(flag, 2);
        }
        const double d = diag[i];
        if (!(d != 0.0) || !isfinite(d)) atomicOr(flag, 4);
    }
}

__global__ void k_compact_count(int n_self, const int *row, const int *col, int *cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_self; i += gridDim.x * blockDim.x) {
        int k = 0;
        for (int p = row[i]; p < row[i + 1]; p++) k += (col[p] < n_self);
        cnt[i] = k;
    }
}

__global__ void k_compact_fill(int n_self, const int *row, const int *col, const double *off,
                               const int *nrow, int *ncol, double *noff) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_self; i += gridDim.x * blockDim.x) {
        int q = nrow[i];
        for (int p = row[i]; p < row[i + 1]; p++) {
            const int c = col[p];
            if (c < n_self) { ncol[q] = c; noff[q] = off[p]; q++; }
        }
    }
}

// ------------------------------------------------------------------ coarsening
// ------------------------------------------------------------------ coarsening
//
// Exact parallel form of CRMatrix::createCoarsening for groupSize 2
// (CRMatrix.h:485-583). In the sequential loop, when row r's turn comes every
// non-boundary row < r is grouped, and a row c > r is grouped iff an earlier
// seed s < r with c in row(s) paired it. With rowT(c) = {s : c in row(s)}, the
// decision of r can only depend on the rows of
//   D(r) = {s < r : s in rowT(r)}              (s could pair r)
//        + {c < r : c in row(r)}               (r needs c's group)
//        + {s < r : s in rowT(c), c in row(r), c > r}   (s could pair c first)
// counted as a multiset. cnt[r] starts at |D(r)| and is decremented when a row
// of D(r) decides; r enters the frontier when it reaches 0 (Kahn's algorithm
// on this DAG). One round decides the whole frontier; all its inputs were
// decided in earlier rounds (state stamped with a round < t), so rows of one
// round never read what another writes, and the result equals the sequential
// one. A row that would join the group g of its strongest grouped neighbour
// also needs every s < r adjacent (through rowT) to a member of g decided
// (s could still join g and change its count); if not, it retries next round.
// Boundary rows never decide and are dropped from all dependency sets.

constexpr int INIT_SUB = 4;    // lanes per row in k_cnt_init
constexpr int ROUND_SUB = 64;  // lanes per frontier row in k_coarsen_round (one wavefront)
constexpr int THIN_SUB = 16;   // the same for levels of at most 10 entries per row
constexpr int SMALL_SUB = 64;  // lanes per frontier row in k_coarsen_small (one wavefront)

// append x to the queue when want; must be reached by every active lane
__device__ __forceinline__ void push_agg(bool want, int x, int *q, int *qcnt) {
    const unsigned long long m = __ballot(want);
    if (!m) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(qcnt, __popcll(m));
    base = __shfl(base, leader, 64);
    if (want) q[base + __popcll(m & ((1ull << lane) - 1ull))] = x;
}

// Integer atomics of the agglomeration rounds. WG = false: device scope (rounds
// spread over the whole chip); WG = true: workgroup scope, for the rounds that one
// workgroup runs alone (every access to the counters comes from its own
// wavefronts, so the atomics need not be coherent beyond its CU).
template <bool WG>
__device__ __forceinline__ int atom_add(int *p, int v) {
    if (WG) return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return atomicAdd(p, v);
}

// Rows that became ready, buffered per lane and appended to the queue with one
// atomic per wavefront (the queue tail is a single hot counter).
constexpr int PB = 4;
constexpr int FANB = 4;   // dependents per lane whose decrements are in flight together

struct PushBuf {
    int v[PB];
    int n;
    int *lb, *ln;   // the block's queue in LDS and its size (nullptr: append to q directly)
    int lcap;       // entries the LDS queue holds
};

template <bool WG>
__device__ __forceinline__ void pb_flush(PushBuf &pb, int *q, int *qcnt) {
    unsigned long long m[PB];
    int before[PB];
    int tot = 0;
#pragma unroll
    for (int k = 0; k < PB; k++) {
        m[k] = __ballot(pb.n > k);
        before[k] = tot;
        tot += __popcll(m[k]);
    }
    if (tot == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)__ballot(1)) - 1;
    int base = 0;
    if (pb.lb) {
        // the block's LDS queue, appended to q once per block; what does not fit
        // goes to q at once
        if (lane == leader) base = __hip_atomic_fetch_add(pb.ln, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        base = __shfl(base, leader, 64);
        const int lo = max(base, pb.lcap);   // first slot past the LDS queue
        const int spill = base + tot - lo;
        int gb = 0;
        if (spill > 0) {
            if (lane == leader) gb = atom_add<WG>(qcnt, spill);
            gb = __shfl(gb, leader, 64);
        }
        const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
        for (int k = 0; k < PB; k++)
            if (pb.n > k) {
                const int idx = base + before[k] + __popcll(m[k] & lt);
                if (idx < pb.lcap) pb.lb[idx] = pb.v[k];
                else q[gb + idx - lo] = pb.v[k];
            }
        pb.n = 0;
        return;
    }
    if (lane == leader) base = atom_add<WG>(qcnt, tot);
    base = __shfl(base, leader, 64);
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < PB; k++)
        if (pb.n > k) q[base + before[k] + __popcll(m[k] & lt)] = pb.v[k];
    pb.n = 0;
}

// must be reached by every active lane
template <bool WG>
__device__ __forceinline__ void pb_push(PushBuf &pb, bool want, int x, int *q, int *qcnt) {
    if (__ballot(pb.n == PB)) pb_flush<WG>(pb, q, qcnt);
#pragma unroll
    for (int k = 0; k < PB; k++)
        if (want && pb.n == k) pb.v[k] = x;
    pb.n += want ? 1 : 0;
}

// Waiter lists of the agglomeration rounds: a blocked join waits in the list of
// the row it waits for (RowRec::wh, linked through wnext) instead of returning to
// the frontier every round; the row closes its list when it decides and puts the
// waiters on the next frontier. A closed list (the row decided in this round)
// sends the waiter to the next frontier itself, so no wake-up is lost.
// Every access to a head or a link is an atomic read-modify-write, and a waiter
// awaits its link's exchange before it moves the head. A waker that finds the
// waiter in the list therefore reads the link with an exchange of its own after
// the waiter's was performed, with no release or acquire fence (which would write
// back or invalidate a whole L2 cache).
constexpr int WCLOSED = -2;
constexpr int WTAKEN = -3;   // a link already read by its waker

template <bool WG>
__device__ __forceinline__ int atom_xchg(int *p, int v) {
    if (WG) return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Row r joins the waiters of row b (one lane). False: b's list is closed (b has
// decided in this round) or there are no lists, and r must go to the next frontier.
template <bool WG>
__device__ __forceinline__ bool park(int r, int *wh, int *wnext) {
    if (!wnext) return false;
    int h = WG ? __hip_atomic_load(wh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
               : __hip_atomic_load(wh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (h != WCLOSED) {
        const int o = atom_xchg<WG>(&wnext[r], h);
        asm volatile("" ::"v"(o) : "memory");   // the link is performed before the head moves
        const bool ok = WG ? __hip_atomic_compare_exchange_strong(wh, &h, r, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_WORKGROUP)
                           : __hip_atomic_compare_exchange_strong(wh, &h, r, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT);
        if (ok) return true;
    }
    return false;
}

// Flags of the entry p = (d, c) of row d (one byte each):
// MAYPAIR: d may pair with c. It is 0 when d can never pair with c.
// d pairs its strongest ungrouped neighbour c only if w(d,c) > theta * (strongest
// grouped weight). If some neighbour e != c has w(d,e) > w(d,c) and
// theta * w(d,e) >= w(d,c), then whenever c is the strongest ungrouped
// neighbour, e (stronger) is grouped, the strongest grouped weight is w(d,e) =
// the row's maximum W1, and the test fails: d joins or starts a group instead.
// That is, MAYPAIR = !(W1 > w && theta * W1 >= w), with w computed exactly as
// the decision computes it.
// IGNORES: d's decision is the same whether c is grouped or not. That holds when
// d's strongest lower neighbour (grouped by the time d decides, as every lower
// row is) has weight L1 > w(d,c) and theta * L1 >= w(d,c): then c is never the
// strongest grouped neighbour, and were c the strongest ungrouped one the pair
// test would fail, so d joins or starts a group either way, with the same target.
// A row waits on a lower seed s because s could pair a higher neighbour c of the
// row first; that dependency is dropped when s may not pair c or the row ignores
// c. Rows still wait for every lower neighbour. This shortens the rounds'
// dependency chains; the decisions do not change.
// STRONG (entries (d, c) with c < d): c is one of d's strongest lower neighbours.
// Every lower row is grouped by the time d decides (the sequential loop has
// processed it), so d reads of a lower neighbour c only its group, and only when
// c is d's strongest grouped neighbour, which a lower neighbour weaker than the
// strongest lower one never is. So d waits for a lower neighbour c only if c may
// pair d or c is STRONG; the decision counts every lower neighbour as grouped, and
// a join whose target is still undecided retries in the next round.
constexpr uint8_t MAYPAIR = 1, IGNORES = 2, STRONG = 4;
// MP lanes per row (a row's entries spread over them).
constexpr int MP = 8;
__global__ void k_may_pair(int n, const int *row, const int *col, const double *off, const double *diag,
                           const uint8_t *bnd, double theta, int lower, uint8_t *mayp) {
    const int lane = threadIdx.x & (MP - 1);
    const int sub = (blockIdx.x * blockDim.x + threadIdx.x) / MP, nsub = (gridDim.x * blockDim.x) / MP;
    for (int d = sub; d < n; d += nsub) {   // uniform over each sub-group
        const int rb = row[d], re = row[d + 1];
        const bool bd = isb(bnd, d);        // never decides: no pair at all
        const double d0 = fabs(diag[d]);
        double w1 = -1.0, l1 = -1.0;        // strongest neighbour, strongest lower neighbour
        for (int p = rb + lane; p < re; p += MP) {
            const int c = col[p];
            if (isb(bnd, c)) continue;
            const double d1 = fabs(diag[c]);
            const double w = fabs(fabs(off[p]) / (d0 < d1 ? d1 : d0));
            if (w > w1) w1 = w;
            if (c < d && w > l1) l1 = w;
        }
        for (int o = MP / 2; o > 0; o >>= 1) {
            const double w2 = __shfl_xor(w1, o, MP), l2 = __shfl_xor(l1, o, MP);
            if (w2 > w1) w1 = w2;
            if (l2 > l1) l1 = l2;
        }
        for (int p = rb + lane; p < re; p += MP) {
            const int c = col[p];
            if (bd || isb(bnd, c)) { mayp[p] = 0; continue; }
            const double d1 = fabs(diag[c]);
            const double w = fabs(fabs(off[p]) / (d0 < d1 ? d1 : d0));
            mayp[p] = ((w1 > w && theta * w1 >= w) ? 0 : MAYPAIR) | ((l1 > w && theta * l1 >= w) ? IGNORES : 0) |
                      ((c < d && (!lower || !(l1 > w))) ? STRONG : 0);
        }
    }
}

// The flags of the transpose pattern: mayT[q] = may_pair of the entry (colT[q], c)
// of row colT[q], for every entry q of row c of the transpose pattern. Its
// position comes from tposT (non-symmetric patterns) or, for a structurally
// symmetric pattern (the transpose pattern is the pattern), from a scan of that
// row (1 if absent, which keeps the dependency).
__global__ void k_may_pair_t(int n, const int *row, const int *col, const int *rowT, const int *colT,
                             const int *tposT, const uint8_t *mayp, uint8_t *mayT) {
    const int lane = threadIdx.x & (MP - 1);
    const int sub = (blockIdx.x * blockDim.x + threadIdx.x) / MP, nsub = (gridDim.x * blockDim.x) / MP;
    for (int c = sub; c < n; c += nsub)
        for (int q = rowT[c] + lane; q < rowT[c + 1]; q += MP) {
            int pos = -1;
            if (tposT) {
                pos = tposT[q];
            } else {
                const int s = colT[q];
                for (int p = row[s]; p < row[s + 1]; p++)
                    if (col[p] == c) { pos = p; break; }
            }
            mayT[q] = pos < 0 ? (MAYPAIR | STRONG) : mayp[pos];   // absent: keep the dependencies
        }
}

__global__ void k_cnt_init(int n, const int *row, const int *col, const int *rowT, const int *colT,
                           int sym, const uint8_t *bnd, const uint8_t *mayp, const uint8_t *mayT, int *cnt,
                           int *dround, int *ldr, int *memr, int *mcnt, int *q, int *qcnt) {
    const long long gt = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int r = (int)(gt / INIT_SUB);
    const int lane = threadIdx.x & (INIT_SUB - 1);
    const bool valid = r < n;
    int k = 0;
    bool b = true;
    if (valid) {
        b = isb(bnd, r);
        if (!b) {
            // A: s < r may pair r (entry (s, r), the transposed flag)
            if (!sym)
                for (int qq = rowT[r] + lane; qq < rowT[r + 1]; qq += INIT_SUB) {
                    const int s = colT[qq];
                    k += (s < r && !isb(bnd, s) && (!mayT || (mayT[qq] & MAYPAIR)));
                }
            for (int p = row[r] + lane; p < row[r + 1]; p += INIT_SUB) {
                const int c = col[p];
                if (isb(bnd, c)) continue;
                if (c < r) {
                    // B: r may need c's group (c STRONG); symmetric patterns: A as well (c may
                    // pair r: entry (c, r), the transposed flag of the same index)
                    if (!mayp) k += sym ? 2 : 1;
                    else k += ((mayp[p] & STRONG) ? 1 : 0) + (sym && (mayT[p] & MAYPAIR) ? 1 : 0);
                } else if (!mayp || !(mayp[p] & IGNORES)) {
                    // s could pair c first (unless s may not pair c)
                    for (int qq = rowT[c]; qq < rowT[c + 1]; qq++) {
                        const int s = colT[qq];
                        k += (s < r && !isb(bnd, s) && (!mayT || (mayT[qq] & MAYPAIR)));
                    }
                }
            }
        }
    }
    for (int o = INIT_SUB / 2; o > 0; o >>= 1) k += __shfl_xor(k, o, 64);
    const bool head = valid && lane == 0;
    if (head) {
        cnt[r] = k;
        ldr[r] = -1;
        dround[r] = b ? -1 : INT_MAX;
        mcnt[r] = 0;
        for (int z = 0; z < MEMW; z++) memr[(size_t)r * MEMW + z] = INT_MAX;
    }
    push_agg(head && !b && k == 0, r, q, qcnt);
}

// Mask of the lanes of this lane's sub-group of S lanes within the wavefront.
template <int S>
__device__ __forceinline__ unsigned long long sub_mask() {
    if (S == 64) return ~0ull;
    return ((1ull << (S & 63)) - 1ull) << (__lane_id() & ~(S - 1));
}

// Walk the concatenation of the per-lane ranges [beg, beg + len) of the S lanes
// of a sub-group, S elements per step: f(owner, idx) runs on every lane, with the
// lane that owns element idx (idx = -1 past the end). The ranges are flattened
// with a prefix sum across lanes, so a set of short lists costs a few
// lane-parallel steps instead of one chain per list.
template <int S, class F>
__device__ __forceinline__ void walk_ranges(int beg, int len, F &&f) {
    const int lane = threadIdx.x & (S - 1);
    int inc = len;   // inclusive prefix over the S lanes
    for (int o = 1; o < S; o <<= 1) {
        const int v = __shfl_up(inc, o, S);
        if (lane >= o) inc += v;
    }
    const int total = __shfl(inc, S - 1, S);
    const int exc = inc - len;
    for (int k0 = 0; k0 < total; k0 += S) {
        const int k = k0 + lane;
        // element k belongs to the lane j with exc_j <= k < inc_j
        int lo = 0, hi = S - 1;
        for (int it = 1; it < S; it <<= 1) {
            const int mid = (lo + hi) >> 1;
            const int v = __shfl(inc, mid, S);
            if (v > k) hi = mid;
            else lo = mid + 1;
        }
        const int bj = __shfl(beg, lo, S);
        const int ej = __shfl(exc, lo, S);
        f(lo, k < total ? bj + (k - ej) : -1);
    }
}

// Cross-lane steps of a decision without the LDS crossbar: DPP moves inside rows
// of 16 lanes (gfx9 encodings: pairs within a quad, quads within a half-row and
// half-rows within a row by mirroring; for a max/min every lane ends with its
// group's result whatever the pairing), readlane across the rows of a wavefront.
constexpr int DPP_QUAD_XOR1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141;  // lane i <-> 7 - i within 8
constexpr int DPP_MIRROR = 0x140;       // lane i <-> 15 - i within 16

template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ int rl_int(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double rl_dbl(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)rl_int((int)b, l), hi = (unsigned)rl_int((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Candidate neighbour of a decision: weight and CSR position (INT_MAX: none).
struct Cand {
    double w;
    int p;
};
// b replaces a when b exists and a does not, or b is stronger, or as strong and
// first in CSR order (the sequential loop's strict '>' keeps the first maximum)
__device__ __forceinline__ Cand better(Cand a, Cand b) {
    return (b.p != INT_MAX && (a.p == INT_MAX || b.w > a.w || (b.w == a.w && b.p < a.p))) ? b : a;
}
template <int CTRL>
__device__ __forceinline__ Cand dpp_cand(Cand c) {
    const long long b = __double_as_longlong(c.w);
    const unsigned lo = (unsigned)dpp_mov<CTRL>((int)b), hi = (unsigned)dpp_mov<CTRL>((int)(b >> 32));
    Cand o;
    o.w = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    o.p = dpp_mov<CTRL>(c.p);
    return o;
}
// The best candidate of the S lanes of each sub-group, on all of them.
template <int S>
__device__ __forceinline__ Cand reduce_cand(Cand c) {
    if (S >= 2) c = better(c, dpp_cand<DPP_QUAD_XOR1>(c));
    if (S >= 4) c = better(c, dpp_cand<DPP_QUAD_XOR2>(c));
    if (S >= 8) c = better(c, dpp_cand<DPP_HALF_MIRROR>(c));
    if (S >= 16) c = better(c, dpp_cand<DPP_MIRROR>(c));
    if (S == 32) {
        Cand o;
        o.w = __shfl_xor(c.w, 16, 64);
        o.p = __shfl_xor(c.p, 16, 64);
        c = better(c, o);
    }
    if (S == 64) {   // every row holds its result: combine the four rows' results
        Cand r0{rl_dbl(c.w, 0), rl_int(c.p, 0)}, r1{rl_dbl(c.w, 16), rl_int(c.p, 16)};
        Cand r2{rl_dbl(c.w, 32), rl_int(c.p, 32)}, r3{rl_dbl(c.w, 48), rl_int(c.p, 48)};
        c = better(better(r0, r1), better(r2, r3));
    }
    return c;
}
// v of lane l (0..S-1) of this lane's sub-group
template <int S>
__device__ __forceinline__ int sub_get(int v, int l) {
    if (S == 64) return rl_int(v, l);
    return __shfl(v, l, S);
}

// Visit the rows adjacent (through the transpose pattern) to the members of a
// group, WB * S at a time: lane j < MEMW of the sub-group holds member j's range
// [tb, tb + tl) of colT (tl = 0: no member). The ranges are gathered to every
// lane at once and each lane finds its elements' owners by comparisons, so a
// step costs one cross-lane latency instead of a scan and a search; f gets WB
// element indices per lane (-1: none), whose loads it can issue together.
constexpr int JWB = 4;
template <int S, int WB, class F>
__device__ __forceinline__ void walk_members(int tb, int tl, F &&f) {
    const int lane = threadIdx.x & (S - 1);
    int mb[MEMW], me[MEMW];
    int total = 0;
#pragma unroll
    for (int j = 0; j < MEMW; j++) {
        mb[j] = sub_get<S>(tb, j) - total;   // element k of member j is colT[mb[j] + k]
        total += sub_get<S>(tl, j);
        me[j] = total;
    }
    for (int k0 = 0; k0 < total; k0 += WB * S) {
        int idx[WB];
#pragma unroll
        for (int w = 0; w < WB; w++) {
            const int k = k0 + w * S + lane;
            int id = -1;
#pragma unroll
            for (int j = MEMW - 1; j >= 0; j--)
                if (k < me[j]) id = mb[j] + k;
            idx[w] = k < total ? id : -1;
        }
        f(idx);
    }
}

// Dependents lists: for every row d, the rows x whose D-set holds d, one entry
// per decrement of cnt[x] that the decision of d makes (bit 31 set: a decrement
// by 2, the A and B parts of a symmetric pattern being the same set). They
// depend on the pattern only, so they are built once per level; a decision then
// finds its dependents with one contiguous load instead of walking row, rowT
// and colT. The parts, as counted from the other side in k_cnt_init:
//   A: x in row(d), x > d                       (d in rowT(x))
//   B: x in rowT(d), x > d, non-symmetric only  (d in row(x))
//   C: c in row(d), c > d: x in rowT(c), d < x < c
// Boundary rows neither decide nor wait, and are in no list.

// Dependents lists with S lanes per row: FILL false counts the entries of each row,
// FILL true writes them (any order inside a row is fine: decrements commute) as
// contiguous runs. The (c, x) pairs of the C part are flattened over the lanes, and
// each step's kept entries are compacted with a ballot, so one store instruction
// writes one run.
template <int S, bool FILL>
__global__ void k_dep_fill(int n, const int *row, const int *col, const int *rowT, const int *colT, int sym,
                           const uint8_t *bnd, const uint8_t *mayp, const uint8_t *mayT, const int *dptr,
                           int *out) {
    const int lane = threadIdx.x & (S - 1);
    const int sub = (blockIdx.x * blockDim.x + threadIdx.x) / S;
    const int nsub = (gridDim.x * blockDim.x) / S;
    const unsigned long long smask = sub_mask<S>();
    const unsigned long long below = (1ull << __lane_id()) - 1ull;
    const unsigned two = sym ? 0x80000000u : 0u;
    for (int d = sub; d < n; d += nsub) {
        if (isb(bnd, d)) {   // the same for every lane of the sub-group
            if (!FILL && lane == 0) out[d] = 0;
            continue;
        }
        int pos = FILL ? dptr[d] : 0;   // count pass (FILL false): out[d] = the number of entries
        const int rb = row[d], re = row[d + 1];
        for (int b0 = rb; b0 < re; b0 += S) {
            const int p = b0 + lane;
            int c = -1, tb = 0, tl = 0;
            bool a_ok = false;
            unsigned a_two = two;
            if (p < re) {
                c = col[p];
                if (c > d && !isb(bnd, c)) {
                    tb = rowT[c];
                    tl = (!mayp || (mayp[p] & MAYPAIR)) ? rowT[c + 1] - tb : 0;   // C part only if d may pair c
                    if (!mayp) {
                        a_ok = true;
                    } else {
                        // A: d may pair c; symmetric patterns: B as well (c needs d's group:
                        // entry (c, d), the transposed flag of the same index)
                        const int a = (mayp[p] & MAYPAIR) ? 1 : 0;
                        const int b = (sym && (mayT[p] & STRONG)) ? 1 : 0;
                        a_ok = a + b > 0;
                        a_two = a + b == 2 ? 0x80000000u : 0u;
                    }
                }
            }
            // A: x in row(d), x > d
            const unsigned long long ma = __ballot(a_ok) & smask;
            if (FILL && a_ok) out[pos + __popcll(ma & below)] = (int)((unsigned)c | a_two);
            pos += __popcll(ma);
            // C: x in rowT(c) with d < x < c
            walk_ranges<S>(tb, tl, [&](int owner, int idx) {
                const int cj = __shfl(c, owner, S);
                bool w = false;
                int x = -1;
                if (idx >= 0) {
                    x = colT[idx];
                    // (and unless x's decision ignores c: flag of the entry (x, c))
                    w = x > d && x < cj && !isb(bnd, x) && !(mayT && (mayT[idx] & IGNORES));
                }
                const unsigned long long mc = __ballot(w) & smask;
                if (FILL && w) out[pos + __popcll(mc & below)] = x;
                pos += __popcll(mc);
            });
        }
        // B: x in rowT(d), x > d (non-symmetric patterns)
        if (!sym) {
            const int tb = rowT[d], te = rowT[d + 1];
            for (int b0 = tb; b0 < te; b0 += S) {
                const int p = b0 + lane;
                bool w = false;
                int x = -1;
                if (p < te) {
                    x = colT[p];
                    // x needs d's group only if d is one of x's strongest lower neighbours
                    w = x > d && !isb(bnd, x) && (!mayT || (mayT[p] & STRONG));
                }
                const unsigned long long mb = __ballot(w) & smask;
                if (FILL && w) out[pos + __popcll(mb & below)] = x;
                pos += __popcll(mb);
            }
        }
        if (!FILL && lane == 0) out[d] = pos;
    }
}

// Per-row state of the agglomeration rounds in one 32-byte record, so that reading
// a neighbour (its diagonal, state, group and dependents range) touches one cache
// line instead of four: the decisions of the large levels are bound by such random
// line fetches.
struct alignas(32) RowRec {
    double ad;      // |diagonal|
    int dround;     // round the row was grouped in (INT_MAX: not yet, -1: boundary)
    int ldr;        // leader (seed) of its group, -1 while ungrouped
    int db, de;     // range of its dependents list
    int blk;        // the undecided row b a blocked join of this row waits for: b for the
                    // row whose group it joins, -2 - b for a row that could still join it
                    // (-1: none)
    int wh;         // first row waiting for this one to decide (-1: none, WCLOSED: decided)
};

__global__ void k_rec_pack(int n, const double *diag, const int *dptr, const int *dround, const int *ldr,
                           RowRec *rec) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        RowRec r;
        r.ad = fabs(diag[i]);
        r.dround = dround[i];
        r.ldr = ldr[i];
        r.db = dptr[i];
        r.de = dptr[i + 1];
        r.blk = -1;
        r.wh = -1;
        rec[i] = r;
    }
}

__global__ void k_rec_unpack(int n, const RowRec *rec, int *dround, int *ldr) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        dround[i] = rec[i].dround;
        ldr[i] = rec[i].ldr;
    }
}

// The next waiter after x (read by lane 0 of the sub-group, given to all its lanes).
template <int S, bool WG>
__device__ __forceinline__ int next_waiter(int x, int *wnext) {
    int nx = 0;
    if ((threadIdx.x & (S - 1)) == 0) nx = atom_xchg<WG>(&wnext[x], WTAKEN);
    return sub_get<S>(nx, 0);
}

// Decide one frontier row with the S lanes of its sub-group. Returns via the
// queue (fan-out) and the shared state arrays. Loads that a later step may need
// (the first dependents of r, the dependents range of every neighbour, the group
// of every neighbour) go out with the loads that the decision itself needs, so
// a decision costs few round trips to memory.
template <int S, bool WG>
__device__ void decide_row(int r, int t, double theta, const int *row, const int *col, const int *rowT,
                           const int *colT, const double *off, const uint8_t *bnd, RowRec *rec,
                           const int *dep, int *mem, int *memr, int *mcnt, int *cnt, int *q, int *qcnt,
                           int *nretry, int *wnext, bool wall, int *lb, int *ln, int lcap) {
    const int lane = threadIdx.x & (S - 1);
    // state, row range, dependents range and diagonal of r in one step
    const RowRec rr = rec[r];
    const int rb = row[r], re = row[r + 1];
    const int dr = rr.dround, db = rr.db, de = rr.de;
    const double d0 = rr.ad;
    if (dr < t) return;   // already paired by an earlier seed (its fan-out is done)
    const int rb0 = rr.blk >= 0 ? rr.blk : -2 - rr.blk;
    if (rr.blk != -1 && !(rec[rb0].dround < t)) {
        // a blocked join whose blocker has not decided yet: waits again (its list,
        // or the next round), without re-evaluating the decision
        bool parked = false;
        if (lane == 0) {
            if (nretry) atom_add<WG>(nretry, 1);   // statistics, when asked for
            parked = (wall || rr.blk >= 0) && park<WG>(r, &rec[rb0].wh, wnext);
        }
        PushBuf pb;
        pb.n = 0;
        pb.lb = lb;
        pb.ln = ln;
        pb.lcap = lcap;
        pb_push<WG>(pb, lane == 0 && !parked, r, q, qcnt);
        pb_flush<WG>(pb, q, qcnt);
        return;
    }
    const int pre = db + lane < de ? dep[db + lane] : 0;   // first S dependents of r
    // r's waiter list is closed now, while the neighbours load, and opened again
    // with the same waiters if r turns out blocked; rows that find it closed in
    // between go to the next frontier themselves
    int w0 = -1, w1 = -1;
    if (wnext && lane == 0) w0 = atom_xchg<WG>(&rec[r].wh, WCLOSED);
    // strongest ungrouped (u) and grouped (g) neighbours; ties: first in CSR order
    Cand bu{-1.0, INT_MAX}, bg{-1.0, INT_MAX};
    int cu = -1, cg = -1, ub = 0, ue = 0, lg = -1, rg = -1;
    for (int p = rb + lane; p < re; p += S) {
        const int c = col[p];
        const double o = off[p];
        if (isb(bnd, c)) continue;
        const RowRec rc = rec[c];
        const double d1 = rc.ad;
        const int dc = rc.dround;
        const int cb = rc.db, ce = rc.de, lc = rc.ldr;
        const double w = fabs(fabs(o) / (d0 < d1 ? d1 : d0));
        // grouped: every lower row (the sequential loop has processed it), and a
        // higher one paired in an earlier round
        if (!(c < r) && !(dc < t)) {
            if (bu.p == INT_MAX || w > bu.w) { bu.w = w; bu.p = p; cu = c; ub = cb; ue = ce; }
        } else {
            if (bg.p == INT_MAX || w > bg.w) { bg.w = w; bg.p = p; cg = c; lg = lc; rg = dc; }
        }
    }
    // the sub-group's best of each kind; the lane that holds it (position p is
    // taken by lane (p - rb) mod S) hands out the neighbour's data
    bu = reduce_cand<S>(bu);
    bg = reduce_cand<S>(bg);
    const double mu = bu.w, mg = bg.w;
    if (bu.p != INT_MAX) {
        const int l = (bu.p - rb) & (S - 1);
        cu = sub_get<S>(cu, l); ub = sub_get<S>(ub, l); ue = sub_get<S>(ue, l);
    } else {
        cu = -1;
    }
    if (bg.p != INT_MAX) {
        const int l = (bg.p - rb) & (S - 1);
        cg = sub_get<S>(cg, l); lg = sub_get<S>(lg, l); rg = sub_get<S>(rg, l);
    } else {
        cg = -1;
    }
    int action;  // 0 pair, 1 new singleton group, 2 join cg's group, 3 retry next round
    int g = -1, slot = 0, blocker = -1;
    bool longw = false;   // blocked by the row whose group r joins (not one that could join it)
    if (cu != -1 && (cg == -1 || mu > theta * mg)) {
        action = 0;
    } else if (cg == -1) {
        action = 1;
    } else if (!(rg < t)) {
        action = 3;   // cg (a lower row) has not decided yet: its group is not known
        blocker = cg;
        longw = true;
    } else {
        g = lg;   // ldr[cg], final since cg was grouped in an earlier round
        // members of g placed before round t (one lane each), and every s < r
        // adjacent to one of them through rowT: r must wait while such an s is
        // undecided (it could still join g and change its count)
        int tb = 0, tl = 0;
        bool is_mem = false;
        if (lane == 0) slot = mcnt[g];
        if (lane < MEMW) {
            const size_t e = (size_t)g * MEMW + lane;
            const int mr = memr[e], m = mem[e];   // both at once; m is used only for a placed member
            if (mr < t) {
                is_mem = true;
                tb = rowT[m];
                tl = rowT[m + 1] - tb;
            }
        }
        const int count = __popcll(__ballot(is_mem) & sub_mask<S>());
        bool blocked = false;
        int bs = -1;
        // (a boundary row never decides and holds round -1, so it never blocks)
        walk_members<S, JWB>(tb, tl, [&](const int (&idx)[JWB]) {
            int sv[JWB], dv[JWB];
#pragma unroll
            for (int w = 0; w < JWB; w++) sv[w] = idx[w] >= 0 ? colT[idx[w]] : -1;
#pragma unroll
            for (int w = 0; w < JWB; w++) dv[w] = (sv[w] >= 0 && sv[w] < r) ? rec[sv[w]].dround : -1;
#pragma unroll
            for (int w = 0; w < JWB; w++)
                if (!(dv[w] < t)) {
                    blocked = true;
                    bs = sv[w];
                }
        });
        // every lane of the sub-group takes part in the vote; the first blocked lane
        // names the row to wait for
        const unsigned long long bm = __ballot(blocked) & sub_mask<S>();
        if (bm) blocker = sub_get<S>(bs, (__ffsll((long long)bm) - 1) & (S - 1));
        blocked = bm != 0;
        action = blocked ? 3 : ((count > 2 + 2) ? 1 : 2);   // coarseCount > groupSize+2
    }
    PushBuf pb;
    pb.n = 0;
    pb.lb = lb;
    pb.ln = ln;
    pb.lcap = lcap;
    if (action == 3) {
        bool parked = false;
        if (lane == 0) {
            if (nretry) atom_add<WG>(nretry, 1);   // statistics: blocked joins
            rec[r].blk = longw ? blocker : -2 - blocker;
            if (wnext) atom_xchg<WG>(&rec[r].wh, w0);   // open r's list again
            // a row that could still join the group mostly decides within a round or
            // two: waiting in its list pays off only with wall (retries are cheaper)
            parked = (wall || longw) && park<WG>(r, &rec[blocker].wh, wnext);
        }
        pb_push<WG>(pb, lane == 0 && !parked, r, q, qcnt);
        pb_flush<WG>(pb, q, qcnt);
        return;
    }
    // the partner's list is closed before its state is written
    if (wnext && lane == 0 && action == 0) w1 = atom_xchg<WG>(&rec[cu].wh, WCLOSED);
    if (lane == 0) {
        if (action == 0) {
            rec[r].ldr = r; rec[cu].ldr = r;
            mem[(size_t)r * MEMW] = r; memr[(size_t)r * MEMW] = t;
            mem[(size_t)r * MEMW + 1] = cu; memr[(size_t)r * MEMW + 1] = t;
            mcnt[r] = 2;
            rec[cu].dround = t; rec[r].dround = t;
        } else if (action == 1) {
            rec[r].ldr = r;
            mem[(size_t)r * MEMW] = r; memr[(size_t)r * MEMW] = t;
            mcnt[r] = 1;
            rec[r].dround = t;
        } else {
            mem[(size_t)g * MEMW + slot] = r; memr[(size_t)g * MEMW + slot] = t;
            mcnt[g] = slot + 1;
            rec[r].ldr = g;
            rec[r].dround = t;
        }
    }
    // the waiters of r and of its partner are walked after the fan-out
    // fan-out: the dependents of r and, for a pair, of its partner. A lane issues
    // the decrements of up to FANB dependents before it waits for any, so a long
    // list costs one atomic round trip per FANB * S entries instead of per S.
    const int l0 = de - db, lt = l0 + (action == 0 ? ue - ub : 0);
    for (int base = 0; base < lt; base += FANB * S) {
        int x[FANB], dec[FANB], old[FANB];
#pragma unroll
        for (int u = 0; u < FANB; u++) {
            const int k = base + u * S + lane;
            x[u] = -1;
            dec[u] = 0;
            if (k < lt) {
                const int e = k < l0 ? (k < S ? pre : dep[db + k]) : dep[ub + (k - l0)];
                x[u] = e & 0x7fffffff;
                dec[u] = 1 + (int)((unsigned)e >> 31);
            }
        }
#pragma unroll
        for (int u = 0; u < FANB; u++) old[u] = x[u] >= 0 ? atom_add<WG>(&cnt[x[u]], -dec[u]) : 0;
#pragma unroll
        for (int u = 0; u < FANB; u++) pb_push<WG>(pb, x[u] >= 0 && old[u] == dec[u], x[u], q, qcnt);
    }
    if (wnext) {
        w0 = sub_get<S>(w0, 0);
        w1 = sub_get<S>(w1, 0);
        for (int x = w0; x >= 0; x = next_waiter<S, WG>(x, wnext)) pb_push<WG>(pb, lane == 0, x, q, qcnt);
        for (int x = w1; x >= 0; x = next_waiter<S, WG>(x, wnext)) pb_push<WG>(pb, lane == 0, x, q, qcnt);
    }
    pb_flush<WG>(pb, q, qcnt);
}

// One round over the frontier (multi-block), one wavefront per frontier row.
// Reads frontier t from cur, appends frontier t+1 to nxt; zeroes the counter of round t+2.
template <int S>
__global__ void __launch_bounds__(1024)
k_coarsen_round(int t, int n, const int *row, const int *col, const int *rowT, const int *colT,
                const double *off, const uint8_t *bnd, RowRec *rec, const int *dep, double theta, int *mem,
                int *memr, int *mcnt, int *cnt, const int *cur, const int *cur_cnt, int *nxt, int *nxt_cnt,
                int *zero_cnt, int *nretry, int *wnext, int wall, int lcap) {
    extern __shared__ int s_lq[];   // lcap entries
    __shared__ int s_ln, s_base, s_nf;
    // one sub-group of S lanes per frontier row
    const int sub = (blockIdx.x * blockDim.x + threadIdx.x) / S;
    const int nsub = (gridDim.x * blockDim.x) / S;
    // this sub-group's first entry, loaded with the frontier size (the list holds n
    // entries; an entry past the frontier is not used)
    const int r0 = sub < n ? cur[sub] : 0;
    if (threadIdx.x == 0) {
        s_ln = 0;
        // the frontier size, loaded once per block: every wavefront loading the
        // same word queues thousands of requests on one L2 channel
        s_nf = *cur_cnt;
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) *zero_cnt = 0;
    const int nf = s_nf;
    // the loop bound is uniform over each wavefront's sub-groups only up to the
    // last pass: sub-groups past the frontier simply skip the row
    for (int base = sub - (sub % (64 / S)); base < nf; base += nsub) {
        const int i = base + (sub % (64 / S));
        if (i < nf)
            decide_row<S, false>(i == sub ? r0 : cur[i], t, theta, row, col, rowT, colT, off, bnd, rec, dep,
                                 mem, memr, mcnt, cnt, nxt, nxt_cnt, nretry, wnext, wall != 0, s_lq, &s_ln, lcap);
    }
    // the block's queue entries go to the next frontier with one atomic
    __syncthreads();
    const int c = min(s_ln, lcap);
    if (threadIdx.x == 0) s_base = c > 0 ? atomicAdd(nxt_cnt, c) : 0;
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += blockDim.x) nxt[s_base + i] = s_lq[i];
}

// All rounds of a level in one workgroup: rounds are separated by workgroup
// barriers instead of kernel launches. The barrier orders the workgroup's own
// global stores and loads (all its wavefronts share one CU), so no device-scope
// fence is needed, and every atomic is a workgroup-scope one: the frontier
// sizes are LDS counters, the dependency counters stay in global memory.
// S lanes per frontier row (64 / S rows per wavefront).
template <int S>
__global__ void __launch_bounds__(1024)
k_coarsen_small(int n, int max_rounds, const int *row, const int *col, const int *rowT, const int *colT,
                const double *off, const uint8_t *bnd, RowRec *rec, const int *dep, double theta, int *mem,
                int *memr, int *mcnt, int *cnt, int *qa, int *qb, const int *qcnt0, int *rounds_out,
                int *nretry, int *wnext, int wall) {
    __shared__ int s_q[2];   // s_q[t & 1]: size of the frontier of round t
    int *cur = qa, *nxt = qb;
    if (threadIdx.x == 0) {
        s_q[1] = *qcnt0;     // frontier of round 1, counted by the initialisation kernel
        s_q[0] = 0;
    }
    __syncthreads();
    int t = 1;
    for (; t <= max_rounds; t++) {
        const int nf = s_q[t & 1];
        // every row reachable has decided; nf > n cannot happen (each row is queued
        // at most once per round) and only guards the reads of the list
        if (nf == 0 || nf > n) break;
        const int per_wave = 64 / S;
        const int wave = threadIdx.x >> 6;
        const int nwaves = blockDim.x >> 6;
        for (int base = wave * per_wave; base < nf; base += nwaves * per_wave) {
            const int i = base + (__lane_id() / S);
            if (i < nf)
                decide_row<S, true>(cur[i], t, theta, row, col, rowT, colT, off, bnd, rec, dep, mem, memr, mcnt,
                                    cnt, nxt, &s_q[(t + 1) & 1], nretry, wnext, wall != 0, nullptr, nullptr, 0);
        }
        __syncthreads();
        if (threadIdx.x == 0) s_q[t & 1] = 0;   // every thread has read it: the counter of round t + 2
        __syncthreads();
        int *tmp = cur; cur = nxt; nxt = tmp;
    }
    if (threadIdx.x == 0) *rounds_out = t;
}

// rows that should have decided but have not (0 after a complete agglomeration)
__global__ void k_count_undecided(int n, const uint8_t *bnd, const int *dround, int *out) {
    int c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        c += (!isb(bnd, i) && dround[i] == INT_MAX);
    const unsigned long long m = __ballot(c != 0);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if (__lane_id() == 0 && m) atomicAdd(out, c);
}

// transpose pattern (order inside a row is irrelevant to its users)
__global__ void k_tr_count(int n, const int *row, const int *col, int *cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int p = row[i]; p < row[i + 1]; p++) atomicAdd(&cnt[col[p]], 1);
}

__global__ void k_tr_fill(int n, const int *row, const int *col, const int *rowT, int *fill, int *colT,
                          int *tpos) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int p = row[i]; p < row[i + 1]; p++) {
            const int c = col[p];
            const int q = rowT[c] + atomicAdd(&fill[c], 1);
            colT[q] = i;
            tpos[q] = p;   // the entry (i, c) this transposed entry stands for
        }
}

// flag != 0 if some entry (i,c) has no mirror (c,i)
__global__ void k_sym_check(int n, const int *row, const int *col, int *flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int p = row[i]; p < row[i + 1]; p++) {
            const int c = col[p];
            bool found = false;
            for (int q = row[c]; q < row[c + 1]; q++)
                if (col[q] == i) { found = true; break; }
            if (!found) { atomicOr(flag, 1); return; }
        }
}

__global__ void k_leader_flags(int n, const int *ldr, const uint8_t *bnd, int *flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        flag[i] = (!isb(bnd, i) && ldr[i] == i) ? 1 : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[n] = 0;
}

__global__ void k_assign_ci(int n, const int *ldr, const uint8_t *bnd, const int *gid, int *ci) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        ci[i] = (isb(bnd, i) || ldr[i] < 0) ? -1 : gid[ldr[i]];
}

// ------------------------------------------------------------------ coarse-to-fine
__global__ void k_c2f_count(int n, const int *ci, int *cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (ci[i] >= 0) atomicAdd(&cnt[ci[i]], 1);
}

__global__ void k_c2f_fill(int n, const int *ci, const int *c2f_row, int *fill, int *c2f_col) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int I = ci[i];
        if (I >= 0) c2f_col[c2f_row[I] + atomicAdd(&fill[I], 1)] = i;
    }
}

// sort each segment of v ascending by v, or by key[v] when key is given
__global__ void k_segsort(int nseg, const int *ptr, int *v, const int *key) {
    for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < nseg; I += gridDim.x * blockDim.x) {
        const int b = ptr[I], e = ptr[I + 1];
        for (int a = b + 1; a < e; a++) {
            const int x = v[a];
            const int kx = key ? key[x] : x;
            int j = a - 1;
            while (j >= b && (key ? key[v[j]] : v[j]) > kx) { v[j + 1] = v[j]; j--; }
            v[j + 1] = x;
        }
    }
}

// Sort key of the coloured layout: the colour, then (coarse levels) rows by
// decreasing length inside windows of SIGMA consecutive rows, so that the 64
// rows of a SELL slice have similar lengths while rows stay near their index
// order (locality of the gathers). Rows of one colour never read each other, so
// their order changes no result.
constexpr int SIGMA = 256;
__global__ void k_color_key(int n, const int *color, const int *row, int by_len, int wbits,
                            unsigned long long *key) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long k = (unsigned long long)color[i];
        if (by_len) {
            k = (k << wbits) | (unsigned long long)(i / SIGMA);
            k = (k << 8) | (unsigned long long)(255 - min(row[i + 1] - row[i], 255));
        }
        key[i] = k;
    }
}

// ------------------------------------------------------------------ coarse operator
// Coarse operator with one sub-group of S lanes per coarse row I (the same results as one
// thread per row, bit for bit). The entries of I's fine rows are flattened over the
// lanes in the reference's order: for each fine row (ascending), a diagonal entry and
// then its CSR entries. Distinct coarse columns are kept in order of first appearance
// (CRMatrix.h:616-640); each value is summed in flattened order by broadcasting the
// entries one by one (CRMatrix.h:722-755), so every sum adds the same terms in the
// same order as the sequential loop.
constexpr int GCH = 4;   // chunks of a coarse row's entries held in registers by k_gal

struct GalEntry {
    int J;        // coarse column of an off-diagonal entry (-1: diagonal entry or dropped)
    int isd;      // 1: diagonal entry of a fine row, or an entry inside the aggregate (J == I)
    double v;
};

__device__ __forceinline__ GalEntry gal_entry(int K, int m, int nf, const int *fi, const int *fo,
                                              const int *fb, int I, const int *col, const double *diag,
                                              const double *off, const int *ci) {
    GalEntry e{-1, 0, 0.0};
    if (K >= m) return e;
    // owner fine row: the last t with fo[t] <= K (nf <= MEMW); selected without
    // indexing the arrays at run time, so that they stay in registers
    int ofi = fi[0], ofo = fo[0], ofb = fb[0];
#pragma unroll
    for (int t = 1; t < MEMW; t++)
        if (t < nf && fo[t] <= K) { ofi = fi[t]; ofo = fo[t]; ofb = fb[t]; }
    if (K == ofo) {
        e.isd = 1;
        e.v = diag[ofi];
    } else {
        const int p = ofb + (K - ofo - 1);
        const int J = ci[col[p]];
        e.v = off[p];
        if (J == I) e.isd = 1;
        else if (J >= 0) e.J = J;
    }
    return e;
}

// Flattened entries of each coarse row (fine rows' lengths plus one diagonal each):
// a bound on its distinct columns, used to place the rows of the one-pass product.
__global__ void k_gal_bound(int nc, const int *c2f_row, const int *c2f_col, const int *row, int *ub) {
    for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < nc; I += gridDim.x * blockDim.x) {
        const int q0 = c2f_row[I], nf = min(c2f_row[I + 1] - q0, MEMW);
        int m = 0;
        for (int t = 0; t < nf; t++) {
            const int f = c2f_col[q0 + t];
            m += 1 + row[f + 1] - row[f];
        }
        ub[I] = m;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ub[nc] = 0;
}

// One pass: row I's distinct columns and sums go to tcol/tval from toff[I] on (a
// bound-sized slot), its count to cnt[I] and its diagonal to cdiag[I].
template <int S>
__global__ void k_gal(int nc, const int *c2f_row, const int *c2f_col, const int *row, const int *col,
                      const double *diag, const double *off, const int *ci, int *cnt, const int *toff,
                      int *tcol, double *cdiag, double *tval) {
    const int lane = threadIdx.x & (S - 1);
    const int sub = (blockIdx.x * blockDim.x + threadIdx.x) / S;
    const int nsub = (gridDim.x * blockDim.x) / S;
    for (int I = sub; I < nc; I += nsub) {
        // fine rows of I (at most MEMW: one lane each), their rows and flattened offsets
        const int q0 = c2f_row[I];
        const int nf = min(c2f_row[I + 1] - q0, MEMW);
        int fi_l = 0, fb_l = 0, len_l = 0;
        if (lane < nf) {
            fi_l = c2f_col[q0 + lane];
            fb_l = row[fi_l];
            len_l = row[fi_l + 1] - fb_l;
        }
        int fi[MEMW], fo[MEMW], fb[MEMW];
        int m = 0;
#pragma unroll
        for (int t = 0; t < MEMW; t++) {
            fi[t] = __shfl(fi_l, t, S);
            fb[t] = __shfl(fb_l, t, S);
            const int ln = __shfl(len_l, t, S);
            fo[t] = m;
            if (t < nf) m += 1 + ln;
        }
        int ndist = 0;   // distinct columns in the chunks before this one
        double d = 0.0;
        if (m <= GCH * S) {
            // up to GCH chunks of S entries held in registers: one pass over all the
            // entries, broadcast one by one in flattened order, finds every first
            // appearance and sums every column
            GalEntry me[GCH];
            bool first[GCH];
            double acc[GCH];
#pragma unroll
            for (int c = 0; c < GCH; c++) {
                me[c] = gal_entry(c * S + lane, m, nf, fi, fo, fb, I, col, diag, off, ci);
                first[c] = me[c].J >= 0;
                acc[c] = 0.0;
            }
#pragma unroll
            for (int ce = 0; ce < GCH; ce++) {
                if (ce * S >= m) break;   // uniform over the sub-group
                for (int k = 0; k < S; k++) {
                    // every lane takes part in each shuffle
                    const int Jo = __shfl(me[ce].J, k, S);
                    const double vo = __shfl(me[ce].v, k, S);
                    const int isdo = __shfl(me[ce].isd, k, S);
#pragma unroll
                    for (int c = 0; c < GCH; c++)
                        if (me[c].J >= 0 && Jo == me[c].J) {
                            if (ce * S + k < c * S + lane) first[c] = false;   // an earlier entry
                            acc[c] += vo;
                        }
                    if (lane == 0 && isdo) d += vo;
                }
            }
#pragma unroll
            for (int c = 0; c < GCH; c++) {
                if (c * S >= m) break;
                const unsigned long long bal = __ballot(first[c]) & sub_mask<S>();
                const int below = __popcll(bal & ((1ull << __lane_id()) - 1ull));
                if (first[c]) {
                    const int pos = toff[I] + ndist + below;
                    tcol[pos] = me[c].J;
                    tval[pos] = acc[c];
                }
                ndist += __popcll(bal);
            }
            if (lane == 0) {
                cdiag[I] = d;
                cnt[I] = ndist;
            }
            continue;
        }
        // longer rows: pass over the chunks of S entries, distinct columns in
        // first-appearance order
        for (int c0 = 0; c0 < m; c0 += S) {
            const GalEntry me = gal_entry(c0 + lane, m, nf, fi, fo, fb, I, col, diag, off, ci);
            // first appearance: no earlier entry (earlier chunks, then earlier lanes) with the same J
            bool first = me.J >= 0;
            for (int e0 = 0; e0 < c0; e0 += S) {
                const GalEntry o = gal_entry(e0 + lane, m, nf, fi, fo, fb, I, col, diag, off, ci);
                for (int k = 0; k < S; k++) {
                    const int Jo = __shfl(o.J, k, S);
                    if (Jo >= 0 && Jo == me.J) first = false;
                }
            }
            for (int k = 0; k < S; k++) {
                const int Jo = __shfl(me.J, k, S);
                if (k < lane && Jo >= 0 && Jo == me.J) first = false;
            }
            const unsigned long long bal = __ballot(first) & sub_mask<S>();
            const int below = __popcll(bal & ((1ull << __lane_id()) - 1ull));
            // the value of a first appearance: every entry of its column, in flattened order
            double acc = 0.0;
            for (int e0 = 0; e0 < m; e0 += S) {
                const GalEntry o = gal_entry(e0 + lane, m, nf, fi, fo, fb, I, col, diag, off, ci);
                for (int k = 0; k < S; k++) {
                    // every lane takes part in each shuffle
                    const int Jo = __shfl(o.J, k, S);
                    const double vo = __shfl(o.v, k, S);
                    const int isdo = __shfl(o.isd, k, S);
                    if (first && Jo == me.J) acc += vo;
                    // the diagonal: lane 0, during the first chunk's pass
                    if (c0 == 0 && lane == 0 && isdo) d += vo;
                }
            }
            if (first) {
                const int pos = toff[I] + ndist + below;
                tcol[pos] = me.J;
                tval[pos] = acc;
            }
            ndist += __popcll(bal);
        }
        if (lane == 0) {
            cdiag[I] = d;
            cnt[I] = ndist;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[nc] = 0;
}

// Packs the bound-sized slots of the one-pass product into the coarse CSR.
template <int S>
__global__ void k_gal_compact(int nc, const int *toff, const int *crow, const int *tcol, const double *tval,
                              int *ccol, double *coff) {
    const int lane = threadIdx.x & (S - 1);
    const int sub = (blockIdx.x * blockDim.x + threadIdx.x) / S;
    const int nsub = (gridDim.x * blockDim.x) / S;
    for (int I = sub; I < nc; I += nsub) {
        const int b = toff[I], o = crow[I], c = crow[I + 1] - o;
        for (int k = lane; k < c; k += S) {
            ccol[o + k] = tcol[b + k];
            coff[o + k] = tval[b + k];
        }
    }
}


// ------------------------------------------------------------------ GS level schedule
// Rows i < j are ordered by the sweep whenever j is in row(i) or i is in row(j):
// the later row reads the earlier one's new value, or the earlier row must read
// the later one's old value. Forward sweep (dir 0): edges go from lower to
// higher index; reverse sweep (dir 1): from higher to lower. With a
// non-symmetric pattern both row and rowT are walked (an edge present in both
// is counted twice on both sides, which keeps the in-degree bookkeeping exact).
__global__ void k_sched_init(int n, int dir, const int *row, const int *col, const int *rowT,
                             const int *colT, int *indeg, int *rows, int *tail) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int k = 0;
        for (int p = row[i]; p < row[i + 1]; p++) k += dir ? (col[p] > i) : (col[p] < i);
        if (rowT)
            for (int p = rowT[i]; p < rowT[i + 1]; p++) k += dir ? (colT[p] > i) : (colT[p] < i);
        indeg[i] = k;
        if (k == 0) rows[atomicAdd(tail, 1)] = i;
    }
}

__global__ void k_sched_round(int lvl, int dir, const int *row, const int *col, const int *rowT,
                              const int *colT, int *indeg, int *rows, int *ptr, int *tail,
                              unsigned *done_blocks) {
    const int b = ptr[lvl], e = ptr[lvl + 1];
    for (int idx = b + blockIdx.x * blockDim.x + threadIdx.x; idx < e;
         idx += gridDim.x * blockDim.x) {
        const int i = rows[idx];
        for (int p = row[i]; p < row[i + 1]; p++) {
            const int c = col[p];
            if (dir ? (c < i) : (c > i))
                if (atomicSub(&indeg[c], 1) == 1) rows[atomicAdd(tail, 1)] = c;
        }
        if (rowT)
            for (int p = rowT[i]; p < rowT[i + 1]; p++) {
                const int c = colT[p];
                if (dir ? (c < i) : (c > i))
                    if (atomicSub(&indeg[c], 1) == 1) rows[atomicAdd(tail, 1)] = c;
            }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(done_blocks, 1u);
        if (prev == gridDim.x - 1) {
            ptr[lvl + 2] = atomicAdd(tail, 0);
            *done_blocks = 0u;
        }
    }
}

__global__ void k_set_ptr01(int *ptr, const int *tail) {
    ptr[0] = 0;
    ptr[1] = *tail;
}

// ------------------------------------------------------------------ colouring
__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

// Jones-Plassmann rounds with first-fit colours; a row is coloured in round t
// if its (hash, index) priority beats every neighbour not coloured before t.
// first row of the segment (system of a batch) holding row i; 0 without segments
// index of the segment holding row i; 0 without segments
__device__ __forceinline__ int seg_index(const int *off, int K, int i) {
    if (!off) return 0;
    int lo = 0, hi = K;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int seg_base(const int *off, int K, int i) {
    return off ? off[seg_index(off, K, i)] : 0;
}

// Colouring priority of each row: larger row first (largest-first ordering needs
// fewer colours), then the hash of the row's index inside its segment, so each
// system of a batch gets the colouring it gets alone (neighbours always share the
// segment). Computed once per level, so that a round reads one word per neighbour.
__global__ void k_color_prio(int n, const int *row, const int *seg_off, int nseg, unsigned long long *prio) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int base = seg_base(seg_off, nseg, i);
        prio[i] = ((unsigned long long)(row[i + 1] - row[i]) << 32) | hash32((unsigned)(i - base));
    }
}

// Jones-Plassmann round over the rows still uncoloured (list cur): a row whose
// priority beats every uncoloured neighbour takes the first colour its coloured
// neighbours leave free; the others go to the next list, through a per-block LDS
// queue appended with one atomic (what does not fit is appended at once). The
// decisions of a round read only colours of earlier rounds, so the order of the
// list does not matter.
constexpr int COL_LQ = 2048;

__global__ void k_color_round(int t, const int *row, const int *col, const int *rowT, const int *colT,
                              const unsigned long long *prio, int *color, int *cround, const int *cur,
                              const int *cur_cnt, int *nxt, int *nxt_cnt, int *zero_cnt) {
    __shared__ int s_nf, s_ln, s_base;
    __shared__ int s_lq[COL_LQ];
    if (threadIdx.x == 0) {
        s_nf = *cur_cnt;   // the list size, loaded once per block
        s_ln = 0;
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) *zero_cnt = 0;
    const int nf = s_nf;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += gridDim.x * blockDim.x) {
        const int i = cur[k];
        const unsigned long long pi = prio[i];
        bool best = true;
        unsigned long long used[4] = {0, 0, 0, 0};
        for (int pass = 0; pass < 2 && best; pass++) {
            const int *rp = pass ? rowT : row;
            const int *cp = pass ? colT : col;
            if (!rp) break;
            for (int p = rp[i]; p < rp[i + 1]; p++) {
                const int j = cp[p];
                if (cround[j] < t) {
                    const int cj = color[j];
                    if (cj < 256) used[cj >> 6] |= 1ull << (cj & 63);
                } else {
                    const unsigned long long pj = prio[j];
                    if (pj > pi || (pj == pi && j > i)) { best = false; break; }
                }
            }
        }
        if (!best) {
            const int q = atomicAdd(&s_ln, 1);
            if (q < COL_LQ) s_lq[q] = i;
            else nxt[atomicAdd(nxt_cnt, 1)] = i;
            continue;
        }
        int c = 0;
        while (c < 256 && ((used[c >> 6] >> (c & 63)) & 1ull)) c++;
        color[i] = c;
        cround[i] = t;
    }
    __syncthreads();
    const int c = min(s_ln, COL_LQ);
    if (threadIdx.x == 0) s_base = c > 0 ? atomicAdd(nxt_cnt, c) : 0;
    __syncthreads();
    for (int k = threadIdx.x; k < c; k += blockDim.x) nxt[s_base + k] = s_lq[k];
}

// Two-colouring by breadth-first parity (a graph without odd cycles, e.g. the
// cell graph of a hexahedral mesh: red-black order). Seeds: the first row of
// every segment (system of a batch), colour 0, so each system gets the colouring
// it gets alone. A row takes the colour opposite to the frontier row that
// reaches it first; the check below rejects the colouring if any edge joins
// equal colours (odd cycle, or a row no seed reached).
__global__ void k_bfs_seed(int n, const int *seg_off, int nseg, int *color, int *q, int *qcnt) {
    const int ns = seg_off ? nseg : 1;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        const int r = seg_off ? seg_off[s] : 0;
        const int e = seg_off ? seg_off[s + 1] : n;
        if (r < e) {
            color[r] = 0;
            q[atomicAdd(qcnt, 1)] = r;
        }
    }
}

constexpr int BFS_LQ = 2048;   // entries of a breadth-first round's per-block LDS queue

__global__ void k_bfs_round(const int *row, const int *col, const int *rowT, const int *colT, int *color,
                            const int *cur, const int *curcnt, int *nxt, int *nxtcnt, int *zero_cnt) {
    // the frontier size, loaded once per block; the rows reached go to an LDS queue
    // that the block appends to the next frontier with one atomic (what does not
    // fit is appended at once)
    __shared__ int s_nf, s_ln, s_base;
    __shared__ int s_lq[BFS_LQ];
    if (threadIdx.x == 0) {
        s_nf = *curcnt;
        s_ln = 0;
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) *zero_cnt = 0;
    const int nf = s_nf;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nf; t += gridDim.x * blockDim.x) {
        const int i = cur[t];
        const int want = 1 - color[i];
        for (int pass = 0; pass < 2; pass++) {
            const int *rp = pass ? rowT : row;
            const int *cp = pass ? colT : col;
            if (!rp) break;
            for (int p = rp[i]; p < rp[i + 1]; p++) {
                const int j = cp[p];
                if (atomicCAS(&color[j], -1, want) == -1) {
                    const int k = atomicAdd(&s_ln, 1);
                    if (k < BFS_LQ) s_lq[k] = j;
                    else nxt[atomicAdd(nxtcnt, 1)] = j;
                }
            }
        }
    }
    __syncthreads();
    const int c = min(s_ln, BFS_LQ);
    if (threadIdx.x == 0) s_base = c > 0 ? atomicAdd(nxtcnt, c) : 0;
    __syncthreads();
    for (int k = threadIdx.x; k < c; k += blockDim.x) nxt[s_base + k] = s_lq[k];
}

// bad[s] != 0: segment s is not two-coloured (each system of a batch decides alone)
__global__ void k_bipartite_check(int n, const int *row, const int *col, const int *color, const int *seg_off,
                                  int nseg, int *bad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int ci = color[i];
        bool ok = ci >= 0;
        for (int p = row[i]; p < row[i + 1] && ok; p++) ok = color[col[p]] != ci;
        if (!ok) atomicOr(&bad[seg_index(seg_off, nseg, i)], 1);
    }
}

// rows of two-coloured segments take their breadth-first colour, the others keep
// the greedy one (no edge joins two segments, so the union is a valid colouring)
__global__ void k_merge_colours(int n, const int *seg_off, int nseg, const int *bad, const int *two, int *color) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (!bad[seg_index(seg_off, nseg, i)]) color[i] = two[i];
}

__global__ void k_fill_int(int n, int *v, int val) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = val;
}

__global__ void k_iota(int n, int *v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}

__global__ void k_invert_perm(int n, const int *perm, int *iperm) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        iperm[perm[p]] = p;
}

// colour histogram: per-block counts in LDS, one global atomic per colour and block
// rows and off-diagonal entries of each colour
__global__ void k_color_hist(int n, const int *color, const int *row, int *hist, unsigned long long *hnnz) {
    __shared__ int h[257];
    __shared__ unsigned long long z[257];
    for (int c = threadIdx.x; c < 257; c += blockDim.x) { h[c] = 0; z[c] = 0; }
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        atomicAdd(&h[color[i]], 1);
        atomicAdd(&z[color[i]], (unsigned long long)(row[i + 1] - row[i]));
    }
    __syncthreads();
    for (int c = threadIdx.x; c < 257; c += blockDim.x)
        if (h[c]) { atomicAdd(&hist[c], h[c]); atomicAdd(&hnnz[c], z[c]); }
}

// ------------------------------------------------------------------ SELL-64
__global__ void k_sell_width(int n, int nslice, const int *perm, const int *row, int *w64) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nslice; s += gridDim.x * blockDim.x) {
        int w = 0;
        for (int l = 0; l < 64; l++) {
            const int p = s * 64 + l;
            if (p >= n) break;
            const int i = perm[p];
            w = max(w, row[i + 1] - row[i]);
        }
        w64[s] = w * 64;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) w64[nslice] = 0;
}

__global__ void k_sell_fill(int n, const int *perm, const int *iperm, const int *row,
                            const int *col, const double *off, const double *diag,
                            const int *soff, int *scol, double *sval, double *pdiag) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int s = p >> 6, lane = p & 63;
        const int w = (soff[s + 1] - soff[s]) >> 6;
        const int i = perm[p];
        const int b = row[i], len = row[i + 1] - b;
        for (int k = 0; k < w; k++) {
            const size_t e = (size_t)soff[s] + (size_t)k * 64 + lane;
            if (k < len) { scol[e] = iperm[col[b + k]]; sval[e] = off[b + k]; }
            else { scol[e] = p; sval[e] = 0.0; }
        }
        pdiag[p] = diag[i];
    }
}

// rows of the last slice beyond n: make their padding entries harmless
__global__ void k_sell_tail(int n, int nslice, const int *soff, int *scol, double *sval) {
    const int s = nslice - 1;
    const int w = (soff[s + 1] - soff[s]) >> 6;
    for (int lane = threadIdx.x; lane < 64; lane += blockDim.x) {
        const int p = s * 64 + lane;
        if (p < n) continue;
        for (int k = 0; k < w; k++) {
            const size_t e = (size_t)soff[s] + (size_t)k * 64 + lane;
            scol[e] = 0; sval[e] = 0.0;
        }
    }
}

__global__ void k_perm_ci(int n, const int *perm, const int *ci, const int *iperm_next, int *pci) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int I = ci[perm[p]];
        pci[p] = I >= 0 ? iperm_next[I] : -1;
    }
}

// ------------------------------------------------------------------ mixed precision
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    const unsigned lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

__global__ void k_to_f32(long long n, const double *a, float *b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

__global__ void k_to_bf16(long long n, const double *a, uint16_t *b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] = f32_to_bf16_rne((float)a[i]);
}

// FP32 diagonal of a level whose off-diagonals are stored in BF16: the diagonal takes
// the rounding error of its row's off-diagonals, so the stored operator keeps the
// row sums (A times the constant vector) of the FP64 operator. Without it, rounding
// can tip a nearly singular coarse diffusion row out of diagonal dominance.
__device__ __forceinline__ double bf16_err(double v) {
    const uint16_t h = f32_to_bf16_rne((float)v);
    return v - (double)__uint_as_float(((unsigned)h) << 16);
}

__global__ void k_diag_comp_csr(int n, const int *row, const double *off, const double *diag, float *d32) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double d = diag[i];
        for (int e = row[i]; e < row[i + 1]; e++) d += bf16_err(off[e]);
        d32[i] = (float)d;
    }
}

__global__ void k_diag_comp_sell(int n, const int *soff, const double *sval, const double *pdiag, float *d32) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int o = soff[p >> 6], w = (soff[(p >> 6) + 1] - o) >> 6, lane = p & 63;
        double d = pdiag[p];
        for (int k = 0; k < w; k++) d += bf16_err(sval[o + k * 64 + lane]);   // padding holds 0
        d32[p] = (float)d;
    }
}

// ------------------------------------------------------------------ dense inverse
// Gauss-Jordan with partial pivoting of a dense n x n FP64 matrix (n <= 256) in one block.
__global__ void k_dense_from_csr(int n, const int *row, const int *col, const double *diag,
                                 const double *off, double *A) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] = 0.0;
        A[(size_t)i * n + i] = diag[i];
        for (int p = row[i]; p < row[i + 1]; p++) A[(size_t)i * n + col[p]] += off[p];
    }
}

__global__ void k_gauss_jordan(int n, double *A, double *Inv, int *status) {
    __shared__ int piv;
    __shared__ double pval;
    __shared__ double s_v[16];
    __shared__ int s_i[16];
    __shared__ double s_prow[256], s_pinv[256], s_f[256];
    for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x)
        Inv[idx] = ((idx / n) == (idx % n)) ? 1.0 : 0.0;
    __syncthreads();
    for (int k = 0; k < n; k++) {
        // pivot: the first row i >= k of largest |A[i][k]| (rows below k with a NaN
        // are skipped; a NaN on the diagonal keeps k and stops below), by a
        // block-wide argmax (n <= blockDim.x)
        {
            const int i = threadIdx.x;
            double v = -1.0;
            int bi = INT_MAX;
            if (i >= k && i < n) {
                const double a = fabs(A[(size_t)i * n + k]);
                if (i == k) v = isnan(a) ? INFINITY : a;
                else v = isnan(a) ? -1.0 : a;
                bi = i;
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double v2 = __shfl_xor(v, o, 64);
                const int i2 = __shfl_xor(bi, o, 64);
                if (v2 > v || (v2 == v && i2 < bi)) { v = v2; bi = i2; }
            }
            const int wv = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) { s_v[wv] = v; s_i[wv] = bi; }
            __syncthreads();
            if (threadIdx.x == 0) {
                double bv = s_v[0];
                int best = s_i[0];
                for (int w = 1; w < (int)(blockDim.x >> 6); w++)
                    if (s_v[w] > bv || (s_v[w] == bv && s_i[w] < best)) { bv = s_v[w]; best = s_i[w]; }
                piv = best;
                pval = fabs(A[(size_t)best * n + k]);
            }
        }
        __syncthreads();
        if (!(pval > 0.0) || !isfinite(pval)) {
            if (threadIdx.x == 0) *status = 1;
            return;
        }
        if (piv != k) {
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                double a = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)piv * n + j]; A[(size_t)piv * n + j] = a;
                double b = Inv[(size_t)k * n + j]; Inv[(size_t)k * n + j] = Inv[(size_t)piv * n + j]; Inv[(size_t)piv * n + j] = b;
            }
        }
        __syncthreads();
        const double inv_p = 1.0 / A[(size_t)k * n + k];
        __syncthreads();   // every thread has read the pivot before row k is rescaled
        // scaled pivot row and the multipliers of column k, staged in LDS: the
        // elimination then reads only its own elements from memory
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const double a = A[(size_t)k * n + j] * inv_p, v = Inv[(size_t)k * n + j] * inv_p;
            s_prow[j] = a; s_pinv[j] = v;
            A[(size_t)k * n + j] = a; Inv[(size_t)k * n + j] = v;
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) s_f[i] = A[(size_t)i * n + k];
        __syncthreads();
        // eight elements per thread at a time: all loads issued before the first store
        constexpr int B = 8;
        for (int base = threadIdx.x; base < n * n; base += B * blockDim.x) {
            double a[B], v[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                const int idx = base + u * blockDim.x;
                if (idx < n * n) { a[u] = A[idx]; v[u] = Inv[idx]; }
            }
#pragma unroll
            for (int u = 0; u < B; u++) {
                const int idx = base + u * blockDim.x;
                if (idx >= n * n) continue;
                const int i = idx / n, j = idx - i * n;
                if (i == k) continue;
                const double f = s_f[i];
                A[idx] = j == k ? 0.0 : a[u] - f * s_prow[j];
                Inv[idx] = v[u] - f * s_pinv[j];
            }
        }
        __syncthreads();
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Phase timing of the setup, printed to stderr when FVM_AMD_SETUP_TIMING is set.
struct PhaseTimer {
    fvm_amg *h;
    bool on;
    double t;
    explicit PhaseTimer(fvm_amg *hh) : h(hh), on(getenv("FVM_AMD_SETUP_TIMING") != nullptr), t(0) {
        if (on) { (void)hipStreamSynchronize(h->stream); t = now_s(); }
    }
    void mark(const char *what, int level) {
        if (!on) return;
        (void)hipStreamSynchronize(h->stream);
        const double n = now_s();
        fprintf(stderr, "[setup] level %2d %-18s %9.3f ms\n", level, what, 1e3 * (n - t));
        t = n;
    }
};

// Structural symmetry test and, for non-symmetric patterns, the transpose pattern.
void ensure_transpose(fvm_amg *h, Level &L) {
    if (L.sym >= 0) return;
    hipStream_t s = h->stream;
    const int n = L.n;
    if (n == 0) { L.sym = 1; return; }
    DArr<int> sflag(1);
    FVM_HIP(hipMemsetAsync(sflag.p, 0, sizeof(int), s));
    k_sym_check<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.row.p, L.col.p, sflag.p);
    L.sym = read_scalar(sflag.p, s) == 0 ? 1 : 0;
    if (L.sym) return;
    DArr<int> cnt((size_t)n + 1), fill((size_t)n + 1);
    FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    FVM_HIP(hipMemsetAsync(fill.p, 0, fill.bytes(), s));
    k_tr_count<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.row.p, L.col.p, cnt.p);
    L.rowT.alloc((size_t)n + 1);
    h->ws.scan.exclusive(cnt.p, L.rowT.p, n + 1, s);
    L.colT.alloc(std::max<long long>(L.nnz, 1));
    L.tposT.alloc(std::max<long long>(L.nnz, 1));
    k_tr_fill<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.row.p, L.col.p, L.rowT.p, fill.p, L.colT.p, L.tposT.p);
    FVM_HIP(hipGetLastError());
}

constexpr int CTR_STRIDE = 64;   // ints between the agglomeration's counters (256 bytes)

// Agglomerate level L (reference order); fills L.ci, L.nc and returns the rounds used.
int coarsen_level(fvm_amg *h, Level &L, double theta, int *retries = nullptr) {
    hipStream_t s = h->stream;
    const int n = L.n;
    ensure_transpose(h, L);
    const int sym = L.sym;
    const int *rowT = sym ? L.row.p : L.rowT.p;
    const int *colT = sym ? L.col.p : L.colT.p;
    DArr<int> ldr(n), dround(n), mem((size_t)n * MEMW), memr((size_t)n * MEMW), mcnt(n), cnt(n),
        lists(2 * (size_t)n + 2), ctr(6 * CTR_STRIDE);
    FVM_HIP(hipMemsetAsync(ctr.p, 0, ctr.bytes(), s));
    // counters one cache line apart: every queue append of a round is an atomic on
    // one of them, and atomics on a shared line wait for each other
    int *qc = ctr.p;             // qc[0], qc[CTR_STRIDE], qc[2 * CTR_STRIDE]: rotating frontier counters
    int *rounds_out = ctr.p + 3 * CTR_STRIDE;
    int *nretry = retries ? ctr.p + 4 * CTR_STRIDE : nullptr;
    const uint8_t *bnd = L.bnd.p;
    int *q0 = lists.p, *q1 = lists.p + n;
    // entries (d, c) along which d can never pair c, and for the counts the position
    // of the entry each transposed entry stands for
    DArr<uint8_t> mayp, mayT;
    const uint8_t *mp = nullptr, *mt = nullptr;
    if (h->coarsen_pair_filter && L.nnz > 0) {
        mayp.alloc(L.nnz);
        mayT.alloc(L.nnz);
        const unsigned gm = grid_for((long long)n * MP, BS, 65536);
        k_may_pair<<<gm, BS, 0, s>>>(n, L.row.p, L.col.p, L.off.p, L.diag.p, bnd, theta, h->coarsen_lower_filter,
                                     mayp.p);
        k_may_pair_t<<<gm, BS, 0, s>>>(n, L.row.p, L.col.p, rowT, colT, sym ? nullptr : L.tposT.p, mayp.p,
                                       mayT.p);
        FVM_HIP(hipGetLastError());
        mp = mayp.p;
        mt = mayT.p;
    }
    // initial counts; the frontier of round 1 goes to q1 (round t reads list t&1)
    {
        const long long threads = (long long)n * INIT_SUB;
        const unsigned g = (unsigned)((threads + 255) / 256);
        k_cnt_init<<<g, 256, 0, s>>>(n, L.row.p, L.col.p, rowT, colT, sym, bnd, mp, mt, cnt.p, dround.p,
                                     ldr.p, memr.p, mcnt.p, q1, qc + CTR_STRIDE);
        FVM_HIP(hipGetLastError());
    }
    // dependents lists (count, scan, fill)
    DArr<int> dptr((size_t)n + 1);
    {
        DArr<int> dcnt((size_t)n + 1);
        FVM_HIP(hipMemsetAsync(dcnt.p + n, 0, sizeof(int), s));
        const unsigned gc = (unsigned)std::max<long long>(1, std::min<long long>(((long long)n * 16 + 255) / 256, 65536));
        k_dep_fill<16, false><<<gc, 256, 0, s>>>(n, L.row.p, L.col.p, rowT, colT, sym, bnd, mp, mt, nullptr,
                                                 dcnt.p);
        h->ws.scan.exclusive(dcnt.p, dptr.p, n + 1, s);
    }
    const int ndep = read_scalar(dptr.p + n, s);
    FVM_CHECK(ndep >= 0, FVM_E_INVALID, "dependents lists exceed 2^31 entries");
    DArr<int> dep((size_t)std::max(ndep, 1));
    {
        const unsigned gf = (unsigned)std::max<long long>(1, std::min<long long>(((long long)n * 16 + 255) / 256, 65536));
        k_dep_fill<16, true><<<gf, 256, 0, s>>>(n, L.row.p, L.col.p, rowT, colT, sym, bnd, mp, mt, dptr.p, dep.p);
        FVM_HIP(hipGetLastError());
    }
    DArr<RowRec> rec(std::max(n, 1));
    k_rec_pack<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.diag.p, dptr.p, dround.p, ldr.p, rec.p);
    FVM_HIP(hipGetLastError());
    // links of the waiter lists (nullptr: blocked joins retry every round)
    DArr<int> wlink;
    if (h->coarsen_wait_lists) wlink.alloc(std::max(n, 1));
    int *wn = h->coarsen_wait_lists ? wlink.p : nullptr;
    const int wall = h->coarsen_wait_lists >= 2;   // also joins blocked by a row that could join
    int t = 1;
    const long long cap = 4LL * n + 64;
    const char *path;
    // lanes per row (measured on the 200^3 hierarchy): 8 for rows of at most 7
    // entries (eight rows per wavefront), 16 up to 10 entries, 32 for large levels
    // of moderate rows (their frontiers are thousands of rows wide), else one
    // wavefront per row; the single-workgroup rounds take the same widths
    int sub = L.nnz <= 7LL * n ? 8
            : L.nnz <= 10LL * n ? THIN_SUB
            : (L.nnz <= 16LL * n && n >= (1 << 20)) ? 32 : ROUND_SUB;
    if (h->coarsen_lanes) sub = h->coarsen_lanes;   // tuning override (8, 16, 32 or 64)
    if (n <= h->coarsen_small_rows) {
        path = "single-workgroup";
        auto launch_small = [&](auto kern) {
            kern<<<1, 1024, 0, s>>>(n, (int)std::min<long long>(cap, INT_MAX - 1), L.row.p, L.col.p, rowT, colT,
                                    L.off.p, bnd, rec.p, dep.p, theta, mem.p, memr.p, mcnt.p, cnt.p, q1, q0,
                                    qc + CTR_STRIDE, rounds_out, nretry, wn, wall);
        };
        const int ssub = h->coarsen_small_lanes ? h->coarsen_small_lanes : sub;
        if (ssub == 8) launch_small(k_coarsen_small<8>);
        else if (ssub == THIN_SUB) launch_small(k_coarsen_small<THIN_SUB>);
        else if (ssub == 32) launch_small(k_coarsen_small<32>);
        else launch_small(k_coarsen_small<SMALL_SUB>);
        FVM_HIP(hipGetLastError());
        t = read_scalar(rounds_out, s);
    } else {
        // rounds until the frontier is empty, checked every `batch` launches; the
        // check of batch b runs while batch b + 1 is queued (a round on an empty
        // frontier only zeroes a counter)
        path = "multi-block";
        // blocks per round launch: grid-stride loops over the frontier, so any
        // size is correct; "coarsen_grid" overrides the default
        // threads per block ("coarsen_block"): each block appends its part of the
        // next frontier with one atomic
        const int bt = h->coarsen_block;
        const unsigned g_round = h->coarsen_grid ? (unsigned)h->coarsen_grid
                                                 : (sub < ROUND_SUB ? 2048u : 1024u) * 256u / (unsigned)bt;
        // entries of each block's LDS queue (at most 32 KiB of LDS)
        const int lcap = std::min(h->coarsen_lds_entries * bt, 8192);
        const size_t lds = std::max<size_t>((size_t)lcap * sizeof(int), sizeof(int));
        const int batch = h->coarsen_batch;
        h->probe.init();
        int t_prev = t;
        for (int b = 0;; b++) {
            for (int k = 0; k < batch; k++, t++) {
                auto launch = [&](auto kern) {
                    kern<<<g_round, bt, lds, s>>>(
                        t, n, L.row.p, L.col.p, rowT, colT, L.off.p, bnd, rec.p, dep.p, theta, mem.p, memr.p,
                        mcnt.p, cnt.p, (t & 1) ? q1 : q0, qc + (t % 3) * CTR_STRIDE, (t & 1) ? q0 : q1,
                        qc + ((t + 1) % 3) * CTR_STRIDE, qc + ((t + 2) % 3) * CTR_STRIDE, nretry, wn, wall, lcap);
                };
                if (sub == 8) launch(k_coarsen_round<8>);
                else if (sub == THIN_SUB) launch(k_coarsen_round<THIN_SUB>);
                else if (sub == 32) launch(k_coarsen_round<32>);
                else launch(k_coarsen_round<ROUND_SUB>);
            }
            FVM_HIP(hipGetLastError());
            h->probe.post(b, qc + (t % 3) * CTR_STRIDE, s);   // frontier of round t (the next one)
            if (b > 0 && h->probe.wait(b - 1) == 0) { t = t_prev; break; }
            t_prev = t;
            if (t >= cap) break;
        }
    }
    k_rec_unpack<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, rec.p, dround.p, ldr.p);
    {
        DArr<int> und(1);
        FVM_HIP(hipMemsetAsync(und.p, 0, sizeof(int), s));
        k_count_undecided<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, bnd, dround.p, und.p);
        const int rem = read_scalar(und.p, s);
        FVM_CHECK(rem == 0, FVM_E_NUMERIC,
                  std::string("coarsening did not terminate (") + path + " path, n=" + std::to_string(n) +
                      ", rounds=" + std::to_string(t) + ", undecided=" + std::to_string(rem) + ")");
    }
    // leaders -> contiguous ids in seed order
    DArr<int> flag((size_t)n + 1), gid((size_t)n + 1);
    k_leader_flags<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, ldr.p, bnd, flag.p);
    h->ws.scan.exclusive(flag.p, gid.p, n + 1, s);
    L.ci.alloc(n);
    k_assign_ci<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, ldr.p, bnd, gid.p, L.ci.p);
    FVM_HIP(hipGetLastError());
    L.nc = read_scalar(gid.p + n, s);
    if (retries) *retries = read_scalar(nretry, s);
    return t - 1;
}

void coarse_to_fine(fvm_amg *h, Level &L) {
    hipStream_t s = h->stream;
    const int n = L.n, nc = L.nc;
    DArr<int> cnt((size_t)nc + 1), fill((size_t)nc + 1);
    FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    FVM_HIP(hipMemsetAsync(fill.p, 0, fill.bytes(), s));
    k_c2f_count<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.ci.p, cnt.p);
    L.c2f_row.alloc((size_t)nc + 1);
    h->ws.scan.exclusive(cnt.p, L.c2f_row.p, nc + 1, s);
    L.c2f_col.alloc(std::max(n, 1));
    k_c2f_fill<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.ci.p, L.c2f_row.p, fill.p, L.c2f_col.p);
    k_segsort<<<grid_for(nc, BS, 4096), BS, 0, s>>>(nc, L.c2f_row.p, L.c2f_col.p, nullptr);
    FVM_HIP(hipGetLastError());
}

void coarse_operator(fvm_amg *h, Level &F, Level &C) {
    hipStream_t s = h->stream;
    const int nc = F.nc;
    // lanes per coarse row from the mean number of flattened entries per coarse row
    // (16 lanes hold up to four chunks of a row in registers: measured fastest on
    // the 200^3 hierarchy, whose rows average 14-36 entries)
    const double mavg = nc ? (double)(F.nnz + F.n) / nc : 0.0;
    const int S = h->gal_lanes ? h->gal_lanes : (mavg <= 40 ? 16 : (mavg <= 56 ? 32 : 64));
    const unsigned g = (unsigned)std::min<long long>(((long long)nc * S + 255) / 256, 65536);
    // slots of the one-pass product: row I gets as many as it has flattened entries
    // (their total is at most nnz + n of the fine level)
    DArr<int> ub((size_t)nc + 1), toff((size_t)nc + 1), cnt((size_t)nc + 1);
    k_gal_bound<<<grid_for(nc, BS, 4096), BS, 0, s>>>(nc, F.c2f_row.p, F.c2f_col.p, F.row.p, ub.p);
    h->ws.scan.exclusive(ub.p, toff.p, nc + 1, s);
    const size_t tsize = (size_t)std::max<long long>(F.nnz + F.n, 1);
    DArr<int> tcol(tsize);
    DArr<double> tval(tsize);
    C.diag.alloc(std::max(nc, 1));
    auto launch = [&](auto kern) {
        kern<<<std::max(g, 1u), 256, 0, s>>>(nc, F.c2f_row.p, F.c2f_col.p, F.row.p, F.col.p, F.diag.p, F.off.p,
                                             F.ci.p, cnt.p, toff.p, tcol.p, C.diag.p, tval.p);
        FVM_HIP(hipGetLastError());
    };
    if (S == 16) launch(k_gal<16>);
    else if (S == 32) launch(k_gal<32>);
    else launch(k_gal<64>);
    C.n = nc;
    C.row.alloc((size_t)nc + 1);
    h->ws.scan.exclusive(cnt.p, C.row.p, nc + 1, s);
    C.nnz = read_scalar(C.row.p + nc, s);
    C.col.alloc(std::max<long long>(C.nnz, 1));
    C.off.alloc(std::max<long long>(C.nnz, 1));
    k_gal_compact<16><<<std::max((unsigned)std::min<long long>(((long long)nc * 16 + 255) / 256, 65536), 1u), 256, 0,
                        s>>>(nc, toff.p, C.row.p, tcol.p, tval.p, C.col.p, C.off.p);
    FVM_HIP(hipGetLastError());
    // a structurally symmetric fine pattern gives a symmetric coarse one: (I, J) is an
    // entry iff some a_ij with i in I, j in J is, and a_ji is then one too (rows
    // outside every aggregate drop out of both sides alike)
    if (F.sym == 1) C.sym = 1;
}

void alloc_vectors(Level &L) {
    const size_t n = std::max(L.n, 1);
    L.b.alloc(n); L.x.alloc(n); L.r.alloc(n); L.t.alloc(n);
}

}  // namespace

// Build the reference hierarchy (AMG::createCoarseLevels, AMG.cpp:149-210) from a
// device CSR of the finest level; ghost rows/columns are dropped first.
// Colouring and SELL layout of levels as soon as they are agglomerated: a host
// thread drives them on the handle's second stream while the setup thread goes
// on with the next levels (agglomeration rounds are latency-bound and leave the
// device mostly idle). A level is handed over with an event of the main stream
// after which its matrix and transpose pattern are final; the worker touches
// nothing else of the hierarchy. The level vector is reserved beforehand, so
// the addresses handed over stay valid.
class LevelPipe {
  public:
    explicit LevelPipe(fvm_amg *h) : h_(h) {
        if (!h_->stream2) FVM_HIP(hipStreamCreateWithFlags(&h_->stream2, hipStreamDefault));
        // blocks of pool2 freed by the last setup may still be read by earlier
        // work of the main stream: the second stream starts after it
        hipEvent_t e;
        FVM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        FVM_HIP(hipEventRecord(e, h_->stream));
        FVM_HIP(hipStreamWaitEvent(h_->stream2, e, 0));
        FVM_HIP(hipEventDestroy(e));
        th_ = std::thread([this] { run(); });
    }
    ~LevelPipe() { stop(); }
    void push(Level *L, int l) {
        hipEvent_t e;
        FVM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        FVM_HIP(hipEventRecord(e, h_->stream));
        {
            std::lock_guard<std::mutex> g(mu_);
            q_.push_back(Item{L, l, e});
        }
        cv_.notify_one();
    }
    // waits for every level handed over; the main stream then follows the second
    void finish() {
        stop();
        if (err_) std::rethrow_exception(err_);
        hipEvent_t e;
        FVM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        FVM_HIP(hipEventRecord(e, h_->stream2));
        FVM_HIP(hipStreamWaitEvent(h_->stream, e, 0));
        FVM_HIP(hipEventDestroy(e));
    }

  private:
    struct Item {
        Level *L;
        int l;
        hipEvent_t ev;
    };
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            done_ = true;
        }
        cv_.notify_one();
        if (th_.joinable()) th_.join();
        for (auto &it : q_) (void)hipEventDestroy(it.ev);   // left over after an error
        q_.clear();
    }
    void run() {
        (void)hipSetDevice(h_->device);
        PoolScope ps(&h_->pool2);
        for (;;) {
            Item it;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [this] { return done_ || !q_.empty(); });
                if (q_.empty()) return;
                it = q_.front();
                q_.pop_front();
            }
            try {
                if (!err_) {
                    FVM_HIP(hipStreamWaitEvent(h_->stream2, it.ev, 0));
                    prepare_fast_level(h_, *it.L, it.l, h_->stream2, h_->ws2, nullptr, 0);
                    FVM_HIP(hipStreamSynchronize(h_->stream2));
                }
            } catch (...) {
                err_ = std::current_exception();
            }
            (void)hipEventDestroy(it.ev);
        }
    }
    fvm_amg *h_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Item> q_;
    bool done_ = false;
    std::exception_ptr err_;
};

// Signature of a setup's inputs: the matrix arrays and the parameters that shape
// the hierarchy (or what the setup prepares for the cycle).
unsigned long long setup_signature(fvm_amg *h, int n_self, int n_total, const int *d_row, const int *d_col,
                                   const double *d_diag, const double *d_off, const uint8_t *d_bnd) {
    hipStream_t s = h->stream;
    const long long nnz = n_total > 0 ? (long long)read_scalar(d_row + n_total, s) : 0;
    DArr<unsigned long long> acc(1);
    FVM_HIP(hipMemsetAsync(acc.p, 0, sizeof(unsigned long long), s));
    k_signature<int><<<grid_for((long long)n_total + 1, BS, 4096), BS, 0, s>>>((long long)n_total + 1, d_row, 1, acc.p);
    if (nnz > 0) {
        k_signature<int><<<grid_for(nnz, BS, 4096), BS, 0, s>>>(nnz, d_col, 2ull << 40, acc.p);
        k_signature<double><<<grid_for(nnz, BS, 4096), BS, 0, s>>>(nnz, d_off, 3ull << 40, acc.p);
    }
    if (n_total > 0) {
        k_signature<double><<<grid_for(n_total, BS, 4096), BS, 0, s>>>(n_total, d_diag, 4ull << 40, acc.p);
        if (d_bnd) k_signature<uint8_t><<<grid_for(n_total, BS, 4096), BS, 0, s>>>(n_total, d_bnd, 5ull << 40, acc.p);
    }
    FVM_HIP(hipGetLastError());
    unsigned long long sig = read_scalar(acc.p, s);
    // host side: sizes and parameters, FNV-1a over their bytes
    auto fold = [&sig](const void *p, size_t nb) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t k = 0; k < nb; k++) sig = (sig ^ c[k]) * 0x100000001B3ull;
    };
    const int ints[] = {n_self, n_total, d_bnd ? 1 : 0, h->prm.maxCoarseLevels, h->prm.coarseGroupSize,
                        h->prm.parallelStopRule, h->prm.gsOrdering, h->prm.precision, h->prm.smootherType,
                        h->bipartite_levels};
    fold(ints, sizeof(ints));
    fold(&nnz, sizeof(nnz));
    fold(&h->prm.weightRatioThreshold, sizeof(double));
    return sig;
}

void build_hierarchy(fvm_amg *h, int n_self, int n_total, const int *d_row, const int *d_col,
                     const double *d_diag, const double *d_off, const uint8_t *d_bnd) {
    const double t0 = now_s();
    hipStream_t s = h->stream;
    FVM_CHECK(n_self >= 0 && n_total >= n_self, FVM_E_INVALID, "need 0 <= n_self <= n_total");
    FVM_CHECK(h->prm.coarseGroupSize == 2 || n_self == 0, FVM_E_INVALID,
              "the GPU agglomeration supports coarseGroupSize == 2 only");
    // opt-in reuse: the same matrix and parameters as the hierarchy in place give
    // the same hierarchy (the setup is deterministic), so it is kept
    unsigned long long sig = 0;
    if (h->reuse_setup) {
        sig = setup_signature(h, n_self, n_total, d_row, d_col, d_diag, d_off, d_bnd);
        if ((h->ready || h->parked) && !h->batch && h->setup_sig_valid && sig == h->setup_sig) {
            h->ready = true;
            h->parked = false;
            h->setup_reused++;
            h->setup_seconds = now_s() - t0;
            return;
        }
    }
    h->setup_sig_valid = false;
    h->parked = false;
    h->lev.clear();
    h->ready = false;
    h->batch = nullptr;
    h->direct_level = -1;
    h->direct_tried = false;
    h->tail_prec = -1;
    h->tail_desc.release();
    invalidate_cycle_graphs(h);
    h->n_self = n_self;
    h->n_total = n_total;
    const long long nnz_in = n_total > 0 ? (long long)read_scalar(d_row + n_total, s) : 0;
    {
        DArr<int> flag(1);
        FVM_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
        if (n_total > 0) {
            k_validate<<<grid_for(n_total, BS, 4096), BS, 0, s>>>(n_self, n_total, nnz_in, d_row,
                                                                  d_col, d_diag, flag.p);
            FVM_HIP(hipGetLastError());
        }
        const int f = read_scalar(flag.p, s);
        FVM_CHECK(!(f & 1), FVM_E_INVALID, "row pointers are not a valid CSR");
        FVM_CHECK(!(f & 2), FVM_E_INVALID, "column index out of range or on the diagonal");
        FVM_CHECK(!(f & 4), FVM_E_NUMERIC, "zero or non-finite diagonal on a self row");
    }
    h->lev.emplace_back();
    Level &F = h->lev[0];
    F.n = n_self;
    {
        DArr<int> cnt((size_t)n_self + 1);
        FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
        if (n_self > 0)
            k_compact_count<<<grid_for(n_self, BS, 4096), BS, 0, s>>>(n_self, d_row, d_col, cnt.p);
        F.row.alloc((size_t)n_self + 1);
        h->ws.scan.exclusive(cnt.p, F.row.p, n_self + 1, s);
        F.nnz = read_scalar(F.row.p + n_self, s);
        F.col.alloc(std::max<long long>(F.nnz, 1));
        F.off.alloc(std::max<long long>(F.nnz, 1));
        F.diag.alloc(std::max(n_self, 1));
        if (n_self > 0) {
            k_compact_fill<<<grid_for(n_self, BS, 4096), BS, 0, s>>>(n_self, d_row, d_col, d_off,
                                                                     F.row.p, F.col.p, F.off.p);
            FVM_HIP(hipMemcpyAsync(F.diag.p, d_diag, sizeof(double) * n_self,
                                   hipMemcpyDeviceToDevice, s));
        }
        if (d_bnd && n_self > 0) {
            F.bnd.alloc(n_self);
            FVM_HIP(hipMemcpyAsync(F.bnd.p, d_bnd, n_self, hipMemcpyDeviceToDevice, s));
        }
        FVM_HIP(hipGetLastError());
    }
    PhaseTimer pt(h);
    pt.mark("validate+compact", 0);
    int fine_size = n_total;  // getLocalSize() of the finest level counts ghosts
    // coloured layout: levels are coloured on the second stream as they are finished
    std::unique_ptr<LevelPipe> pipe;
    int handed = 0;
    if (h->prm.gsOrdering == FVM_GS_MULTICOLOR && n_self > 0) {
        h->lev.reserve((size_t)std::max(h->prm.maxCoarseLevels, 0) + 2);
        pipe.reset(new LevelPipe(h));
    }
    for (int k = 0; k < h->prm.maxCoarseLevels && n_self > 0; k++) {
        Level &A = h->lev.back();
        int retries = 0;
        const int rounds = coarsen_level(h, A, h->prm.weightRatioThreshold, pt.on ? &retries : nullptr);
        if (pipe) { pipe->push(&A, k); handed = k + 1; }   // its pattern is analysed and final
        pt.mark("coarsen", k);
        if (pt.on) fprintf(stderr, "[setup] level %2d n=%d nnz=%lld rounds=%d blocked-joins=%d\n", k, A.n, A.nnz, rounds, retries);
        if (k == 0) h->finest_rounds = rounds;
        const int nc = A.nc;
        if (nc == fine_size) { A.ci.release(); A.nc = 0; break; }       // AMG.cpp:163-168
        if (!h->prm.parallelStopRule && nc <= 3) { A.ci.release(); A.nc = 0; break; }
        coarse_to_fine(h, A);
        pt.mark("coarse-to-fine", k);
        Level C;
        coarse_operator(h, A, C);
        pt.mark("galerkin", k);
        h->lev.push_back(std::move(C));
        if (h->prm.verbosity > 1)
            printf("Created coarse level %d of size %d\n", k, nc);
        fine_size = nc;
        if (h->prm.parallelStopRule && nc <= 3) break;                  // AMG.cpp:180-181
    }
    if (pipe) {
        for (int l = handed; l < (int)h->lev.size(); l++) {
            ensure_transpose(h, h->lev[l]);
            pipe->push(&h->lev[l], l);
        }
        pipe->finish();
        pipe.reset();
        prepare_fast_maps(h);
        pt.mark("colour+sell (overlapped)", -1);
    }
    for (auto &L : h->lev) alloc_vectors(L);
    h->ws.partials.ensure(4096);
    h->ws.scal.ensure(64);
    FVM_HIP(hipStreamSynchronize(s));
    h->ready = true;
    pt.mark("vectors", -1);
    // mode-specific preparation
    if (h->prm.smootherType == FVM_GAUSS_SEIDEL && h->prm.gsOrdering == FVM_GS_EXACT)
        for (auto &L : h->lev) prepare_exact_schedule(h, L);
    pt.mark("gs-schedule", -1);
    if (h->prm.gsOrdering == FVM_GS_MULTICOLOR) prepare_fast(h);
    pt.mark("colour+sell", -1);
    if (h->prm.precision == FVM_PREC_MIXED) {
        prepare_mixed(h, h->prm.gsOrdering == FVM_GS_MULTICOLOR);
        pt.mark("mixed", -1);
        prepare_direct(h);
        pt.mark("direct", -1);
    }
    FVM_HIP(hipStreamSynchronize(s));
    if (h->reuse_setup) {
        h->setup_sig = sig;
        h->setup_sig_valid = true;
    }
    h->setup_seconds = now_s() - t0;
}

void prepare_exact_schedule(fvm_amg *h, Level &L) {
    if (L.sched_ready) return;
    hipStream_t s = h->stream;
    const int n = L.n;
    if (n == 0) { L.sched_ready = true; return; }
    ensure_transpose(h, L);
    for (int dir = 0; dir < 2; dir++) {
        DArr<int> indeg(n), ctr(2);
        DArr<int> &rows = dir ? L.bwd_rows : L.fwd_rows;
        DArr<int> &ptr = dir ? L.bwd_ptr : L.fwd_ptr;
        rows.alloc(n);
        const int batch = 32;
        ptr.alloc((size_t)n + batch + 4);
        FVM_HIP(hipMemsetAsync(ctr.p, 0, ctr.bytes(), s));
        int *tail = ctr.p;
        unsigned *done_blocks = (unsigned *)(ctr.p + 1);
        const int *rT = L.sym ? nullptr : L.rowT.p;
        const int *cT = L.sym ? nullptr : L.colT.p;
        k_sched_init<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, dir, L.row.p, L.col.p, rT, cT, indeg.p,
                                                          rows.p, tail);
        k_set_ptr01<<<1, 1, 0, s>>>(ptr.p, tail);
        FVM_HIP(hipGetLastError());
        int lvl = 0;
        const unsigned g = grid_for(n, BS, 1024);
        for (;;) {
            for (int k = 0; k < batch; k++, lvl++) {
                FVM_CHECK((size_t)lvl + 2 < ptr.n, FVM_E_NUMERIC, "GS schedule overflow");
                k_sched_round<<<g, BS, 0, s>>>(lvl, dir, L.row.p, L.col.p, rT, cT, indeg.p, rows.p,
                                               ptr.p, tail, done_blocks);
            }
            FVM_HIP(hipGetLastError());
            if (read_scalar(tail, s) >= n) break;
            FVM_CHECK(lvl <= n + 1, FVM_E_NUMERIC, "GS schedule did not terminate");
            // grow the pointer array if the schedule is deep
            if ((size_t)lvl + batch + 3 >= ptr.n) {
                DArr<int> np(ptr.n * 2);
                FVM_HIP(hipMemcpyAsync(np.p, ptr.p, ptr.bytes(), hipMemcpyDeviceToDevice, s));
                ptr = std::move(np);
            }
        }
        std::vector<int> hp(lvl + 2);
        FVM_HIP(hipMemcpyAsync(hp.data(), ptr.p, sizeof(int) * (lvl + 2), hipMemcpyDeviceToHost, s));
        FVM_HIP(hipStreamSynchronize(s));
        int nl = 0, maxw = 0;
        while (nl + 1 < (int)hp.size() && hp[nl] < n) { maxw = std::max(maxw, hp[nl + 1] - hp[nl]); nl++; }
        hp.resize(nl + 1);
        if (dir) { L.bwd_nlev = nl; L.bwd_maxw = maxw; L.h_bwd_ptr = hp; }
        else { L.fwd_nlev = nl; L.fwd_maxw = maxw; L.h_fwd_ptr = hp; }
    }
    L.runs_width = -1;
    L.sched_ready = true;
}

// Graph colouring and coloured SELL-64 layout of every level.
// Colouring and SELL-64 layout of one level, on stream s with workspace ws
// (arrays from the calling thread's pool). The level's transpose pattern must
// exist (ensure_transpose).
void prepare_fast_level(fvm_amg *h, Level &L, int l, hipStream_t s, Workspace &ws, const int *seg_off, int nseg) {
    const int n = L.n;
    if (n == 0) return;
    FVM_CHECK(L.sym >= 0, FVM_E_INVALID, "level pattern not analysed before colouring");
    const int *rT = L.sym ? nullptr : L.rowT.p;
    const int *cT = L.sym ? nullptr : L.colT.p;
    // blocks of a colouring round (grid-stride loops): this stream runs beside the
    // agglomeration rounds, which are latency-bound and need few CUs but prompt ones
    const unsigned cgrid = (unsigned)h->colour_grid;
    DArr<int> color(n), cround(n);
    // segments (systems) of the level: with every one two-coloured the greedy
    // colouring is skipped; with some, it runs and the others keep their red-black
    const int nsys = seg_off ? nseg : 1;
    int n_two = 0;
    DArr<int> bfs_color, bad;
    if (l < h->bipartite_levels) {
        // red-black order when the level's graph has no odd cycle
        DArr<int> qa(n), qb(n), cnt(3);
        bfs_color.alloc(n);
        bad.alloc(nsys);
        k_fill_int<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, bfs_color.p, -1);
        FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
        k_bfs_seed<<<grid_for(std::max(nseg, 1), BS), BS, 0, s>>>(n, seg_off, nseg, bfs_color.p, qa.p, cnt.p);
        int t = 0;
        for (;;) {
            for (int k = 0; k < 16; k++, t++) {
                const int *cur = (t & 1) ? qb.p : qa.p;
                int *nxt = (t & 1) ? qa.p : qb.p;
                k_bfs_round<<<cgrid, BS, 0, s>>>(L.row.p, L.col.p, rT, cT, bfs_color.p, cur, cnt.p + (t % 3), nxt,
                                                cnt.p + ((t + 1) % 3), cnt.p + ((t + 2) % 3));
            }
            FVM_HIP(hipGetLastError());
            if (read_scalar(cnt.p + (t % 3), s) == 0) break;
            FVM_CHECK(t < 2 * n + 64, FVM_E_NUMERIC, "breadth-first colouring did not terminate");
        }
        FVM_HIP(hipMemsetAsync(bad.p, 0, bad.bytes(), s));
        k_bipartite_check<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.row.p, L.col.p, bfs_color.p, seg_off, nseg,
                                                               bad.p);
        std::vector<int> hb(nsys);
        FVM_HIP(hipMemcpyAsync(hb.data(), bad.p, bad.bytes(), hipMemcpyDeviceToHost, s));
        FVM_HIP(hipStreamSynchronize(s));
        for (int v : hb) n_two += v == 0;
    }
    if (n_two == nsys) {
        FVM_HIP(hipMemcpyAsync(color.p, bfs_color.p, color.bytes(), hipMemcpyDeviceToDevice, s));
    } else {
        DArr<unsigned long long> prio(n);
        k_color_prio<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.row.p, seg_off, nseg, prio.p);
        FVM_HIP(hipMemsetAsync(color.p, 0, color.bytes(), s));
        k_fill_int<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, cround.p, INT_MAX);
        // lists of the uncoloured rows (round t reads list t & 1), sizes in rotating counters
        DArr<int> la(n), lb(n), lcnt(3);
        k_iota<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, la.p);
        FVM_HIP(hipMemsetAsync(lcnt.p, 0, lcnt.bytes(), s));
        k_fill_int<<<1, 1, 0, s>>>(1, lcnt.p + 1, n);
        int t = 1;
        for (;;) {
            for (int k = 0; k < 8; k++, t++)
                k_color_round<<<std::min(grid_for(n, BS, 4096), cgrid), BS, 0, s>>>(
                    t, L.row.p, L.col.p, rT, cT, prio.p, color.p, cround.p, (t & 1) ? la.p : lb.p, lcnt.p + (t % 3),
                    (t & 1) ? lb.p : la.p, lcnt.p + ((t + 1) % 3), lcnt.p + ((t + 2) % 3));
            FVM_HIP(hipGetLastError());
            if (read_scalar(lcnt.p + (t % 3), s) == 0) break;   // the list of round t (the next one)
            FVM_CHECK(t < 4 * n + 64, FVM_E_NUMERIC, "colouring did not terminate");
        }
        if (n_two > 0)
            k_merge_colours<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, seg_off, nseg, bad.p, bfs_color.p, color.p);
    }
    // stable sort of rows by colour (coarse levels: then by decreasing length) -> perm
    DArr<int> idx(n);
    DArr<unsigned long long> keys_in(n), keys_out(n);
    L.perm.alloc(n);
    L.iperm.alloc(n);
    k_iota<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, idx.p);
    const int by_len = l > 0 ? 1 : 0;
    int wbits = 1;
    while ((1LL << wbits) <= (long long)(n / SIGMA)) wbits++;
    const int key_bits = by_len ? 9 + wbits + 8 : 9;
    k_color_key<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, color.p, L.row.p, by_len, wbits, keys_in.p);
    size_t bytes = 0;
    FVM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in.p, keys_out.p, idx.p,
                                               L.perm.p, n, 0, key_bits, s));
    ws.sort_tmp.ensure(bytes + 16);
    FVM_HIP(hipcub::DeviceRadixSort::SortPairs(ws.sort_tmp.p, bytes, keys_in.p, keys_out.p,
                                               idx.p, L.perm.p, n, 0, key_bits, s));
    k_invert_perm<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.perm.p, L.iperm.p);
    DArr<int> hist(257);
    DArr<unsigned long long> hnnz(257);
    FVM_HIP(hipMemsetAsync(hist.p, 0, hist.bytes(), s));
    FVM_HIP(hipMemsetAsync(hnnz.p, 0, hnnz.bytes(), s));
    k_color_hist<<<grid_for(n, BS, 1024), BS, 0, s>>>(n, color.p, L.row.p, hist.p, hnnz.p);
    std::vector<int> hh(257);
    std::vector<unsigned long long> hz(257);
    FVM_HIP(hipMemcpyAsync(hh.data(), hist.p, hist.bytes(), hipMemcpyDeviceToHost, s));
    FVM_HIP(hipMemcpyAsync(hz.data(), hnnz.p, hnnz.bytes(), hipMemcpyDeviceToHost, s));
    FVM_HIP(hipStreamSynchronize(s));
    L.color_ptr.assign(1, 0);
    L.color_nnz.clear();
    for (int c = 0; c < 257 && L.color_ptr.back() < n; c++) {
        L.color_ptr.push_back(L.color_ptr.back() + hh[c]);
        L.color_nnz.push_back((long long)hz[c]);
    }
    L.ncolor = (int)L.color_ptr.size() - 1;
    L.d_color_ptr.alloc(L.color_ptr.size());
    FVM_HIP(hipMemcpyAsync(L.d_color_ptr.p, L.color_ptr.data(), sizeof(int) * L.color_ptr.size(),
                           hipMemcpyHostToDevice, s));
    FVM_HIP(hipStreamSynchronize(s));
    if (getenv("FVM_AMD_SETUP_TIMING"))
        fprintf(stderr, "[setup] level %2d n=%d nnz=%lld colours=%d\n", l, n, L.nnz, L.ncolor);
    // SELL-64
    const int nslice = (n + 63) / 64;
    DArr<int> w64((size_t)nslice + 1);
    k_sell_width<<<grid_for(nslice, BS, 4096), BS, 0, s>>>(n, nslice, L.perm.p, L.row.p, w64.p);
    L.sell_off.alloc((size_t)nslice + 1);
    ws.scan.exclusive(w64.p, L.sell_off.p, nslice + 1, s);
    const int total = read_scalar(L.sell_off.p + nslice, s);
    L.sell_total = total;
    {
        const double wmean = (double)total / (64.0 * nslice);
        L.sell_u = wmean > 16 ? 32 : (wmean > 8 ? 16 : 8);
        if (getenv("FVM_AMD_SETUP_TIMING"))
            fprintf(stderr, "[setup] level %2d SELL-64 entries %d for nnz %lld (%.2fx), mean width %.1f\n", l,
                    total, L.nnz, L.nnz ? (double)total / (double)L.nnz : 0.0, wmean);
    }
    L.sell_col.alloc(std::max(total, 1));
    L.sell_val.alloc(std::max(total, 1));
    L.pdiag.alloc(n);
    k_sell_fill<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.perm.p, L.iperm.p, L.row.p, L.col.p,
                                                     L.off.p, L.diag.p, L.sell_off.p,
                                                     L.sell_col.p, L.sell_val.p, L.pdiag.p);
    k_sell_tail<<<1, 64, 0, s>>>(n, nslice, L.sell_off.p, L.sell_col.p, L.sell_val.p);
    FVM_HIP(hipGetLastError());
}

// Coloured-index transfer maps between consecutive levels (every level coloured).
void prepare_fast_maps(fvm_amg *h) {
    hipStream_t s = h->stream;
    const int nl = (int)h->lev.size();
    for (int l = 0; l + 1 < nl; l++) {
        Level &L = h->lev[l];
        Level &C = h->lev[l + 1];
        const int n = L.n, nc = C.n;
        L.pci.alloc(std::max(n, 1));
        k_perm_ci<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.perm.p, L.ci.p, C.iperm.p, L.pci.p);
        DArr<int> cnt((size_t)nc + 1), fill((size_t)nc + 1);
        FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
        FVM_HIP(hipMemsetAsync(fill.p, 0, fill.bytes(), s));
        k_c2f_count<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.pci.p, cnt.p);
        L.pc2f_row.alloc((size_t)nc + 1);
        h->ws.scan.exclusive(cnt.p, L.pc2f_row.p, nc + 1, s);
        L.pc2f_col.alloc(std::max(n, 1));
        k_c2f_fill<<<grid_for(n, BS, 4096), BS, 0, s>>>(n, L.pci.p, L.pc2f_row.p, fill.p,
                                                        L.pc2f_col.p);
        // fine rows of a coarse row in their original order: the restriction sums
        // in the same order whatever the layout of the level
        k_segsort<<<grid_for(nc, BS, 4096), BS, 0, s>>>(nc, L.pc2f_row.p, L.pc2f_col.p, L.perm.p);
        FVM_HIP(hipGetLastError());
    }
    for (auto &L : h->lev) L.fast_ready = true;
    FVM_HIP(hipStreamSynchronize(s));
}

void prepare_fast(fvm_amg *h) {
    const int nl = (int)h->lev.size();
    bool all = true;
    for (auto &L : h->lev) all = all && L.fast_ready;
    if (all) return;
    for (int l = 0; l < nl; l++) {
        Level &L = h->lev[l];
        if (L.n == 0) continue;
        ensure_transpose(h, L);
        // batched hierarchy: colour each system as it is coloured alone
        const int *seg_off = nullptr;
        int nseg = 0;
        if (h->batch && l < (int)h->batch->d_off.size()) {
            seg_off = h->batch->d_off[l].p;
            nseg = h->batch->K;
        }
        prepare_fast_level(h, L, l, h->stream, h->ws, seg_off, nseg);
    }
    prepare_fast_maps(h);
}

// FP32 values on the finest level, BF16 off-diagonals (FP32 diagonal that keeps the
// FP64 row sums) on coarse levels.
void prepare_mixed(fvm_amg *h, bool fast) {
    hipStream_t s = h->stream;
    for (size_t l = 0; l < h->lev.size(); l++) {
        Level &L = h->lev[l];
        const int n = std::max(L.n, 1);
        if (!L.b32.p) { L.b32.alloc(n); L.x32.alloc(n); L.r32.alloc(n); L.t32.alloc(n); }
        if (!fast && !L.mixed_ready_exact) {
            L.diag32.alloc(n);
            if (l == 0) {
                k_to_f32<<<grid_for(L.n, BS, 4096), BS, 0, s>>>(L.n, L.diag.p, L.diag32.p);
                L.off32.alloc(std::max<long long>(L.nnz, 1));
                k_to_f32<<<grid_for(L.nnz, BS, 4096), BS, 0, s>>>(L.nnz, L.off.p, L.off32.p);
            } else {
                k_diag_comp_csr<<<grid_for(L.n, BS, 4096), BS, 0, s>>>(L.n, L.row.p, L.off.p, L.diag.p,
                                                                       L.diag32.p);
                L.off16.alloc(std::max<long long>(L.nnz, 1));
                k_to_bf16<<<grid_for(L.nnz, BS, 4096), BS, 0, s>>>(L.nnz, L.off.p, L.off16.p);
            }
            L.mixed_ready_exact = true;
        }
        if (fast && !L.mixed_ready_fast && L.n > 0) {
            const int total = L.sell_total;
            L.pdiag32.alloc(n);
            if (l == 0) {
                k_to_f32<<<grid_for(L.n, BS, 4096), BS, 0, s>>>(L.n, L.pdiag.p, L.pdiag32.p);
                L.sell_val32.alloc(std::max(total, 1));
                k_to_f32<<<grid_for(total, BS, 4096), BS, 0, s>>>(total, L.sell_val.p, L.sell_val32.p);
            } else {
                k_diag_comp_sell<<<grid_for(L.n, BS, 4096), BS, 0, s>>>(L.n, L.sell_off.p, L.sell_val.p,
                                                                        L.pdiag.p, L.pdiag32.p);
                L.sell_bf16.alloc(std::max(total, 1));
                k_to_bf16<<<grid_for(total, BS, 4096), BS, 0, s>>>(total, L.sell_val.p, L.sell_bf16.p);
            }
            L.mixed_ready_fast = true;
        }
        FVM_HIP(hipGetLastError());
    }
    FVM_HIP(hipStreamSynchronize(s));
}

// FP64 inverse of the first coarse level with at most 256 rows (coarsest solve of the
// mixed-precision cycle). Falls back to smoothing if that matrix is singular.
void prepare_direct(fvm_amg *h) {
    hipStream_t s = h->stream;
    h->direct_level = -1;
    h->direct_tried = true;   // a singular level is not factored again until the next setup
    int lv = -1;
    for (size_t l = 1; l < h->lev.size(); l++)
        if (h->lev[l].n <= 256 && h->lev[l].n > 0) { lv = (int)l; break; }
    if (lv < 0) return;
    Level &L = h->lev[lv];
    const int n = L.n;
    DArr<double> A((size_t)n * n);
    h->dense_inv.alloc((size_t)n * n);
    DArr<int> st(1);
    FVM_HIP(hipMemsetAsync(st.p, 0, sizeof(int), s));
    k_dense_from_csr<<<1, 256, 0, s>>>(n, L.row.p, L.col.p, L.diag.p, L.off.p, A.p);
    k_gauss_jordan<<<1, 1024, 0, s>>>(n, A.p, h->dense_inv.p, st.p);
    FVM_HIP(hipGetLastError());
    h->direct_attempts++;
    if (read_scalar(st.p, s) == 0) h->direct_level = lv;
}

// ====================================================================== batched hierarchy
namespace {

// system of combined row i: the k with off[k] <= i < off[k+1] (empty systems skipped)
__device__ __forceinline__ int find_sys(const int *off, int K, int i) {
    int lo = 0, hi = K;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_batch_validate(int total, int K, const BatchSys *sys, const int *off, int *flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int k = find_sys(off, K, i);
        const BatchSys S = sys[k];
        const int li = i - off[k];
        const int b = S.row[li], e = S.row[li + 1];
        if (b < 0 || e < b) { atomicOr(flag, 1); continue; }
        for (int p = b; p < e; p++) {
            const int c = S.col[p];
            if (c < 0 || c >= S.n_total || c == li) atomicOr(flag, 2);
        }
        const double d = S.diag[li];
        if (!(d != 0.0) || !isfinite(d)) atomicOr(flag, 4);
    }
}

__global__ void k_batch_compact_count(int total, int K, const BatchSys *sys, const int *off, int *cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int k = find_sys(off, K, i);
        const BatchSys S = sys[k];
        const int li = i - off[k];
        int c = 0;
        for (int p = S.row[li]; p < S.row[li + 1]; p++) c += (S.col[p] < S.n_self);
        cnt[i] = c;
    }
}

__global__ void k_batch_compact_fill(int total, int K, const BatchSys *sys, const int *off, const int *nrow,
                                     int *ncol, double *noff, double *ndiag, uint8_t *nbnd) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int k = find_sys(off, K, i);
        const BatchSys S = sys[k];
        const int li = i - off[k];
        int q = nrow[i];
        for (int p = S.row[li]; p < S.row[li + 1]; p++) {
            const int c = S.col[p];
            if (c < S.n_self) { ncol[q] = c + off[k]; noff[q] = S.off[p]; q++; }
        }
        ndiag[i] = S.diag[li];
        if (nbnd) nbnd[i] = S.bnd ? S.bnd[li] : 0;
    }
}

// rows of systems that no longer coarsen take no part in the agglomeration
__global__ void k_batch_mark(int n, int K, const int *off, const uint8_t *stopped, uint8_t *bnd) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (stopped[find_sys(off, K, i)]) bnd[i] = 1;
}

__global__ void k_batch_group_range(int n, int K, const int *off, const int *ci, int *gmin, int *gmax) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int I = ci[i];
        if (I < 0) continue;
        const int k = find_sys(off, K, i);
        atomicMin(&gmin[k], I);
        atomicMax(&gmax[k], I);
    }
}

__global__ void k_batch_renumber(int n, int K, const int *off, const uint8_t *drop, const int *shift, int *ci) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int I = ci[i];
        if (I < 0) continue;
        const int k = find_sys(off, K, i);
        ci[i] = drop[k] ? -1 : I - shift[k];
    }
}

}  // namespace

// Block-diagonal hierarchy of B.K systems with each system's own stop rule
// (AMG.cpp:149-210 applied per system). Level l holds the level-l rows of every
// system whose hierarchy reaches l, in system order.
void build_hierarchy_batched(fvm_amg *h, BatchInfo &B) {
    const double t0 = now_s();
    hipStream_t s = h->stream;
    const int K = B.K;
    FVM_CHECK(K > 0, FVM_E_INVALID, "empty batch");
    FVM_CHECK(h->prm.coarseGroupSize == 2, FVM_E_INVALID, "the GPU agglomeration supports coarseGroupSize == 2 only");
    h->lev.clear();
    h->ready = false;
    h->batch = nullptr;
    h->direct_level = -1;
    h->direct_tried = false;
    h->tail_prec = -1;
    h->tail_desc.release();
    invalidate_cycle_graphs(h);
    std::vector<int> off0(K + 1, 0);
    bool any_bnd = false;
    for (int k = 0; k < K; k++) {
        const BatchSys &S = B.sys[k];
        FVM_CHECK(S.n_self >= 0 && S.n_total >= S.n_self, FVM_E_INVALID,
                  "system " + std::to_string(k) + ": need 0 <= n_self <= n_total");
        FVM_CHECK(S.n_self == 0 || (S.row && S.col && S.diag && S.off), FVM_E_INVALID,
                  "system " + std::to_string(k) + ": null array");
        FVM_CHECK((long long)off0[k] + S.n_self < INT_MAX, FVM_E_INVALID, "batch exceeds 2^31 rows");
        off0[k + 1] = off0[k] + S.n_self;
        any_bnd = any_bnd || S.bnd;
    }
    const int total = off0[K];
    h->n_self = total;
    h->n_total = total;
    B.d_sys.alloc(K);
    FVM_HIP(hipMemcpyAsync(B.d_sys.p, B.sys.data(), sizeof(BatchSys) * K, hipMemcpyHostToDevice, s));
    B.d_off0.alloc(K + 1);
    FVM_HIP(hipMemcpyAsync(B.d_off0.p, off0.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice, s));
    {
        DArr<int> flag(1);
        FVM_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
        if (total > 0)
            k_batch_validate<<<grid_for(total, BS, 4096), BS, 0, s>>>(total, K, B.d_sys.p, B.d_off0.p, flag.p);
        FVM_HIP(hipGetLastError());
        const int f = read_scalar(flag.p, s);
        FVM_CHECK(!(f & 1), FVM_E_INVALID, "row pointers are not a valid CSR");
        FVM_CHECK(!(f & 2), FVM_E_INVALID, "column index out of range or on the diagonal");
        FVM_CHECK(!(f & 4), FVM_E_NUMERIC, "zero or non-finite diagonal on a self row");
    }
    h->lev.emplace_back();
    {
        Level &F = h->lev[0];
        F.n = total;
        DArr<int> cnt((size_t)total + 1);
        FVM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
        if (total > 0)
            k_batch_compact_count<<<grid_for(total, BS, 4096), BS, 0, s>>>(total, K, B.d_sys.p, B.d_off0.p, cnt.p);
        F.row.alloc((size_t)total + 1);
        h->ws.scan.exclusive(cnt.p, F.row.p, total + 1, s);
        F.nnz = read_scalar(F.row.p + total, s);
        F.col.alloc(std::max<long long>(F.nnz, 1));
        F.off.alloc(std::max<long long>(F.nnz, 1));
        F.diag.alloc(std::max(total, 1));
        if (any_bnd) F.bnd.alloc(std::max(total, 1));
        if (total > 0)
            k_batch_compact_fill<<<grid_for(total, BS, 4096), BS, 0, s>>>(total, K, B.d_sys.p, B.d_off0.p, F.row.p,
                                                                         F.col.p, F.off.p, F.diag.p,
                                                                         any_bnd ? F.bnd.p : nullptr);
        FVM_HIP(hipGetLastError());
    }
    B.off.assign(1, off0);
    B.depth.assign(K, 1);
    std::vector<long long> fine_size(K);
    std::vector<char> active(K);
    for (int k = 0; k < K; k++) {
        fine_size[k] = B.sys[k].n_total;   // getLocalSize() of the finest level counts ghosts
        active[k] = B.sys[k].n_self > 0;
    }
    DArr<int> d_off(K + 1), gmin(K), gmax(K), shift(K);
    DArr<uint8_t> flags(K);
    std::vector<int> hmin(K), hmax(K), hshift(K);
    std::vector<uint8_t> hflag(K);
    for (int lvl = 0; lvl < h->prm.maxCoarseLevels; lvl++) {
        if (std::none_of(active.begin(), active.end(), [](char a) { return a != 0; })) break;
        Level &A = h->lev.back();
        const std::vector<int> offl = B.off[lvl];
        FVM_HIP(hipMemcpyAsync(d_off.p, offl.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice, s));
        bool any_stopped = false;
        for (int k = 0; k < K; k++) {
            hflag[k] = (!active[k] && offl[k + 1] > offl[k]) ? 1 : 0;
            any_stopped = any_stopped || hflag[k];
        }
        if (any_stopped) {
            if (!A.bnd.p) {
                A.bnd.alloc(std::max(A.n, 1));
                FVM_HIP(hipMemsetAsync(A.bnd.p, 0, A.bnd.bytes(), s));
            }
            FVM_HIP(hipMemcpyAsync(flags.p, hflag.data(), K, hipMemcpyHostToDevice, s));
            k_batch_mark<<<grid_for(A.n, BS, 4096), BS, 0, s>>>(A.n, K, d_off.p, flags.p, A.bnd.p);
        }
        coarsen_level(h, A, h->prm.weightRatioThreshold);
        if (A.nc == 0) { A.ci.release(); break; }
        k_fill_int<<<grid_for(K, BS), BS, 0, s>>>(K, gmin.p, INT_MAX);
        k_fill_int<<<grid_for(K, BS), BS, 0, s>>>(K, gmax.p, -1);
        k_batch_group_range<<<grid_for(A.n, BS, 4096), BS, 0, s>>>(A.n, K, d_off.p, A.ci.p, gmin.p, gmax.p);
        FVM_HIP(hipGetLastError());
        FVM_HIP(hipMemcpyAsync(hmin.data(), gmin.p, sizeof(int) * K, hipMemcpyDeviceToHost, s));
        FVM_HIP(hipMemcpyAsync(hmax.data(), gmax.p, sizeof(int) * K, hipMemcpyDeviceToHost, s));
        FVM_HIP(hipStreamSynchronize(s));
        std::vector<int> nck(K, 0);
        bool any_drop = false;
        int dropped = 0;
        for (int k = 0; k < K; k++) {
            nck[k] = hmax[k] >= 0 ? hmax[k] - hmin[k] + 1 : 0;
            hflag[k] = 0;
            hshift[k] = dropped;
            if (!active[k]) continue;
            if (nck[k] == fine_size[k] || (!h->prm.parallelStopRule && nck[k] <= 3)) {   // AMG.cpp:163-168
                hflag[k] = 1;
                any_drop = true;
                dropped += nck[k];
                active[k] = 0;
            }
        }
        if (any_drop) {
            FVM_HIP(hipMemcpyAsync(flags.p, hflag.data(), K, hipMemcpyHostToDevice, s));
            FVM_HIP(hipMemcpyAsync(shift.p, hshift.data(), sizeof(int) * K, hipMemcpyHostToDevice, s));
            k_batch_renumber<<<grid_for(A.n, BS, 4096), BS, 0, s>>>(A.n, K, d_off.p, flags.p, shift.p, A.ci.p);
            FVM_HIP(hipGetLastError());
            A.nc -= dropped;
        }
        if (A.nc == 0) { A.ci.release(); break; }
        coarse_to_fine(h, A);
        Level C;
        coarse_operator(h, A, C);
        h->lev.push_back(std::move(C));
        std::vector<int> offn(K + 1, 0);
        for (int k = 0; k < K; k++) {
            const bool kept = active[k] && !hflag[k];
            offn[k + 1] = offn[k] + (kept ? nck[k] : 0);
            if (kept) {
                fine_size[k] = nck[k];
                B.depth[k] = lvl + 2;
                if (h->prm.parallelStopRule && nck[k] <= 3) active[k] = 0;   // AMG.cpp:180-181
            }
        }
        FVM_CHECK(offn[K] == h->lev.back().n, FVM_E_INVALID, "batched coarse level size mismatch");
        B.off.push_back(offn);
    }
    for (auto &L : h->lev) alloc_vectors(L);
    h->ws.partials.ensure(4096);
    h->ws.scal.ensure(64);
    B.d_off.clear();
    for (size_t l = 0; l < B.off.size(); l++) {
        B.d_off.emplace_back(K + 1);
        FVM_HIP(hipMemcpyAsync(B.d_off.back().p, B.off[l].data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice, s));
    }
    FVM_HIP(hipStreamSynchronize(s));
    h->ready = true;
    h->batch = &B;
    if (h->prm.gsOrdering == FVM_GS_MULTICOLOR) {
        prepare_fast(h);
    } else if (h->prm.smootherType == FVM_GAUSS_SEIDEL) {
        for (auto &L : h->lev) prepare_exact_schedule(h, L);
    }
    FVM_HIP(hipStreamSynchronize(s));
    h->setup_seconds = now_s() - t0;
}

}  // namespace fvm
