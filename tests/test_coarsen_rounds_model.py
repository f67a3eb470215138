"""Model of the round-based exact coarsening (k_cnt_init / k_coarsen_round /
decide_row in fvm_amd/csrc/setup.hip), emulated on the CPU with the kernels'
snapshot semantics, dependency counters and frontier. It pins the algorithm itself: on
symmetric and non-symmetric patterns, with boundary rows, the parallel rounds
give exactly the sequential greedy map of the oracle (CRMatrix.h:468-586), and
no two decisions of one round write the same state. With filt, the dependencies
that k_may_pair proves idle are dropped (a seed that can never pair the
neighbour, a row whose decision ignores the neighbour's grouping): the map must
not change, on matrices with tied weights too.
"""
import os

import numpy as np
import pytest

import matgen
import oracle_lib as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

INF = 1 << 30


def weight(off, d0, d1):
    return abs(abs(off) / (d1 if d0 < d1 else d0))


def emulate(m, bnd=None, theta=0.65, filt=False, lower=False, waits=False):
    """Count-based rounds as in k_cnt_init / k_coarsen_round / decide_row.

    waits: a blocked join waits in the list of the row it waits for, and that row
    puts it on the next frontier when it decides (the kernels' waiter lists),
    instead of returning to every frontier."""
    n = m["n_self"]
    row0, col0, off0 = m["row"], m["col"], m["off"]
    R = [0]; C = []; V = []
    for i in range(n):
        for p in range(row0[i], row0[i + 1]):
            if col0[p] < n:
                C.append(int(col0[p])); V.append(off0[p])
        R.append(len(C))
    row = R; col = C; off = V; diag = m["diag"]
    isb = lambda i: bnd is not None and bool(bnd[i])
    rowT = [[] for _ in range(n)]
    for i in range(n):
        for p in range(row[i], row[i + 1]):
            rowT[col[p]].append(i)
    sym = all(sorted(rowT[i]) == sorted(col[row[i]:row[i + 1]]) for i in range(n))
    # k_may_pair: per entry (d, c) of row d, may d pair c / does d's decision ignore c
    pos = {}
    may = [True] * len(col)
    ign = [False] * len(col)
    strong = [True] * len(col)   # (x, d), d < x: d is one of x's strongest lower neighbours
    for d in range(n):
        for p in range(row[d], row[d + 1]):
            pos[(d, col[p])] = p
        ws = [(col[p], weight(off[p], abs(diag[d]), abs(diag[col[p]]))) for p in range(row[d], row[d + 1])
              if not isb(col[p])]
        w1 = max([w for _, w in ws], default=-1.0)
        l1 = max([w for c, w in ws if c < d], default=-1.0)
        for p in range(row[d], row[d + 1]):
            c = col[p]
            if isb(d) or isb(c):
                may[p] = False
                continue
            w = weight(off[p], abs(diag[d]), abs(diag[c]))
            may[p] = not (w1 > w and theta * w1 >= w)
            ign[p] = l1 > w and theta * l1 >= w
            strong[p] = not (l1 > w)
    may_sc = lambda s, c: (not filt) or may[pos[(s, c)]]
    ign_xc = lambda x, c: filt and ign[pos[(x, c)]]
    # lower: x waits for a lower neighbour d only if d may pair x (entry (d, x)) or d
    # is one of x's strongest lower neighbours (entry (x, d)); x counts every lower
    # neighbour as grouped
    need_a = lambda d, x: (not lower) or may[pos[(d, x)]]      # d could pair x
    need_b = lambda x, d: (not lower) or strong[pos[(x, d)]]   # x may need d's group
    cnt = [0] * n
    dr = [(-1 if isb(i) else INF) for i in range(n)]
    ldr = [-1] * n
    mem = [[None] * 5 for _ in range(n)]; memr = [[INF] * 5 for _ in range(n)]; mcnt = [0] * n
    frontier = []
    remaining = 0
    for r in range(n):
        if isb(r):
            continue
        remaining += 1
        k = 0
        if not sym:
            k += sum(1 for s in rowT[r] if s < r and not isb(s) and need_a(s, r))
        for p in range(row[r], row[r + 1]):
            c = col[p]
            if isb(c):
                continue
            if c < r:
                k += need_b(r, c) + (need_a(c, r) if sym else 0)
            elif not ign_xc(r, c):
                k += sum(1 for s in rowT[c] if s < r and not isb(s) and may_sc(s, c))
        cnt[r] = k
        if k == 0:
            frontier.append(r)
    t = 0
    evals = 0

    def fan_out(d, nxt):
        for p in range(row[d], row[d + 1]):
            x = col[p]
            if x > d and not isb(x):
                # A: d could pair x (x in row(d)); symmetric patterns: B as well (d in row(x))
                dec = int(need_a(d, x)) + (int(need_b(x, d)) if sym else 0)
                if dec:
                    cnt[x] -= dec
                    if cnt[x] == 0:
                        nxt.append(x)
        if not sym:
            for x in rowT[d]:
                if x > d and not isb(x) and need_b(x, d):   # B: d in row(x)
                    cnt[x] -= 1
                    if cnt[x] == 0:
                        nxt.append(x)
        for p in range(row[d], row[d + 1]):
            c = col[p]
            if c <= d or isb(c) or not may_sc(d, c):
                continue
            for x in rowT[c]:
                if d < x < c and not isb(x) and not ign_xc(x, c):
                    cnt[x] -= 1
                    if cnt[x] == 0:
                        nxt.append(x)

    waiters = {}

    def park(r, b, nxt):
        assert not dr[b] < t and not isb(b)
        if waits:
            waiters.setdefault(b, []).append(r)
        else:
            nxt.append(r)

    def wake(d, nxt):
        nxt.extend(waiters.pop(d, []))

    while remaining > 0:
        t += 1
        assert t < 4 * n + 64, "no termination"
        nxt = []
        decisions = []
        for r in frontier:
            evals += 1
            if dr[r] < t:
                continue
            d0 = abs(diag[r]); cu = cg = -1; mu = mg = -1.0
            for p in range(row[r], row[r + 1]):
                c = col[p]
                if isb(c):
                    continue
                d1 = abs(diag[c]); w = abs(abs(off[p]) / (d1 if d0 < d1 else d0))
                if not (c < r or dr[c] < t):   # every lower row is grouped by r's turn
                    if cu == -1 or w > mu:
                        cu, mu = c, w
                else:
                    if cg == -1 or w > mg:
                        cg, mg = c, w
            if cu != -1 and (cg == -1 or mu > theta * mg):
                act = 0
            elif cg == -1:
                act = 1
            else:
                if not dr[cg] < t:   # the group of cg is not known yet
                    park(r, cg, nxt)
                    continue
                g = ldr[cg]; count = 0; blocked = False
                for k in range(5):
                    if not memr[g][k] < t:
                        continue
                    count += 1
                    for s in rowT[mem[g][k]]:
                        if s < r and not isb(s) and not dr[s] < t and not blocked:
                            blocked = True
                            blocker = s
                if blocked:
                    park(r, blocker, nxt)
                    continue
                act = 1 if count > 4 else 2
            decisions.append((r, act, cu, cg))
        touched = set()
        for (r, act, cu, cg) in decisions:
            if act == 0:
                assert r not in touched and cu not in touched
                touched |= {r, cu}
                ldr[r] = r; ldr[cu] = r; mem[r][0] = r; memr[r][0] = t; mem[r][1] = cu; memr[r][1] = t
                mcnt[r] = 2; dr[r] = t; dr[cu] = t; remaining -= 2
            elif act == 1:
                assert r not in touched
                touched.add(r)
                ldr[r] = r; mem[r][0] = r; memr[r][0] = t; mcnt[r] = 1; dr[r] = t; remaining -= 1
            else:
                g = ldr[cg]
                assert ("g", g) not in touched
                touched.add(("g", g))
                k = mcnt[g]; mem[g][k] = r; memr[g][k] = t; mcnt[g] = k + 1; ldr[r] = g; dr[r] = t
                remaining -= 1
        for (r, act, cu, cg) in decisions:
            fan_out(r, nxt)
            wake(r, nxt)
            if act == 0:
                fan_out(cu, nxt)
                wake(cu, nxt)
        assert len(nxt) == len(set(nxt)), "a row queued twice"
        frontier = nxt
    gid = {}
    g = 0
    for i in range(n):
        if not isb(i) and ldr[i] == i:
            gid[i] = g
            g += 1
    ci = [(-1 if isb(i) or ldr[i] < 0 else gid[ldr[i]]) for i in range(n)]
    assert not waiters, "rows left waiting"
    return g, np.array(ci), t, evals


def _check(m, bnd=None):
    nc, ci = O.coarsen(m, is_bnd=bnd)
    rounds = []
    for filt, lower, waits in ((False, False, False), (True, False, False), (True, True, False),
                               (True, False, True), (True, True, True)):
        g, ci2, t, ev = emulate(m, bnd, filt=filt, lower=lower, waits=waits)
        assert g == nc, (filt, lower, waits)
        assert np.array_equal(ci[: m["n_self"]], ci2), (filt, lower, waits)
        rounds.append(t)
    return rounds


def test_rounds_model_kat226_all_levels():
    m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    a = O.OracleAMG()
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    _check(m)
    for l in range(1, a.num_levels()):
        _check(a.level(l))


@pytest.mark.parametrize("trial", range(12))
def test_rounds_model_random_nonsymmetric(trial):
    rng = np.random.default_rng(100 + trial)
    n = int(rng.integers(5, 200))
    dens = rng.uniform(0.005, 0.08)
    rows = []
    for i in range(n):
        k = int(rng.binomial(n, dens))
        rows.append([int(j) for j in rng.choice(n, size=min(k, n - 1), replace=False) if j != i])
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.int32)
    vals = rng.choice([0.5, 1.0, 2.0], size=len(col)) if trial % 2 else rng.uniform(0.1, 2, len(col))
    diag = -np.array([np.abs(vals[row[i]:row[i + 1]]).sum() + rng.uniform(0, 1) for i in range(n)])
    m = dict(n_self=n, n_total=n, row=row, col=col, diag=diag, off=np.array(vals, np.float64))
    bnd = (rng.uniform(size=n) < 0.1).astype(np.uint8) if trial % 3 == 0 else None
    _check(m, bnd)


def test_rounds_model_grid_with_ghosts():
    _check(matgen.grid3d(7, seed=3))


@pytest.mark.parametrize("seed", range(3))
def test_rounds_model_filters_shorten_coarse_levels(seed):
    """A lognormal-coefficient grid and its coarse levels: the filtered
    dependencies give the oracle's maps in no more rounds, and fewer on the
    coarse levels, where chains of 'could pair first' dependencies dominate."""
    m = matgen.grid3d(9, seed=40 + seed, contrast=2.5, ghosts=False)
    a = O.OracleAMG()
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    fewer = 0
    for l in range(min(4, a.num_levels())):
        full, filtered, lower, _, _ = _check(a.level(l))
        print("level", l, "rounds", full, filtered, lower)
        assert filtered <= full
        fewer += filtered < full
    assert fewer >= 1


@pytest.mark.parametrize("trial", range(8))
def test_rounds_model_filters_with_tied_symmetric_weights(trial):
    """Symmetric values drawn from {0.5, 1, 2} (many exactly tied weights, where
    the first-in-CSR-order rule decides) on a random symmetric pattern."""
    rng = np.random.default_rng(300 + trial)
    n = int(rng.integers(20, 160))
    edges = {}
    for _ in range(int(n * rng.uniform(1.5, 4))):
        i, j = (int(v) for v in rng.choice(n, 2, replace=False))
        edges[(min(i, j), max(i, j))] = float(rng.choice([0.5, 1.0, 2.0]))
    rows = [[] for _ in range(n)]
    for (i, j), v in edges.items():
        rows[i].append((j, -v)); rows[j].append((i, -v))
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c, _ in r], np.int32)
    off = np.array([v for r in rows for _, v in r], np.float64)
    diag = np.array([sum(-v for _, v in r) + float(rng.choice([0.0, 0.5])) + 1e-3 for r in rows])
    m = dict(n_self=n, n_total=n, row=row, col=col, diag=diag, off=off)
    bnd = (rng.uniform(size=n) < 0.1).astype(np.uint8) if trial % 2 else None
    _check(m, bnd)
