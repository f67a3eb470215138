"""Model of the round-based exact coarsening (k_coarsen_round in
fvm_amd/csrc/setup.hip), emulated on the CPU with the kernel's snapshot
semantics, worklist and notifications. It pins the algorithm itself: on
symmetric and non-symmetric patterns, with boundary rows, the parallel rounds
give exactly the sequential greedy map of the oracle (CRMatrix.h:468-586), and
no two decisions of one round write the same state.
"""
import os

import numpy as np
import pytest

import matgen
import oracle_lib as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

INF=1<<30
def emulate(m, bnd=None, theta=0.65):
    n=m['n_self']; row=m['row']; col=m['col']; off=m['off']; diag=m['diag']
    # compact ghost cols
    R=[0]; C=[]; V=[]
    for i in range(n):
        for p in range(row[i],row[i+1]):
            if col[p]<n: C.append(col[p]); V.append(off[p])
        R.append(len(C))
    row=np.array(R); col=np.array(C); off=np.array(V)
    isb=lambda i: bnd is not None and bnd[i]
    rowsT=[[] for _ in range(n)]
    for i in range(n):
        for p in range(row[i],row[i+1]): rowsT[col[p]].append(i)
    sym = all(sorted(rowsT[i])==sorted(col[row[i]:row[i+1]].tolist()) for i in range(n))
    T=lambda i: rowsT[i]
    ldr=[-1]*n; dr=[(-1 if isb(i) else INF) for i in range(n)]
    mem=[[None]*5 for _ in range(n)]; memr=[[INF]*5 for _ in range(n)]; mcnt=[0]*n
    stamp=[0]*n; remaining=sum(1 for i in range(n) if not isb(i))
    cand=list(range(n)); t=0; evals=0
    while remaining>0:
        t+=1; nxt=[]; writes=[]
        def push(x):
            if isb(x) or dr[x]<=t or stamp[x]>=t+1: return
            stamp[x]=t+1; nxt.append(x)
        snap_dr=list(dr)
        for r in cand:
            evals+=1
            if snap_dr[r]<t: continue
            ready=True
            if not sym:
                for s in T(r):
                    if s<r and not isb(s) and not snap_dr[s]<t: ready=False
            for p in range(row[r],row[r+1]):
                c=col[p]
                if isb(c): continue
                if c<r:
                    if not snap_dr[c]<t: ready=False
                elif not snap_dr[c]<t:
                    for s in T(c):
                        if s<r and not isb(s) and not snap_dr[s]<t: ready=False
            if not ready: continue
            d0=abs(diag[r]); cu=cg=-1; mu=mg=0.0
            for p in range(row[r],row[r+1]):
                c=col[p]
                if isb(c): continue
                d1=abs(diag[c]); w=abs(abs(off[p])/(d1 if d0<d1 else d0))
                if not snap_dr[c]<t:
                    if cu==-1 or w>mu: cu,mu=c,w
                else:
                    if cg==-1 or w>mg: cg,mg=c,w
            if cu!=-1 and (cg==-1 or mu>theta*mg): act=0
            elif cg==-1: act=1
            else:
                g=ldr[cg]; cnt=0; blocked=False
                for k in range(5):
                    if not memr[g][k]<t: continue
                    cnt+=1; mm=mem[g][k]
                    for s in T(mm):
                        if s<r and not isb(s) and not snap_dr[s]<t: blocked=True
                if blocked:
                    if stamp[r]<t+1: stamp[r]=t+1; nxt.append(r)
                    continue
                act=1 if cnt>4 else 2
            writes.append((r,act,cu,cg))
        # apply (and check write conflicts)
        touched=set()
        for (r,act,cu,cg) in writes:
            if act==0:
                assert r not in touched and cu not in touched; touched|={r,cu}
                ldr[r]=r; ldr[cu]=r; mem[r][0]=r; memr[r][0]=t; mem[r][1]=cu; memr[r][1]=t; mcnt[r]=2; dr[r]=t; dr[cu]=t; remaining-=2
            elif act==1:
                assert r not in touched; touched.add(r)
                ldr[r]=r; mem[r][0]=r; memr[r][0]=t; mcnt[r]=1; dr[r]=t; remaining-=1
            else:
                g=ldr[cg]; assert ('g',g) not in touched; touched.add(('g',g))
                k=mcnt[g]; mem[g][k]=r; memr[g][k]=t; mcnt[g]=k+1; ldr[r]=g; dr[r]=t; remaining-=1
        for (r,act,cu,cg) in writes:
            for d in ([r,cu] if act==0 else [r]):
                for p in range(row[d],row[d+1]):
                    c=col[p]
                    if c>d: push(c)
                    for c2 in T(c):
                        if c2>d: push(c2)
                if not sym:
                    for c in T(d):
                        if c>d: push(c)
        cand=nxt
        assert t<4*n+64, "no termination"
    gid={}; g=0
    for i in range(n):
        if not isb(i) and ldr[i]==i: gid[i]=g; g+=1
    ci=[(-1 if isb(i) or ldr[i]<0 else gid[ldr[i]]) for i in range(n)]
    return g, np.array(ci), t, evals


def _check(m, bnd=None):
    nc, ci = O.coarsen(m, is_bnd=bnd)
    g, ci2, t, ev = emulate(m, bnd)
    assert g == nc
    assert np.array_equal(ci[: m["n_self"]], ci2)
    return t


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
