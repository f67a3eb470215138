"""Synthetic CRMatrix systems for tests (test-side generator).

``grid3d`` builds an N^3 hex-mesh diffusion-like matrix in the reference layout:
interior cells first (lexicographic, x fastest), one ghost row per boundary
face after them, and each cell's columns in ascending face order as
CRConnectivity::multiply produces them (interior x-, y-, z-faces, then the six
boundary groups). Ghost columns of interior rows carry the zero coefficient
that Dirichlet treatment leaves (GenericBCS.h:77-115).
"""
import numpy as np


def _face_lists(N):
    idx = lambda i, j, k: i + N * (j + N * k)
    f0, f1 = [], []
    for k in range(N):
        for j in range(N):
            for i in range(N - 1):
                f0.append(idx(i, j, k)); f1.append(idx(i + 1, j, k))
    for k in range(N):
        for j in range(N - 1):
            for i in range(N):
                f0.append(idx(i, j, k)); f1.append(idx(i, j + 1, k))
    for k in range(N - 1):
        for j in range(N):
            for i in range(N):
                f0.append(idx(i, j, k)); f1.append(idx(i, j, k + 1))
    n_int = len(f0)
    bnd = []
    for (axis, side) in [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)]:
        for a in range(N):
            for c in range(N):
                if axis == 0:
                    cell = idx(0 if side == 0 else N - 1, c, a)
                elif axis == 1:
                    cell = idx(c, 0 if side == 0 else N - 1, a)
                else:
                    cell = idx(c, a, 0 if side == 0 else N - 1)
                bnd.append(cell)
    return np.array(f0), np.array(f1), n_int, np.array(bnd)


def grid3d(N, seed=0, random_coeffs=True, ghosts=True, contrast=1.0):
    rng = np.random.default_rng(seed)
    n = N ** 3
    f0, f1, n_int, bcell = _face_lists(N)
    nb = len(bcell)
    n_total = n + (nb if ghosts else 0)
    k = np.exp(contrast * rng.standard_normal(n)) if random_coeffs else np.ones(n)
    w_int = 2 * k[f0] * k[f1] / (k[f0] + k[f1])
    if random_coeffs:
        w_int = w_int * rng.uniform(0.5, 1.5, size=len(w_int))
    # faces: interior then boundary; face f has cells (c0, c1)
    c0 = np.concatenate([f0, bcell])
    c1 = np.concatenate([f1, n + np.arange(nb)])
    coef = np.concatenate([w_int, 2 * k[bcell]])
    # cellFaces in ascending face order -> neighbours in that order
    nf = len(c0)
    rows = [[] for _ in range(n_total)]
    vals = [[] for _ in range(n_total)]
    diag = np.zeros(n_total)
    for f in range(nf):
        a, b, w = c0[f], c1[f], coef[f]
        if b >= n and not ghosts:
            diag[a] -= w
            continue
        rows[a].append(b); vals[a].append(w if b < n else 0.0)
        rows[b].append(a); vals[b].append(0.0 if b >= n else w)
        diag[a] -= w
        if b < n:
            diag[b] -= w
    for g in range(n, n_total):
        diag[g] = -1.0   # Dirichlet ghost row: identity
    row = np.zeros(n_total + 1, np.int32)
    row[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.int32)
    off = np.array([v for r in vals for v in r], np.float64)
    b = np.zeros(n_total)
    b[:n] = rng.uniform(-1, 1, n)
    return dict(n_self=n, n_total=n_total, row=row, col=col, diag=diag, off=off, b=b)


def chain(n, w=1.0, d=None):
    """1-D chain i -- i+1 with equal weights: pairwise agglomeration gives (0,1),(2,3),..."""
    row = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        cols += nb
        vals += [w] * len(nb)
        row[i + 1] = row[i] + len(nb)
    diag = np.full(n, -2.0 * w if d is None else d)
    return dict(n_self=n, n_total=n, row=row, col=np.array(cols, np.int32), diag=diag,
                off=np.array(vals), b=np.ones(n))


def dense(m):
    n = m["n_self"]
    A = np.zeros((m["n_total"], m["n_total"]))
    A[np.arange(m["n_total"]), np.arange(m["n_total"])] = m["diag"]
    for i in range(m["n_total"]):
        for p in range(m["row"][i], m["row"][i + 1]):
            A[i, m["col"][p]] += m["off"][p]
    return A
