"""Single-block linear system in the reference's CRMatrix layout.

Mirrors what ``LinearSystem`` holds for a scalar model after ``initSolve``
(src/fvm/src/modules/fvmbase/LinearSystem.cpp:40-64): one CSR matrix with the
diagonal kept apart (CRMatrix.h:73-84), the right-hand side ``b`` (the
assembled residual) and the correction ``delta`` (zero after initSolve).
"""
import numpy as np


class LinearSystem:
    def __init__(self, n_self, n_total, row, col, diag, off, b, is_boundary=None, isSymmetric=False):
        self.n_self = int(n_self)
        self.n_total = int(n_total)
        self.row = np.ascontiguousarray(row, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.diag = np.ascontiguousarray(diag, dtype=np.float64)
        self.off = np.ascontiguousarray(off, dtype=np.float64)
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        self.is_boundary = None if is_boundary is None else np.ascontiguousarray(is_boundary, dtype=np.uint8)
        self.isSymmetric = bool(isSymmetric)
        self.delta = np.zeros(self.n_total, dtype=np.float64)
        if self.row.shape != (self.n_total + 1,):
            raise ValueError("row must have n_total+1 entries")
        if self.diag.shape != (self.n_total,) or self.b.shape != (self.n_total,):
            raise ValueError("diag and b must have n_total entries")
        nnz = int(self.row[-1]) if self.n_total else 0
        if self.col.shape[0] < nnz or self.off.shape[0] < nnz:
            raise ValueError("col/off shorter than row[n_total]")

    def residual(self):
        """r = b + A delta over self rows (numpy, for reporting)."""
        n = self.n_self
        rows = np.repeat(np.arange(self.n_total), np.diff(self.row))
        Ax = self.diag * self.delta
        np.add.at(Ax, rows, self.off[: len(rows)] * self.delta[self.col[: len(rows)]])
        return (self.b + Ax)[:n]


def read_matrix_market(mat_path, rhs_path):
    """MatrixMarket coordinate file + rhs -> LinearSystem, as MMReader does
    (src/fvm/src/modules/importers/MMReader.cpp:80-184): off-diagonals kept in
    file order per row, mirrored entries for 'symmetric', b = -rhs."""
    with open(mat_path) as f:
        tokens = f.read().split()
    k = 0
    h = tokens[k]; k += 1
    if h == "%%":
        if tokens[k] != "MatrixMarket":
            raise ValueError("not a MatrixMarket file")
        k += 1
    elif h != "%%MatrixMarket":
        raise ValueError("not a MatrixMarket file")
    mtype, coord, ftype, symm = tokens[k:k + 4]; k += 4
    if mtype != "matrix" or coord != "coordinate" or ftype != "real":
        raise ValueError("not a real sparse MatrixMarket matrix")
    if symm not in ("general", "symmetric"):
        raise ValueError("not symmetric or general matrix")
    sym = symm == "symmetric"
    nr, nc, nz = int(tokens[k]), int(tokens[k + 1]), int(tokens[k + 2]); k += 3
    if nr != nc:
        raise ValueError("not a square matrix")
    ent = np.array(tokens[k:k + 3 * nz], dtype=object).reshape(nz, 3)
    I = ent[:, 0].astype(np.int64) - 1
    J = ent[:, 1].astype(np.int64) - 1
    V = ent[:, 2].astype(np.float64)
    diag = np.zeros(nr)
    isd = I == J
    diag[I[isd]] = V[isd]
    oi, oj, ov = I[~isd], J[~isd], V[~isd]
    if sym:
        # each entry appends (i,j) then (j,i): interleave to keep that order
        oi = np.stack([oi, oj], 1).reshape(-1)
        oj2 = np.stack([J[~isd], I[~isd]], 1).reshape(-1)
        ov = np.repeat(ov, 2)
        oj = oj2
    order = np.argsort(oi, kind="stable")
    col = oj[order].astype(np.int32)
    off = ov[order]
    row = np.zeros(nr + 1, np.int32)
    np.add.at(row, oi + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    with open(rhs_path) as f:
        rhs = np.array(f.read().split()[:nr], dtype=np.float64)
    if rhs.shape[0] < nr:
        rhs = np.concatenate([rhs, np.zeros(nr - rhs.shape[0])])
    return LinearSystem(nr, nr, row, col, diag, off, -rhs, isSymmetric=sym)
