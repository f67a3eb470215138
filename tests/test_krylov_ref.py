"""Pins the BiCGStab restatement of BCGStab::solve (BCGStab.cpp:27-171) in
oracle/krylov_ref.py:
- against scipy's independent BiCGSTAB with the same preconditioner: scipy solves
  A x = -b with M = -smooth, which is the reference's delta-form iteration with every
  vector negated, so the iterates must agree;
- by the invariant that the recurrence residual equals the true residual b + A x;
- on KAT-226 with the oracle AMG as preconditioner: it converges to the stop rule.
A wrong sign of alpha or omega, a swapped dot operand or a dropped beta fails the
first two."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import LinearOperator, bicgstab as scipy_bicgstab

import matgen
import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import krylov_ref as K  # noqa: E402

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _csr(m):
    n = m["n_self"]
    rows = np.repeat(np.arange(n), np.diff(m["row"][:n + 1]))
    cols = m["col"][:m["row"][n]]
    keep = cols < n
    A = sp.csr_matrix((m["off"][:m["row"][n]][keep], (rows[keep], cols[keep])), shape=(n, n))
    return (A + sp.diags(m["diag"][:n])).tocsr()


def _system(kind):
    if kind == "kat226":
        m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    else:
        m = matgen.grid3d(8, seed=4, ghosts=False, contrast=1.5)
    oa = O.OracleAMG(n_max_iterations=1)
    oa.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    n = m["n_self"]
    A = _csr(m)
    smooth = lambda b: oa.smooth(b, np.zeros(n))[:n]
    return m, A, smooth


@pytest.mark.parametrize("kind", ["kat226", "grid"])
@pytest.mark.parametrize("k", [1, 2, 5])
def test_matches_scipy_bicgstab(kind, k):
    m, A, smooth = _system(kind)
    n = m["n_self"]
    b = m["b"][:n]
    ref = K.bcgstab(lambda v: A @ v, smooth, b, np.zeros(n), n_max=k, rel_tol=0.0)
    assert ref["iterations"] == k
    M = LinearOperator((n, n), matvec=lambda v: -smooth(np.asarray(v).ravel()))
    xs, info = scipy_bicgstab(A, -b, x0=np.zeros(n), rtol=1e-300, atol=0.0, maxiter=k, M=M)
    scale = np.abs(xs).max()
    assert np.abs(ref["x"] - xs).max() <= 1e-9 * scale


@pytest.mark.parametrize("kind", ["kat226", "grid"])
def test_recurrence_residual_is_true_residual(kind):
    m, A, smooth = _system(kind)
    n = m["n_self"]
    b = m["b"][:n]
    for k in (1, 3, 6):
        ref = K.bcgstab(lambda v: A @ v, smooth, b, np.zeros(n), n_max=k, rel_tol=0.0)
        # the two differ by rounding only: measured against r0, not against the small residual
        assert abs(ref["history"][-1] - ref["rfinal"]) <= 1e-12 * ref["r0"]
        assert ref["history"][-1] < 0.5 * ref["r0"]
        np.testing.assert_allclose(ref["rfinal"], np.abs(b + A @ ref["x"]).sum(), rtol=1e-14)


def test_kat226_converges_with_oracle_amg():
    m, A, smooth = _system("kat226")
    n = m["n_self"]
    ref = K.bcgstab(lambda v: A @ v, smooth, m["b"][:n], np.zeros(n), n_max=100, rel_tol=1e-8)
    assert ref["history"][-1] / ref["r0"] < 1e-8
    assert ref["iterations"] < 40
    assert ref["rfinal"] / ref["r0"] < 1e-7


@pytest.mark.parametrize("kind", ["kat226", "grid"])
@pytest.mark.parametrize("k", [1, 2, 5])
def test_cg_matches_scipy_cg(kind, k):
    """CG::solve (CG.cpp:25-137) against scipy's CG with M = -smooth on A x = -b."""
    from scipy.sparse.linalg import cg as scipy_cg
    m, A, smooth = _system(kind)
    n = m["n_self"]
    b = m["b"][:n]
    ref = K.cg(lambda v: A @ v, smooth, b, np.zeros(n), n_max=k, rel_tol=0.0)
    assert ref["iterations"] == k
    M = LinearOperator((n, n), matvec=lambda v: -smooth(np.asarray(v).ravel()))
    xs, info = scipy_cg(A, -b, x0=np.zeros(n), rtol=1e-300, atol=0.0, maxiter=k, M=M)
    assert np.abs(ref["x"] - xs).max() <= 1e-9 * np.abs(xs).max()
    assert abs(ref["history"][-1] - ref["rfinal"]) <= 1e-12 * ref["r0"]
