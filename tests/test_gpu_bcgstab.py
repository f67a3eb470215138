"""Device BiCGStab (fvm_bcgstab_solve_device, BCGStab.cpp:27-171) and CG
(fvm_cg_solve_device, CG.cpp:25-137) against their restatements in oracle/krylov_ref.py, with the oracle AMG's smooth as its
preconditioner. In the reference-order FP64 mode the preconditioner is bit-identical
on both sides; the dot products and 1-norms are summed in a different order, so the
bar is: the same iteration count, every history entry within 1e-7 relative, and the
solution within 1e-7 of its max-norm. The coloured and mixed-precision
preconditioners are different smoothers: they must reach the same stop rule."""
import os
import sys

import numpy as np
import pytest

import matgen
import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import krylov_ref as K  # noqa: E402

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _system(kind):
    if kind == "kat226":
        return O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    if kind == "cav32":
        return O.load_mm(os.path.join(G, "cav32_matrix_mesh0.mat"), os.path.join(G, "cav32_matrix.rhs"))
    return matgen.grid3d(12, seed=6, ghosts=False, contrast=1.5)


def _dev(a, dt):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda:0")


def _gpu(m, gs, prec):
    import torch
    import fvm_amd
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.gsOrdering = gs
    amg.precision = prec
    amg.setup_device(m["n_self"], m["n_total"], _dev(m["row"], torch.int32), _dev(m["col"], torch.int32),
                     _dev(m["diag"], torch.float64), _dev(m["off"], torch.float64))
    return amg


def _oracle(m, method=None):
    method = method or K.bcgstab
    n = m["n_self"]
    oa = O.OracleAMG(n_max_iterations=1)
    oa.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    smooth = lambda b: oa.smooth(np.concatenate([b, np.zeros(m["n_total"] - n)]), np.zeros(m["n_total"]))[:n]
    matvec = lambda v: O.residual(m, np.concatenate([v, np.zeros(m["n_total"] - n)]), np.zeros(n))
    return method(matvec, smooth, m["b"][:n], np.zeros(n), n_max=100, rel_tol=1e-8)


@pytest.mark.parametrize("kind", ["kat226", "cav32", "grid"])
@pytest.mark.parametrize("method", ["bcgstab", "cg"])
def test_reference_order_matches_oracle(kind, method):
    import torch
    import fvm_amd
    m = _system(kind)
    n = m["n_self"]
    ref = _oracle(m, K.bcgstab if method == "bcgstab" else K.cg)
    amg = _gpu(m, 0, 0)
    solver = (fvm_amd.BCGStab if method == "bcgstab" else fvm_amd.CG)(amg)
    x = torch.zeros(m["n_total"], dtype=torch.float64, device="cuda:0")
    r0, rf, it = solver.solve_device(_dev(m["b"], torch.float64), x)
    print(kind, method, "iterations", it, ref["iterations"], "history", solver.history(), ref["history"], flush=True)
    np.testing.assert_allclose(r0, ref["r0"], rtol=1e-12)
    assert it == ref["iterations"]
    np.testing.assert_allclose(solver.history(), ref["history"], rtol=1e-7)
    xg = x.cpu().numpy()[:n]
    assert np.abs(xg - ref["x"]).max() <= 1e-7 * np.abs(ref["x"]).max()
    # the final true residual is small: compared against r0
    assert abs(rf - ref["rfinal"]) <= 1e-10 * r0
    assert solver.history()[-1] / r0 < 1e-8 and rf / r0 < 1e-7


@pytest.mark.parametrize("prec", [0, 1])
def test_bcgstab_coloured_reaches_stop_rule(prec):
    import torch
    import fvm_amd
    m = _system("grid")
    amg = _gpu(m, 1, prec)
    solver = fvm_amd.BCGStab(amg)
    # the FP32 cycle is a slightly inexact preconditioner: a looser goal in mixed precision
    tol = 1e-9 if prec == 0 else 1e-6
    solver.relativeTolerance = tol
    x = torch.zeros(m["n_total"], dtype=torch.float64, device="cuda:0")
    r0, rf, it = solver.solve_device(_dev(m["b"], torch.float64), x)
    h = solver.history()
    print("prec", prec, "iterations", it, "history", h, "true", rf / r0, flush=True)
    assert len(h) == it + 1 and h[-1] / r0 < tol
    assert it < 40
    assert rf / r0 < 10 * tol
    # a second solve on the same handle gives the same bits
    x2 = torch.zeros_like(x)
    r0b, rfb, itb = solver.solve_device(_dev(m["b"], torch.float64), x2)
    assert (r0b, rfb, itb) == (r0, rf, it) and torch.equal(x, x2)


@pytest.mark.parametrize("cls", ["BCGStab", "CG"])
def test_host_solve_and_reference_methods(cls):
    """solve(ls) on a host LinearSystem gives the device solve's result; the
    reference's getTotalIterations (BCGStab.h:31) adds up, cleanup reaches the
    preconditioner, and smooth refuses (BCGStab.cpp:172-176)."""
    import torch
    import fvm_amd
    m = _system("kat226")
    ls = fvm_amd.LinearSystem(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"], m["b"])
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    solver = getattr(fvm_amd, cls)(amg)
    r0 = solver.solve(ls)
    it1 = solver.lastIterations
    amg2 = _gpu(m, 0, 0)
    ref = getattr(fvm_amd, cls)(amg2)
    x = torch.zeros(m["n_total"], dtype=torch.float64, device="cuda:0")
    r0d, _, itd = ref.solve_device(_dev(m["b"], torch.float64), x)
    assert r0 == r0d and it1 == itd
    assert np.array_equal(ls.delta[:m["n_self"]], x.cpu().numpy()[:m["n_self"]])
    ls.delta[:] = 0
    solver.solve(ls)
    assert solver.getTotalIterations() == 2 * it1
    solver.cleanup()
    with pytest.raises(RuntimeError):
        solver.smooth(ls)
