"""ThermalModel::advance (ThermalModel_impl.h:424-456) on the device against the
oracle's loop (oracle/model_ref.py around the C oracle's assembly and AMG): the
temperatures after the outer iterations are bit-identical (assembly, the
reference-order AMG, postSolve's solveBoundary and the update all keep the
reference's arithmetic), and the norms agree within 1e-12 relative (parallel
vs sequential 1-norm sums)."""
import os
import sys

import numpy as np
import pytest

import hexmesh
import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import model_ref as MR  # noqa: E402
from test_model_ref import BC_T, BC_V, _oracle_parts, dirichlet_ghosts  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,jit,seed", [(5, 0.2, 3), (7, 0.1, 9)])
def test_advance_matches_oracle(N, jit, seed):
    import torch
    import fvm_amd
    from fvm_amd.thermal import Mesh, ThermalModel
    mesh = hexmesh.hex_mesh(N, jit, seed)
    nct = mesh["n_cells"] + mesh["n_ghost"]
    rng = np.random.default_rng(seed)
    k = np.exp(rng.normal(0, 1.0, nct))
    s = rng.uniform(0, 1, nct)
    assemble, solve = _oracle_parts(mesh, k, s)
    xo, norms_o = MR.thermal_advance(assemble, solve, np.zeros(nct), niter=6, rel_tol=1e-8, abs_tol=1e-20,
                                     dirichlet=dirichlet_ghosts(mesh))
    m = Mesh.hex(N, jit, seed)
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.nMaxIterations = 400
    amg.relativeTolerance = 1e-12
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda:0")
    x = torch.zeros(nct, dtype=torch.float64, device="cuda:0")
    model = ThermalModel(m, dev(k), dev(s), x, BC_T, BC_V, amg, relative_tolerance=1e-8, absolute_tolerance=1e-20)
    norms = model.advance(6)
    print("N", N, "norms", norms, norms_o, flush=True)
    assert len(norms) == len(norms_o) == model.niters
    np.testing.assert_allclose(norms, norms_o, rtol=1e-12)
    assert np.array_equal(x.cpu().numpy(), xo)


def test_advance_with_bcgstab_around_the_amg():
    """The same loop with the reference's BiCGStab around the AMG: not the same
    arithmetic as the oracle's AMG-only loop, so the bar is the physics: the outer
    loop (deferred non-orthogonal correction) converges in the same number of
    linearizations, and the temperatures agree with the oracle's to the solver
    tolerance."""
    import torch
    import fvm_amd
    from fvm_amd.thermal import Mesh, ThermalModel
    N, jit, seed = 6, 0.2, 5
    mesh = hexmesh.hex_mesh(N, jit, seed)
    nct = mesh["n_cells"] + mesh["n_ghost"]
    rng = np.random.default_rng(seed)
    k = np.exp(rng.normal(0, 1.0, nct))
    s = rng.uniform(0, 1, nct)
    assemble, solve = _oracle_parts(mesh, k, s)
    xo, norms_o = MR.thermal_advance(assemble, solve, np.zeros(nct), niter=40, rel_tol=1e-8, abs_tol=1e-20,
                                     dirichlet=dirichlet_ghosts(mesh))
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    solver = fvm_amd.BCGStab(amg)
    solver.relativeTolerance = 1e-12
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda:0")
    x = torch.zeros(nct, dtype=torch.float64, device="cuda:0")
    model = ThermalModel(Mesh.hex(N, jit, seed), dev(k), dev(s), x, BC_T, BC_V, solver)
    norms = model.advance(40)
    assert norms[-1] < 1e-8 * norms[0] and abs(len(norms) - len(norms_o)) <= 1
    xg = x.cpu().numpy()
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
