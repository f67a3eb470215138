"""Pins oracle/model_ref.py (LinearSystem::postSolve -> CRMatrix::solveBoundary,
ThermalModel::advance) with the C oracle's assembly and AMG: solveBoundary leaves
every boundary row's equation satisfied; on a jittered mesh the advance loop
converges geometrically (the secondary flux is a deferred correction from the
current field's gradient), the self rows' residual at the final temperature is
small, and Dirichlet ghosts hold their boundary values."""
import os
import sys

import numpy as np

import hexmesh
import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import model_ref as MR  # noqa: E402

D, NM = 0, 1
BC_T = [0, D, D, NM, NM, NM, NM]
BC_V = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def _oracle_parts(mesh, k, s):
    def assemble(x):
        return O.fv_thermal(mesh, k, s, x, BC_T, BC_V)

    def solve(sys):
        oa = O.OracleAMG(n_max_iterations=400, relative_tolerance=1e-12)
        oa.setup(sys["n_self"], sys["n_total"], sys["row"], sys["col"], sys["diag"], sys["off"], sys["is_boundary"])
        res = oa.solve(sys["b"])
        return res["r0"], res["x"]
    return assemble, solve


def dirichlet_ghosts(mesh):
    gp = mesh["group_ptr"]
    fc = mesh["face_cells"].reshape(-1, 2)
    idx, val = [], []
    for g in range(1, len(gp) - 1):
        if BC_T[g] == D:
            idx.append(fc[gp[g]:gp[g + 1], 1])
            val.append(np.full(gp[g + 1] - gp[g], BC_V[g]))
    return np.concatenate(idx), np.concatenate(val)


def test_advance_dirichlet_ghosts_hold_boundary_values():
    N, jit, seed = 5, 0.2, 3
    mesh = hexmesh.hex_mesh(N, jit, seed)
    nc, ng = mesh["n_cells"], mesh["n_ghost"]
    rng = np.random.default_rng(seed)
    k = np.exp(rng.normal(0, 1.0, nc + ng))
    s = rng.uniform(0, 1, nc + ng)
    x0 = np.zeros(nc + ng)
    assemble, solve = _oracle_parts(mesh, k, s)
    x, norms = MR.thermal_advance(assemble, solve, x0, niter=40, rel_tol=1e-8, abs_tol=1e-20,
                                  dirichlet=dirichlet_ghosts(mesh))
    # the non-orthogonal secondary flux enters b from the gradient of the current
    # field (deferred correction): the outer norms fall geometrically to the tolerance
    assert 2 < len(norms) < 40 and norms[-1] < 1e-8 * norms[0]
    assert all(norms[i + 1] < 0.5 * norms[i] for i in range(len(norms) - 1))
    gp = mesh["group_ptr"]
    fc = mesh["face_cells"].reshape(-1, 2)
    for g in range(1, len(gp) - 1):
        if BC_T[g] != D:
            continue
        ghosts = fc[gp[g]:gp[g + 1], 1]
        assert ghosts.min() >= nc
        np.testing.assert_allclose(x[ghosts], BC_V[g], rtol=0, atol=1e-12)
    # the self rows' residual at the final temperature is small
    sys = assemble(x)
    assert np.abs(sys["b"][:nc]).sum() < 1e-7 * norms[0]


def test_solve_boundary_is_the_ghost_row_equation():
    N, jit, seed = 4, 0.1, 8
    mesh = hexmesh.hex_mesh(N, jit, seed)
    nct = mesh["n_cells"] + mesh["n_ghost"]
    rng = np.random.default_rng(1)
    sys = O.fv_thermal(mesh, np.exp(rng.normal(0, 1, nct)), rng.uniform(0, 1, nct), rng.normal(0, 1, nct),
                       BC_T, BC_V)
    x = rng.normal(0, 1, nct)
    x[sys["n_self"]:] = 0.0
    MR.solve_boundary(sys, x)
    for nr in range(sys["n_self"], sys["n_total"]):
        if sys["is_boundary"][nr]:
            lo, hi = sys["row"][nr], sys["row"][nr + 1]
            eq = sys["b"][nr] + sys["diag"][nr] * x[nr] + np.dot(sys["off"][lo:hi], x[sys["col"][lo:hi]])
            assert abs(eq) <= 1e-12 * (abs(sys["b"][nr]) + abs(sys["diag"][nr] * x[nr]) + 1.0)
