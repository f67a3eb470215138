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


# ----------------------------------------------------------------- gray phonon model
def _octants():
    vg = np.array([[sx, sy, sz] for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)], float) / np.sqrt(3.0)
    return vg


def _phonon_parts(mesh):
    def assemble(vg, tau, e0, e):
        return O.fv_phonon(mesh, vg, tau, e0, e)

    def solve(sys):
        oa = O.OracleAMG(n_max_iterations=400, relative_tolerance=1e-12)
        oa.setup(sys["n_self"], sys["n_total"], sys["row"], sys["col"], sys["diag"], sys["off"], sys["is_boundary"])
        res = oa.solve(sys["b"])
        return res["r0"], res["x"]
    return assemble, solve


def test_phonon_newton_update_closed_form():
    """Gray modes make sum_k e0_k(T) dk3_k / tau_k linear in T: Newton lands on
    T = (e_sum + sum_k tref_k cp_k dk3_k / tau_k) / sum_k cp_k dk3_k / tau_k."""
    rng = np.random.default_rng(3)
    K, n = 7, 50
    e = rng.uniform(0, 2, (K, n))
    dk3 = rng.uniform(0.1, 1, K); tau = rng.uniform(0.1, 10, K)
    cp = rng.uniform(0.5, 2, K); tref = rng.uniform(-1, 1, K)
    TL = np.full(n, 300.0)
    MR.phonon_update_TL(e, dk3, tau, cp, tref, TL)
    a = (cp * dk3 / tau).sum()
    want = ((e * (dk3 / tau)[:, None]).sum(0) + (tref * cp * dk3 / tau).sum()) / a
    np.testing.assert_allclose(TL, want, rtol=1e-12)


def test_phonon_equilibrium_is_a_fixed_point():
    mesh = hexmesh.hex_mesh(3, 0.15, 4)
    nct = mesh["n_cells"] + mesh["n_ghost"]
    vg = _octants(); K = len(vg)
    dk3 = np.full(K, 4 * np.pi / K); tau = np.linspace(0.5, 2, K); cp = np.ones(K); tref = np.zeros(K)
    wall = [0.0] + [300.0] * 6
    st = MR.phonon_init(nct, K, cp, tref, 300.0)
    g = O.fv_metrics(mesh)
    assemble, solve = _phonon_parts(mesh)
    norms, conv = MR.phonon_advance(assemble, solve, mesh, g, vg, dk3, tau, cp, tref, wall, st, 3, 1e-30, 0.0)
    assert max(norms) < 1e-10 * 300 * mesh["n_cells"]
    np.testing.assert_allclose(st["TL"], 300.0, rtol=1e-13)
    np.testing.assert_allclose(st["e"], 300.0, rtol=1e-13)


def test_phonon_iteration_norm_is_the_sum_of_direction_residuals():
    """PhononModel.h:557-572: the iteration's norm is the sum over the directions
    of the initial 1-norm each direction's AMG::solve returns, i.e. of ||b_k||_1."""
    mesh = hexmesh.hex_mesh(3, 0.1, 5)
    nc = mesh["n_cells"]; nct = nc + mesh["n_ghost"]
    vg = _octants(); K = len(vg)
    dk3 = np.full(K, 4 * np.pi / K); tau = np.full(K, 1.0); cp = np.ones(K); tref = np.zeros(K)
    wall = [0.0, 310.0, 290.0, 300.0, 300.0, 300.0, 300.0]
    st = MR.phonon_init(nct, K, cp, tref, 300.0)
    g = O.fv_metrics(mesh)
    MR.phonon_walls(mesh, g["fA"], g["fAm"], vg, cp, tref, wall, st["e"])
    want = sum(np.abs(O.fv_phonon(mesh, vg[k], tau[k], st["e0"][k][:nc], st["e"][k])["b"][:nc]).sum()
               for k in range(K))
    assemble, solve = _phonon_parts(mesh)
    norms, _ = MR.phonon_advance(assemble, solve, mesh, g, vg, dk3, tau, cp, tref, wall, st, 1, 1e-30, 0.0)
    np.testing.assert_allclose(norms[0], want, rtol=1e-13)


def test_phonon_slab_maximum_principle_and_mirror_symmetry():
    """Walls x0 = 310, x1 = 290, the others 300, a mirror-symmetric direction set on
    a uniform mesh: the converged lattice temperature lies between the wall
    temperatures and T - 300 is odd under x -> 1 - x."""
    N = 4
    mesh = hexmesh.hex_mesh(N)
    nc = mesh["n_cells"]; nct = nc + mesh["n_ghost"]
    vg = _octants(); K = len(vg)
    dk3 = np.full(K, 4 * np.pi / K); tau = np.full(K, 0.05); cp = np.ones(K); tref = np.zeros(K)
    wall = [0.0, 310.0, 290.0, 300.0, 300.0, 300.0, 300.0]
    st = MR.phonon_init(nct, K, cp, tref, 300.0)
    g = O.fv_metrics(mesh)
    MR.phonon_walls(mesh, g["fA"], g["fAm"], vg, cp, tref, wall, st["e"])
    assemble, solve = _phonon_parts(mesh)
    norms, conv = MR.phonon_advance(assemble, solve, mesh, g, vg, dk3, tau, cp, tref, wall, st, 200, 1e-10, 0.0)
    assert conv and norms[-1] < 1e-10 * max(norms[:5])
    T = st["TL"][:nc].reshape(N, N, N)          # [k][j][i]
    assert T.min() > 290 and T.max() < 310
    np.testing.assert_allclose(T - 300, -(T[:, :, ::-1] - 300), rtol=0, atol=1e-7)
    assert T[:, :, 0].mean() > 300 > T[:, :, -1].mean()
