"""Pins of the oracle's pieces that the known-answer tests do not isolate:
closed forms, textbook definitions and hand-derived maps, chosen so that a
dropped term, a wrong sign or index, or a transposed operand fails."""
import numpy as np
import pytest

import matgen
import oracle_lib as O


def test_residual_is_b_plus_Ax():
    m = matgen.grid3d(5, seed=1)
    A = matgen.dense(m)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(m["n_total"])
    x[m["n_self"]:] = 0.0
    r = O.residual(m, x, m["b"])
    ref = (m["b"] + A @ x)[: m["n_self"]]
    np.testing.assert_allclose(r, ref, rtol=1e-13, atol=1e-13)


def test_forward_and_reverse_gs_are_triangular_solves():
    # forward GS: x_new = -(D+L)^{-1} (b + U x_old); reverse: -(D+U)^{-1}(b + L x_old)
    m = matgen.grid3d(4, seed=2, ghosts=False)
    A = matgen.dense(m)
    n = m["n_self"]
    rng = np.random.default_rng(5)
    x0 = rng.standard_normal(n)
    L = np.tril(A, -1); U = np.triu(A, 1); D = np.diag(np.diag(A))
    xf = O.gs(m, x0, m["b"])
    np.testing.assert_allclose(xf, -np.linalg.solve(D + L, m["b"] + U @ x0), rtol=1e-11, atol=1e-12)
    xr = O.gs(m, x0, m["b"], reverse=True)
    np.testing.assert_allclose(xr, -np.linalg.solve(D + U, m["b"] + L @ x0), rtol=1e-11, atol=1e-12)


def test_chain_pairs_equal_weights():
    m = matgen.chain(9)
    nc, ci = O.coarsen(m)
    # (0,1),(2,3),(4,5),(6,7) and 8 joins its grouped neighbour's group (count 2 <= 4)
    assert nc == 4
    assert ci.tolist() == [0, 0, 1, 1, 2, 2, 3, 3, 3]


def test_star_crowded_group_rule():
    # star: centre 0 with leaves 1..7, strong centre-leaf weights.
    # seed 0 pairs leaf 1; each later leaf sees only the grouped centre and joins
    # while the group count <= groupSize+2 = 4, then starts singletons.
    n = 8
    row = [0]; col = []; off = []
    for i in range(n):
        nb = list(range(1, n)) if i == 0 else [0]
        col += nb; off += [1.0] * len(nb); row.append(len(col))
    m = dict(n_self=n, n_total=n, row=np.array(row, np.int32), col=np.array(col, np.int32),
             diag=np.full(n, -7.0), off=np.array(off))
    nc, ci = O.coarsen(m)
    assert ci.tolist() == [0, 0, 0, 0, 0, 1, 2, 3]
    assert nc == 4


def test_boundary_rows_and_ghosts_stay_unassigned():
    m = matgen.grid3d(4, seed=4)
    isb = np.zeros(m["n_total"], np.uint8)
    isb[[0, 5, 17]] = 1
    nc, ci = O.coarsen(m, is_bnd=isb)
    assert (ci[m["n_self"]:] == -1).all()
    assert ci[0] == ci[5] == ci[17] == -1
    sel = np.ones(m["n_self"], bool); sel[[0, 5, 17]] = False
    c = ci[: m["n_self"]][sel]
    assert (c >= 0).all() and c.max() == nc - 1
    counts = np.bincount(c, minlength=nc)
    assert counts.min() >= 1 and counts.max() <= 2 + 3
    # ids appear in increasing order of their first (seed) row
    first = [np.argmax(ci == g) for g in range(nc)]
    assert first == sorted(first)


@pytest.mark.parametrize("seed", [0, 1])
def test_galerkin_equals_dense_ptap(seed):
    m = matgen.grid3d(5, seed=seed)
    a = O.OracleAMG(max_coarse_levels=1)
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    L0, L1 = a.level(0), a.level(1)
    ci = L0["ci"]
    n, nc = m["n_self"], L1["n_self"]
    P = np.zeros((m["n_total"], nc))
    for i in range(n):
        if ci[i] >= 0:
            P[i, ci[i]] = 1.0
    Ac = P.T @ matgen.dense(m) @ P
    C = matgen.dense(dict(L1, n_total=nc))
    np.testing.assert_allclose(C, Ac, rtol=1e-12, atol=1e-12)
    # first-appearance column order inside coarse rows
    for I in range(nc):
        seen = []
        for i in np.nonzero(ci[:n] == I)[0]:
            for p in range(m["row"][i], m["row"][i + 1]):
                J = ci[m["col"][p]]
                if J >= 0 and J != I and J not in seen:
                    seen.append(J)
        assert L1["col"][L1["row"][I]:L1["row"][I + 1]].tolist() == seen


def test_two_level_cycle_matches_dense_algebra():
    # one V(0,1) cycle with two levels: x1 = P A_c^{-1}-free form computed densely
    m = matgen.grid3d(3, seed=7, ghosts=False)
    a = O.OracleAMG(max_coarse_levels=1, parallel_stop_rule=1)
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    assert a.num_levels() == 2
    L1 = a.level(1)
    ci = a.level(0)["ci"]
    A = matgen.dense(m); n = m["n_self"]; nc = L1["n_self"]
    b = m["b"]
    # finest: x=0, r=b; coarse: b_c = P^T r; coarse level (coarsest) gets 0+1 sweeps of symmetric GS from 0
    P = np.zeros((n, nc)); P[np.arange(n), ci] = 1.0
    Ac = P.T @ A @ P
    bc = P.T @ b
    Dc = np.diag(np.diag(Ac)); Lc = np.tril(Ac, -1); Uc = np.triu(Ac, 1)
    xc = -np.linalg.solve(Dc + Lc, bc)
    xc = -np.linalg.solve(Dc + Uc, bc + Lc @ xc)
    x = P @ xc
    D = np.diag(np.diag(A)); L = np.tril(A, -1); U = np.triu(A, 1)
    x = -np.linalg.solve(D + L, b + U @ x)
    x = -np.linalg.solve(D + U, b + L @ x)
    got = a.smooth(b, np.zeros(n))
    np.testing.assert_allclose(got, x, rtol=1e-10, atol=1e-12)


def test_converged_solution_solves_system():
    m = matgen.grid3d(6, seed=9)
    a = O.OracleAMG(relative_tolerance=1e-12, n_max_iterations=2000)
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    res = a.solve(m["b"])
    A = matgen.dense(m)
    n = m["n_self"]
    xs = -np.linalg.solve(A[:n, :n], m["b"][:n])
    np.testing.assert_allclose(res["x"][:n], xs, rtol=1e-9, atol=1e-10)
