"""Closed-form pins of the diffusion assembly oracle (oracle/fv_oracle.c) that a
plausible mistake in the face conductivity or in the non-orthogonal metric fails.

1. Layered conductivity (DiffusionDiscretization.h:19-27 harmonicAverage, :172-185):
   a uniform N^3 mesh with k constant on each x-layer of cells, T = 1 on x = 0,
   T = 0 on x = 1, zero flux on the other walls. The finite-volume system is
   then a chain of series resistances and its cell-centre temperatures equal
   the exact 1-D solution T(x_i) = 1 - q (sum_{j<i} h/k_j + h/(2 k_i)),
   q = 1 / sum_j h/k_j. Only the harmonic face mean reproduces it: an
   arithmetic mean of the two cell conductivities gives another network.
2. Affinely sheared mesh (DiffusionDiscretization.h:187-189,
   MeshMetricsCalculator_impl.h:239-305): the image of the uniform mesh under
   x -> M x has parallelepiped cells. An interior face normal to axis a has the
   area vector A = h^2 cof(M) e_a and the centroid step ds = h M e_a, so the
   coefficient k |A|^2 / (A . ds) is k h det(M) |M^-T e_a|^2, derived here from
   M alone. A Dirichlet face has half the step and twice the coefficient. The
   shortcut |A| / |ds| agrees with it only on orthogonal meshes.
"""
import numpy as np
import pytest

import hexmesh
import oracle_lib as O

D, NM = 0, 1
BC_T = [0, D, D, NM, NM, NM, NM]          # x0, x1 Dirichlet; y, z zero-flux
BC_V = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def _ghost_copy(mesh, v):
    fcl = mesh["face_cells"].reshape(-1, 2)
    for f in range(mesh["group_ptr"][1], fcl.shape[0]):
        v[fcl[f, 1]] = v[fcl[f, 0]]
    return v


def _dense_interior(S):
    nc = S["n_self"]
    A = np.zeros((nc, nc))
    A[np.arange(nc), np.arange(nc)] = S["diag"][:nc]
    for c in range(nc):
        for p in range(S["row"][c], S["row"][c + 1]):
            if S["col"][p] < nc:
                A[c, S["col"][p]] += S["off"][p]
    return A


def _entry(S, i, j):
    for p in range(S["row"][i], S["row"][i + 1]):
        if S["col"][p] == j:
            return S["off"][p]
    raise KeyError((i, j))


@pytest.mark.parametrize("klayers", [[1.0, 10.0, 0.1, 3.0, 0.5, 7.0], [2.0, 2.0, 1e-3, 1e3, 4.0, 0.25]])
def test_layered_conductivity_series_solution(klayers):
    N = len(klayers)
    h = 1.0 / N
    m = hexmesh.hex_mesh(N)
    nc = m["n_cells"]
    nct = nc + m["n_ghost"]
    i_of = np.arange(nc) % N
    k = np.zeros(nct)
    k[:nc] = np.asarray(klayers)[i_of]
    _ghost_copy(m, k)
    S = O.fv_thermal(m, k, np.zeros(nct), np.zeros(nct), BC_T, BC_V)
    # delta form: A T = -b with T = delta (the current field is zero)
    T = np.linalg.solve(_dense_interior(S), -S["b"][:nc])
    R = h / np.asarray(klayers)                    # resistance of each layer per unit area
    q = 1.0 / R.sum()
    Tx = 1.0 - q * (np.concatenate([[0.0], np.cumsum(R)[:-1]]) + 0.5 * R)
    np.testing.assert_allclose(T, Tx[i_of], rtol=0, atol=1e-12)
    # and the heat flow through every x-section is the same
    assert np.ptp(q * np.ones(N)) == 0


def _sheared(N, M):
    m = hexmesh.hex_mesh(N)
    m["nodes"] = (m["nodes"].reshape(-1, 3) @ np.asarray(M).T).reshape(-1).copy()
    return m


@pytest.mark.parametrize("M", [
    [[1.0, 0.3, 0.2], [0.1, 0.9, 0.25], [-0.05, 0.15, 1.1]],
    [[1.0, 0.6, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
])
def test_sheared_mesh_coefficients_closed_form(M):
    N, kc = 4, 2.5
    h = 1.0 / N
    M = np.asarray(M)
    det = np.linalg.det(M)
    MinvT = np.linalg.inv(M).T
    coef_axis = [kc * h * det * MinvT[:, a].dot(MinvT[:, a]) for a in range(3)]
    m = _sheared(N, M)
    nc = m["n_cells"]; nct = nc + m["n_ghost"]
    # metrics: area vectors and volumes are the affine images
    g = O.fv_metrics(m)
    fA = g["fA"].reshape(-1, 3)
    fcl = m["face_cells"].reshape(-1, 2)
    nfx = (N - 1) * N * N
    cof = det * MinvT
    for f in range(3 * nfx):
        np.testing.assert_allclose(fA[f], h * h * cof[:, f // nfx], rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(g["vol"][:nc], h ** 3 * det, rtol=1e-13)
    k = np.full(nct, kc)
    S = O.fv_thermal(m, k, np.zeros(nct), np.zeros(nct), BC_T, BC_V)
    for f in range(3 * nfx):
        c0, c1 = fcl[f]
        want = coef_axis[f // nfx]
        assert _entry(S, c0, c1) == pytest.approx(want, rel=1e-13)
        assert _entry(S, c1, c0) == pytest.approx(want, rel=1e-13)
    # Dirichlet faces (x0, x1): coefficient 2 k h det |M^-T e_x|^2 in the diagonal and in b
    gp = m["group_ptr"]
    dsum = np.zeros(nc)
    bexp = np.zeros(nc)
    for grp, Tb in ((1, 1.0), (2, 0.0)):
        for f in range(gp[grp], gp[grp + 1]):
            c0 = fcl[f, 0]
            dsum[c0] += 2.0 * coef_axis[0]
            bexp[c0] += 2.0 * coef_axis[0] * Tb
    A = _dense_interior(S)
    rowsum = A.sum(1)
    np.testing.assert_allclose(rowsum, -dsum, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(S["b"][:nc], bexp, rtol=1e-13, atol=1e-15)
