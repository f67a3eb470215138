"""Pins of the assembly oracle (oracle/fv_oracle.c) by closed forms and
identities the mathematics fixes:
- uniform hex mesh: |A| = h^2, V = h^3, interior coefficient k h (harmonic mean
  of equal k), Dirichlet face coefficient 2 k h, zero-flux Neumann faces
  eliminate to no contribution (DiffusionDiscretization.h:200-218,
  GenericBCS.h:77-157, CRMatrix.h:899-944);
- jittered mesh: volumes sum to the domain volume, every closed cell surface
  has zero net area vector, centroids lie inside their cell's node box;
- the assembled operator is symmetric, has zero row sums away from Dirichlet
  faces, and b collects V*s plus 2 k h T_b on Dirichlet faces.
A wrong sign, a dropped Neumann elimination or a transposed face orientation
fails one of these."""
import numpy as np
import pytest

import hexmesh
import oracle_lib as O

D, NM = 0, 1
BC_T = [0, D, D, NM, NM, NM, NM]          # x0, x1 Dirichlet; y, z zero-flux
BC_V = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def test_uniform_metrics_closed_form():
    N = 4
    m = hexmesh.hex_mesh(N)
    g = O.fv_metrics(m)
    h = 1.0 / N
    np.testing.assert_allclose(g["fAm"], h * h, rtol=0, atol=1e-16)
    nc = m["n_cells"]
    np.testing.assert_allclose(g["vol"][:nc], h ** 3, rtol=1e-14)
    # centroid of cell (i,j,k) is ((i+.5)h, (j+.5)h, (k+.5)h)
    cc = g["cc"].reshape(-1, 3)[:nc]
    ijk = np.stack([np.arange(nc) % N, (np.arange(nc) // N) % N, np.arange(nc) // (N * N)], 1)
    np.testing.assert_allclose(cc, (ijk + 0.5) * h, rtol=0, atol=1e-15)


def test_face_orientation_points_from_c0_to_c1():
    m = hexmesh.hex_mesh(3, jitter=0.2, seed=5)
    g = O.fv_metrics(m)
    fA = g["fA"].reshape(-1, 3); cc = g["cc"].reshape(-1, 3)
    fcl = m["face_cells"].reshape(-1, 2)
    ds = cc[fcl[:, 1]] - cc[fcl[:, 0]]
    assert (np.einsum("ij,ij->i", fA, ds) > 0).all()


@pytest.mark.parametrize("seed", [1, 2])
def test_jittered_divergence_identities(seed):
    N = 5
    m = hexmesh.hex_mesh(N, jitter=0.2, seed=seed)
    g = O.fv_metrics(m)
    nc = m["n_cells"]
    assert abs(g["vol"][:nc].sum() - 1.0) < 1e-13
    fA = g["fA"].reshape(-1, 3)
    net = np.zeros((nc + m["n_ghost"], 3))
    fcl = m["face_cells"].reshape(-1, 2)
    np.add.at(net, fcl[:, 0], fA)
    np.add.at(net, fcl[:, 1], -fA)
    assert np.abs(net[:nc]).max() < 1e-15
    # ghost volume copies the interior neighbour; ghost centroid = face centroid
    nf = fcl.shape[0]
    for f in range(m["group_ptr"][1], nf):
        c0, c1 = fcl[f]
        assert g["vol"][c1] == g["vol"][c0]
        assert np.array_equal(g["cc"].reshape(-1, 3)[c1], g["fc"].reshape(-1, 3)[f])


def test_uniform_thermal_matrix_closed_form():
    N = 4
    h = 1.0 / N
    m = hexmesh.hex_mesh(N)
    nct = m["n_cells"] + m["n_ghost"]
    k = np.ones(nct); src = np.zeros(nct); x = np.zeros(nct)
    S = O.fv_thermal(m, k, src, x, BC_T, BC_V)
    nc = m["n_cells"]
    row, col, off, diag, b = S["row"], S["col"], S["off"], S["diag"], S["b"]
    i = np.arange(nc) % N
    for c in range(nc):
        cols = col[row[c]:row[c + 1]]
        vals = off[row[c]:row[c + 1]]
        interior = cols < nc
        np.testing.assert_allclose(vals[interior], k[0] * h, rtol=1e-15)
        assert (vals[~interior] == 0).all()
        ndir = int(i[c] == 0) + int(i[c] == N - 1)
        np.testing.assert_allclose(diag[c], -(interior.sum() * h + ndir * 2 * h), rtol=1e-14)
        np.testing.assert_allclose(b[c], 2 * h * 1.0 if i[c] == 0 else 0.0, rtol=1e-14, atol=1e-16)
    # Neumann ghost rows are flagged boundary, Dirichlet ghost rows are identity
    isb = S["is_boundary"]
    gp = m["group_ptr"]
    fcl = m["face_cells"].reshape(-1, 2)
    for f in range(gp[1], gp[3]):
        assert isb[fcl[f, 1]] == 0 and diag[fcl[f, 1]] == -1.0
    for f in range(gp[3], gp[7]):
        assert isb[fcl[f, 1]] == 1


def test_jittered_thermal_operator_properties():
    N = 5
    m = hexmesh.hex_mesh(N, jitter=0.2, seed=3)
    nc = m["n_cells"]; nct = nc + m["n_ghost"]
    rng = np.random.default_rng(0)
    k = np.exp(rng.standard_normal(nct))
    fcl = m["face_cells"].reshape(-1, 2)
    for f in range(m["group_ptr"][1], fcl.shape[0]):
        k[fcl[f, 1]] = k[fcl[f, 0]]
    src = rng.uniform(0, 1, nct); x = np.zeros(nct)
    S = O.fv_thermal(m, k, src, x, BC_T, BC_V)
    row, col, off, diag, b = S["row"], S["col"], S["off"], S["diag"], S["b"]
    A = np.zeros((nc, nc))
    dir_sum = np.zeros(nc)
    for c in range(nc):
        for p in range(row[c], row[c + 1]):
            if col[p] < nc:
                A[c, col[p]] = off[p]
    np.testing.assert_array_equal(A, A.T)      # diffusion operator is symmetric
    g = O.fv_metrics(m)
    # Dirichlet faces: row sum deficit = diffusion coefficient of the face; b = V s + coeff * T_b
    fA = g["fA"].reshape(-1, 3); cc = g["cc"].reshape(-1, 3)
    expect_b = g["vol"][:nc] * src[:nc]
    gp = m["group_ptr"]
    for grp, Tb in [(1, 1.0), (2, 0.0)]:
        for f in range(gp[grp], gp[grp + 1]):
            c0, c1 = fcl[f]
            ds = cc[c1] - cc[c0]
            coef = k[c0] * (g["fAm"][f] ** 2 / fA[f].dot(ds))
            dir_sum[c0] += coef
            expect_b[c0] += coef * Tb
    rowsum = diag[:nc] + A.sum(1)
    np.testing.assert_allclose(rowsum, -dir_sum, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(b[:nc], expect_b, rtol=1e-12, atol=1e-15)
