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


def _linear_field(mesh, a, g):
    cc = O.fv_metrics(mesh)["cc"].reshape(-1, 3)
    return a + cc @ np.asarray(g)


@pytest.mark.parametrize("kind", ["jittered", "sheared"])
def test_least_squares_gradient_exact_on_linear_fields(kind):
    """GradientModel.h:128-276: the weighted least-squares gradient reproduces a
    linear field's gradient in every cell (the normal matrix is inverted exactly),
    and ghost cells carry their cell's gradient (GradientModel.h:500-515)."""
    if kind == "jittered":
        m = hexmesh.hex_mesh(5, jitter=0.2, seed=4)
    else:
        m = _sheared(4, [[1.0, 0.3, 0.2], [0.1, 0.9, 0.25], [-0.05, 0.15, 1.1]])
    g = np.array([0.7, -1.3, 2.1])
    x = _linear_field(m, 0.4, g)
    grad = O.fv_gradient(m, x)
    np.testing.assert_allclose(grad, np.broadcast_to(g, grad.shape), rtol=0, atol=1e-12 * np.abs(g).max())


def test_secondary_flux_makes_linear_fields_exact_on_sheared_meshes():
    """DiffusionDiscretization.h:187-205: with the secondary term
    gradF . k (A - ds |A|^2 / (A . ds)) the face flux of a linear field is k A . g
    exactly on a non-orthogonal mesh, so a cell surrounded by interior faces has a
    zero residual (its area vectors sum to zero). Without that term the residual
    is of the order of the skew."""
    N, kc = 6, 1.7
    M = [[1.0, 0.45, 0.2], [0.1, 0.9, 0.3], [-0.05, 0.15, 1.1]]
    m = _sheared(N, M)
    nc = m["n_cells"]; nct = nc + m["n_ghost"]
    g = np.array([0.7, -1.3, 2.1])
    x = _linear_field(m, 0.4, g)
    S = O.fv_thermal(m, np.full(nct, kc), np.zeros(nct), x, BC_T, BC_V)
    ijk = np.stack([np.arange(nc) % N, (np.arange(nc) // N) % N, np.arange(nc) // (N * N)], 1)
    inner = np.all((ijk > 0) & (ijk < N - 1), axis=1)
    h = 1.0 / N
    scale = kc * h * h * np.abs(g).sum()
    assert np.abs(S["b"][:nc][inner]).max() < 1e-12 * scale
    # the two-point part alone leaves an O(skew) residual at the same cells
    gs = O.fv_gradient(m, x)
    assert np.abs(gs[:nc] - g).max() < 1e-12


def test_upwind_convection_with_source_closed_form():
    """ConvectionDiscretization.h:160-185 upwind, SourceDiscretization.h:56, pure
    convection (k = 0) along +x (or -x) with face flux F = A . u: every x-line is
    an upwind chain F T_{i-1} - F T_i + V s = 0 from the inflow wall's value, so
    T_i = T_b + (i + 1) h s / |u|; the outflow wall's specified-temperature faces
    extrapolate (ThermalModel_impl.h:306-318)."""
    N, s0 = 5, 0.8
    h = 1.0 / N
    m = hexmesh.hex_mesh(N)
    nc = m["n_cells"]; nct = nc + m["n_ghost"]
    fA = O.fv_metrics(m)["fA"].reshape(-1, 3)
    for u, Tb in ((np.array([1.5, 0.0, 0.0]), 0.3), (np.array([-2.0, 0.0, 0.0]), 0.7)):
        F = fA @ u
        # both x walls at Tb (the outflow one extrapolates); the y and z walls are
        # specified-temperature walls too: with k = 0 and F = 0 they couple nothing
        # (zero-flux walls would leave singular ghost rows without diffusion)
        bc_t = [0, D, D, D, D, D, D]
        bc_v = [0.0, Tb, Tb, 5.0, 5.0, 5.0, 5.0]
        x = np.zeros(nct)
        fcl = m["face_cells"].reshape(-1, 2)
        gp = m["group_ptr"]
        for grp in (1, 2):
            x[fcl[gp[grp]:gp[grp + 1], 1]] = Tb            # Dirichlet ghosts hold their value
        S = O.fv_thermal(m, np.zeros(nct), np.full(nct, s0), x, bc_t, bc_v, conv_flux=F,
                         cont_res=np.zeros(nc))
        T = np.linalg.solve(_dense_interior(S), -S["b"][:nc])
        i = np.arange(nc) % N
        steps = (i + 1) if u[0] > 0 else (N - i)
        np.testing.assert_allclose(T, Tb + steps * h * s0 / abs(u[0]), rtol=1e-12, atol=1e-14)
        # the inflow wall is Dirichlet (identity ghost rows), the outflow wall extrapolates (marked)
        inflow, outflow = (1, 2) if u[0] > 0 else (2, 1)
        assert not S["is_boundary"][fcl[gp[inflow]:gp[inflow + 1], 1]].any()
        assert S["is_boundary"][fcl[gp[outflow]:gp[outflow + 1], 1]].all()


def test_continuity_residual_and_central_branch():
    """ConvectionDiscretization.h:193-197 adds the continuity residual to the
    diagonal of the self cells; the central-difference branch (:131-147) keeps the
    upwind linearization and evaluates the face value as 0.5 (x0 + x0), so with
    every face flux non-negative it is the upwind system exactly, and otherwise it
    differs only in b."""
    N = 4
    m = hexmesh.hex_mesh(N, jitter=0.15, seed=2)
    nc = m["n_cells"]; nct = nc + m["n_ghost"]
    rng = np.random.default_rng(3)
    fA = O.fv_metrics(m)["fA"].reshape(-1, 3)
    k = np.exp(rng.normal(0, 0.5, nct))
    x = rng.normal(0, 1, nct)
    cr = rng.normal(0, 0.1, nc)
    F = fA @ np.array([0.9, 0.4, 0.2])
    base = O.fv_thermal(m, k, np.zeros(nct), x, BC_T, BC_V, conv_flux=F, cont_res=np.zeros(nc))
    withc = O.fv_thermal(m, k, np.zeros(nct), x, BC_T, BC_V, conv_flux=F, cont_res=cr)
    fcl = m["face_cells"].reshape(-1, 2)
    wall_cells = set(fcl[m["group_ptr"][1]:, 0])      # boundary conditions add to these diagonals later
    plain = np.array([c not in wall_cells for c in range(nc)])
    np.testing.assert_array_equal(withc["diag"][:nc][plain], base["diag"][:nc][plain] + cr[plain])
    np.testing.assert_array_equal(withc["off"], base["off"])
    Fpos = np.abs(F)
    up = O.fv_thermal(m, k, np.zeros(nct), x, BC_T, BC_V, conv_flux=Fpos, cont_res=np.zeros(nc))
    ce = O.fv_thermal(m, k, np.zeros(nct), x, BC_T, BC_V, conv_flux=Fpos, cont_res=np.zeros(nc), central=True)
    for key in ("diag", "off", "b", "is_boundary"):
        np.testing.assert_array_equal(up[key], ce[key])
    ce2 = O.fv_thermal(m, k, np.zeros(nct), x, BC_T, BC_V, conv_flux=F, cont_res=np.zeros(nc), central=True)
    np.testing.assert_array_equal(ce2["off"], base["off"])
    np.testing.assert_array_equal(ce2["diag"], base["diag"])
    assert not np.array_equal(ce2["b"], base["b"])


def test_assembly_is_independent_of_interior_face_orientation():
    """Reversing an interior face (swap its cells, reverse its nodes so that A
    flips) describes the same mesh, so the assembled system must not change: the
    harmonic face conductivity, the metric |A|^2 / (A . ds), the volume-weighted
    face gradient (g0 V0 + g1 V1) / (V0 + V1) of the secondary flux and the
    upwind choice are all symmetric in the two cells. A one-sided face gradient
    or conductivity fails this on a non-uniform field."""
    m = hexmesh.hex_mesh(4, jitter=0.2, seed=6)
    nc = m["n_cells"]; nct = nc + m["n_ghost"]
    rng = np.random.default_rng(9)
    k = np.exp(rng.normal(0, 1.0, nct))
    _ghost_copy(m, k)
    x = rng.normal(0, 1, nct)
    src = rng.uniform(0, 1, nct)
    fA = O.fv_metrics(m)["fA"].reshape(-1, 3)
    F = fA @ np.array([0.6, -0.3, 0.8])
    ref = O.fv_thermal(m, k, src, x, BC_T, BC_V, conv_flux=F, cont_res=np.zeros(nc))
    flip = dict(m)
    fcl = m["face_cells"].reshape(-1, 2).copy()
    fn = m["face_nodes"].reshape(-1, 4).copy()
    sel = np.arange(0, m["group_ptr"][1], 2)
    fcl[sel] = fcl[sel][:, ::-1]
    fn[sel] = fn[sel][:, ::-1]
    flip["face_cells"] = fcl.reshape(-1).copy()
    flip["face_nodes"] = fn.reshape(-1).copy()
    Ff = F.copy()
    Ff[sel] = -Ff[sel]
    out = O.fv_thermal(flip, k, src, x, BC_T, BC_V, conv_flux=Ff, cont_res=np.zeros(nc))
    for key in ("row", "col", "is_boundary"):
        assert np.array_equal(out[key], ref[key]), key
    for key in ("diag", "off", "b"):
        np.testing.assert_allclose(out[key], ref[key], rtol=1e-12, atol=1e-13 * np.abs(ref[key]).max())


@pytest.mark.parametrize("jit", [0.0, 0.15])
def test_neumann_heat_flux_linear_solution(jit):
    """GenericBCS.h:123-150: a specified heat flux q enters the domain through its
    wall (the ghost row's residual is q |A| minus the flux the interior sends
    into the domain). T = 0 on x = 0, flux q on x = 1, zero flux on y and z: the
    exact solution T = (q / k) x is linear, which the scheme reproduces on a
    uniform mesh; on a jittered one the heat through every x-section still
    adds up to q."""
    N, kc, q = 5, 1.7, 2.3
    m = hexmesh.hex_mesh(N, jitter=jit, seed=3)
    nc = m["n_cells"]; nct = nc + m["n_ghost"]
    bt = [0, D, NM, NM, NM, NM, NM]
    bv = [0.0, 0.0, q, 0.0, 0.0, 0.0, 0.0]
    S = O.fv_thermal(m, np.full(nct, kc), np.zeros(nct), np.zeros(nct), bt, bv)
    T = np.linalg.solve(_dense_interior(S), -S["b"][:nc])
    cc = O.fv_metrics(m)["cc"].reshape(-1, 3)[:nc]
    if jit == 0.0:
        np.testing.assert_allclose(T, q / kc * cc[:, 0], rtol=1e-12, atol=1e-13)
    else:
        # total heat leaving through the Dirichlet wall equals the heat entering: q * 1
        g = O.fv_metrics(m)
        fcl = m["face_cells"].reshape(-1, 2)
        gp = m["group_ptr"]
        out = 0.0
        for f in range(gp[1], gp[2]):
            c0 = fcl[f, 0]
            ds = g["cc"].reshape(-1, 3)[fcl[f, 1]] - g["cc"].reshape(-1, 3)[c0]
            A = g["fA"].reshape(-1, 3)[f]
            out += kc * (A.dot(A) / A.dot(ds)) * (T[c0] - 0.0)
        np.testing.assert_allclose(out, q * 1.0, rtol=1e-10)
