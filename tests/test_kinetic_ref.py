"""Pins of the kinetic-model oracle (oracle/kinetic_ref.py, oracle/fv_oracle.c
ofv_kinetic) against what the mathematics fixes: the uniform quadrature, moments
of a sampled Maxwellian, the scaled-pivoting inverse against numpy, conservation
by the BGK Newton solve, the collision frequency's power law, zero mass flux
through diffuse and specular walls, the assembly's closed form on a uniform field,
and the rest state as a fixed point of the loop."""
import os
import sys

import numpy as np
import pytest

import hexmesh
import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import kinetic_ref as KR  # noqa: E402


def _opts(**kw):
    o = dict(KR.DEFAULTS)
    o.update(kw)
    return o


def test_quadrature_is_the_uniform_grid():
    N1, N2, N3, clim, T2 = 5, 4, 3, 5.5, 1.3
    cx, cy, cz, w = KR.quadrature_cartesian(N1, N2, N3, clim, T2)
    lim = clim * np.sqrt(0.5 * T2)
    assert len(cx) == N1 * N2 * N3
    gx = np.linspace(-lim, lim, N1); gy = np.linspace(-lim, lim, N2); gz = np.linspace(-lim, lim, N3)
    want = np.array([(a, b, c) for a in gx for b in gy for c in gz])
    np.testing.assert_allclose(np.stack([cx, cy, cz], 1), want, rtol=0, atol=1e-13)
    dx, dy, dz = 2 * lim / (N1 - 1), 2 * lim / (N2 - 1), 2 * lim / (N3 - 1)
    np.testing.assert_allclose(w, dx * dy * dz, rtol=1e-14)


def test_moments_of_a_sampled_maxwellian():
    cx, cy, cz, w = KR.quadrature_cartesian(28, 28, 28, 6.0, 1.0)
    rho = np.array([1.3, 0.7]); u = np.array([[0.2, -0.1, 0.05], [0.0, 0.3, -0.2]]); T = np.array([0.9, 1.4])
    f = KR.maxwellian(rho, u, T, cx, cy, cz)
    d, v, Tm, p = KR.macroparameters(f, cx, cy, cz, w)
    # the continuous distribution's moments, to the grid's accuracy
    np.testing.assert_allclose(d, rho, rtol=1e-5)
    np.testing.assert_allclose(v, u, atol=1e-5)
    np.testing.assert_allclose(Tm, T, rtol=1e-5)
    np.testing.assert_allclose(p, d * Tm, rtol=1e-15)
    # the sampled values' moments from the definition, summed another way
    c = np.stack([cx, cy, cz], 1)
    d2 = w @ f
    v2 = (c * w[:, None]).T @ f / d2
    e2 = ((c * c).sum(1) * w) @ f
    T2 = (e2 - (v2 * v2).sum(0) * d2) / (1.5 * d2)
    np.testing.assert_allclose(d, d2, rtol=1e-13)
    np.testing.assert_allclose(v, v2.T, rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(Tm, T2, rtol=1e-12)


@pytest.mark.parametrize("n", [5, 10])
def test_inverse_gauss_against_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)) + n * np.eye(n)
    np.testing.assert_allclose(KR.inverse_gauss(a), np.linalg.inv(a), rtol=1e-12, atol=1e-14)
    # a zero leading pivot needs the row interchange
    p = rng.normal(size=(n, n))
    p[0, 0] = 0.0
    p[:, 0] *= 1e-3
    np.testing.assert_allclose(KR.inverse_gauss(p) @ p, np.eye(n), atol=1e-10)


def test_newton_bgk_conserves_the_discrete_moments():
    """The BGK equilibrium is built so that its discrete moments [1, c, |c|^2] equal
    the cell's; Newton from the continuous Maxwellian's coefficients gets there in
    a few steps on a coarse grid, where the two differ."""
    cx, cy, cz, w = KR.quadrature_cartesian(8, 8, 8, 4.5, 1.0)
    rng = np.random.default_rng(4)
    n = 12
    rho = rng.uniform(0.5, 2.0, n); u = rng.uniform(-0.3, 0.3, (n, 3)); T = rng.uniform(0.6, 1.6, n)
    coeff = np.zeros((n, 5))
    coeff[:, 0] = rho / np.power(KR.PI * T, 1.5)
    coeff[:, 1] = 1 / T
    o = _opts(ToleranceF=1e-14, ToleranceX=1e-14)
    steps = KR.newton_bgk(rho, u, T, coeff, cx, cy, cz, w, o)
    assert steps.max() <= 8 and steps.min() >= 1
    feq = KR.equilibrium(coeff, u, cx, cy, cz)
    m = np.stack([np.ones_like(cx), cx, cy, cz, cx * cx + cy * cy + cz * cz], 1)
    got = (m * w[:, None]).T @ feq
    want = np.stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2],
                     1.5 * rho * T + rho * (u * u).sum(1)])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)
    # and the coefficients are no longer the continuous Maxwellian's
    assert np.abs(coeff[:, 1] - 1 / T).max() > 1e-4


def test_newton_esbgk_conserves_its_moments_and_reduces_to_bgk_at_prandtl_one():
    """The ES-BGK equilibrium's ten discrete moments equal rho [1, u, u_a u_b + T_ab]
    for the tensor of a non-equilibrium f; at Prandtl 1 the tensor is the BGK
    equilibrium's and the ES-BGK equilibrium is the BGK one."""
    cx, cy, cz, w = KR.quadrature_cartesian(8, 8, 8, 4.5, 1.0)
    rng = np.random.default_rng(12)
    n = 6
    rho = rng.uniform(0.7, 1.5, n); u = rng.uniform(-0.2, 0.2, (n, 3)); T = rng.uniform(0.8, 1.3, n)
    f = KR.maxwellian(rho, u, T, cx, cy, cz) * (1 + 0.1 * np.sin(np.arange(len(cx)))[:, None])
    d, v, Tm, _ = KR.macroparameters(f, cx, cy, cz, w)
    o = _opts(ToleranceF=1e-14, ToleranceX=1e-14)
    coeff = np.zeros((n, 5)); coeff[:, 0] = d / np.power(KR.PI * Tm, 1.5); coeff[:, 1] = 1 / Tm
    KR.newton_bgk(d, v, Tm, coeff, cx, cy, cz, w, o)
    feq = KR.equilibrium(coeff, v, cx, cy, cz)
    m = np.stack([np.ones_like(cx), cx, cy, cz, cx * cx, cy * cy, cz * cz, cx * cy, cy * cz, cz * cx], 1)
    for Pr in (2.0 / 3.0, 1.0):
        T6 = KR.macroparameters_es(f, feq, v, d, cx, cy, cz, w, Pr)
        cg = np.zeros((n, 10))
        cg[:, 0] = coeff[:, 0]; cg[:, 1] = coeff[:, 1]; cg[:, 2] = coeff[:, 2]; cg[:, 3] = coeff[:, 1]
        cg[:, 4] = coeff[:, 3]; cg[:, 5] = coeff[:, 1]; cg[:, 6] = coeff[:, 4]
        steps = KR.newton_esbgk(d, v, T6, cg, cx, cy, cz, w, o)
        assert steps.max() <= 10
        fes = KR.equilibrium_es(cg, v, cx, cy, cz)
        got = (m * w[:, None]).T @ fes
        Txx, Tyy, Tzz, Txy, Tyz, Tzx = T6
        want = np.stack([d, d * v[:, 0], d * v[:, 1], d * v[:, 2], d * (v[:, 0] ** 2 + Txx),
                         d * (v[:, 1] ** 2 + Tyy), d * (v[:, 2] ** 2 + Tzz), d * (v[:, 0] * v[:, 1] + Txy),
                         d * (v[:, 1] * v[:, 2] + Tyz), d * (v[:, 2] * v[:, 0] + Tzx)])
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12)
        if Pr == 1.0:
            np.testing.assert_allclose(fes, feq, rtol=1e-9, atol=1e-13)
        else:
            assert np.abs(fes - feq).max() > 1e-6


def test_collision_frequency_power_law():
    o = _opts(T_init=300.0, Tmuref=300.0, mu_w=0.81, fgamma=1)
    rho = np.array([1.0, 2.0, 1.0]); T = np.array([1.0, 1.0, 2.0])
    nu = KR.collision_frequency(rho, T, o)
    R = 8314.0 / o["molecularWeight"]
    mu0 = o["rho_init"] * R * o["T_init"] * o["nonDimLt"] / np.sqrt(2 * R * o["T_init"])
    # at the reference temperature the viscosity is muref
    np.testing.assert_allclose(nu[0], mu0 / o["muref"], rtol=1e-14)
    # linear in density, T^(1 - mu_w) in temperature
    np.testing.assert_allclose(nu[1] / nu[0], 2.0, rtol=1e-14)
    np.testing.assert_allclose(nu[2] / nu[0], 2.0 ** (1 - 0.81), rtol=1e-13)
    np.testing.assert_allclose(KR.collision_frequency(rho, T, _opts(fgamma=2, T_init=300.0, Tmuref=300.0)),
                               o["Prandtl"] * nu, rtol=1e-15)


def _wall_case(kind):
    mesh = hexmesh.hex_mesh(3, 0.15, 5)
    nc, ng = mesh["n_cells"], mesh["n_ghost"]
    g = O.fv_metrics(mesh)
    quad = KR.quadrature_cartesian(6, 6, 6, 4.5, 1.0)
    cx, cy, cz, w = quad
    rng = np.random.default_rng(9)
    nct = nc + ng
    rho = rng.uniform(0.8, 1.2, nct); u = rng.uniform(-0.2, 0.2, (nct, 3)); T = rng.uniform(0.8, 1.2, nct)
    f = KR.maxwellian(rho, u, T, cx, cy, cz)
    bcs = {gi: (kind, np.array([0.0, 0.1, -0.05]), 1.1) if kind == "wall" else (kind,) for gi in range(1, 7)}
    refl = {}
    gp = mesh["group_ptr"]
    for gi in range(1, 7):
        f0 = gp[gi]
        refl[gi] = KR.reflection_map(g["fA"][3 * f0:3 * f0 + 3] / g["fAm"][f0], cx, cy, cz)
    f_before = f.copy()
    macro = (rho.copy(), u.copy(), T.copy())
    KR.boundary_values(mesh, g["fA"], g["fAm"], bcs, refl, f, macro, cx, cy, cz, w, 1e-50)
    return mesh, g, quad, bcs, refl, f_before, f


def test_diffuse_wall_has_no_net_mass_flux():
    mesh, g, (cx, cy, cz, w), bcs, refl, f0, f = _wall_case("wall")
    fcl = mesh["face_cells"].reshape(-1, 2)
    gp = mesh["group_ptr"]
    for fi in range(gp[1], gp[7]):
        c0, c1 = fcl[fi]
        en = g["fA"][3 * fi:3 * fi + 3] / g["fAm"][fi]
        uw, Tw = bcs[1][1], bcs[1][2]
        rel = cx * en[0] + cy * en[1] + cz * en[2] - uw @ en
        inc = rel < 1e-50
        face_f = np.where(inc, f[:, c1], f[:, c0])
        out = (w * rel * f[:, c0])[~inc].sum()
        assert abs((w * rel * face_f).sum()) <= 1e-13 * out
        # incoming ghost values are the wall Maxwellian, scaled
        fw = np.exp(-((cx - uw[0]) ** 2 + (cy - uw[1]) ** 2 + (cz - uw[2]) ** 2) / Tw)
        ratio = f[inc, c1] / fw[inc]
        np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)
        np.testing.assert_array_equal(f[~inc, c1], f0[~inc, c0])


def test_specular_wall_mirrors_and_has_no_normal_flux():
    mesh, g, (cx, cy, cz, w), bcs, refl, f0, f = _wall_case("symmetry")
    fcl = mesh["face_cells"].reshape(-1, 2)
    gp = mesh["group_ptr"]
    # axis-aligned boundary planes of the cube: the table mirrors one component
    m = refl[1]
    np.testing.assert_array_equal(cx[m], -cx); np.testing.assert_array_equal(cy[m], cy)
    np.testing.assert_array_equal(m[m], np.arange(len(cx)))
    for fi in range(gp[1], gp[2]):
        c0, c1 = fcl[fi]
        en = g["fA"][3 * fi:3 * fi + 3] / g["fAm"][fi]
        cn = cx * en[0] + cy * en[1] + cz * en[2]
        inc = cn < 0.0
        np.testing.assert_array_equal(f[inc, c1], f0[m[inc], c0])
        face_f = np.where(inc, f[:, c1], f[:, c0])
        # the x0 plane is flat: the incoming flux mirrors the outgoing one exactly
        assert abs((w * (-cx) * face_f).sum()) <= 1e-13 * (w * np.abs(cx) * f0[:, c0]).sum()


def test_group_without_rule_keeps_its_ghosts():
    mesh, g, quad, bcs, refl, f0, f = _wall_case("copy")
    np.testing.assert_array_equal(f, f0)


def test_zero_gradient_copies_the_cell():
    mesh, g, quad, bcs, refl, f0, f = _wall_case("zerograd")
    fcl = mesh["face_cells"].reshape(-1, 2)
    gp = mesh["group_ptr"]
    for fi in range(gp[1], gp[7]):
        np.testing.assert_array_equal(f[:, fcl[fi][1]], f0[:, fcl[fi][0]])


def test_kinetic_assembly_closed_form_and_phonon_form():
    """A uniform field: the convection part conserves (interior rows sum to the
    collision diagonal -V nu, b = -V nu (f - feq)); with nu = 1/tau and feq = e0 the
    system is the phonon direction's up to rounding."""
    mesh = hexmesh.hex_mesh(4, 0.2, 3)
    nc, ng = mesh["n_cells"], mesh["n_ghost"]
    g = O.fv_metrics(mesh)
    c = np.array([0.7, -0.4, 0.25])
    rng = np.random.default_rng(2)
    nue = rng.uniform(0.5, 2.0, nc)
    feq = rng.uniform(0.1, 1.0, nc)
    x = np.full(nc + ng, 0.6)
    kind = np.zeros(7, np.int32); wall_u = np.zeros((7, 3))
    S = O.fv_kinetic(mesh, c, nue, feq, x, kind, wall_u)
    N = mesh["N"]
    interior = [i + N * (j + N * k) for k in range(1, N - 1) for j in range(1, N - 1) for i in range(1, N - 1)]
    for cc in interior:
        rs = S["diag"][cc] + S["off"][S["row"][cc]:S["row"][cc + 1]].sum()
        np.testing.assert_allclose(rs, -g["vol"][cc] * nue[cc], rtol=1e-12)
        np.testing.assert_allclose(S["b"][cc], -g["vol"][cc] * nue[cc] * (0.6 - feq[cc]), rtol=1e-12, atol=1e-15)
    tau = 0.8
    P = O.fv_phonon(mesh, c, tau, feq, x)
    K2 = O.fv_kinetic(mesh, c, np.full(nc, 1 / tau), feq, x, kind, wall_u)
    np.testing.assert_allclose(K2["diag"], P["diag"], rtol=1e-13)
    np.testing.assert_allclose(K2["off"], P["off"], rtol=1e-13)
    np.testing.assert_allclose(K2["b"], P["b"], rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(K2["is_boundary"], P["is_boundary"])


def kinetic_parts(mesh, quad, bcs):
    cx, cy, cz, w = quad
    kind = np.zeros(len(mesh["group_ptr"]) - 1, np.int32)
    wall_u = np.zeros((len(kind), 3))
    for gi, bc in bcs.items():
        if bc[0] == "zerograd":
            kind[gi] = 1
        elif bc[0] == "wall":
            wall_u[gi] = bc[1]

    def assemble(j, nu, feq, f):
        return O.fv_kinetic(mesh, [cx[j], cy[j], cz[j]], nu, feq, f, kind, wall_u)

    def solve(sys):
        oa = O.OracleAMG(n_max_iterations=400, relative_tolerance=1e-13)
        oa.setup(sys["n_self"], sys["n_total"], sys["row"], sys["col"], sys["diag"], sys["off"], sys["is_boundary"])
        res = oa.solve(sys["b"])
        return res["r0"], res["x"]
    return assemble, solve


def reflections(mesh, metrics, quad, bcs):
    gp = mesh["group_ptr"]
    out = {}
    for gi, bc in bcs.items():
        if bc[0] == "symmetry":
            f0 = gp[gi]
            out[gi] = KR.reflection_map(metrics["fA"][3 * f0:3 * f0 + 3] / metrics["fAm"][f0], *quad[:3])
    return out


def test_rest_state_is_a_fixed_point():
    mesh = hexmesh.hex_mesh(3, 0.1, 6)
    nct = mesh["n_cells"] + mesh["n_ghost"]
    g = O.fv_metrics(mesh)
    quad = KR.quadrature_cartesian(4, 4, 4, 4.0, 1.0)
    o = _opts()
    bcs = {1: ("wall", np.zeros(3), 1.0), 2: ("wall", np.zeros(3), 1.0), 3: ("symmetry",), 4: ("symmetry",),
           5: ("zerograd",), 6: ("zerograd",)}
    st = KR.init_state(nct, quad, o)
    assemble, solve = kinetic_parts(mesh, quad, bcs)
    f0 = st["f"].copy()
    norms, conv = KR.kinetic_advance(assemble, solve, mesh, g, quad, bcs, reflections(mesh, g, quad, bcs), o, st,
                                     1, 1e-30, 0.0)
    scale = (np.abs(f0[:, :mesh["n_cells"]]) * g["vol"][None, :mesh["n_cells"]]).sum() * st["nu"].max()
    assert norms[0] <= 1e-12 * scale
    np.testing.assert_allclose(st["f"], f0, rtol=1e-12, atol=1e-15)


def test_couette_slab_is_antisymmetric():
    """Walls at x = 0 and x = 1 moving in -y and +y at the same temperature,
    symmetry on the other faces: the velocity field is odd under the half-turn
    about the slab's axis, even under the mirror in z, bounded by the wall
    speeds and increasing across the slab, and the norms fall."""
    N = 4
    mesh = hexmesh.hex_mesh(N)
    nc = mesh["n_cells"]
    nct = nc + mesh["n_ghost"]
    g = O.fv_metrics(mesh)
    quad = KR.quadrature_cartesian(5, 5, 5, 4.0, 1.0)
    o = _opts(rho_init=1e-4)
    bcs = {1: ("wall", np.array([0.0, -0.2, 0.0]), 1.0), 2: ("wall", np.array([0.0, 0.2, 0.0]), 1.0),
           3: ("symmetry",), 4: ("symmetry",), 5: ("symmetry",), 6: ("symmetry",)}
    st = KR.init_state(nct, quad, o)
    assemble, solve = kinetic_parts(mesh, quad, bcs)
    norms, conv = KR.kinetic_advance(assemble, solve, mesh, g, quad, bcs, reflections(mesh, g, quad, bcs), o, st,
                                     8, 1e-30, 0.0)
    assert norms[-1] < norms[1]
    vy = st["v"][:nc, 1].reshape(N, N, N)      # [k, j, i]
    # the half-turn about the slab's z axis swaps the walls and reverses the flow;
    # the mirror in z leaves it alone
    np.testing.assert_allclose(vy[:, ::-1, ::-1], -vy, atol=1e-12)
    np.testing.assert_allclose(vy[::-1], vy, atol=1e-12)
    assert np.all(np.abs(vy) < 0.2)
    assert np.all(np.diff(vy.mean((0, 1))) > 0)
