"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

Restatement of the ESBGK kinetic model loop around the per-direction solves
(esbgkbase/KineticModel.h, KineticBoundaryConditions.h, Quadrature.h,
fvmbase/MatrixOperation.h), Maxwellian, BGK and ES-BGK equilibria (options fgamma 0,
1 and 2),
with the diffuse-wall, zero-gradient and specular (symmetry) boundaries. The
per-direction assembly and the AMG solve are passed in (the C oracle's
`ofv_kinetic` and AMG), so this file holds the quadrature, the moments, the
collision frequency, the equilibria with their Newton solve, the boundary values
and the loop. Every sum over directions runs in direction order, per cell, as the
reference's loops do.

Pinned by tests/test_kinetic_ref.py, against what the mathematics fixes:
- the quadrature's abscissae and weights (closed form of the uniform grid);
- the moments of a sampled Maxwellian (density, velocity and temperature of the
  continuous distribution, to the grid's accuracy; exactly those of the sampled
  values when recomputed from the definition);
- the BGK Newton solve: the discrete moments of the equilibrium it returns equal
  the cell's moments (the conservation it is built for), reached quadratically;
  the ES-BGK solve likewise for its ten moments, and at Prandtl 1 it gives the
  BGK equilibrium back (the temperature tensor is then the BGK one's);
- the scaled-pivoting Gauss inverse against numpy.linalg.inv;
- the diffuse wall: zero net mass flux through the wall, ghost = wall Maxwellian
  for incoming directions; the specular wall: zero normal mass flux and mirrored
  values; the zero-gradient wall: ghost = cell;
- the loop: a uniform resting Maxwellian between walls at its temperature is a
  fixed point (zero norms), and a planted sign error in the collision frequency,
  the Newton Jacobian or the wall flux balance fails one of the above.
"""
import numpy as np

from model_ref import solve_boundary

PI = np.arccos(-1.0)   # KineticBC.h:108 (pi = acos(-1))

# KineticModelOptions defaults (KineticBC.h:61-113)
DEFAULTS = dict(rho_init=9.28e-6, T_init=273.15, nonDimLt=1.0, molecularWeight=40.0, Tmuref=273.15,
                muref=2.117e-5, mu_w=0.81, Prandtl=2.0 / 3.0, fgamma=2, ktrial=50, ToleranceX=1e-8,
                ToleranceF=1e-16, epsilon=1e-50)


def quadrature_cartesian(N1, N2, N3, clim, T2):
    """Quadrature(N1, N2, N3, clim, T2) (Quadrature.h:83-181): uniform abscissae on
    [-clim sqrt(T2/2), clim sqrt(T2/2)] built by repeated addition, weights the
    products of the spacings, directions ordered j1 (x) slowest, j3 (z) fastest."""
    lim = clim * np.sqrt(0.5 * T2)
    ab = []
    sp = []
    for N in (N1, N2, N3):
        cmin, cmax = -lim, lim
        d = (cmax - cmin) / (N - 1.0)
        a = np.empty(N)
        a[0] = cmin
        for j in range(1, N):
            a[j] = a[j - 1] + d
        ab.append(a)
        sp.append(d)
    cx, cy, cz, w = [], [], [], []
    for j1 in range(N1):
        for j2 in range(N2):
            for j3 in range(N3):
                cx.append(ab[0][j1]); cy.append(ab[1][j2]); cz.append(ab[2][j3])
                w.append(sp[0] * sp[1] * sp[2])
    return np.array(cx), np.array(cy), np.array(cz), np.array(w)


def maxwellian(density, v, T, cx, cy, cz):
    """initializeMaxwellian (KineticModel.h:1216-1240): K x ncells."""
    K = len(cx)
    f = np.empty((K, len(density)))
    for j in range(K):
        d0, d1, d2 = cx[j] - v[:, 0], cy[j] - v[:, 1], cz[j] - v[:, 2]
        f[j] = density / np.power(PI * T, 1.5) * np.exp(-(d0 * d0 + d1 * d1 + d2 * d2) / T)
    return f


def macroparameters(f, cx, cy, cz, w):
    """ComputeMacroparameters (KineticModel.h:471-548) over every cell, ghosts
    included: density, velocity, temperature, pressure."""
    nct = f.shape[1]
    density = np.zeros(nct)
    v = np.zeros((nct, 3))
    T = np.zeros(nct)
    for j in range(len(cx)):
        fj = f[j]
        density = density + w[j] * fj
        v[:, 0] = v[:, 0] + (cx[j] * fj) * w[j]
        v[:, 1] = v[:, 1] + (cy[j] * fj) * w[j]
        v[:, 2] = v[:, 2] + (cz[j] * fj) * w[j]
        T = T + (cx[j] * cx[j] + cy[j] * cy[j] + cz[j] * cz[j]) * fj * w[j]
    v[:, 0] = v[:, 0] / density
    v[:, 1] = v[:, 1] / density
    v[:, 2] = v[:, 2] / density
    T = T - (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]) * density
    T = T / (1.5 * density)
    return density, v, T, density * T


def collision_frequency(density, T, o):
    """ComputeCollisionfrequency (KineticModel.h:645-677): power-law viscosity and
    nu = rho T / mu * mu0 (times Prandtl for fgamma 2)."""
    R = 8314.0 / o["molecularWeight"]
    mu0 = o["rho_init"] * R * o["T_init"] * o["nonDimLt"] / np.power(2 * R * o["T_init"], 0.5)
    visc = o["muref"] * np.power(T * o["T_init"] / o["Tmuref"], o["mu_w"])
    nu = density * T / visc * mu0
    if o["fgamma"] == 2:
        nu = o["Prandtl"] * nu
    return nu


def equilibrium(coeff, v, cx, cy, cz):
    """initializeMaxwellianEq / the BGK equilibrium (KineticModel.h:796-803, 944-951):
    feq = a0 exp(-a1 |c-u|^2 + a2 (cx-u) + a3 (cy-v) + a4 (cz-w))."""
    K = len(cx)
    feq = np.empty((K, v.shape[0]))
    for j in range(K):
        d0, d1, d2 = cx[j] - v[:, 0], cy[j] - v[:, 1], cz[j] - v[:, 2]
        feq[j] = coeff[:, 0] * np.exp(-coeff[:, 1] * (d0 * d0 + d1 * d1 + d2 * d2) + coeff[:, 2] * d0
                                      + coeff[:, 3] * d1 + coeff[:, 4] * d2)
    return feq


def inverse_gauss(a):
    """inverseGauss (MatrixOperation.h:200-290): Gauss elimination with scaled partial
    pivoting (rescaling factor = largest |entry| of each row), the multipliers stored
    below the diagonal, then the unit right-hand sides eliminated and back-substituted."""
    a = np.array(a, dtype=np.float64)
    n = a.shape[0]
    b = np.eye(n)
    x = np.zeros((n, n))
    indx = list(range(n))
    c = np.zeros(n)
    for i in range(n):
        c1 = 0.0
        for j in range(n):
            if abs(a[i, j]) > c1:
                c1 = abs(a[i, j])
        c[i] = c1
    for j in range(n - 1):
        pi1 = 0.0
        pk = j
        for i in range(j, n):
            pi = abs(a[indx[i], j]) / c[indx[i]]
            if pi > pi1:
                pi1 = pi
                pk = i
        indx[j], indx[pk] = indx[pk], indx[j]
        for i in range(j + 1, n):
            pj = a[indx[i], j] / a[indx[j], j]
            a[indx[i], j] = pj
            for k in range(j + 1, n):
                a[indx[i], k] = a[indx[i], k] - pj * a[indx[j], k]
    for i in range(n - 1):
        for j in range(i + 1, n):
            for k in range(n):
                b[indx[j], k] = b[indx[j], k] - a[indx[j], i] * b[indx[i], k]
    for col in range(n):
        x[n - 1, col] = b[indx[n - 1], col] / a[indx[n - 1], n - 1]
        for j in range(n - 2, -1, -1):
            x[j, col] = b[indx[j], col]
            for k in range(j + 1, n):
                x[j, col] = x[j, col] - a[indx[j], k] * x[k, col]
            x[j, col] = x[j, col] / a[indx[j], j]
    return x


def newton_bgk(density, v, T, coeff, cx, cy, cz, w, o):
    """NewtonsMethodBGK + setJacobianBGK (KineticModel.h:829-905, 957-990), cell by
    cell: F(a) = [rho, rho u, 1.5 rho T + rho |u|^2] - sum_j w_j feq_j(a) [1, c, |c|^2],
    stop on sum|F| <= ToleranceF (before the step) or sum|step| <= ToleranceX (after
    it), at most ktrial steps. The cells are independent; they are iterated side by
    side here, each with the reference's own sequence of operations (direction sums
    in direction order). coeff (ncells x 5) is updated in place; returns the steps
    taken per cell."""
    K = len(cx)
    mab = np.stack([np.ones(K), cx, cy, cz, cx * cx + cy * cy + cz * cz], 1)   # malphaBGK (Quadrature.h:163-167)
    n = len(density)
    steps = np.zeros(n, int)
    active = np.ones(n, bool)
    for trial in range(o["ktrial"]):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            break
        rho, u, Tc, a = density[idx], v[idx], T[idx], coeff[idx]
        fvec = np.stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2],
                         1.5 * rho * Tc + rho * (u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2])], 1)
        fjac = np.zeros((len(idx), 5, 5))
        for j in range(K):
            d0, d1, d2 = cx[j] - u[:, 0], cy[j] - u[:, 1], cz[j] - u[:, 2]
            Cc = d0 * d0 + d1 * d1 + d2 * d2
            E = a[:, 0] * np.exp(-a[:, 1] * Cc + a[:, 2] * d0 + a[:, 3] * d1 + a[:, 4] * d2) * w[j]
            for row in range(5):
                fvec[:, row] += -E * mab[j, row]
            mexp = (-E / a[:, 0], E * Cc, -E * d0, -E * d1, -E * d2)
            for row in range(5):
                for col in range(5):
                    fjac[:, row, col] += mab[j, row] * mexp[col]
        for q, c in enumerate(idx):
            errf = 0.0
            for row in range(5):
                errf += abs(fvec[q, row])
            if errf <= o["ToleranceF"]:
                active[c] = False
                continue
            inv = inverse_gauss(fjac[q])
            pvec = -fvec[q]
            errx = 0.0
            for row in range(5):
                xr = 0.0
                for col in range(5):
                    xr += inv[row, col] * pvec[col]
                errx += abs(xr)
                coeff[c, row] += xr
            steps[c] += 1
            if errx <= o["ToleranceX"]:
                active[c] = False
    return steps


def macroparameters_es(f, feq, v, density, cx, cy, cz, w, Pr):
    """ComputeMacroparametersESBGK (KineticModel.h:565-631): the temperature tensor
    T_ab = sum_j (c_a - u_a)(c_b - u_b) ((1 - 1/Pr) f_j + 1/Pr feq_j) w_j / rho, with
    feq the BGK equilibrium; returns (Txx, Tyy, Tzz, Txy, Tyz, Tzx) per cell."""
    nct = f.shape[1]
    T6 = [np.zeros(nct) for _ in range(6)]
    for j in range(len(cx)):
        g = (1 - 1 / Pr) * f[j] + 1 / Pr * feq[j]
        d0, d1, d2 = cx[j] - v[:, 0], cy[j] - v[:, 1], cz[j] - v[:, 2]
        T6[0] = T6[0] + d0 * d0 * g * w[j]
        T6[1] = T6[1] + d1 * d1 * g * w[j]
        T6[2] = T6[2] + d2 * d2 * g * w[j]
        T6[3] = T6[3] + d0 * d1 * g * w[j]
        T6[4] = T6[4] + d1 * d2 * g * w[j]
        T6[5] = T6[5] + d2 * d0 * g * w[j]
    return [t / density for t in T6]


def es_exponent(xn, cx, cy, cz, u, j):
    """The exponent of the ES-BGK equilibrium (KineticModel.h:1151-1161, 1196-1200)."""
    c1, c2, c3 = cx[j] - u[:, 0], cy[j] - u[:, 1], cz[j] - u[:, 2]
    return (-xn[:, 1] * (c1 * c1) + xn[:, 2] * c1 - xn[:, 3] * (c2 * c2) + xn[:, 4] * c2
            - xn[:, 5] * (c3 * c3) + xn[:, 6] * c3
            + xn[:, 7] * cx[j] * cy[j] + xn[:, 8] * cy[j] * cz[j] + xn[:, 9] * cz[j] * cx[j])


def equilibrium_es(coeffg, v, cx, cy, cz):
    """The ES-BGK equilibrium (EquilibriumDistributionESBGK, KineticModel.h:1146-1162)."""
    K = len(cx)
    out = np.empty((K, v.shape[0]))
    for j in range(K):
        out[j] = coeffg[:, 0] * np.exp(es_exponent(coeffg, cx, cy, cz, v, j))
    return out


def newton_esbgk(density, v, T6, coeffg, cx, cy, cz, w, o):
    """NewtonsMethodESBGK + setJacobianESBGK (KineticModel.h:1019-1108, 1166-1213):
    the ten coefficients fitted so that the equilibrium's discrete moments
    [1, c, c_x^2, c_y^2, c_z^2, c_x c_y, c_y c_z, c_z c_x] equal
    rho [1, u, u_a u_b + T_ab]; the same stop tests as the BGK fit."""
    K = len(cx)
    mab = np.stack([np.ones(K), cx, cy, cz, cx * cx, cy * cy, cz * cz, cx * cy, cy * cz, cz * cx], 1)
    Txx, Tyy, Tzz, Txy, Tyz, Tzx = T6
    n = len(density)
    steps = np.zeros(n, int)
    active = np.ones(n, bool)
    for trial in range(o["ktrial"]):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            break
        rho, u, a = density[idx], v[idx], coeffg[idx]
        fvec = np.stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2],
                         rho * (u[:, 0] * u[:, 0] + Txx[idx]), rho * (u[:, 1] * u[:, 1] + Tyy[idx]),
                         rho * (u[:, 2] * u[:, 2] + Tzz[idx]), rho * (u[:, 0] * u[:, 1] + Txy[idx]),
                         rho * (u[:, 1] * u[:, 2] + Tyz[idx]), rho * (u[:, 2] * u[:, 0] + Tzx[idx])], 1)
        fjac = np.zeros((len(idx), 10, 10))
        for j in range(K):
            c1, c2, c3 = cx[j] - u[:, 0], cy[j] - u[:, 1], cz[j] - u[:, 2]
            E = a[:, 0] * np.exp(es_exponent(a, cx, cy, cz, u, j)) * w[j]
            for row in range(10):
                fvec[:, row] += -E * mab[j, row]
            mexp = (-E / a[:, 0], E * (c1 * c1), -E * c1, E * (c2 * c2), -E * c2, E * (c3 * c3), -E * c3,
                    -E * cx[j] * cy[j], -E * cy[j] * cz[j], -E * cz[j] * cx[j])
            for row in range(10):
                for col in range(10):
                    fjac[:, row, col] += mab[j, row] * mexp[col]
        for q, c in enumerate(idx):
            errf = 0.0
            for row in range(10):
                errf += abs(fvec[q, row])
            if errf <= o["ToleranceF"]:
                active[c] = False
                continue
            inv = inverse_gauss(fjac[q])
            pvec = -fvec[q]
            errx = 0.0
            for row in range(10):
                xr = 0.0
                for col in range(10):
                    xr += inv[row, col] * pvec[col]
                errx += abs(xr)
                coeffg[c, row] += xr
            steps[c] += 1
            if errx <= o["ToleranceX"]:
                active[c] = False
    return steps


def reflection_map(en, cx, cy, cz):
    """Specular reflection table of a symmetry group (KineticModel.h:290-318): for
    every direction j the direction nearest to c_j - 2 (c_j.n) n (first minimum)."""
    K = len(cx)
    out = np.zeros(K, int)
    for j in range(K):
        cn = cx[j] * en[0] + cy[j] * en[1] + cz[j] * en[2]
        ix, iy, iz = cx[j] - 2.0 * cn * en[0], cy[j] - 2.0 * cn * en[1], cz[j] - 2.0 * cn * en[2]
        best, arg = 1e54, 0
        for js in range(K):
            d = (ix - cx[js]) ** 2 + (iy - cy[js]) ** 2 + (iz - cz[js]) ** 2
            if d < best:
                best, arg = d, js
        out[j] = arg
    return out


def boundary_values(mesh, fA, fAm, bcs, refl, f, macro, cx, cy, cz, w, eps):
    """callBoundaryConditions (KineticModel.h:3271-3330) for the diffuse wall
    (KineticBoundaryConditions.h:65-140), zero gradient (:385-406) and specular
    symmetry (:287-330) groups, face by face in group order. bcs[g] = ("wall", uw, Tw),
    ("zerograd",), ("symmetry",) or ("copy",): a group type callBoundaryConditions has
    no branch for ("CopyBC" in test/ESBGK-tests/Couette/Par_Couette.py:199-206) keeps
    its ghosts. f (K x ncells) and macro (density, v, T) are
    changed in place."""
    density, v, T = macro
    fcl = mesh["face_cells"].reshape(-1, 2)
    gp = mesh["group_ptr"]
    K = len(cx)
    for g in range(1, len(gp) - 1):
        kind = bcs[g][0]
        for fi in range(gp[g], gp[g + 1]):
            c0, c1 = fcl[fi]
            en = fA[3 * fi:3 * fi + 3] / fAm[fi]
            if kind == "copy":
                continue     # a group type without a boundary-value rule: ghosts keep their values
            if kind == "zerograd":
                f[:, c1] = f[:, c0]
            elif kind == "symmetry":
                for j in range(K):
                    cn = cx[j] * en[0] + cy[j] * en[1] + cz[j] * en[2]
                    f[j, c1] = f[refl[g][j], c0] if cn < 0.0 else f[j, c0]
            else:
                uw, Tw = bcs[g][1], bcs[g][2]
                v[c1] = uw
                T[c1] = Tw
                Nmr = 0.0
                Dmr = 0.0
                for j in range(K):
                    cn = cx[j] * en[0] + cy[j] * en[1] + cz[j] * en[2]
                    un = uw[0] * en[0] + uw[1] * en[1] + uw[2] * en[2]
                    d0, d1, d2 = cx[j] - uw[0], cy[j] - uw[1], cz[j] - uw[2]
                    fwall = 1.0 / np.power(PI * Tw, 1.5) * np.exp(-(d0 * d0 + d1 * d1 + d2 * d2) / Tw)
                    if cn - un < eps:
                        Dmr = Dmr - fwall * w[j] * (cn - un)
                    else:
                        Nmr = Nmr + f[j, c0] * w[j] * (cn - un)
                nwall = Nmr / Dmr
                density[c1] = nwall
                for j in range(K):
                    cn = cx[j] * en[0] + cy[j] * en[1] + cz[j] * en[2]
                    un = uw[0] * en[0] + uw[1] * en[1] + uw[2] * en[2]
                    if cn - un < eps:
                        d0, d1, d2 = cx[j] - uw[0], cy[j] - uw[1], cz[j] - uw[2]
                        f[j, c1] = nwall / np.power(PI * Tw, 1.5) * np.exp(-(d0 * d0 + d1 * d1 + d2 * d2) / Tw)
                    else:
                        f[j, c1] = f[j, c0]


def init_state(nct, quad, o):
    """KineticModel constructor (KineticModel.h:81-120): InitializeMacroparameters
    (rest, density 1, temperature 1, coeff of that Maxwellian, :349-390), the
    Maxwellian f, then the moments, collision frequency and equilibrium of it."""
    cx, cy, cz, w = quad
    density = np.ones(nct)
    v = np.zeros((nct, 3))
    T = np.ones(nct)
    coeff = np.zeros((nct, 5))
    coeff[:, 0] = density / np.power(PI * T, 1.5)
    coeff[:, 1] = 1 / T
    coeffg = np.zeros((nct, 10))
    if o["fgamma"] == 2:   # InitializeMacroparameters (KineticModel.h:392-402)
        coeffg[:, 0] = coeff[:, 0]; coeffg[:, 1] = coeff[:, 1]; coeffg[:, 2] = coeff[:, 2]
        coeffg[:, 3] = coeff[:, 1]; coeffg[:, 4] = coeff[:, 3]; coeffg[:, 5] = coeff[:, 1]
        coeffg[:, 6] = coeff[:, 4]
    f = maxwellian(density, v, T, cx, cy, cz)
    density, v, T, p = macroparameters(f, cx, cy, cz, w)
    nu = collision_frequency(density, T, o)
    feq = equilibrium(coeff, v, cx, cy, cz)
    st = dict(f=f, feq=feq, density=density, v=v, T=T, nu=nu, coeff=coeff, coeffg=coeffg, niters=0, initial=None)
    st["feq_lin"] = equilibrium_es(coeffg, v, cx, cy, cz) if o["fgamma"] == 2 else feq
    return st


def kinetic_advance(assemble, solve, mesh, metrics, quad, bcs, refl, o, st, niter, rel_tol, abs_tol):
    """KineticModel::advance (KineticModel.h:3360-3462) with the directions one after
    another: assemble(j, nu, feq_j, f_j) -> system; solve(system) -> (r0, delta).
    Per iteration: every direction linearized, solved (its initial norm added to the
    iteration's), postSolve and updateSolution; the initial norm (first iteration,
    then the largest of the first five); then always the boundary values, moments,
    collision frequency and equilibrium (fgamma 0: the Maxwellian of the initial
    coefficients around the new velocity; 1: the BGK Newton solve; 2: that, then the
    temperature tensor and the ES-BGK Newton solve, KineticModel.h:3448-3452); then
    the stop test. The collision term relaxes to the ES-BGK equilibrium for fgamma 2,
    to the BGK / Maxwellian one otherwise (KineticModel.h:1520-1539). Returns
    (norms, converged)."""
    cx, cy, cz, w = quad
    K = len(cx)
    nc = mesh["n_cells"]
    norms = []
    for _ in range(niter):
        rnorm = None
        for j in range(K):
            sysj = assemble(j, st["nu"][:nc], st["feq_lin"][j][:nc], st["f"][j])
            r0, delta = solve(sysj)
            rnorm = r0 if rnorm is None else rnorm + r0
            solve_boundary(sysj, delta)
            st["f"][j] += delta
        if st["initial"] is None:
            st["initial"] = rnorm
        if st["niters"] < 5:
            st["initial"] = max(st["initial"], rnorm)
        ratio = rnorm / st["initial"]
        norms.append(rnorm)
        st["niters"] += 1
        macro = (st["density"], st["v"], st["T"])
        boundary_values(mesh, metrics["fA"], metrics["fAm"], bcs, refl, st["f"], macro, cx, cy, cz, w,
                        o["epsilon"])
        st["density"], st["v"], st["T"], _ = macroparameters(st["f"], cx, cy, cz, w)
        st["nu"] = collision_frequency(st["density"], st["T"], o)
        if o["fgamma"] >= 1:
            newton_bgk(st["density"], st["v"], st["T"], st["coeff"], cx, cy, cz, w, o)
        st["feq"] = equilibrium(st["coeff"], st["v"], cx, cy, cz)
        st["feq_lin"] = st["feq"]
        if o["fgamma"] == 2:
            T6 = macroparameters_es(st["f"], st["feq"], st["v"], st["density"], cx, cy, cz, w, o["Prandtl"])
            st["T6"] = T6
            newton_esbgk(st["density"], st["v"], T6, st["coeffg"], cx, cy, cz, w, o)
            st["feq_lin"] = equilibrium_es(st["coeffg"], st["v"], cx, cy, cz)
        if rnorm < abs_tol or ratio < rel_tol:
            return norms, True
    return norms, False
