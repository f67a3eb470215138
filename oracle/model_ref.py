"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

Restatement of the model-level drivers around the solver: LinearSystem::postSolve
(LinearSystem.cpp:250-252) -> CRMatrix::solveBoundary (CRMatrix.h:433-454),
ThermalModel::advance (ThermalModel_impl.h:424-456) and the gray
PhononModel::advance (phononbase/PhononModel.h:302-382, :530-600, Kspace.h:425-441,
:546-562, PhononBoundary.h:165-202). Assembly and the AMG solve are passed in
(the C oracle's), so this file holds only the loops, the boundary rows and the
lattice-temperature update. Pinned by tests/test_model_ref.py: the Dirichlet
closed form (after postSolve and the update, a Dirichlet ghost cell holds its
boundary value), the closed form of the gray Newton update (one exact step of
a linear equation), the equilibrium fixed point of the phonon loop, its
per-iteration norm as the sum of independent per-direction solves, and the
maximum principle and mirror symmetry of a slab between two walls.
"""
import numpy as np


def solve_boundary(sys, x):
    """x[nr] = -(b[nr] + sum off*x[col]) / diag[nr] for boundary rows nr >= n_self,
    in ascending row order and CSR order (CRMatrix.h:443-453). x is changed in place."""
    row, col, off, diag, b, isb = sys["row"], sys["col"], sys["off"], sys["diag"], sys["b"], sys["is_boundary"]
    for nr in range(sys["n_self"], sys["n_total"]):
        if isb[nr]:
            s = b[nr]
            for nb in range(row[nr], row[nr + 1]):
                s += off[nb] * x[col[nb]]
            x[nr] = -s / diag[nr]
    return x


def thermal_advance(assemble, solve, x, niter, rel_tol, abs_tol, dirichlet=None):
    """ThermalModel::advance: assemble(x) -> system dict; solve(sys) -> (r0, delta of
    n_total with the self rows solved); dirichlet = (ghost cells, values): the
    linearization also sets those ghosts to their boundary values (GenericBCS.h:104).
    Returns (x, norms)."""
    x = np.array(x, np.float64)
    norms = []
    initial = None
    for _ in range(niter):
        sys = assemble(x)                                  # initLinearization .. initSolve
        if dirichlet is not None:
            x[dirichlet[0]] = dirichlet[1]                 # _x[c1] = bValue
        r_norm, delta = solve(sys)                         # :437
        if initial is None:
            initial = r_norm                               # :439
        ratio = r_norm / initial if initial else 0.0       # :441
        norms.append(r_norm)
        solve_boundary(sys, delta)                         # postSolve, :448
        x = x + delta                                      # updateSolution, :449
        if r_norm < abs_tol or ratio < rel_tol:            # :452-453
            break
    return x, norms


def phonon_update_TL(e, dk3, tau, cp, tref, TL):
    """PhononModel::updateTL (PhononModel.h:302-330): per cell e_sum = sum_k e_k dk3_k
    / tau_k (directions in order), then Kspace::NewtonSolve (Kspace.h:425-441) from
    the current TL with e0tau = sum_k (T - tref_k) cp_k / tau_k dk3_k and its
    derivative (Kspace.h:546-562, gray pmode.h:146-160). e: K x ncells; TL updated
    in place (vectorized over cells; every cell iterates on its own)."""
    K = e.shape[0]
    esum = np.zeros(TL.shape)
    for k in range(K):
        esum += e[k] * dk3[k] / tau[k]
    guess = TL.copy()
    deltaT = np.ones(TL.shape)
    iters = np.zeros(TL.shape, int)
    active = (deltaT > 1e-6) & (iters < 10)
    while active.any():
        e0tau = np.zeros(TL.shape)
        de0 = np.zeros(TL.shape)
        for k in range(K):
            e0tau += (guess - tref[k]) * cp[k] / tau[k] * dk3[k]
            de0 += cp[k] / tau[k] * dk3[k]
        d = (e0tau - esum) / de0
        newguess = guess - d
        dT = np.abs(d / guess)
        guess = np.where(active, newguess, guess)
        deltaT = np.where(active, dT, deltaT)
        iters = np.where(active, iters + 1, iters)
        active = (deltaT > 1e-6) & (iters < 10)
    TL[:] = guess
    return TL


def phonon_walls(mesh, fA, fAm, vg, cp, tref, wall_T, e):
    """PhononBoundary::applyTemperatureWall (PhononBoundary.h:165-202): for every
    boundary face and direction, the ghost takes its cell's value when the
    direction leaves the domain, e0(Twall) otherwise."""
    fcl = mesh["face_cells"].reshape(-1, 2)
    gp = mesh["group_ptr"]
    for k in range(e.shape[0]):
        for g in range(1, len(gp) - 1):
            for f in range(gp[g], gp[g + 1]):
                c0, c1 = fcl[f]
                en = fA[3 * f:3 * f + 3] / fAm[f]
                vdn = vg[k][0] * en[0] + vg[k][1] * en[1] + vg[k][2] * en[2]
                e[k, c1] = e[k, c0] if vdn > 0.0 else (wall_T[g] - tref[k]) * cp[k]
    return e


def phonon_advance(assemble, solve, mesh, metrics, vg, dk3, tau, cp, tref, wall_T, state, niter, rel_tol, abs_tol):
    """PhononModel::advance (PhononModel.h:530-600) with the directions solved one
    after another. assemble(vg_k, tau_k, e0_k, e_k) -> system; solve(system) ->
    (r0, delta). state: dict(e, e0, TL, niters, initial) changed in place.
    Returns (norms of the iterations, converged)."""
    e, e0, TL = state["e"], state["e0"], state["TL"]
    K = e.shape[0]
    nc = mesh["n_cells"]
    norms = []
    for _ in range(niter):
        rnorm = None
        for k in range(K):
            sysk = assemble(vg[k], tau[k], e0[k][:nc], e[k])
            r0, delta = solve(sysk)
            rnorm = r0 if rnorm is None else rnorm + r0       # :557-572
            solve_boundary(sysk, delta)                        # postSolve
            e[k] += delta                                      # updateSolution
        if state.get("initial") is None:
            state["initial"] = rnorm
        if state["niters"] < 5:
            state["initial"] = max(state["initial"], rnorm)    # setMax
        ratio = rnorm / state["initial"]
        norms.append(rnorm)
        state["niters"] += 1
        if rnorm < abs_tol or ratio < rel_tol:
            return norms, True
        phonon_update_TL(e, dk3, tau, cp, tref, TL)
        for k in range(K):
            e0[k] = (TL - tref[k]) * cp[k]                     # updatee0
        phonon_walls(mesh, metrics["fA"], metrics["fAm"], vg, cp, tref, wall_T, e)
    return norms, False


def phonon_init(nct, K, cp, tref, T_init):
    """PhononModel::init (PhononModel.h:115-160)."""
    TL = np.full(nct, float(T_init))
    e0 = np.stack([(TL - tref[k]) * cp[k] for k in range(K)])
    return dict(e=e0.copy(), e0=e0, TL=TL, niters=0, initial=None)
