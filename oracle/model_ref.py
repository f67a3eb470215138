"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

Restatement of the model-level driver around the solver: LinearSystem::postSolve
(LinearSystem.cpp:250-252) -> CRMatrix::solveBoundary (CRMatrix.h:433-454), and
ThermalModel::advance (ThermalModel_impl.h:424-456). Assembly and the AMG solve
are passed in (the C oracle's), so this file holds only the loop and the
boundary rows. Pinned by the Dirichlet closed form: after postSolve and the
update, a Dirichlet ghost cell holds its boundary value (tests/test_model_ref.py).
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
