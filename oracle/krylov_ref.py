"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

FP64 numpy restatement of BCGStab::solve (BCGStab.cpp:27-171), the reference's
BiCGStab loop around a preconditioner's AMG::smooth, used as the checker of
fvm_amd's device BCGStab (fvm_bcgstab_solve_device). It is pinned against
scipy's independent BiCGSTAB (tests/test_krylov_ref.py) with the same
preconditioner, and by the invariant that the recurrence residual equals the
true residual b + A x.

Delta form: r = b + A x. smooth(b) is one preconditioner application from a
zero delta (it approximately solves b + A y = 0, so y ~ -A^{-1} b).
"""
import numpy as np


def bcgstab(matvec, smooth, b, x, n_max=100, rel_tol=1e-8, abs_tol=1e-50):
    """Returns dict(x, r0, rfinal, iterations, history) as BCGStab::solve leaves them.

    history: r0, then ||r||_1 after every completed iteration (the values the
    reference prints at verbosity > 0, BCGStab.cpp:132-135).
    """
    b = np.asarray(b, np.float64)
    x = np.array(x, np.float64)
    r = b + matvec(x)                                  # computeResidual, BCGStab.cpp:35
    r0 = np.abs(r).sum()                               # getOneNorm, :37
    rt = r.copy()                                      # rTilda, :40
    hist = [r0]
    p = None
    rho = alpha = omega = 0.0
    it = 0
    for i in range(n_max):                             # :63
        it = i + 1
        rho_prev = rho
        rho = float(r @ rt)                            # :67
        if p is None:
            p = r.copy()                               # :70-74
        else:
            beta = (rho / rho_prev) * (alpha / omega)  # :77-79
            p = (p - omega * v) * beta + r             # msaxpy, *=, += (:80-82)
        ph = smooth(p)                                 # :85-89
        v = matvec(ph)                                 # :91
        alpha = rho / float(rt @ v)                    # :93-96
        x = x - alpha * ph                             # :98
        r = r - alpha * v                              # :99
        if np.abs(r).sum() < abs_tol:                  # :101-106
            break
        sh = smooth(r)                                 # :108-113
        t = matvec(sh)                                 # :115
        omega = float(t @ r) / float(t @ t)            # :117-123
        x = x - omega * sh                             # :125
        r = r - omega * t                              # :126
        rn = np.abs(r).sum()                           # :128
        hist.append(rn)
        if rn < abs_tol or rn / r0 < rel_tol:          # :130, 142-143
            break
    rf = np.abs(b + matvec(x)).sum()                   # :153-154
    return dict(x=x, r0=r0, rfinal=rf, iterations=it, history=np.array(hist))


def cg(matvec, smooth, b, x, n_max=100, rel_tol=1e-8, abs_tol=1e-50):
    """CG::solve (CG.cpp:25-137); returns what bcgstab returns (history: r0, then
    ||r||_1 after every iteration)."""
    b = np.asarray(b, np.float64)
    x = np.array(x, np.float64)
    r = b + matvec(x)                                  # CG.cpp:34
    r0 = np.abs(r).sum()                               # :36
    hist = [r0]
    p = None
    rho = 0.0
    it = 0
    for i in range(n_max):                             # :61
        it = i + 1
        z = smooth(r)                                  # :63-67
        rho_prev = rho
        rho = float(r @ z)                             # :71
        if p is None:
            p = z.copy()                               # :75-78
        else:
            p = p * (rho / rho_prev) + z               # :81-83
        q = matvec(p)                                  # :86
        alpha = rho / float(p @ q)                     # :88-91
        x = x - alpha * p                              # :93
        r = r - alpha * q                              # :94
        rn = np.abs(r).sum()                           # :96
        hist.append(rn)
        if rn < abs_tol or rn / r0 < rel_tol:          # :99, 111
            break
    rf = np.abs(b + matvec(x)).sum()                   # :122-123
    return dict(x=x, r0=r0, rfinal=rf, iterations=it, history=np.array(hist))
