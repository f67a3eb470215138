"""BGK kinetic model (KineticModel::advance, src/fvm/src/modules/esbgkbase/
KineticModel.h:3360-3462) on the device through libfvm_amd.so: every discrete
velocity of an outer iteration is assembled and all are solved together by one
batched AMG; the moments, collision frequency, equilibrium and boundary values
run as device kernels."""
import ctypes

import numpy as np

from ._lib import AmgParams, KineticOptions, check, lib

_KINDS = {"wall": 0, "zerograd": 1, "symmetry": 2, "copy": 3}


def _vp(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Quadrature:
    """Cartesian velocity quadrature, Quadrature(N1, N2, N3, clim, T2)
    (Quadrature.h:83-181): cx, cy, cz, dcxyz (weights) as numpy arrays."""

    def __init__(self, N1, N2, N3, clim=5.5, T2=1.0):
        K = int(N1) * int(N2) * int(N3)
        self.cx, self.cy, self.cz, self.dcxyz = (np.zeros(K) for _ in range(4))
        check(lib().fvm_quadrature_cartesian(int(N1), int(N2), int(N3), float(clim), float(T2), _vp(self.cx),
                                             _vp(self.cy), _vp(self.cz), _vp(self.dcxyz)))

    def getDirCount(self):
        return len(self.cx)


def default_options():
    """KineticModelOptions defaults (KineticBC.h:61-113): fgamma 2 (ES-BGK)."""
    o = KineticOptions()
    check(lib().fvm_kinetic_default_options(ctypes.byref(o)))
    return o


class KineticModel:
    """bcs: {face group: ("wall", (u, v, w), T) | ("zerograd",) | ("symmetry",) |
    ("copy",)} for every boundary group of the mesh ("copy": a group type without a
    boundary-value rule, whose ghosts keep their values); options by the reference's names (rho_init,
    T_init, mu_w, Tmuref, muref, molecularWeight, nonDimLt, Prandtl, fgamma, ktrial,
    ToleranceX, ToleranceF, epsilon). The per-direction AMG's parameters follow the
    reference's AMG names (set_solver)."""

    def __init__(self, mesh, quadrature, bcs, **options):
        q = quadrature
        self.K = q.getDirCount()
        self.mesh = mesh                      # borrowed by the library: kept alive here
        ng = mesh.n_groups
        kind = np.zeros(ng, np.int32)
        wall_u = np.zeros((ng, 3))
        wall_T = np.ones(ng)
        for g in range(1, ng):
            if g not in bcs:
                raise ValueError("face group %d has no boundary condition" % g)
            bc = bcs[g]
            kind[g] = _KINDS[bc[0]]
            if bc[0] == "wall":
                wall_u[g] = bc[1]
                wall_T[g] = bc[2]
        o = default_options()
        for k, v in options.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        self.options = o
        self._keep = [np.ascontiguousarray(a, np.float64) for a in (q.cx, q.cy, q.cz, q.dcxyz)]
        self._keep += [kind, np.ascontiguousarray(wall_u.reshape(-1)), wall_T]
        h = ctypes.c_void_p()
        check(lib().fvm_kinetic_create(ctypes.byref(h), mesh._h, self.K, *[_vp(a) for a in self._keep[:4]],
                                       ctypes.byref(o), *[_vp(a) for a in self._keep[4:]]))
        self._h = h
        p = AmgParams()
        lib().fvm_amg_default_params(ctypes.byref(p))
        p.verbosity = 0
        # KineticModelOptions' linear solver (COMETBC.h:166-177): AMG to 1e-15, 100 cycles
        p.relativeTolerance = 1e-15
        p.nMaxIterations = 100
        self._p = p
        self.norms = []

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_kinetic_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_solver(self, **kw):
        """Per-direction AMG parameters by the reference's names (AMG.h:73-81)."""
        for k, v in kw.items():
            setattr(self._p, k, v)
        check(lib().fvm_kinetic_set_params(self._h, ctypes.byref(self._p)))

    def advance(self, niter, relTolerance=1e-14, absTolerance=1e-22):
        """Up to niter outer iterations; returns True when the stop test ended them
        (the reference's return value)."""
        check(lib().fvm_kinetic_set_params(self._h, ctypes.byref(self._p)))
        it, conv = ctypes.c_int(), ctypes.c_int()
        buf = np.zeros(max(niter, 1))
        check(lib().fvm_kinetic_advance(self._h, int(niter), float(relTolerance), float(absTolerance),
                                        ctypes.byref(it), _vp(buf), int(niter), ctypes.byref(conv)))
        self.norms.extend(buf[: it.value].tolist())
        return bool(conv.value)

    def state(self):
        """dict of host copies: f, feq (K x (n_cells + n_ghost)), density, temperature,
        nu (n_cells + n_ghost), velocity ((n_cells + n_ghost) x 3)."""
        nct = self.mesh.n_total
        f = np.zeros((self.K, nct)); feq = np.zeros((self.K, nct))
        rho = np.zeros(nct); T = np.zeros(nct); nu = np.zeros(nct); v = np.zeros((3, nct))
        check(lib().fvm_kinetic_get(self._h, _vp(f), _vp(feq), _vp(rho), _vp(v), _vp(T), _vp(nu)))
        return dict(f=f, feq=feq, density=rho, velocity=v.T.copy(), temperature=T, nu=nu)


def assemble_direction(mesh, c, nu, feq, x, bc_kind, wall_u, eps=1e-50):
    """fvm_kinetic_assemble for one velocity c on device arrays (torch tensors or raw
    pointers) -> fvm_amd.thermal.System."""
    from .thermal import System
    ptr = lambda t: ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())
    cv = np.ascontiguousarray(c, np.float64)
    kind = np.ascontiguousarray(bc_kind, np.int32)
    wu = np.ascontiguousarray(wall_u, np.float64).reshape(-1)
    out = ctypes.c_void_p()
    check(lib().fvm_kinetic_assemble(mesh._h, _vp(cv), ptr(nu), ptr(feq), ptr(x), _vp(kind), _vp(wu), float(eps),
                                     ctypes.byref(out)))
    return System(out)
