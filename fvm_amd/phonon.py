"""Gray phonon BTE model (PhononModel::advance, src/fvm/src/modules/phononbase/
PhononModel.h:530-600) on the device through libfvm_amd.so: all directions of an
outer iteration are assembled and solved together by one batched AMG."""
import ctypes

import numpy as np

from ._lib import AmgParams, check, lib


def _vp(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class PhononModel:
    """Directions vg (K x 3), weights dk3, relaxation times tau, gray equilibrium
    e0_k(T) = cp_k (T - tref_k); wall_T: one temperature per face group of the
    mesh (entry 0, the interior group, unused). The attribute names of the
    per-direction AMG follow the reference's AMG (set them before advance)."""

    def __init__(self, mesh, vg, dk3, tau, cp, tref, wall_T, T_init=300.0):
        vg = np.ascontiguousarray(vg, np.float64).reshape(-1, 3)
        K = vg.shape[0]
        arr = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (K,)))
        self.K = K
        self.mesh = mesh                      # borrowed by the library: kept alive here
        self._keep = [vg, arr(dk3), arr(tau), arr(cp), arr(tref), np.ascontiguousarray(wall_T, np.float64)]
        if self._keep[5].size != mesh.n_groups:
            raise ValueError("need one wall temperature per face group")
        h = ctypes.c_void_p()
        check(lib().fvm_phonon_create(ctypes.byref(h), mesh._h, K, *[_vp(a) for a in self._keep], float(T_init)))
        self._h = h
        p = AmgParams()
        lib().fvm_amg_default_params(ctypes.byref(p))
        p.verbosity = 0
        object.__setattr__(self, "_p", p)
        self.norms = []

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_phonon_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_solver(self, **kw):
        """Per-direction AMG parameters by the reference's names (AMG.h:73-81)."""
        for k, v in kw.items():
            setattr(self._p, k, v)
        check(lib().fvm_phonon_set_params(self._h, ctypes.byref(self._p)))

    def callBoundaryConditions(self):
        check(lib().fvm_phonon_apply_walls(self._h))

    def advance(self, niter, relTolerance=1e-8, absTolerance=1e-20):
        """Up to niter outer iterations; returns True while not converged (the
        reference's keepGoing)."""
        check(lib().fvm_phonon_set_params(self._h, ctypes.byref(self._p)))
        it, conv = ctypes.c_int(), ctypes.c_int()
        buf = np.zeros(max(niter, 1))
        check(lib().fvm_phonon_advance(self._h, int(niter), float(relTolerance), float(absTolerance),
                                       ctypes.byref(it), _vp(buf), int(niter), ctypes.byref(conv)))
        self.norms.extend(buf[: it.value].tolist())
        return not conv.value

    def state(self):
        """(TL, e, e0) host copies; e and e0 are K x (n_cells + n_ghost)."""
        nct = self.mesh.n_total
        TL = np.zeros(nct)
        e = np.zeros((self.K, nct))
        e0 = np.zeros((self.K, nct))
        check(lib().fvm_phonon_get(self._h, _vp(TL), _vp(e), _vp(e0)))
        return TL, e, e0

    def set_field(self, e):
        e = np.ascontiguousarray(e, np.float64)
        check(lib().fvm_phonon_set_field(self._h, _vp(e)))
