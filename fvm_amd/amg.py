"""Python mirror of the reference ``AMG`` / ``LinearSolver`` (and a flexible GMRES).

Attribute names, enum values and defaults follow
src/fvm/src/modules/fvmbase/AMG.h:27-110, AMG.cpp:15-36 and LinearSolver.h:15-35,
so driver scripts written against ``fvmbaseExt.AMG()`` read the same:

    solver = AMG()
    solver.relativeTolerance = 1e-9
    solver.nMaxIterations = 200000
    rnorm0 = solver.solve(ls)        # ls.delta receives the correction
    solver.cleanup()

Every call goes through the C ABI of libfvm_amd.so; errors raise ``FvmError``
(the reference raises ``CException``, seen from Python as RuntimeError).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import AmgParams, LinSolverParams, FgmresParams, check, lib


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _dptr(t):
    """Device address of a tensor-like object (``data_ptr()``) or a raw integer address."""
    return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())


class AMG:
    # AMG::CycleType / AMG::SmootherType (AMG.h:31-42)
    V_CYCLE, W_CYCLE, F_CYCLE = 0, 1, 2
    GAUSS_SEIDEL, JACOBI = 0, 1
    # extensions
    GS_EXACT, GS_MULTICOLOR = 0, 1
    PREC_FP64, PREC_MIXED = 0, 1

    _FIELDS = [f for f, _ in AmgParams._fields_]

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        check(L.fvm_amg_create(ctypes.byref(h), int(device)))
        self._h = h
        p = AmgParams()
        check(L.fvm_amg_default_params(ctypes.byref(p)))
        object.__setattr__(self, "_p", p)
        self._finest = None
        self._print_path = None
        self.fieldName = "x"          # the bracketed name of the convergence lines
        self.printStyle = 0           # 0: every cycle (serial build), 1: first and last (parallel build)
        self.device = int(device)

    # parameter attributes map onto the C struct
    def __getattr__(self, name):
        if name in AMG._FIELDS:
            return getattr(self._p, name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in AMG._FIELDS:
            setattr(self._p, name, value)
        else:
            object.__setattr__(self, name, value)
            if name in ("fieldName", "printStyle") and "printStyle" in self.__dict__:
                self._push_output()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_amg_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _push_params(self):
        check(lib().fvm_amg_set_params(self._h, ctypes.byref(self._p)))

    # ------------------------------------------------------------------ setup
    def setup(self, ls):
        """Build the hierarchy for ``ls`` (the first-call branch of AMG::solve)."""
        self.isSymmetric = int(ls.isSymmetric)
        self._push_params()
        check(lib().fvm_amg_setup(self._h, ls.n_self, ls.n_total, _ptr(ls.row), _ptr(ls.col),
                                  _ptr(ls.diag), _ptr(ls.off), _ptr(ls.is_boundary)))
        self._finest = ls

    def setup_device(self, n_self, n_total, row, col, diag, off, is_boundary=None, isSymmetric=False):
        """Build from device tensors (torch, on this handle's device)."""
        self.isSymmetric = int(isSymmetric)
        self._push_params()
        dp = lambda t: None if t is None else _dptr(t)
        check(lib().fvm_amg_setup_device(self._h, int(n_self), int(n_total), dp(row), dp(col), dp(diag),
                                         dp(off), dp(is_boundary)))
        self._finest = "device"

    # ------------------------------------------------------------------ LinearSolver API
    def solve(self, ls):
        """AMG::solve (AMG.cpp:219-282). Returns the initial residual 1-norm."""
        if self._finest is not ls:
            self.setup(ls)
        self._push_params()
        r0 = ctypes.c_double()
        rf = ctypes.c_double()
        cyc = ctypes.c_int()
        check(lib().fvm_amg_solve(self._h, _ptr(ls.b), _ptr(ls.delta), ctypes.byref(r0), ctypes.byref(rf),
                                  ctypes.byref(cyc)))
        self.lastCycles = cyc.value
        self.lastResidual = rf.value
        return r0.value

    def solve_device(self, b, x):
        """Solve on device tensors (delta form); returns (r0, r_final, cycles)."""
        self._push_params()
        r0 = ctypes.c_double()
        rf = ctypes.c_double()
        cyc = ctypes.c_int()
        check(lib().fvm_amg_solve_device(self._h, _dptr(b), _dptr(x), ctypes.byref(r0), ctypes.byref(rf),
                                         ctypes.byref(cyc)))
        self.lastCycles = cyc.value
        self.lastResidual = rf.value
        return r0.value, rf.value, cyc.value

    def smooth(self, ls):
        """AMG::smooth (AMG.cpp:285-298): one cycle on ls.delta."""
        if self._finest is not ls:
            self.setup(ls)
        self._push_params()
        check(lib().fvm_amg_smooth(self._h, _ptr(ls.b), _ptr(ls.delta)))

    def cleanup(self):
        """AMG::cleanup (AMG.cpp:212-217)."""
        check(lib().fvm_amg_cleanup(self._h))
        self._finest = None

    def getTotalIterations(self):
        return int(lib().fvm_amg_total_iterations(self._h))

    def setMergeLevelSize(self, size):
        """Merging coarse levels onto one MPI rank (AMG.h:53-65) is meaningless on a
        single device; the value is recorded and has no effect."""
        self.mergeLevelSize = int(size)

    def _push_output(self):
        path = None if self._print_path is None else self._print_path.encode()
        check(lib().fvm_amg_set_output(self._h, str(self.fieldName).encode(), path, int(self.printStyle)))

    def redirectPrintToFile(self, fname):
        """AMG.cpp:323-330: convergence lines go to ``fname`` instead of stdout."""
        self._print_path = str(fname)
        self._push_output()

    def redirectPrintToScreen(self):
        """AMG.cpp:333-337"""
        self._print_path = None
        self._push_output()

    # ------------------------------------------------------------------ introspection
    def history(self, cap=1 << 20):
        buf = np.zeros(cap)
        k = check(lib().fvm_amg_history(self._h, _ptr(buf), cap))
        return buf[:k].copy()

    def num_levels(self):
        return lib().fvm_amg_num_levels(self._h)

    def level(self, l):
        n = ctypes.c_int()
        nnz = ctypes.c_int()
        check(lib().fvm_amg_level_info(self._h, l, ctypes.byref(n), ctypes.byref(nnz)))
        row = np.zeros(n.value + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        diag = np.zeros(max(n.value, 1))
        off = np.zeros(max(nnz.value, 1))
        ci = np.zeros(max(n.value, 1), np.int32)
        check(lib().fvm_amg_level_get(self._h, l, _ptr(row), _ptr(col), _ptr(diag), _ptr(off), _ptr(ci)))
        return dict(n=n.value, nnz=nnz.value, row=row, col=col[:nnz.value], diag=diag[:n.value],
                    off=off[:nnz.value], ci=ci[:n.value])

    def level_colours(self, l):
        """Colour of every row of level l (reference order; -1 without colouring) and
        the number of colours (fvm_amg_level_colours)."""
        n = ctypes.c_int()
        check(lib().fvm_amg_level_info(self._h, l, ctypes.byref(n), None))
        col = np.zeros(max(n.value, 1), np.int32)
        k = ctypes.c_int()
        check(lib().fvm_amg_level_colours(self._h, l, _ptr(col), ctypes.byref(k)))
        return col[:n.value], k.value

    def direct_level(self, with_inverse=False):
        """Coarsest dense solve of the mixed cycle: dict(level, n, attempts[, inv])."""
        lv, n, att = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().fvm_amg_direct_level(self._h, ctypes.byref(lv), ctypes.byref(n), ctypes.byref(att), None))
        out = dict(level=lv.value, n=n.value, attempts=att.value)
        if with_inverse and n.value > 0:
            inv = np.zeros((n.value, n.value))
            check(lib().fvm_amg_direct_level(self._h, None, None, None, _ptr(inv)))
            out["inv"] = inv
        return out

    def direct_apply(self, b):
        """x = -A^{-1} b on the direct level (the cycle's dense-apply kernel)."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        check(lib().fvm_amg_direct_apply(self._h, _ptr(b), _ptr(x)))
        return x

    def set_tuning(self, key, value):
        """Performance knob that does not change results (see fvm_amg_set_tuning)."""
        check(lib().fvm_amg_set_tuning(self._h, key.encode(), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_longlong()
        check(lib().fvm_amg_get_tuning(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, reps=3):
        """Time the result-preserving launch shapes on this hierarchy and keep the
        fastest (fvm_amg_autotune); returns the report (candidates, choice)."""
        import json
        self._push_params()
        buf = ctypes.create_string_buffer(8192)
        check(lib().fvm_amg_autotune(self._h, int(reps), buf, len(buf)))
        return json.loads(buf.value.decode())

    def setup_stats(self):
        t = ctypes.c_double()
        r = ctypes.c_int()
        check(lib().fvm_amg_setup_stats(self._h, ctypes.byref(t), ctypes.byref(r)))
        return dict(setup_seconds=t.value, finest_rounds=r.value)



class BCGStab:
    """BiCGStab with an AMG preconditioner (BCGStab.cpp:27-171), FP64 on the device.

    (``CG`` below is the reference's CG.cpp:25-137 with the same interface.)

    Fields are LinearSolver's (LinearSolver.h:19-24); ``preconditioner`` is an AMG
    whose hierarchy is built, applied as AMG::smooth (one cycle from a zero delta).
    ``solve_device`` returns (r0, final true-residual 1-norm, iterations); the
    per-iteration 1-norms are in ``history()``.
    """

    _FIELDS = [f for f, _ in LinSolverParams._fields_]
    _ENTRY = "fvm_bcgstab_solve_device"

    def __init__(self, preconditioner=None):
        self.preconditioner = preconditioner
        p = LinSolverParams()
        check(lib().fvm_linsolver_default_params(ctypes.byref(p)))
        self._p = p
        self._hist = np.zeros(0)
        self._total = 0

    def __getattr__(self, name):
        if name in BCGStab._FIELDS:
            return getattr(self.__dict__["_p"], name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in BCGStab._FIELDS:
            setattr(self._p, name, value)
        else:
            object.__setattr__(self, name, value)

    def solve_device(self, b, x, hist_cap=1 << 16):
        pc = self.preconditioner
        if pc is None:
            raise ValueError("BCGStab needs a preconditioner")
        pc._push_params()
        r0 = ctypes.c_double()
        rf = ctypes.c_double()
        it = ctypes.c_int()
        buf = np.zeros(hist_cap)
        k = check(getattr(lib(), self._ENTRY)(pc._h, ctypes.byref(self._p), _dptr(b), _dptr(x), ctypes.byref(r0),
                                               ctypes.byref(rf), ctypes.byref(it), _ptr(buf), hist_cap))
        self._hist = buf[:k].copy()
        self.lastIterations = it.value
        self.lastResidual = rf.value
        self._total += it.value
        return r0.value, rf.value, it.value

    def solve(self, ls):
        """Host LinearSystem: ls.delta receives the correction; returns the initial
        1-norm, as the reference's solve does."""
        import torch
        pc = self.preconditioner
        if pc is None:
            raise ValueError("%s needs a preconditioner" % type(self).__name__)
        if pc._finest is not ls:
            pc.setup(ls)
        b = torch.from_numpy(ls.b).to("cuda:%d" % pc.device)
        x = torch.from_numpy(ls.delta).to("cuda:%d" % pc.device)
        r0, _, _ = self.solve_device(b, x)
        ls.delta[:ls.n_self] = x.cpu().numpy()[:ls.n_self]
        return r0

    def history(self):
        return self._hist.copy()

    def getTotalIterations(self):
        """Iterations over all solves (BCGStab.h:31)."""
        return self._total

    def cleanup(self):
        """The preconditioner's cleanup (BCGStab.cpp:21-24)."""
        if self.preconditioner is not None:
            self.preconditioner.cleanup()

    def smooth(self, ls):
        raise RuntimeError("cannot use %s as preconditioner" % type(self).__name__)   # BCGStab.cpp:172-176


class CG(BCGStab):
    """Preconditioned CG with an AMG preconditioner (CG.cpp:25-137), FP64 on the device."""

    _ENTRY = "fvm_cg_solve_device"


class FGMRES:
    """FP64 flexible GMRES(restart), right-preconditioned by an AMG handle's cycle.

    Plays the role of the reference's Krylov wrappers around ``AMG::smooth``
    (BCGStab.cpp:27-171): ``preconditioner`` is an AMG whose hierarchy is built.
    """

    _FIELDS = [f for f, _ in FgmresParams._fields_]

    def __init__(self, preconditioner):
        L = lib()
        self.preconditioner = preconditioner
        h = ctypes.c_void_p()
        check(L.fvm_fgmres_create(ctypes.byref(h), preconditioner._h))
        self._h = h
        p = FgmresParams()
        check(L.fvm_fgmres_default_params(ctypes.byref(p)))
        self._p = p

    def __getattr__(self, name):
        if name in FGMRES._FIELDS:
            return getattr(self.__dict__["_p"], name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in FGMRES._FIELDS:
            setattr(self._p, name, value)
        else:
            object.__setattr__(self, name, value)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_fgmres_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def solve(self, ls):
        pc = self.preconditioner
        if pc._finest is not ls:
            pc.setup(ls)
        pc._push_params()
        check(lib().fvm_fgmres_set_params(self._h, ctypes.byref(self._p)))
        r0 = ctypes.c_double()
        rf = ctypes.c_double()
        it = ctypes.c_int()
        check(lib().fvm_fgmres_solve(self._h, _ptr(ls.b), _ptr(ls.delta), ctypes.byref(r0), ctypes.byref(rf),
                                     ctypes.byref(it)))
        self.lastIterations = it.value
        self.lastResidual = rf.value
        return r0.value

    def solve_device(self, b, x):
        pc = self.preconditioner
        pc._push_params()
        check(lib().fvm_fgmres_set_params(self._h, ctypes.byref(self._p)))
        r0 = ctypes.c_double()
        rf = ctypes.c_double()
        it = ctypes.c_int()
        check(lib().fvm_fgmres_solve_device(self._h, _dptr(b), _dptr(x), ctypes.byref(r0), ctypes.byref(rf),
                                            ctypes.byref(it)))
        return r0.value, rf.value, it.value

    def history(self, cap=1 << 16):
        buf = np.zeros(cap)
        k = check(lib().fvm_fgmres_history(self._h, _ptr(buf), cap))
        return buf[:k].copy()

    def set_learned(self, net, stall_ratio=0.9):
        """Correct every preconditioner application with a trained LearnedCorrection
        (None detaches); the network is switched off for the rest of a solve when the
        residual estimate falls by less than stall_ratio over 5 iterations."""
        self._net = net
        check(lib().fvm_fgmres_set_learned(self._h, None if net is None else net._h, float(stall_ratio)))

    def learned_stats(self):
        u = ctypes.c_int(); f = ctypes.c_int()
        check(lib().fvm_fgmres_learned_stats(self._h, ctypes.byref(u), ctypes.byref(f)))
        return dict(iterations_used=u.value, fallback_at=f.value)


class BatchAMG:
    """Many independent systems through one handle (the per-direction solves of
    PhononModel::advance, PhononModel.h:530-600): per system the same result as
    an ``AMG`` solve with the same parameters (exact Gauss-Seidel order, FP64).
    Parameter attributes are those of ``AMG``."""

    _FIELDS = AMG._FIELDS

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        check(L.fvm_batch_create(ctypes.byref(h), int(device)))
        self._h = h
        p = AmgParams()
        check(L.fvm_amg_default_params(ctypes.byref(p)))
        p.verbosity = 0
        object.__setattr__(self, "_p", p)
        self.K = 0

    def __getattr__(self, name):
        if name in BatchAMG._FIELDS:
            return getattr(self.__dict__["_p"], name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in BatchAMG._FIELDS:
            setattr(self._p, name, value)
        else:
            object.__setattr__(self, name, value)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _push_params(self):
        check(lib().fvm_batch_set_params(self._h, ctypes.byref(self._p)))

    def setup_device(self, systems):
        """systems: list of dicts with n_self, n_total and device tensors row, col,
        diag, off and optionally is_boundary (e.g. ``System.device_pointers()``)."""
        K = len(systems)
        ns = np.array([s["n_self"] for s in systems], np.int32)
        nt = np.array([s["n_total"] for s in systems], np.int32)

        def parr(key):
            a = (ctypes.c_void_p * K)()
            for k, s in enumerate(systems):
                v = s.get(key)
                a[k] = None if v is None else (v if isinstance(v, int) else v.data_ptr())
            return a
        self._keep = [parr(k) for k in ("row", "col", "diag", "off", "is_boundary")]
        self._push_params()
        check(lib().fvm_batch_setup_device(self._h, K, _ptr(ns), _ptr(nt), *self._keep))
        self.K = K

    def solve_device(self, bs, xs):
        """bs, xs: lists of device tensors (or addresses); returns (r0, r_final, cycles) arrays."""
        K = self.K
        pb = (ctypes.c_void_p * K)(*[_dptr(b).value for b in bs])
        px = (ctypes.c_void_p * K)(*[_dptr(x).value for x in xs])
        r0 = np.zeros(K); rf = np.zeros(K); cyc = np.zeros(K, np.int32)
        self._push_params()
        check(lib().fvm_batch_solve_device(self._h, pb, px, _ptr(r0), _ptr(rf), _ptr(cyc)))
        return r0, rf, cyc

    def level_sizes(self, k, cap=64):
        buf = np.zeros(cap, np.int32)
        d = check(lib().fvm_batch_level_sizes(self._h, int(k), _ptr(buf), cap))
        return buf[:d].copy()

    def set_tuning(self, key, value):
        check(lib().fvm_batch_set_tuning(self._h, key.encode(), int(value)))

    def setup_stats(self):
        t = ctypes.c_double(); n = ctypes.c_int(); lv = ctypes.c_int()
        check(lib().fvm_batch_setup_stats(self._h, ctypes.byref(t), ctypes.byref(n), ctypes.byref(lv)))
        return dict(setup_seconds=t.value, total_rows=n.value, levels=lv.value)
