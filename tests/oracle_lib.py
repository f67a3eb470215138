"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure only.

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module; the product package never does.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


class OmParams(ctypes.Structure):
    _fields_ = [
        ("max_coarse_levels", _i), ("n_pre_sweeps", _i), ("n_post_sweeps", _i),
        ("coarse_group_size", _i), ("weight_ratio_threshold", _d), ("cycle_type", _i),
        ("smoother_type", _i), ("scale_corrections", _i), ("n_max_iterations", _i),
        ("verbosity", _i), ("relative_tolerance", _d), ("absolute_tolerance", _d),
        ("parallel_stop_rule", _i), ("is_symmetric", _i),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library not built: %s (run __graft_entry__.build())" % ORACLE_SO)
        L = ctypes.CDLL(ORACLE_SO)
        L.om_amg_create.restype = _p
        L.om_amg_destroy.argtypes = [_p]
        L.om_amg_set_params.argtypes = [_p, ctypes.POINTER(OmParams)]
        L.om_default_params.argtypes = [ctypes.POINTER(OmParams)]
        L.om_amg_setup.argtypes = [_p, _i, _i, _p, _p, _p, _p, _p]
        L.om_amg_solve.argtypes = [_p, _p, _p, _p, _p, _p, _p, _i]
        L.om_amg_smooth.argtypes = [_p, _p, _p]
        L.om_amg_num_levels.argtypes = [_p]
        L.om_amg_level_info.argtypes = [_p, _i, _p, _p, _p]
        L.om_amg_level_get.argtypes = [_p, _i, _p, _p, _p, _p, _p]
        L.om_amg_total_iterations.argtypes = [_p]
        L.om_amg_total_iterations.restype = ctypes.c_longlong
        L.om_mm_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _p, _p, _p, _p, _p, _p, _p, _p]
        L.om_residual_flat.argtypes = [_i, _p, _p, _p, _p, _p, _p, _p]
        L.om_gs_flat.argtypes = [_i, _p, _p, _p, _p, _p, _p, _i]
        L.om_coarsen_flat.argtypes = [_i, _i, _p, _p, _p, _p, _p, _i, _d, _p]
        mesh_args = [_i, _i, _p, _p, _p, _p, _i, _i, _i, _p]
        L.ofv_metrics.argtypes = mesh_args + [_p, _p, _p, _p, _p]
        L.ofv_thermal_nnz.argtypes = [_i, _p, _i, _i]
        L.ofv_gradient.argtypes = mesh_args + [_p, _p]
        L.ofv_thermal.argtypes = mesh_args + [_p, _p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p]
        L.ofv_phonon.argtypes = mesh_args + [_p, _d, _p, _p, _p, _p, _p, _p, _p, _p]
        L.ofv_kinetic.argtypes = mesh_args + [_p, _p, _p, _p, _p, _p, _d, _p, _p, _p, _p, _p, _p]
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data


def load_mm(mat_path, rhs_path):
    """MatrixMarket + rhs -> dict(row, col, diag, off, b, sym) following MMReader order."""
    L = lib()
    n = ctypes.c_int(); nnz = ctypes.c_int(); sym = ctypes.c_int()
    rc = L.om_mm_load(mat_path.encode(), rhs_path.encode(), ctypes.byref(n), ctypes.byref(nnz),
                      ctypes.byref(sym), None, None, None, None, None)
    if rc:
        raise RuntimeError("om_mm_load failed %d" % rc)
    N, Z = n.value, nnz.value
    row = np.zeros(N + 1, np.int32); col = np.zeros(max(Z, 1), np.int32)
    diag = np.zeros(N, np.float64); off = np.zeros(max(Z, 1), np.float64); b = np.zeros(N, np.float64)
    rc = L.om_mm_load(mat_path.encode(), rhs_path.encode(), ctypes.byref(n), ctypes.byref(nnz),
                      ctypes.byref(sym), ptr(row), ptr(col), ptr(diag), ptr(off), ptr(b))
    return dict(n_self=N, n_total=N, row=row, col=col[:Z], diag=diag, off=off[:Z], b=b,
                sym=bool(sym.value))


class OracleAMG:
    """Mirror of the reference AMG attribute names (AMG.h:73-81, LinearSolver.h:31-34)."""

    def __init__(self, **kw):
        self._L = lib()
        self._h = self._L.om_amg_create()
        p = OmParams()
        self._L.om_default_params(ctypes.byref(p))
        self._p = p
        names = {f for f, _ in OmParams._fields_}
        for k, v in kw.items():
            if k not in names:
                raise TypeError("unknown oracle parameter %r" % k)
            setattr(p, k, v)

    def __del__(self):
        try:
            self._L.om_amg_destroy(self._h)
        except Exception:
            pass

    def setup(self, n_self, n_total, row, col, diag, off, is_bnd=None):
        self._L.om_amg_set_params(self._h, ctypes.byref(self._p))
        self._keep = [np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                      np.ascontiguousarray(diag, np.float64), np.ascontiguousarray(off, np.float64),
                      None if is_bnd is None else np.ascontiguousarray(is_bnd, np.uint8)]
        r, c, d, o, ib = self._keep
        rc = self._L.om_amg_setup(self._h, n_self, n_total, ptr(r), ptr(c), ptr(d), ptr(o), ptr(ib))
        if rc:
            raise RuntimeError("oracle setup failed")
        self.n_self, self.n_total = n_self, n_total

    def solve(self, b, x=None, hist_cap=100000):
        self._L.om_amg_set_params(self._h, ctypes.byref(self._p))
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n_total) if x is None else np.array(x, np.float64)
        r0 = ctypes.c_double(); rf = ctypes.c_double(); cyc = ctypes.c_int()
        hist = np.zeros(hist_cap)
        self._L.om_amg_solve(self._h, ptr(b), ptr(x), ctypes.byref(r0), ctypes.byref(rf),
                             ctypes.byref(cyc), ptr(hist), hist_cap)
        return dict(x=x, r0=r0.value, rfinal=rf.value, cycles=cyc.value,
                    history=hist[:cyc.value + 1].copy())

    def smooth(self, b, x):
        self._L.om_amg_set_params(self._h, ctypes.byref(self._p))
        b = np.ascontiguousarray(b, np.float64)
        x = np.array(x, np.float64)
        self._L.om_amg_smooth(self._h, ptr(b), ptr(x))
        return x

    def num_levels(self):
        return self._L.om_amg_num_levels(self._h)

    def level(self, l):
        ns = ctypes.c_int(); nt = ctypes.c_int(); nz = ctypes.c_int()
        self._L.om_amg_level_info(self._h, l, ctypes.byref(ns), ctypes.byref(nt), ctypes.byref(nz))
        row = np.zeros(nt.value + 1, np.int32); col = np.zeros(max(nz.value, 1), np.int32)
        diag = np.zeros(nt.value); off = np.zeros(max(nz.value, 1)); ci = np.zeros(nt.value, np.int32)
        self._L.om_amg_level_get(self._h, l, ptr(row), ptr(col), ptr(diag), ptr(off), ptr(ci))
        return dict(n_self=ns.value, n_total=nt.value, row=row, col=col[:nz.value], diag=diag,
                    off=off[:nz.value], ci=ci)

    def total_iterations(self):
        return self._L.om_amg_total_iterations(self._h)


def residual(m, x, b):
    r = np.zeros(m["n_self"])
    lib().om_residual_flat(m["n_self"], ptr(m["row"]), ptr(m["col"]), ptr(m["diag"]),
                           ptr(m["off"]), ptr(np.ascontiguousarray(x, np.float64)),
                           ptr(np.ascontiguousarray(b, np.float64)), ptr(r))
    return r


def gs(m, x, b, reverse=False):
    x = np.array(x, np.float64)
    lib().om_gs_flat(m["n_self"], ptr(m["row"]), ptr(m["col"]), ptr(m["diag"]), ptr(m["off"]),
                     ptr(x), ptr(np.ascontiguousarray(b, np.float64)), int(reverse))
    return x


def coarsen(m, group_size=2, theta=0.65, is_bnd=None):
    ci = np.zeros(m["n_total"], np.int32)
    ib = None if is_bnd is None else np.ascontiguousarray(is_bnd, np.uint8)
    nc = lib().om_coarsen_flat(m["n_self"], m["n_total"], ptr(m["row"]), ptr(m["col"]),
                               ptr(m["diag"]), ptr(m["off"]), ptr(ib), group_size, theta, ptr(ci))
    return nc, ci


def _mesh_args(mesh):
    """(dim, n_faces, nodes, fnptr, face_nodes, face_cells, n_cells, n_ghost, n_groups, gptr)
    of a mesh dict (tests/hexmesh.py layout; "dim" defaults to 3 and "face_nptr" to
    four nodes per face)."""
    nf = mesh["face_cells"].size // 2
    fnptr = mesh.get("face_nptr")
    if fnptr is None:
        fnptr = np.arange(nf + 1, dtype=np.int32) * 4
    keep = [np.ascontiguousarray(fnptr, np.int32), np.ascontiguousarray(mesh["nodes"], np.float64),
            np.ascontiguousarray(mesh["face_nodes"], np.int32), np.ascontiguousarray(mesh["face_cells"], np.int32),
            np.ascontiguousarray(mesh["group_ptr"], np.int32)]
    fp, nodes, fn, fcl, gp = keep
    args = [int(mesh.get("dim", 3)), nf, ptr(nodes), ptr(fp), ptr(fn), ptr(fcl), mesh["n_cells"], mesh["n_ghost"],
            gp.size - 1, ptr(gp)]
    return args, keep


def fv_metrics(mesh):
    """Oracle mesh metrics of a mesh dict (tests/hexmesh.py layout)."""
    nf = mesh["face_cells"].size // 2
    nct = mesh["n_cells"] + mesh["n_ghost"]
    out = dict(fA=np.zeros(3 * nf), fAm=np.zeros(nf), fc=np.zeros(3 * nf), cc=np.zeros(3 * nct),
               vol=np.zeros(nct))
    args, keep = _mesh_args(mesh)
    lib().ofv_metrics(*args, ptr(out["fA"]), ptr(out["fAm"]), ptr(out["fc"]), ptr(out["cc"]), ptr(out["vol"]))
    return out


def fv_gradient(mesh, x):
    """Oracle least-squares gradient (GradientModel) of a cell field -> (n_total, 3)."""
    nct = mesh["n_cells"] + mesh["n_ghost"]
    g = np.zeros(3 * nct)
    args, keep = _mesh_args(mesh)
    lib().ofv_gradient(*args, ptr(np.ascontiguousarray(x, np.float64)), ptr(g))
    return g.reshape(-1, 3)


def fv_thermal(mesh, k, src, x, bc_type, bc_value, conv_flux=None, cont_res=None, central=False):
    """Oracle thermal (convection-diffusion) assembly -> CRMatrix system dict."""
    L = lib()
    nf = mesh["face_cells"].size // 2
    nc, ng = mesh["n_cells"], mesh["n_ghost"]
    nct = nc + ng
    nnz = L.ofv_thermal_nnz(nf, ptr(np.ascontiguousarray(mesh["face_cells"], np.int32)), nc, ng)
    row = np.zeros(nct + 1, np.int32); col = np.zeros(max(nnz, 1), np.int32)
    diag = np.zeros(nct); off = np.zeros(max(nnz, 1)); b = np.zeros(nct); isb = np.zeros(nct, np.uint8)
    bt = np.ascontiguousarray(bc_type, np.int32); bv = np.ascontiguousarray(bc_value, np.float64)
    k = np.ascontiguousarray(k, np.float64); src = np.ascontiguousarray(src, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    cf = None if conv_flux is None else np.ascontiguousarray(conv_flux, np.float64)
    cr = None if cont_res is None else np.ascontiguousarray(cont_res, np.float64)
    args, keep = _mesh_args(mesh)
    L.ofv_thermal(*args, ptr(bt), ptr(bv), ptr(k), ptr(src), ptr(x), ptr(cf), ptr(cr), int(bool(central)),
                  ptr(row), ptr(col), ptr(diag), ptr(off), ptr(b), ptr(isb))
    return dict(n_self=nc, n_total=nct, row=row, col=col[:nnz], diag=diag, off=off[:nnz], b=b,
                is_boundary=isb)


def fv_kinetic(mesh, c, nue, feq, x, bc_kind, wall_u, eps=1e-50):
    """Oracle assembly of one ESBGK velocity direction (collision + upwind convection +
    kinetic boundary faces) -> system dict. bc_kind[g]: 1 zero gradient, 0 wall-like;
    wall_u: n_groups x 3 wall velocities."""
    L = lib()
    nf = mesh["face_cells"].size // 2
    nc, ng = mesh["n_cells"], mesh["n_ghost"]
    nct = nc + ng
    nnz = L.ofv_thermal_nnz(nf, ptr(np.ascontiguousarray(mesh["face_cells"], np.int32)), nc, ng)
    row = np.zeros(nct + 1, np.int32); col = np.zeros(max(nnz, 1), np.int32)
    diag = np.zeros(nct); off = np.zeros(max(nnz, 1)); b = np.zeros(nct); isb = np.zeros(nct, np.uint8)
    cv = np.ascontiguousarray(c, np.float64)
    nue = np.ascontiguousarray(nue, np.float64)
    feq = np.ascontiguousarray(feq, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    kind = np.ascontiguousarray(bc_kind, np.int32)
    wu = np.ascontiguousarray(wall_u, np.float64).reshape(-1)
    args, keep = _mesh_args(mesh)
    L.ofv_kinetic(*args, ptr(cv), ptr(nue), ptr(feq), ptr(x), ptr(kind), ptr(wu), ctypes.c_double(eps),
                  ptr(row), ptr(col), ptr(diag), ptr(off), ptr(b), ptr(isb))
    return dict(n_self=nc, n_total=nct, row=row, col=col[:nnz], diag=diag, off=off[:nnz], b=b,
                is_boundary=isb)


def fv_phonon(mesh, vg, tau, e0, x):
    """Oracle phonon-direction assembly (upwind convection + relaxation) -> system dict."""
    L = lib()
    nf = mesh["face_cells"].size // 2
    nc, ng = mesh["n_cells"], mesh["n_ghost"]
    nct = nc + ng
    nnz = L.ofv_thermal_nnz(nf, ptr(np.ascontiguousarray(mesh["face_cells"], np.int32)), nc, ng)
    row = np.zeros(nct + 1, np.int32); col = np.zeros(max(nnz, 1), np.int32)
    diag = np.zeros(nct); off = np.zeros(max(nnz, 1)); b = np.zeros(nct); isb = np.zeros(nct, np.uint8)
    v = np.ascontiguousarray(vg, np.float64)
    e0 = np.ascontiguousarray(e0, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    args, keep = _mesh_args(mesh)
    L.ofv_phonon(*args, ptr(v), ctypes.c_double(tau), ptr(e0), ptr(x), ptr(row), ptr(col),
                 ptr(diag), ptr(off), ptr(b), ptr(isb))
    return dict(n_self=nc, n_total=nct, row=row, col=col[:nnz], diag=diag, off=off[:nnz], b=b,
                is_boundary=isb)
