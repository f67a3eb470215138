"""Device mesh and thermal finite-volume assembly (ThermalModel linearization).

Mirrors the steps of ThermalModel<T>::Impl::linearize + LinearSystem::initSolve
(src/fvm/src/modules/fvmbase/ThermalModel_impl.h:236-398, LinearSystem.cpp:40-64):
mesh metrics, DiffusionDiscretization, SourceDiscretization, Dirichlet
('SpecifiedTemperature') and Neumann ('SpecifiedHeatFlux') boundary conditions
and boundary-row elimination, all on the GPU through libfvm_amd.so.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .linear_system import LinearSystem

FVM_BC_DIRICHLET, FVM_BC_NEUMANN = 0, 1   # enum of fvm_amd.h

BC_DIRICHLET = 0
BC_NEUMANN = 1


def _vp(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Mesh:
    """Quad-face mesh in device memory (interior cells + one ghost per boundary face)."""

    def __init__(self, handle, device=0):
        self._h = handle
        self.device = device
        nc, ng, nf, nn, ngr = (ctypes.c_int() for _ in range(5))
        check(lib().fvm_mesh_info(self._h, ctypes.byref(nc), ctypes.byref(ng), ctypes.byref(nf),
                                  ctypes.byref(nn), ctypes.byref(ngr)))
        self.n_cells, self.n_ghost, self.n_faces, self.n_nodes, self.n_groups = (
            nc.value, ng.value, nf.value, nn.value, ngr.value)
        dim, nfn = ctypes.c_int(), ctypes.c_longlong()
        check(lib().fvm_mesh_topology(self._h, ctypes.byref(dim), ctypes.byref(nfn), None))
        self.dim, self.n_face_nodes = dim.value, nfn.value

    @classmethod
    def hex(cls, n, jitter=0.0, seed=0, device=0):
        h = ctypes.c_void_p()
        check(lib().fvm_mesh_create_hex(ctypes.byref(h), int(device), int(n), float(jitter), int(seed)))
        return cls(h, device)

    @classmethod
    def from_arrays(cls, nodes, face_nodes, face_cells, n_cells, n_ghost, group_ptr, device=0, face_nptr=None,
                    dim=3):
        """Mesh from host arrays: nodes (3 per node), face_nodes (a CSR over face_nptr;
        four per face when face_nptr is None), face_cells (2 per face), group_ptr."""
        nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1)
        face_nodes = np.ascontiguousarray(face_nodes, np.int32).reshape(-1)
        face_cells = np.ascontiguousarray(face_cells, np.int32).reshape(-1)
        group_ptr = np.ascontiguousarray(group_ptr, np.int32)
        nf = face_cells.size // 2
        if face_nptr is None:
            face_nptr = np.arange(nf + 1, dtype=np.int32) * 4
        face_nptr = np.ascontiguousarray(face_nptr, np.int32)
        h = ctypes.c_void_p()
        check(lib().fvm_mesh_create_general(ctypes.byref(h), int(device), int(dim), nodes.size // 3, _vp(nodes),
                                            nf, _vp(face_nptr), _vp(face_nodes), _vp(face_cells), int(n_cells),
                                            int(n_ghost), group_ptr.size - 1, _vp(group_ptr)))
        return cls(h, device)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_mesh_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def n_total(self):
        return self.n_cells + self.n_ghost

    def arrays(self):
        nodes = np.zeros(3 * self.n_nodes)
        fn = np.zeros(self.n_face_nodes, np.int32)
        fc = np.zeros(2 * self.n_faces, np.int32)
        gp = np.zeros(self.n_groups + 1, np.int32)
        fp = np.zeros(self.n_faces + 1, np.int32)
        check(lib().fvm_mesh_get(self._h, _vp(nodes), _vp(fn), _vp(fc), _vp(gp)))
        check(lib().fvm_mesh_topology(self._h, None, None, _vp(fp)))
        return dict(nodes=nodes, face_nodes=fn, face_cells=fc, group_ptr=gp, face_nptr=fp, dim=self.dim,
                    n_cells=self.n_cells, n_ghost=self.n_ghost)

    def gradient(self, x, out=None):
        """Least-squares gradient (GradientModel::compute) of a device cell field:
        torch float64 tensor of n_total entries -> (n_total, 3) tensor."""
        import torch
        if out is None:
            out = torch.empty((self.n_total, 3), dtype=torch.float64, device=x.device)
        check(lib().fvm_mesh_gradient(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr())))
        return out

    def metrics(self):
        nf, nct = self.n_faces, self.n_total
        fA = np.zeros(3 * nf); fAm = np.zeros(nf); fc = np.zeros(3 * nf)
        cc = np.zeros(3 * nct); vol = np.zeros(nct)
        check(lib().fvm_mesh_metrics(self._h, _vp(fA), _vp(fAm), _vp(fc), _vp(cc), _vp(vol)))
        return dict(fA=fA, fAm=fAm, fc=fc, cc=cc, vol=vol)

    def metrics_device(self):
        """Compute connectivity and metrics on the device without copying them out."""
        check(lib().fvm_mesh_metrics(self._h, None, None, None, None, None))

    def lognormal_field(self, out, sigma, block=1, seed=0):
        """Fill a device tensor (torch, float64, n_total entries)."""
        check(lib().fvm_mesh_lognormal_field(self._h, float(sigma), int(block), int(seed),
                                             ctypes.c_void_p(out.data_ptr())))
        return out

    def uniform_field(self, out, lo=0.0, hi=1.0, seed=0):
        check(lib().fvm_mesh_uniform_field(self._h, float(lo), float(hi), int(seed),
                                           ctypes.c_void_p(out.data_ptr())))
        return out


class System:
    """Assembled CRMatrix system in device memory (owned by the library)."""

    def __init__(self, handle):
        self._h = handle
        ns, nt = ctypes.c_int(), ctypes.c_int()
        nz = ctypes.c_longlong()
        check(lib().fvm_system_info(self._h, ctypes.byref(ns), ctypes.byref(nt), ctypes.byref(nz)))
        self.n_self, self.n_total, self.nnz = ns.value, nt.value, nz.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_system_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def device_pointers(self):
        ptrs = [ctypes.c_void_p() for _ in range(6)]
        check(lib().fvm_system_device(self._h, *[ctypes.byref(p) for p in ptrs]))
        return dict(zip(["row", "col", "diag", "off", "b", "is_boundary"], [p.value for p in ptrs]))

    def to_host(self):
        nt, nz = self.n_total, self.nnz
        row = np.zeros(nt + 1, np.int32); col = np.zeros(max(nz, 1), np.int32)
        diag = np.zeros(nt); off = np.zeros(max(nz, 1)); b = np.zeros(nt); isb = np.zeros(nt, np.uint8)
        check(lib().fvm_system_get(self._h, _vp(row), _vp(col), _vp(diag), _vp(off), _vp(b), _vp(isb)))
        return dict(n_self=self.n_self, n_total=nt, row=row, col=col[:nz], diag=diag, off=off[:nz], b=b,
                    is_boundary=isb)

    def post_solve(self, delta):
        """LinearSystem::postSolve (LinearSystem.cpp:250-252): the boundary rows'
        entries of the device vector delta (n_total) from the self rows' ones."""
        check(lib().fvm_system_post_solve(self._h, ctypes.c_void_p(delta.data_ptr())))

    def linear_system(self):
        h = self.to_host()
        return LinearSystem(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["b"],
                            is_boundary=h["is_boundary"])


def assemble(mesh, k, source, x, bc_types, bc_values, flux=None, continuity_residual=None, central=False):
    """ThermalModel linearization on the device: k, source, x are torch float64
    tensors with mesh.n_total entries; bc_types/bc_values per face group (index 0
    unused). flux (n_faces, optional): the convecting face flux of the
    ConvectionDiscretization term; continuity_residual (n_cells, optional) is
    added to the diagonal; central selects the reference's central-difference branch."""
    t = np.ascontiguousarray(bc_types, np.int32)
    v = np.ascontiguousarray(bc_values, np.float64)
    if t.size != mesh.n_groups or v.size != mesh.n_groups:
        raise ValueError("need one boundary condition per face group")
    dp = lambda a: None if a is None else ctypes.c_void_p(a.data_ptr())
    if flux is not None and flux.numel() != mesh.n_faces:
        raise ValueError("flux needs one entry per face")
    if continuity_residual is not None and continuity_residual.numel() < mesh.n_cells:
        raise ValueError("continuity_residual needs one entry per cell")
    h = ctypes.c_void_p()
    check(lib().fvm_convdiff_assemble(mesh._h, dp(k), dp(source), dp(x), dp(flux), dp(continuity_residual),
                                      int(bool(central)), _vp(t), _vp(v), ctypes.byref(h)))
    return System(h)


def assemble_phonon(mesh, vg, tau, e0, x):
    """One discrete-ordinate direction of the gray phonon BTE on the device
    (PhononModel::linearizePhononModel, PhononModel.h:437-523): vg is the group
    velocity (3 floats), tau the relaxation time, e0 (n_cells) and x
    (mesh.n_total, ghosts hold boundary values) torch float64 tensors."""
    v = np.ascontiguousarray(vg, np.float64)
    if v.size != 3:
        raise ValueError("vg needs 3 components")
    h = ctypes.c_void_p()
    check(lib().fvm_phonon_assemble(mesh._h, _vp(v), float(tau), ctypes.c_void_p(e0.data_ptr()),
                                    ctypes.c_void_p(x.data_ptr()), ctypes.byref(h)))
    return System(h)


class ThermalModel:
    """ThermalModel::advance (ThermalModel_impl.h:424-456) on the device.

    Each outer iteration: linearize around the current temperature x (assemble,
    with initSolve's elimination), solve for the delta with ``solver`` (an
    fvm_amd.AMG, or a BCGStab / CG / FGMRES around one; the hierarchy is rebuilt per
    iteration as after the reference's cleanup), postSolve (boundary rows), x += delta. The stop rule is the
    reference's: the solver's returned initial norm against the first one.
    """

    def __init__(self, mesh, k, source, x, bc_types, bc_values, solver, relative_tolerance=1e-8,
                 absolute_tolerance=1e-20):
        self.mesh, self.k, self.source, self.x = mesh, k, source, x
        self.bc_types, self.bc_values = bc_types, bc_values
        self.solver = solver
        self.relativeTolerance = relative_tolerance      # ThermalModelOptions
        self.absoluteTolerance = absolute_tolerance
        self.niters = 0
        self.initial_norm = None
        self.norms = []
        # Dirichlet ghost cells: the linearization sets them to their boundary values
        # (GenericBCS.h:104); the assembly itself reads x only
        import torch
        a = mesh.arrays()
        gp = np.asarray(a["group_ptr"])
        fc = np.asarray(a["face_cells"]).reshape(-1, 2)
        idx, val = [], []
        for g in range(1, len(gp) - 1):
            if int(bc_types[g]) == FVM_BC_DIRICHLET:
                idx.append(fc[gp[g]:gp[g + 1], 1])
                val.append(np.full(gp[g + 1] - gp[g], float(bc_values[g])))
        self._dir_idx = torch.tensor(np.concatenate(idx) if idx else np.zeros(0, np.int64), dtype=torch.int64,
                                     device=x.device)
        self._dir_val = torch.tensor(np.concatenate(val) if val else np.zeros(0), dtype=torch.float64,
                                     device=x.device)

    def advance(self, niter):
        import torch
        from ._lib import check as _check
        for _ in range(niter):
            S = assemble(self.mesh, self.k, self.source, self.x, self.bc_types, self.bc_values)
            if self._dir_idx.numel():
                self.x[self._dir_idx] = self._dir_val          # _x[c1] = bValue
            torch.cuda.synchronize(self.x.device)              # torch's stream, then the library's
            pc = getattr(self.solver, "preconditioner", self.solver)   # an AMG, or a Krylov solver's AMG
            pc._push_params()
            _check(lib().fvm_amg_setup_system(pc._h, S._h))
            delta = torch.zeros_like(self.x)
            r_norm = self.solver.solve_device(S.device_pointers()["b"], delta)[0]
            if self.initial_norm is None:
                self.initial_norm = r_norm
            ratio = r_norm / self.initial_norm if self.initial_norm else 0.0
            self.norms.append(r_norm)
            S.post_solve(delta)
            self.x += delta                                # updateSolution
            torch.cuda.synchronize(self.x.device)
            self.niters += 1
            if r_norm < self.absoluteTolerance or ratio < self.relativeTolerance:
                break
        return self.norms
