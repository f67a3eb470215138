"""ctypes binding of libfvm_amd.so (declared in include/fvm_amd.h).

The library must have been built (``python -m fvm_amd.build`` or
``__graft_entry__.build()``); importing the solver without it raises, there is
no CPU fallback.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfvm_amd.so")

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p
_ll = ctypes.c_longlong

FVM_OK = 0
FVM_E_INVALID = -1
FVM_E_HIP = -2
FVM_E_NOT_SETUP = -3
FVM_E_NUMERIC = -4
FVM_E_NOMEM = -5


class AmgParams(ctypes.Structure):
    _fields_ = [
        ("maxCoarseLevels", _i), ("nPreSweeps", _i), ("nPostSweeps", _i), ("coarseGroupSize", _i),
        ("weightRatioThreshold", _d), ("cycleType", _i), ("smootherType", _i),
        ("scaleCorrections", _i), ("nMaxIterations", _i), ("verbosity", _i),
        ("relativeTolerance", _d), ("absoluteTolerance", _d),
        ("isSymmetric", _i), ("parallelStopRule", _i), ("gsOrdering", _i), ("precision", _i),
    ]


class FgmresParams(ctypes.Structure):
    _fields_ = [("restart", _i), ("nMaxIterations", _i), ("verbosity", _i),
                ("relativeTolerance", _d), ("absoluteTolerance", _d)]


class LinSolverParams(ctypes.Structure):
    _fields_ = [("nMaxIterations", _i), ("verbosity", _i), ("relativeTolerance", _d),
                ("absoluteTolerance", _d)]


class KineticOptions(ctypes.Structure):
    _fields_ = [("rho_init", _d), ("T_init", _d), ("nonDimLt", _d), ("molecularWeight", _d), ("Tmuref", _d),
                ("muref", _d), ("mu_w", _d), ("Prandtl", _d), ("fgamma", _i), ("ktrial", _i),
                ("ToleranceX", _d), ("ToleranceF", _d), ("epsilon", _d)]


# name -> (restype, argtypes)
SIGNATURES = {
    "fvm_version": (_i, []),
    "fvm_last_error": (ctypes.c_char_p, []),
    "fvm_amg_default_params": (_i, [ctypes.POINTER(AmgParams)]),
    "fvm_amg_create": (_i, [ctypes.POINTER(_p), _i]),
    "fvm_amg_destroy": (_i, [_p]),
    "fvm_amg_set_params": (_i, [_p, ctypes.POINTER(AmgParams)]),
    "fvm_amg_get_params": (_i, [_p, ctypes.POINTER(AmgParams)]),
    "fvm_amg_setup": (_i, [_p, _i, _i, _p, _p, _p, _p, _p]),
    "fvm_amg_setup_device": (_i, [_p, _i, _i, _p, _p, _p, _p, _p]),
    "fvm_amg_solve": (_i, [_p, _p, _p, _p, _p, _p]),
    "fvm_amg_solve_device": (_i, [_p, _p, _p, _p, _p, _p]),
    "fvm_amg_smooth": (_i, [_p, _p, _p]),
    "fvm_amg_smooth_device": (_i, [_p, _p, _p]),
    "fvm_amg_cleanup": (_i, [_p]),
    "fvm_amg_total_iterations": (_ll, [_p]),
    "fvm_amg_history": (_i, [_p, _p, _i]),
    "fvm_amg_num_levels": (_i, [_p]),
    "fvm_amg_level_info": (_i, [_p, _i, _p, _p]),
    "fvm_amg_level_get": (_i, [_p, _i, _p, _p, _p, _p, _p]),
    "fvm_amg_level_colours": (_i, [_p, _i, _p, _p]),
    "fvm_amg_setup_stats": (_i, [_p, _p, _p]),
    "fvm_amg_direct_level": (_i, [_p, _p, _p, _p, _p]),
    "fvm_amg_set_output": (_i, [_p, ctypes.c_char_p, ctypes.c_char_p, _i]),
    "fvm_amg_direct_apply": (_i, [_p, _p, _p]),
    "fvm_amg_set_kernel_timing": (_i, [_p, _i]),
    "fvm_amg_set_tuning": (_i, [_p, ctypes.c_char_p, _ll]),
    "fvm_amg_get_tuning": (_i, [_p, ctypes.c_char_p, ctypes.POINTER(_ll)]),
    "fvm_amg_kernel_timing": (_i, [_p, _i, _p, _p]),
    "fvm_amg_kernel_bytes": (_i, [_p, _i, _p]),
    "fvm_fgmres_default_params": (_i, [ctypes.POINTER(FgmresParams)]),
    "fvm_fgmres_create": (_i, [ctypes.POINTER(_p), _p]),
    "fvm_fgmres_destroy": (_i, [_p]),
    "fvm_fgmres_set_params": (_i, [_p, ctypes.POINTER(FgmresParams)]),
    "fvm_fgmres_solve_device": (_i, [_p, _p, _p, _p, _p, _p]),
    "fvm_fgmres_solve": (_i, [_p, _p, _p, _p, _p, _p]),
    "fvm_fgmres_history": (_i, [_p, _p, _i]),
    "fvm_linsolver_default_params": (_i, [ctypes.POINTER(LinSolverParams)]),
    "fvm_bcgstab_solve_device": (_i, [_p, ctypes.POINTER(LinSolverParams), _p, _p, _p, _p, _p, _p, _i]),
    "fvm_cg_solve_device": (_i, [_p, ctypes.POINTER(LinSolverParams), _p, _p, _p, _p, _p, _p, _i]),
    "fvm_mesh_create_hex": (_i, [ctypes.POINTER(_p), _i, _i, _d, ctypes.c_uint64]),
    "fvm_mesh_create": (_i, [ctypes.POINTER(_p), _i, _i, _p, _i, _p, _p, _i, _i, _i, _p]),
    "fvm_mesh_create_general": (_i, [ctypes.POINTER(_p), _i, _i, _i, _p, _i, _p, _p, _p, _i, _i, _i, _p]),
    "fvm_mesh_destroy": (_i, [_p]),
    "fvm_mesh_topology": (_i, [_p, _p, _p, _p]),
    "fvm_mesh_gradient": (_i, [_p, _p, _p]),
    "fvm_convdiff_assemble": (_i, [_p, _p, _p, _p, _p, _p, _i, _p, _p, ctypes.POINTER(_p)]),
    "fvm_convdiff_reassemble_noalloc": (_i, [_p, _p, _p, _p, _p, _p, _i, _p, _p, _p]),
    "fvm_mesh_info": (_i, [_p, _p, _p, _p, _p, _p]),
    "fvm_mesh_get": (_i, [_p, _p, _p, _p, _p]),
    "fvm_mesh_metrics": (_i, [_p, _p, _p, _p, _p, _p]),
    "fvm_mesh_lognormal_field": (_i, [_p, _d, _i, ctypes.c_uint64, _p]),
    "fvm_mesh_uniform_field": (_i, [_p, _d, _d, ctypes.c_uint64, _p]),
    "fvm_thermal_assemble": (_i, [_p, _p, _p, _p, _p, _p, ctypes.POINTER(_p)]),
    "fvm_phonon_assemble": (_i, [_p, _p, _d, _p, _p, ctypes.POINTER(_p)]),
    "fvm_amg_autotune": (_i, [_p, _i, ctypes.c_char_p, _i]),
    "fvm_mpnn_num_params": (_i, []),
    "fvm_mpnn_create": (_i, [ctypes.POINTER(_p), _i, ctypes.c_uint64]),
    "fvm_mpnn_destroy": (_i, [_p]),
    "fvm_mpnn_set_params": (_i, [_p, _p]),
    "fvm_mpnn_get_params": (_i, [_p, _p]),
    "fvm_mpnn_set_graph": (_i, [_p, _i, _p, _p, _p, _p]),
    "fvm_mpnn_forward": (_i, [_p, _p, _p, _i]),
    "fvm_mpnn_loss_grad": (_i, [_p, _p, _p]),
    "fvm_mpnn_loss_grad_pair": (_i, [_p, _p, _p, _p]),
    "fvm_mpnn_get_grad": (_i, [_p, _p, _i]),
    "fvm_mpnn_adam_step": (_i, [_p, ctypes.c_float]),
    "fvm_fgmres_set_learned": (_i, [_p, _p, _d]),
    "fvm_fgmres_learned_stats": (_i, [_p, _p, _p]),
    "fvm_batch_create": (_i, [ctypes.POINTER(_p), _i]),
    "fvm_batch_destroy": (_i, [_p]),
    "fvm_batch_set_params": (_i, [_p, ctypes.POINTER(AmgParams)]),
    "fvm_batch_setup_device": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _p]),
    "fvm_batch_solve_device": (_i, [_p, _p, _p, _p, _p, _p]),
    "fvm_batch_level_sizes": (_i, [_p, _i, _p, _i]),
    "fvm_batch_set_tuning": (_i, [_p, ctypes.c_char_p, _ll]),
    "fvm_batch_setup_stats": (_i, [_p, _p, _p, _p]),
    "fvm_system_info": (_i, [_p, _p, _p, _p]),
    "fvm_system_device": (_i, [_p, _p, _p, _p, _p, _p, _p]),
    "fvm_system_get": (_i, [_p, _p, _p, _p, _p, _p, _p]),
    "fvm_system_post_solve": (_i, [_p, _p]),
    "fvm_system_destroy": (_i, [_p]),
    "fvm_amg_setup_system": (_i, [_p, _p]),
    "fvm_phonon_create": (_i, [ctypes.POINTER(_p), _p, _i, _p, _p, _p, _p, _p, _p, _d]),
    "fvm_phonon_destroy": (_i, [_p]),
    "fvm_phonon_set_params": (_i, [_p, ctypes.POINTER(AmgParams)]),
    "fvm_phonon_apply_walls": (_i, [_p]),
    "fvm_phonon_advance": (_i, [_p, _i, _d, _d, _p, _p, _i, _p]),
    "fvm_phonon_get": (_i, [_p, _p, _p, _p]),
    "fvm_phonon_set_field": (_i, [_p, _p]),
    "fvm_kinetic_assemble": (_i, [_p, _p, _p, _p, _p, _p, _p, _d, ctypes.POINTER(_p)]),
    "fvm_quadrature_cartesian": (_i, [_i, _i, _i, _d, _d, _p, _p, _p, _p]),
    "fvm_kinetic_default_options": (_i, [ctypes.POINTER(KineticOptions)]),
    "fvm_kinetic_create": (_i, [ctypes.POINTER(_p), _p, _i, _p, _p, _p, _p, ctypes.POINTER(KineticOptions),
                                _p, _p, _p]),
    "fvm_kinetic_destroy": (_i, [_p]),
    "fvm_kinetic_set_params": (_i, [_p, ctypes.POINTER(AmgParams)]),
    "fvm_kinetic_advance": (_i, [_p, _i, _d, _d, _p, _p, _i, _p]),
    "fvm_kinetic_get": (_i, [_p, _p, _p, _p, _p, _p, _p]),
}

_lib = None


class FvmError(RuntimeError):
    """Raised for a non-zero status of the C ABI (the reference raises CException)."""

    def __init__(self, code, msg):
        super().__init__("fvm_amd error %d: %s" % (code, msg))
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libfvm_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        _share_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _share_hip_runtime():
    """Load torch's HIP runtime first when torch is importable.

    torch ships its own libamdhip64 with the same soname (libamdhip64.so.7) as
    /opt/rocm's. Whichever is loaded first is the one both users bind to; if
    this library loaded /opt/rocm's and torch later loaded its own copy, the
    process would hold two HIP runtimes and torch would see no device.
    """
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def check(rc):
    if rc < 0:
        raise FvmError(rc, lib().fvm_last_error().decode(errors="replace"))
    return rc
