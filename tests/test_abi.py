"""The C-ABI library loads and exports every function include/fvm_amd.h declares
(no GPU calls: only symbol lookups and argument-free queries)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fvm_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_surface():
    names = declared_functions()
    for must in ["fvm_amg_create", "fvm_amg_setup", "fvm_amg_solve", "fvm_amg_smooth", "fvm_amg_cleanup",
                 "fvm_amg_total_iterations", "fvm_last_error", "fvm_version", "fvm_fgmres_solve"]:
        assert must in names


def test_library_exports_all_declared_symbols():
    from fvm_amd import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # python binding covers the whole surface
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_version_and_defaults_without_gpu():
    from fvm_amd import _lib
    L = _lib.lib()
    assert L.fvm_version() == 10000
    p = _lib.AmgParams()
    assert L.fvm_amg_default_params(ctypes.byref(p)) == 0
    # AMG.cpp:15-36, LinearSolver.h:19-24
    assert (p.maxCoarseLevels, p.nPreSweeps, p.nPostSweeps, p.coarseGroupSize) == (30, 0, 1, 2)
    assert p.weightRatioThreshold == 0.65 and p.cycleType == 0 and p.smootherType == 0
    assert p.scaleCorrections == 1 and p.nMaxIterations == 100 and p.verbosity == 2
    assert p.relativeTolerance == 1e-8 and p.absoluteTolerance == 1e-50
    f = _lib.FgmresParams()
    assert L.fvm_fgmres_default_params(ctypes.byref(f)) == 0 and f.restart == 30
