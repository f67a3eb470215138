"""Failure handling on the GPU: a non-finite residual ends a solve with
FVM_E_NUMERIC instead of returning NaN norms (AMG::solve would print nan and
go on; here the caller gets an error it can act on)."""
import numpy as np
import pytest

import matgen

pytestmark = pytest.mark.gpu


def _dev(a, dt):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda:0")


def _amg(m, prec):
    import torch
    import fvm_amd
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.gsOrdering = 1
    amg.precision = prec
    amg.nMaxIterations = 20
    amg.setup_device(m["n_self"], m["n_total"], _dev(m["row"], torch.int32), _dev(m["col"], torch.int32),
                     _dev(m["diag"], torch.float64), _dev(m["off"], torch.float64))
    return amg


@pytest.mark.parametrize("prec", [0, 1])
def test_non_finite_rhs_is_an_error(prec):
    import torch
    import fvm_amd
    from fvm_amd._lib import FVM_E_NUMERIC, FvmError
    m = matgen.grid3d(12, seed=5, ghosts=False, contrast=1.5)
    b = m["b"].copy()
    b[7] = np.nan
    amg = _amg(m, prec)
    x = torch.zeros(m["n_self"], dtype=torch.float64, device="cuda:0")
    with pytest.raises(FvmError) as e:
        amg.solve_device(_dev(b, torch.float64), x)
    assert e.value.code == FVM_E_NUMERIC
    fg = fvm_amd.FGMRES(amg)
    fg.nMaxIterations = 20
    x = torch.zeros(m["n_self"], dtype=torch.float64, device="cuda:0")
    with pytest.raises(FvmError) as e:
        fg.solve_device(_dev(b, torch.float64), x)
    assert e.value.code == FVM_E_NUMERIC
    # the handle stays usable: a finite right-hand side solves
    x = torch.zeros(m["n_self"], dtype=torch.float64, device="cuda:0")
    r0, rf, it = amg.solve_device(_dev(m["b"], torch.float64), x)
    assert np.isfinite(rf) and rf < r0
