"""Mixed-precision coarse levels keep the FP64 row sums: the BF16 off-diagonals'
rounding error goes into the FP32 diagonal. Without it, the W-cycle on the 18-level
hierarchy of the 8M-cell (200^3) diffusion system turned NaN after one cycle: the
rounding tipped nearly singular coarse rows out of diagonal dominance, and the
W-cycle's many coarse visits amplified the resulting growth. With it, the mixed
W-cycle's residual history follows the FP64 W-cycle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BC_T = [0, 0, 0, 1, 1, 1, 1]
BC_V = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def _system(N, seed):
    import torch
    from fvm_amd.thermal import Mesh, assemble
    mesh = Mesh.hex(N, 0.2, seed)
    nt = mesh.n_total
    k = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    s = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    x = torch.zeros(nt, dtype=torch.float64, device="cuda:0")
    mesh.lognormal_field(k, 2.5, 8, seed + 1)
    mesh.uniform_field(s, 0.0, 1.0, seed + 2)
    return mesh, assemble(mesh, k, s, x, BC_T, BC_V)


def _w_history(S, prec, its):
    """Residual 1-norms of a stationary W-cycle solve: r0, then its - 1 cycles."""
    import torch
    import fvm_amd
    from fvm_amd._lib import check, lib
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.gsOrdering = 1
    amg.precision = prec
    amg.cycleType = 1
    amg.relativeTolerance = 1e-30
    amg.nMaxIterations = its
    amg._push_params()
    check(lib().fvm_amg_setup_system(amg._h, S._h))
    x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
    amg.solve_device(S.device_pointers()["b"], x)
    assert bool(torch.isfinite(x).all())
    return np.asarray(amg.history(), dtype=np.float64), amg.num_levels()


@pytest.mark.parametrize("N", [24, 200])
def test_mixed_w_cycle_follows_fp64(N):
    _, S = _system(N, 20261202)
    h64, nl = _w_history(S, 0, 5)
    h32, nl32 = _w_history(S, 1, 5)
    print("N", N, "levels", nl, "fp64", h64, "mixed", h32, flush=True)
    assert nl == nl32
    assert np.all(np.isfinite(h32)) and len(h32) == len(h64) == 5
    assert h32[0] == h64[0]
    # each cycle reduces the residual as the FP64 cycle does, within 1 %
    assert np.all(np.abs(h32[1:] / h64[1:] - 1.0) < 1e-2)
    assert h32[-1] < 1e-2 * h32[0]
