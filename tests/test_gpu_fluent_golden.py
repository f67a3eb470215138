"""The reference's mesh-based goldens on the GPU: meshes read by the Fluent reader
(2-D quads and triangles, 3-D hexahedra and tetrahedra) go to the device; the
device metrics, least-squares gradient and assembled system are bit-identical to
the oracle's, the device assembly reproduces the THERMAL_MATRIX golden, and one
ThermalModel::advance with the reference-order AMG reproduces the
SOLVER_AMG temperatures within the reference's 1e-5 (delta bit-identical to the
oracle's)."""
import numpy as np
import pytest

import fluent_golden as FG
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


@pytest.mark.parametrize("case", ["cav32.cas", "tri_894.cas", "cav_hexa.cas", "cav_tetra.cas"])
def test_general_mesh_metrics_gradient_and_system_bit_identical(case):
    from fvm_amd.thermal import assemble
    print("mesh", case, flush=True)
    m = FG.read_mesh(case)
    a = m.arrays()
    dm = m.to_device()
    assert dm.dim == m.dim and dm.n_face_nodes == m.face_nodes.size
    back = dm.arrays()
    for key in ("face_nptr", "face_nodes", "face_cells", "group_ptr"):
        assert np.array_equal(back[key], a[key]), key
    g = dm.metrics()
    o = O.fv_metrics(a)
    for key in ("fA", "fAm", "fc", "cc", "vol"):
        assert np.array_equal(g[key], o[key]), key
    rng = np.random.default_rng(7)
    nct = m.n_cells + m.n_ghost
    x = 300.0 + rng.normal(0, 10, nct)
    assert np.array_equal(dm.gradient(_dev(x)).cpu().numpy(), O.fv_gradient(a, x))
    bt, bv = FG.thermal_bcs(m)
    k = np.exp(rng.normal(0, 0.5, nct))
    S = assemble(dm, _dev(k), _dev(np.zeros(nct)), _dev(x), bt, bv).to_host()
    ref = O.fv_thermal(a, k, np.zeros(nct), x, bt, bv)
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(S[key], ref[key]), key


def test_thermal_matrix_golden_on_device():
    from fvm_amd.thermal import assemble
    m = FG.read_mesh("cav32.cas")
    dm = m.to_device()
    nct = m.n_cells + m.n_ghost
    bt, bv = FG.thermal_bcs(m)
    S = assemble(dm, _dev(np.ones(nct)), _dev(np.zeros(nct)), _dev(np.full(nct, FG.T_INIT)), bt, bv).to_host()
    rows, rhs = FG.read_matrix_golden()
    mine = FG.dumped_rows(S)
    for i, (a, b) in enumerate(zip(mine, rows)):
        assert [c for c, _ in a] == [c for c, _ in b], i
        np.testing.assert_allclose([v for _, v in a], [v for _, v in b], rtol=0, atol=1e-8)
    np.testing.assert_allclose(-S["b"][:1024], rhs, rtol=0, atol=1e-8)


@pytest.mark.parametrize("case,name", FG.SOLVER_CASES)
def test_solver_amg_thermal_goldens_on_device(case, name):
    import torch
    import fvm_amd
    from fvm_amd.thermal import ThermalModel
    print("advance", case, flush=True)
    m = FG.read_mesh(case)
    dm = m.to_device()
    nct = m.n_cells + m.n_ghost
    bt, bv = FG.thermal_bcs(m)
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.relativeTolerance = 1e-9
    amg.nMaxIterations = 2000
    amg.maxCoarseLevels = 20
    x = torch.full((nct,), FG.T_INIT, dtype=torch.float64, device="cuda:0")
    model = ThermalModel(dm, _dev(np.ones(nct)), _dev(np.zeros(nct)), x, bt, bv, amg)
    model.advance(1)
    Tg, _ = FG.read_temperature_golden(name)
    T = x.cpu().numpy()[: m.n_cells]
    print(case, "cycles", amg.lastCycles, "max |T - golden|", np.abs(T - Tg).max(), flush=True)
    np.testing.assert_allclose(T, Tg, rtol=0, atol=1e-5)
    # the same solve in the oracle: bit-identical delta
    a = m.arrays()
    S = O.fv_thermal(a, np.ones(nct), np.zeros(nct), np.full(nct, FG.T_INIT), bt, bv)
    oa = O.OracleAMG(relative_tolerance=1e-9, n_max_iterations=2000, max_coarse_levels=20)
    oa.setup(S["n_self"], S["n_total"], S["row"], S["col"], S["diag"], S["off"], S["is_boundary"])
    res = oa.solve(S["b"])
    assert amg.lastCycles == res["cycles"]
    assert np.array_equal(T, FG.T_INIT + res["x"][: m.n_cells])
