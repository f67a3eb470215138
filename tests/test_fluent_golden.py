"""The reference's mesh-based goldens, on the CPU: the Fluent reader
(fvm_amd/fluent.py) and the assembly oracle reproduce
- THERMAL_MATRIX/GOLDEN/matrix_mesh0.mat and matrix.rhs of the 32 x 32 cavity
  (ASCII case file, 2-D quads): same rows, same column order, same values
  (the golden's six decimals are exact here);
- PARALLEL_TESTS/SOLVER_AMG/ThermalSolver/*/proc1 temperatures after one
  advance with the oracle's AMG on quads (binary case), triangles, hexahedra
  and tetrahedra, within the reference's 1e-5, in the golden's cell order (the
  cell-centroid y column matches too).
"""
import numpy as np
import pytest

import fluent_golden as FG
import oracle_lib as O


def _assemble(mesh):
    a = mesh.arrays()
    nct = mesh.n_cells + mesh.n_ghost
    bt, bv = FG.thermal_bcs(mesh)
    return O.fv_thermal(a, np.ones(nct), np.zeros(nct), np.full(nct, FG.T_INIT), bt, bv)


def test_reader_layout_of_the_cavity():
    m = FG.read_mesh("cav32.cas")
    assert (m.dim, m.n_cells, m.n_ghost, m.n_faces) == (2, 1024, 128, 2112)
    assert list(m.group_ptr) == [0, 1984, 2016, 2048, 2080, 2112]
    assert m.group_ids == [0, 3, 4, 5, 6] and m.group_types == ["interior"] + ["wall"] * 4
    assert np.all(np.diff(m.face_nptr) == 2)
    # binary double-precision copy of the same mesh: identical topology and nodes
    mb = FG.read_mesh("cav32_thermal.cas")
    assert np.array_equal(mb.face_cells, m.face_cells) and np.array_equal(mb.face_nodes, m.face_nodes)
    assert np.array_equal(mb.nodes, m.nodes)
    assert mb.group_types[1] == "velocity-inlet"


@pytest.mark.parametrize("case", ["cav32.cas", "tri_894.cas", "cav_hexa.cas", "cav_tetra.cas"])
def test_reader_metrics_are_consistent(case):
    m = FG.read_mesh(case)
    g = O.fv_metrics(m.arrays())
    nc = m.n_cells
    assert np.all(g["vol"][:nc] > 0)
    assert abs(g["vol"][:nc].sum() - 1.0) < 1e-12          # unit square / cube
    # closed cells: the outward area vectors of every cell sum to zero
    fA = g["fA"].reshape(-1, 3)
    fcl = m.face_cells.reshape(-1, 2)
    net = np.zeros((nc + m.n_ghost, 3))
    np.add.at(net, fcl[:, 0], fA)
    np.add.at(net, fcl[:, 1], -fA)
    assert np.abs(net[:nc]).max() < 1e-14
    # every area vector points from c0 to c1
    cc = g["cc"].reshape(-1, 3)
    assert np.all(np.einsum("ij,ij->i", fA, cc[fcl[:, 1]] - cc[fcl[:, 0]]) > 0)


def test_thermal_matrix_golden():
    rows, rhs = FG.read_matrix_golden()
    S = _assemble(FG.read_mesh("cav32.cas"))
    mine = FG.dumped_rows(S)
    assert len(mine) == len(rows) == 1024
    for i, (a, b) in enumerate(zip(mine, rows)):
        assert [c for c, _ in a] == [c for c, _ in b], i
        np.testing.assert_allclose([v for _, v in a], [v for _, v in b], rtol=0, atol=1e-8)
    np.testing.assert_allclose(-S["b"][:1024], rhs, rtol=0, atol=1e-8)


@pytest.mark.parametrize("case,name", FG.SOLVER_CASES)
def test_solver_amg_thermal_goldens(case, name):
    m = FG.read_mesh(case)
    S = _assemble(m)
    oa = O.OracleAMG(relative_tolerance=1e-9, n_max_iterations=2000, max_coarse_levels=20)
    oa.setup(S["n_self"], S["n_total"], S["row"], S["col"], S["diag"], S["off"], S["is_boundary"])
    res = oa.solve(S["b"])
    nc = m.n_cells
    T = FG.T_INIT + res["x"][:nc]
    Tg, cyg = FG.read_temperature_golden(name)
    assert Tg.size == nc
    np.testing.assert_allclose(T, Tg, rtol=0, atol=1e-5)
    cy = O.fv_metrics(m.arrays())["cc"].reshape(-1, 3)[:nc, 1]
    np.testing.assert_allclose(cy, cyg, rtol=0, atol=1e-10)
