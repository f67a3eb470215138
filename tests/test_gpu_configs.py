"""GPU runs of the benchmark configurations against the oracle, at their full sizes.

- 32^3 uniform mesh, k = 1, T = 1 / 0 on the x walls, zero flux elsewhere,
  initial field 0, plain FP64 AMG in the reference's row order with the
  reference defaults and nMaxIterations 10000: the device assembly equals the
  oracle's, every level's map and Galerkin matrix equal the oracle's, and the
  solve to 1e-8 takes the same number of cycles with a bit-identical delta.
- 128^3 jittered mesh, lognormal k (sigma 1.5 per cell), source U[0,1]: every
  level of the hierarchy (fine-to-coarse map, CSR, Galerkin diagonal and
  off-diagonals) equals the oracle's (CRMatrix.h:468-758) for every lane width
  of the agglomeration rounds, and three reference-order cycles give a
  bit-identical delta.
- 200^3 jittered mesh, lognormal k (sigma 2.5 in 8^3 blocks): the hierarchy the
  mixed-precision solver builds equals the oracle's at every level.
- The FP64 dense inverse of the coarsest level (mixed cycle) against
  numpy.linalg on that level's matrix, and the smoothing fallback when that
  level is singular.
"""
import os
import sys

import numpy as np
import pytest

import hexmesh
import matgen
import oracle_lib as O

pytestmark = pytest.mark.gpu

SEED = 20261202
D, NM = 0, 1
BC_T = [0, D, D, NM, NM, NM, NM]
BC_V = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def _amg(**kw):
    import fvm_amd
    a = fvm_amd.AMG()
    a.verbosity = 0
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def _setup_system(amg, S):
    from fvm_amd._lib import check, lib
    amg._push_params()
    check(lib().fvm_amg_setup_system(amg._h, S._h))


def compact0(h):
    """Reference-order CSR of the self rows with ghost columns dropped (vectorized)."""
    n = h["n_self"]
    row = h["row"][: n + 1].astype(np.int64)
    col = h["col"][: row[n]]
    off = h["off"][: row[n]]
    keep = col < n
    ck = np.concatenate([[0], np.cumsum(keep)])
    return ck[row].astype(np.int32), col[keep], off[keep]


def compare_hierarchy(ga, oa, h, lane_note=""):
    L = oa.num_levels()
    assert ga.num_levels() == L, lane_note
    row0, col0, off0 = compact0(h)
    for l in range(L):
        g = ga.level(l)
        o = oa.level(l)
        if l == 0:
            assert np.array_equal(g["row"], row0), lane_note
            assert np.array_equal(g["col"], col0), lane_note
            assert np.array_equal(g["off"], off0), lane_note
            assert np.array_equal(g["diag"], h["diag"][: h["n_self"]]), lane_note
            assert np.array_equal(g["ci"], o["ci"][: h["n_self"]]), (lane_note, l)
        else:
            assert g["n"] == o["n_self"], (lane_note, l)
            for key in ("row", "col", "diag", "off", "ci"):
                assert np.array_equal(g[key], o[key]), (lane_note, l, key)
    return L


def _torch_fields(mesh, sigma, block, seed, k_const=None, source=True):
    import torch
    nt = mesh.n_total
    k = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    if k_const is None:
        mesh.lognormal_field(k, sigma, block, seed + 1)
    else:
        k.fill_(k_const)
    s = torch.zeros(nt, dtype=torch.float64, device="cuda:0")
    if source:
        mesh.uniform_field(s, 0.0, 1.0, seed + 2)
    x = torch.zeros(nt, dtype=torch.float64, device="cuda:0")
    return k, s, x


def test_uniform_32cube_plain_fp64_amg_exact():
    import torch
    from fvm_amd.thermal import Mesh, assemble
    N = 32
    print("32^3 uniform, k=1, plain FP64 reference-order AMG", flush=True)
    mesh = Mesh.hex(N, 0.0, 0)
    k, s, x = _torch_fields(mesh, 0, 1, 0, k_const=1.0, source=False)
    S = assemble(mesh, k, s, x, BC_T, BC_V)
    h = S.to_host()
    ref = O.fv_thermal(hexmesh.hex_mesh(N), np.ones(mesh.n_total), np.zeros(mesh.n_total),
                       np.zeros(mesh.n_total), BC_T, BC_V)
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(h[key], ref[key]), key
    amg = _amg(nMaxIterations=10000, relativeTolerance=1e-8)
    _setup_system(amg, S)
    b = torch.from_numpy(h["b"]).cuda()
    xd = torch.zeros(h["n_total"], dtype=torch.float64, device="cuda:0")
    r0, rf, cyc = amg.solve_device(b, xd)
    oa = O.OracleAMG(n_max_iterations=10000, relative_tolerance=1e-8)
    oa.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["is_boundary"])
    ores = oa.solve(h["b"])
    L = compare_hierarchy(amg, oa, h)
    print("levels", L, "cycles", cyc, "oracle", ores["cycles"], "ratio", rf / r0, flush=True)
    assert cyc == ores["cycles"]
    assert rf / r0 < 1e-8
    np.testing.assert_allclose(r0, ores["r0"], rtol=1e-12)
    np.testing.assert_allclose(amg.history(), ores["history"], rtol=1e-12)
    assert np.array_equal(xd.cpu().numpy(), ores["x"])


@pytest.fixture(scope="module")
def cfg2_system():
    import torch
    from fvm_amd.thermal import Mesh, assemble
    print("128^3 jittered, lognormal sigma 1.5 per cell: assembling", flush=True)
    mesh = Mesh.hex(128, 0.2, SEED)
    k, s, x = _torch_fields(mesh, 1.5, 1, SEED)
    S = assemble(mesh, k, s, x, BC_T, BC_V)
    torch.cuda.synchronize()
    h = S.to_host()
    oa = O.OracleAMG(n_max_iterations=4, relative_tolerance=1e-30)
    oa.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["is_boundary"])
    return dict(mesh=mesh, S=S, h=h, oa=oa)


@pytest.mark.parametrize("lanes,small", [(0, 8192), (8, 0), (16, 0), (32, 0), (64, 0)])
def test_128cube_hierarchy_bit_exact(cfg2_system, lanes, small):
    c = cfg2_system
    print("128^3 hierarchy: coarsen_lanes %d coarsen_small_rows %d" % (lanes, small), flush=True)
    amg = _amg()
    amg.set_tuning("coarsen_lanes", lanes)
    amg.set_tuning("coarsen_small_rows", small)
    _setup_system(amg, c["S"])
    L = compare_hierarchy(amg, c["oa"], c["h"], "lanes %d small %d" % (lanes, small))
    assert L >= 15


def test_128cube_three_exact_cycles(cfg2_system):
    import torch
    c = cfg2_system
    h = c["h"]
    print("128^3: three reference-order cycles", flush=True)
    amg = _amg(nMaxIterations=4, relativeTolerance=1e-30)
    _setup_system(amg, c["S"])
    b = torch.from_numpy(h["b"]).cuda()
    xd = torch.zeros(h["n_total"], dtype=torch.float64, device="cuda:0")
    r0, rf, cyc = amg.solve_device(b, xd)
    ores = c["oa"].solve(h["b"])
    assert cyc == ores["cycles"] == 3
    np.testing.assert_allclose(amg.history(), ores["history"], rtol=1e-12)
    assert np.array_equal(xd.cpu().numpy(), ores["x"])


def test_200cube_mixed_hierarchy_bit_exact():
    import torch
    from fvm_amd.thermal import Mesh, assemble
    print("200^3 jittered, sigma 2.5 in 8^3 blocks: mixed-precision hierarchy vs oracle", flush=True)
    mesh = Mesh.hex(200, 0.2, SEED)
    k, s, x = _torch_fields(mesh, 2.5, 8, SEED)
    S = assemble(mesh, k, s, x, BC_T, BC_V)
    torch.cuda.synchronize()
    h = S.to_host()
    amg = _amg(precision=1, gsOrdering=1)
    _setup_system(amg, S)
    oa = O.OracleAMG()
    oa.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["is_boundary"])
    L = compare_hierarchy(amg, oa, h)
    print("levels", L, flush=True)
    assert L >= 17


def _level_dense(lv):
    n = lv["n"]
    A = np.zeros((n, n))
    A[np.arange(n), np.arange(n)] = lv["diag"]
    for i in range(n):
        for p in range(lv["row"][i], lv["row"][i + 1]):
            A[i, lv["col"][p]] += lv["off"][p]
    return A


@pytest.mark.parametrize("N,order", [(12, 1), (14, 0)])
def test_coarsest_dense_inverse_against_numpy(N, order):
    import fvm_amd
    m = matgen.grid3d(N, seed=31, contrast=1.5)
    amg = _amg(precision=1, gsOrdering=order)
    ls = fvm_amd.LinearSystem(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"], m["b"])
    amg.setup(ls)
    d = amg.direct_level(with_inverse=True)
    assert d["level"] >= 1 and 0 < d["n"] <= 256 and d["attempts"] == 1
    # the first coarse level with at most 256 rows
    assert amg.level(d["level"] - 1)["n"] > 256
    A = _level_dense(amg.level(d["level"]))
    cond = np.linalg.cond(A)
    inv_np = np.linalg.inv(A)
    # both factorizations are backward stable: the inverses agree to a small
    # multiple of cond(A) * eps (a wrong pivot, row or multiplier gives O(1))
    tol = 1e-13 * cond
    err = np.abs(d["inv"] - inv_np).max() / np.abs(inv_np).max()
    print("inverse: cond %.3g, relative max error %.3g, bar %.3g" % (cond, err, tol), flush=True)
    assert err <= tol
    rng = np.random.default_rng(N)
    for _ in range(3):
        b = rng.uniform(-1, 1, d["n"])
        x = amg.direct_apply(b)
        xs = np.linalg.solve(A, -b)
        assert np.abs(x - xs).max() <= tol * np.abs(xs).max()
    print("direct level", d["level"], "rows", d["n"], "cond %.3g" % cond, flush=True)
    # solving again reuses the inverse
    amg.solve(ls)
    assert amg.direct_level()["attempts"] == 1


def _with_singular_block(m):
    """Block-diagonal composite of a grid matrix (no ghosts) and a 4-row chain
    Laplacian [[-1,1],[1,-2,1],[1,-2,1],[1,-1]] whose pairs (0,1), (2,3) give the
    exactly singular coarse block [[-1,1],[1,-1]]."""
    n = m["n_self"]
    brow = [0, 1, 3, 5, 6]
    bcol = [1, 0, 2, 1, 3, 2]
    boff = [1.0] * 6
    bdiag = [-1.0, -2.0, -2.0, -1.0]
    row = np.concatenate([m["row"], m["row"][-1] + np.array(brow[1:])]).astype(np.int32)
    col = np.concatenate([m["col"], n + np.array(bcol)]).astype(np.int32)
    off = np.concatenate([m["off"], boff])
    diag = np.concatenate([m["diag"], bdiag])
    b = np.concatenate([m["b"], np.zeros(4)])
    return dict(n_self=n + 4, n_total=n + 4, row=row, col=col, diag=diag, off=off, b=b)


def test_singular_coarsest_level_falls_back_to_smoothing():
    import fvm_amd
    m = _with_singular_block(matgen.grid3d(7, seed=32, ghosts=False))
    # the oracle's first coarse level holds the exactly singular 2x2 block
    oa = O.OracleAMG(max_coarse_levels=1)
    oa.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    c = oa.level(1)
    assert c["n_self"] <= 256
    A = np.zeros((c["n_self"],) * 2)
    A[np.arange(c["n_self"]), np.arange(c["n_self"])] = c["diag"]
    for i in range(c["n_self"]):
        for p in range(c["row"][i], c["row"][i + 1]):
            A[i, c["col"][p]] += c["off"][p]
    assert np.array_equal(A[-2:, -2:], [[-1.0, 1.0], [1.0, -1.0]])
    assert np.linalg.matrix_rank(A) == c["n_self"] - 1
    amg = _amg(precision=1, gsOrdering=1, maxCoarseLevels=1, relativeTolerance=1e-9, nMaxIterations=2000)
    ls = fvm_amd.LinearSystem(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"], m["b"])
    r0 = amg.solve(ls)
    d = amg.direct_level()
    assert d["level"] == -1 and d["attempts"] == 1
    # the grid part converges with the coarsest level smoothed; the block stays at zero
    assert amg.lastResidual / r0 < 1e-9
    assert np.array_equal(ls.delta[-4:], np.zeros(4))
    xs = np.linalg.solve(matgen.dense(m)[:-4, :-4], -m["b"][:-4])
    np.testing.assert_allclose(ls.delta[:-4], xs, rtol=0, atol=1e-8 * np.abs(xs).max())
    # a second solve of the same system keeps the hierarchy and does not factor the
    # singular level again
    first = ls.delta.copy()
    ls.delta[:] = 0.0
    amg.solve(ls)
    assert amg.direct_level()["attempts"] == 1
    assert np.array_equal(ls.delta, first)


def test_200cube_mixed_solution_agrees_with_fp64_amg():
    """The mixed-precision solver (FP64 FGMRES(30), FP32/BF16 V-cycle) and the plain
    FP64 AMG, both driven to 1e-13 on the 8M-cell system, agree within 1e-9
    relative (max-norm): the bar for converged solutions."""
    import torch
    import fvm_amd
    sys_path = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if sys_path not in sys.path:
        sys.path.insert(0, sys_path)
    import bench
    print("200^3: mixed FGMRES and plain FP64 AMG to 1e-13", flush=True)
    mesh, S, _ = bench.build_workload(200, torch)
    b = S.device_pointers()["b"]
    mixed = bench.make_amg(1, 1)
    bench.setup_on(mixed, S)
    fg = fvm_amd.FGMRES(mixed)
    fg.relativeTolerance = 1e-13
    fg.nMaxIterations = 3000
    xm = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
    r0, rf, it = fg.solve_device(b, xm)
    plain = bench.make_amg(1, 0)
    plain.relativeTolerance = 1e-13
    plain.nMaxIterations = 20000
    bench.setup_on(plain, S)
    xp = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
    r0p, rfp, cyc = plain.solve_device(b, xp)
    n = S.n_self
    err = (xm[:n] - xp[:n]).abs().max().item() / xp[:n].abs().max().item()
    print("FGMRES %d iterations (ratio %.2e), FP64 AMG %d cycles (ratio %.2e), difference %.3g"
          % (it, rf / r0, cyc, rfp / r0p, err), flush=True)
    assert rf / r0 < 1e-13 and rfp / r0p < 1e-13
    assert err < 1e-9
