"""GPU parity of the HIP AMG (through the C ABI) against the CPU oracle.

Bar: coarsening maps, coarse matrices and (in the exact GS layout) the
correction delta are bit-identical to the oracle; residual 1-norms agree to
1e-12 relative (the GPU sums the norm in a fixed parallel order, the reference
sequentially) and cycle counts are equal. The coloured Gauss-Seidel layout and
the mixed-precision cycle are different smoothers: they must reach the same
tolerance and agree with the oracle's converged solution within 1e-9 relative
(FP64) or the stated tolerance (mixed).
"""
import os

import numpy as np
import pytest

import matgen
import oracle_lib as O

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fvm():
    import fvm_amd
    return fvm_amd


def ls_from(m, sym=False, is_bnd=None):
    f = _fvm()
    return f.LinearSystem(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"], m["b"],
                          is_boundary=is_bnd, isSymmetric=sym)


def oracle_for(m, is_bnd=None, **kw):
    a = O.OracleAMG(**kw)
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"], is_bnd)
    return a


def gpu_amg(**kw):
    f = _fvm()
    a = f.AMG()
    a.verbosity = 0
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def compact_level0(m):
    """Reference-order CSR of the self rows with ghost columns removed."""
    n = m["n_self"]
    row = [0]; col = []; off = []
    for i in range(n):
        for p in range(m["row"][i], m["row"][i + 1]):
            if m["col"][p] < n:
                col.append(m["col"][p]); off.append(m["off"][p])
        row.append(len(col))
    return np.array(row), np.array(col), np.array(off)


def check_hierarchy(ga, oa, m):
    assert ga.num_levels() == oa.num_levels()
    for l in range(oa.num_levels()):
        g = ga.level(l)
        o = oa.level(l)
        if l == 0:
            row, col, off = compact_level0(m)
            assert np.array_equal(g["row"], row)
            assert np.array_equal(g["col"], col)
            assert np.array_equal(g["off"], off)
            assert np.array_equal(g["ci"], o["ci"][: m["n_self"]])
        else:
            assert g["n"] == o["n_self"]
            assert np.array_equal(g["row"], o["row"]), l
            assert np.array_equal(g["col"], o["col"]), l
            assert np.array_equal(g["diag"], o["diag"]), l      # bit-identical Galerkin sums
            assert np.array_equal(g["off"], o["off"]), l
            assert np.array_equal(g["ci"], o["ci"]), l


def run_pair(m, gkw=None, okw=None, sym=False, is_bnd=None):
    gkw = dict(gkw or {}); okw = dict(okw or {})
    oa = oracle_for(m, is_bnd=is_bnd, is_symmetric=int(sym), **okw)
    ores = oa.solve(m["b"])
    ga = gpu_amg(**gkw)
    ls = ls_from(m, sym=sym, is_bnd=is_bnd)
    r0 = ga.solve(ls)
    return ga, oa, ls, r0, ores


def test_kat226_exact():
    m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    ga, oa, ls, r0, ores = run_pair(m)
    assert [ga.level(l)["n"] for l in range(1, ga.num_levels())] == [108, 48, 20, 8, 3]
    check_hierarchy(ga, oa, m)
    np.testing.assert_allclose(r0, ores["r0"], rtol=1e-12)   # parallel vs sequential 1-norm sum
    assert ga.lastCycles == ores["cycles"] == 40
    np.testing.assert_allclose(ga.history(), ores["history"], rtol=1e-12)
    assert np.array_equal(ls.delta, ores["x"])
    assert "%.6g" % ga.lastResidual == "5.32223e-05"


def test_kat_cav32_exact():
    m = O.load_mm(os.path.join(G, "cav32_matrix_mesh0.mat"), os.path.join(G, "cav32_matrix.rhs"))
    kw = dict(relativeTolerance=1e-9, maxCoarseLevels=20, nMaxIterations=200000)
    okw = dict(relative_tolerance=1e-9, max_coarse_levels=20, n_max_iterations=200000)
    ga, oa, ls, r0, ores = run_pair(m, kw, okw)
    check_hierarchy(ga, oa, m)
    assert r0 == 63200.0
    assert ga.lastCycles == 56 == ores["cycles"]
    assert np.array_equal(ls.delta, ores["x"])
    assert "%.6g" % ga.lastResidual == "5.75812e-05"


@pytest.mark.parametrize("N,seed,random_coeffs,ghosts", [(12, 0, True, True), (16, 1, True, False),
                                                          (10, 2, False, True), (20, 3, True, True)])
def test_grid_exact(N, seed, random_coeffs, ghosts):
    m = matgen.grid3d(N, seed=seed, random_coeffs=random_coeffs, ghosts=ghosts)
    kw = dict(nMaxIterations=60)
    ga, oa, ls, r0, ores = run_pair(m, kw, dict(n_max_iterations=60))
    check_hierarchy(ga, oa, m)
    assert ga.lastCycles == ores["cycles"]
    assert np.array_equal(ls.delta[: m["n_self"]], ores["x"][: m["n_self"]])
    assert np.array_equal(ls.delta[m["n_self"]:], np.zeros(m["n_total"] - m["n_self"]))  # ghosts untouched
    np.testing.assert_allclose(ga.history(), ores["history"], rtol=1e-12)


def test_boundary_rows_exact():
    m = matgen.grid3d(10, seed=5)
    isb = np.zeros(m["n_total"], np.uint8)
    isb[[3, 50, 51, 400]] = 1
    ga, oa, ls, r0, ores = run_pair(m, dict(nMaxIterations=40), dict(n_max_iterations=40), is_bnd=isb)
    check_hierarchy(ga, oa, m)
    assert np.array_equal(ls.delta, ores["x"])


@pytest.mark.parametrize("cycle", [1, 2])
def test_w_and_f_cycles_exact(cycle):
    m = matgen.grid3d(10, seed=6)
    ga, oa, ls, r0, ores = run_pair(m, dict(cycleType=cycle, nMaxIterations=30),
                                    dict(cycle_type=cycle, n_max_iterations=30))
    assert ga.lastCycles == ores["cycles"]
    assert np.array_equal(ls.delta, ores["x"])


def test_jacobi_exact_both_layouts():
    m = matgen.grid3d(10, seed=7)
    okw = dict(smoother_type=1, n_max_iterations=40)
    for order in (0, 1):
        ga, oa, ls, r0, ores = run_pair(m, dict(smootherType=1, gsOrdering=order, nMaxIterations=40), okw)
        assert ga.lastCycles == ores["cycles"]
        if order == 0:
            assert np.array_equal(ls.delta, ores["x"])
        else:   # same arithmetic per row, different reduction order of the coarse sums only
            np.testing.assert_allclose(ls.delta, ores["x"], rtol=1e-12, atol=1e-14)


def test_pre_sweeps_and_symmetric_scaling_exact():
    m = matgen.grid3d(9, seed=8, ghosts=False)
    ga, oa, ls, r0, ores = run_pair(m, dict(nPreSweeps=1, nPostSweeps=2, nMaxIterations=25),
                                    dict(n_pre_sweeps=1, n_post_sweeps=2, n_max_iterations=25), sym=True)
    assert ga.lastCycles == ores["cycles"]
    assert np.array_equal(ls.delta, ores["x"])


def test_smooth_is_one_cycle():
    m = matgen.grid3d(8, seed=9)
    oa = oracle_for(m)
    x0 = np.zeros(m["n_total"])
    xo = oa.smooth(m["b"], x0)
    ga = gpu_amg()
    ls = ls_from(m)
    ga.smooth(ls)
    assert np.array_equal(ls.delta, xo)


def _solution(m):
    oa = oracle_for(m, relative_tolerance=1e-14, n_max_iterations=5000)
    return oa.solve(m["b"])["x"][: m["n_self"]]


def test_multicolor_reaches_tolerance_and_solution():
    m = matgen.grid3d(16, seed=10, contrast=1.5)
    xs = _solution(m)
    ga = gpu_amg(gsOrdering=1, relativeTolerance=1e-14, nMaxIterations=3000)
    ls = ls_from(m)
    ga.solve(ls)
    h = ga.history()
    assert h[-1] / h[0] < 1e-13
    np.testing.assert_allclose(ls.delta[: m["n_self"]], xs, rtol=1e-9, atol=1e-9 * np.abs(xs).max())


@pytest.mark.parametrize("order", [0, 1])
def test_mixed_precision_refinement(order):
    m = matgen.grid3d(16, seed=11, contrast=1.5)
    xs = _solution(m)
    ga = gpu_amg(precision=1, gsOrdering=order, relativeTolerance=1e-13, nMaxIterations=3000)
    ls = ls_from(m)
    ga.solve(ls)
    h = ga.history()
    assert h[-1] / h[0] < 1e-13
    np.testing.assert_allclose(ls.delta[: m["n_self"]], xs, rtol=0, atol=1e-9 * np.abs(xs).max())


@pytest.mark.parametrize("prec,order", [(0, 0), (0, 1), (1, 1)])
def test_fgmres(prec, order):
    f = _fvm()
    m = matgen.grid3d(16, seed=12, contrast=2.0)
    xs = _solution(m)
    pc = gpu_amg(precision=prec, gsOrdering=order)
    kr = f.FGMRES(pc)
    kr.relativeTolerance = 1e-13
    ls = ls_from(m)
    r0 = kr.solve(ls)
    h = kr.history()
    assert h[0] == r0 and h[-1] / r0 < 1e-13
    r = ls.residual()
    assert np.abs(r).sum() / r0 < 1e-12
    np.testing.assert_allclose(ls.delta[: m["n_self"]], xs, rtol=0, atol=1e-9 * np.abs(xs).max())
    # fewer preconditioner applications than plain AMG cycles
    pa = gpu_amg(precision=0, gsOrdering=order, relativeTolerance=1e-13, nMaxIterations=5000)
    ls2 = ls_from(m)
    pa.solve(ls2)
    assert kr.lastIterations < pa.lastCycles


def test_errors_are_reported():
    f = _fvm()
    from fvm_amd._lib import FvmError, FVM_E_NUMERIC, FVM_E_INVALID, FVM_E_NOT_SETUP
    m = matgen.grid3d(4, seed=13)
    bad = dict(m); bad["diag"] = m["diag"].copy(); bad["diag"][5] = 0.0
    a = gpu_amg()
    with pytest.raises(FvmError) as e:
        a.solve(ls_from(bad))
    assert e.value.code == FVM_E_NUMERIC
    bad = dict(m); bad["col"] = m["col"].copy(); bad["col"][3] = m["n_total"] + 7
    with pytest.raises(FvmError) as e:
        gpu_amg().solve(ls_from(bad))
    assert e.value.code == FVM_E_INVALID
    a = gpu_amg(coarseGroupSize=3)
    with pytest.raises(FvmError) as e:
        a.solve(ls_from(m))
    assert e.value.code == FVM_E_INVALID
    b = np.zeros(m["n_total"]); x = np.zeros(m["n_total"])
    import ctypes
    from fvm_amd._lib import lib
    a = gpu_amg()
    rc = lib().fvm_amg_solve(a._h, ctypes.c_void_p(b.ctypes.data), ctypes.c_void_p(x.ctypes.data), None, None, None)
    assert rc == FVM_E_NOT_SETUP


def test_small_and_degenerate_systems():
    f = _fvm()
    # 1x1 system
    ls = f.LinearSystem(1, 1, np.array([0, 0]), np.zeros(0, np.int32), np.array([-2.0]), np.zeros(0), np.array([4.0]))
    a = gpu_amg()
    r0 = a.solve(ls)
    assert r0 == 4.0 and ls.delta[0] == 2.0
    # zero right-hand side: r0 = 0 < absTol fails (0 < 1e-50 is true) -> no cycles
    m = matgen.grid3d(5, seed=14)
    m0 = dict(m); m0["b"] = np.zeros(m["n_total"])
    a = gpu_amg()
    ls = ls_from(m0)
    assert a.solve(ls) == 0.0 and a.lastCycles == 0
    # iteration cap: nMaxIterations-1 cycles
    a = gpu_amg(nMaxIterations=7, relativeTolerance=1e-30)
    a.solve(ls_from(m))
    assert a.lastCycles == 6 and a.getTotalIterations() == 6


def test_device_entry_points_match_host():
    import torch
    f = _fvm()
    m = matgen.grid3d(12, seed=15)
    a = gpu_amg()
    ls = ls_from(m)
    a.solve(ls)
    d = torch.device("cuda:0")
    t = lambda v, dt: torch.from_numpy(np.ascontiguousarray(v)).to(device=d, dtype=dt)
    b = gpu_amg()
    b.setup_device(m["n_self"], m["n_total"], t(m["row"], torch.int32), t(m["col"], torch.int32),
                   t(m["diag"], torch.float64), t(m["off"], torch.float64))
    x = torch.zeros(m["n_total"], dtype=torch.float64, device=d)
    r0, rf, cyc = b.solve_device(t(m["b"], torch.float64), x)
    torch.cuda.synchronize()
    assert cyc == a.lastCycles
    assert np.array_equal(x.cpu().numpy(), ls.delta)


@pytest.mark.parametrize("small,kat,lanes,block", [(0, 0, 0, 512), (1 << 30, 0, 0, 512), (0, 1, 0, 512),
                                                    (1 << 30, 1, 0, 512), (1 << 30, 0, 8, 512),
                                                    (1 << 30, 0, 16, 512), (1 << 30, 0, 32, 512),
                                                    (1 << 30, 1, 8, 512), (0, 0, 8, 256), (0, 1, 16, 1024),
                                                    (0, 0, 0, 1024), (0, 1, 0, 256)])
def test_coarsening_paths_agree(small, kat, lanes, block):
    # launched rounds (every block size) and single-workgroup rounds, with every lane
    # width per row, give the oracle's maps (symmetric grid and the non-symmetric KAT-226)
    if kat:
        m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    else:
        m = matgen.grid3d(14, seed=21, contrast=2.0)
    oa = oracle_for(m)
    ga = gpu_amg()
    ga.set_tuning("coarsen_small_rows", small)
    ga.set_tuning("coarsen_lanes", lanes)
    ga.set_tuning("coarsen_block", block)
    # 1 or 0 queue entries per thread in LDS: most or all appends spill to the global queue
    ga.set_tuning("coarsen_lds_entries", {256: 1, 1024: 0}.get(block, 8))
    ls = ls_from(m)
    ga.setup(ls)
    check_hierarchy(ga, oa, m)


@pytest.mark.parametrize("pair,lower,waits", [(0, 0, 2), (1, 0, 0), (1, 0, 2), (1, 1, 0), (1, 1, 1), (1, 1, 2)])
@pytest.mark.parametrize("case", ["grid", "kat", "grid_small", "kat_small", "grid32"])
def test_coarsening_filters_and_wait_lists_agree(pair, lower, waits, case):
    # the dependency filters and the blocked joins' waiter lists change the rounds,
    # never the maps: every combination gives the oracle's hierarchy
    if case.startswith("kat"):
        m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    elif case == "grid32":
        m = matgen.grid3d(32, seed=23, contrast=2.5)
    else:
        m = matgen.grid3d(14, seed=21, contrast=2.0)
    oa = oracle_for(m)
    ga = gpu_amg()
    ga.set_tuning("coarsen_small_rows", 1 << 30 if case.endswith("small") else 0)
    ga.set_tuning("coarsen_pair_filter", pair)
    ga.set_tuning("coarsen_lower_filter", lower)
    ga.set_tuning("coarsen_wait_lists", waits)
    ls = ls_from(m)
    ga.setup(ls)
    check_hierarchy(ga, oa, m)


@pytest.mark.parametrize("prec,smoother,npre", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_tail_kernel_is_bit_identical(prec, smoother, npre):
    # neither the single-workgroup tail nor replaying levels >= 1 from a captured
    # graph may change a single bit of the coloured cycle
    m = matgen.grid3d(18, seed=22, contrast=1.5)
    res = []
    # tail_lds: the tail's levels that fit in LDS are swept from there (1) or not (0)
    for tail, graph, lds in ((0, 0, 1), (1 << 30, 0, 1), (0, 1, 1), (1 << 30, 1, 1), (2500, 1, 1), (1000, 0, 1),
                             (1 << 30, 0, 0)):
        ga = gpu_amg(gsOrdering=1, precision=prec, smootherType=smoother, nPreSweeps=npre, nMaxIterations=12,
                     relativeTolerance=1e-30)
        ga.set_tuning("tail_max_rows", tail)
        ga.set_tuning("cycle_graph", graph)
        ga.set_tuning("tail_lds", lds)
        ls = ls_from(m)
        ga.solve(ls)
        res.append((ls.delta.copy(), ga.history()))
        # a second solve on the same handle replays the captured graph
        ls = ls_from(m)
        ga.solve(ls)
        res.append((ls.delta.copy(), ga.history()))
    for d, hist in res[1:]:
        assert np.array_equal(res[0][0], d)
        assert np.array_equal(res[0][1], hist)


@pytest.mark.parametrize("prec", [0, 1])
def test_autotune_keeps_results(prec):
    # every launch shape the autotuner may pick gives the same bits
    m = matgen.grid3d(20, seed=23, contrast=1.5)
    ref = gpu_amg(gsOrdering=1, precision=prec, nMaxIterations=15, relativeTolerance=1e-30)
    ls0 = ls_from(m)
    ref.solve(ls0)
    ga = gpu_amg(gsOrdering=1, precision=prec, nMaxIterations=15, relativeTolerance=1e-30)
    ls1 = ls_from(m)
    ga.setup(ls1)
    rep = ga.autotune(reps=2)
    assert len(rep["candidates"]) == 60 and len(rep["spmv_candidates"]) == 6
    ch = rep["chosen"]
    assert any(c["sell_block"] == ch["sell_block"] and c["tail_max_rows"] == ch["tail_max_rows"]
               for c in rep["candidates"])
    assert ga.get_tuning("sell_block") == ch["sell_block"]
    assert ga.get_tuning("tail_max_rows") == ch["tail_max_rows"]
    assert ga.get_tuning("cycle_graph") == ch["cycle_graph"]
    ga.solve(ls1)
    assert np.array_equal(ls0.delta, ls1.delta)
    assert np.array_equal(ref.history(), ga.history())
    # and every block size and block order by hand
    for bs, xm in ((64, 0), (128, 1), (512, 0), (256, 1)):
        gb = gpu_amg(gsOrdering=1, precision=prec, nMaxIterations=15, relativeTolerance=1e-30)
        gb.set_tuning("sell_block", bs)
        gb.set_tuning("spmv_block", bs)
        gb.set_tuning("sell_xcd", xm)
        gb.set_tuning("spmv_xcd", 1 - xm)
        ls2 = ls_from(m)
        gb.solve(ls2)
        assert np.array_equal(ls0.delta, ls2.delta), bs


@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_reference_order_sweep_in_one_workgroup_is_bit_identical(cycle):
    """The reference-order smoother runs as one dataflow launch per sweep
    direction (exact_flow 1: each row updates once the rows it reads new values
    of have published theirs), or level by level (exact_flow 0): each stretch of
    narrow dependency levels (at most exact_wg_width rows) in one workgroup and
    wider levels one launch each. Every variant gives the oracle's bits, so all
    give the same delta and history (widths 8, 64, 200: sweeps that mix both
    launch kinds)."""
    m = matgen.grid3d(20, seed=29, contrast=1.5)
    out = []
    for flow, width in ((1, 256), (0, 0), (0, 8), (0, 64), (0, 200), (0, 4096), (0, 1 << 30)):
        ga = gpu_amg(gsOrdering=0, cycleType=cycle, nMaxIterations=12, relativeTolerance=1e-30)
        ga.set_tuning("exact_flow", flow)
        ga.set_tuning("exact_wg_width", width)
        ls = ls_from(m)
        ga.solve(ls)
        out.append((ls.delta.copy(), ga.history()))
    want = oracle_for(m, n_max_iterations=12, relative_tolerance=1e-30, cycle_type=cycle).solve(m["b"])
    for d, hist in out:
        assert np.array_equal(d, want["x"])
        assert np.array_equal(hist, out[0][1])


@pytest.mark.parametrize("key,good,bad", [
    ("coarsen_pair_filter", [0, 1], []), ("coarsen_lower_filter", [0, 1], []),
    ("coarsen_wait_lists", [0, 1, 2], [3, -1]), ("coarsen_block", [256, 512, 1024], [128, 300]),
    ("coarsen_lds_entries", [0, 8, 16], [17, -1]), ("coarsen_batch", [4, 32, 256], [2, 257]),
    ("tail_lds", [0, 1], []), ("gal_lanes", [0, 16, 32, 64], [8, 48]), ("colour_grid", [16, 256, 4096], [8, 5000])])
def test_setup_tuning_keys_round_trip_and_reject_bad_values(key, good, bad):
    # the agglomeration, coarse-operator, colouring and tail knobs: a value set is the
    # value read, and one out of range is refused with the previous value kept
    from fvm_amd._lib import FvmError
    ga = gpu_amg()
    for v in good:
        ga.set_tuning(key, v)
        assert ga.get_tuning(key) == v
    keep = ga.get_tuning(key)
    for v in bad:
        with pytest.raises(FvmError):
            ga.set_tuning(key, v)
        assert ga.get_tuning(key) == keep


@pytest.mark.parametrize("case", ["grid", "kat", "grid32"])
def test_every_level_colouring_is_proper(case):
    # the coloured smoother's colours (red-black on a bipartite finest level, the
    # greedy rounds over the uncoloured rows elsewhere): no entry joins two rows of
    # one colour, every row has a colour, and the colours are 0..k-1
    if case == "kat":
        m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    elif case == "grid32":
        m = matgen.grid3d(32, seed=24, contrast=2.5)
    else:
        m = matgen.grid3d(14, seed=21, contrast=2.0)
    ga = gpu_amg(gsOrdering=1)
    ls = ls_from(m)
    ga.setup(ls)
    checked = 0
    for l in range(ga.num_levels()):
        colour, k = ga.level_colours(l)
        if k == 0:
            continue
        g = ga.level(l)
        assert colour.min() == 0 and colour.max() == k - 1
        row, col = g["row"], g["col"]
        rows = np.repeat(np.arange(g["n"]), np.diff(row))
        keep = col < g["n"]   # entries to rows of this level (ghost columns have no colour)
        assert not np.any(colour[rows[keep]] == colour[col[keep]]), "entry inside one colour on level %d" % l
        checked += 1
    assert checked > 0
    if case.startswith("grid"):
        assert ga.level_colours(0)[1] == 2   # red-black on the hexahedral finest level


def test_opt_in_setup_reuse_gives_the_rebuilt_results():
    """reuse_setup 1: a setup of the same matrix (a new system object, even after
    cleanup, which the reference's models call after every solve) keeps the
    hierarchy, recognised by a signature of the matrix and the hierarchy's
    parameters; the solve equals a fresh handle's bit for bit. A changed value or
    parameter rebuilds."""
    m = matgen.grid3d(12, seed=31, contrast=1.5)
    kw = dict(nMaxIterations=20, relativeTolerance=1e-30)
    fresh = gpu_amg(**kw)
    want = ls_from(m)
    fresh.solve(want)
    ga = gpu_amg(**kw)
    ga.set_tuning("reuse_setup", 1)
    ls = ls_from(m)
    ga.solve(ls)
    assert ga.get_tuning("setups_reused") == 0
    ga.cleanup()
    for k in range(2):
        ls = ls_from(m)
        ga.solve(ls)
        assert ga.get_tuning("setups_reused") == k + 1
        assert np.array_equal(ls.delta, want.delta)
        np.testing.assert_array_equal(ga.history(), fresh.history())
    # one changed coefficient: rebuilt, and equal to a fresh handle on that matrix
    m2 = dict(m)
    m2["off"] = m["off"].copy()
    m2["off"][7] *= 1.5
    ls2 = ls_from(m2)
    ga.solve(ls2)
    assert ga.get_tuning("setups_reused") == 2
    f2 = gpu_amg(**kw)
    want2 = ls_from(m2)
    f2.solve(want2)
    assert np.array_equal(ls2.delta, want2.delta)
    # a hierarchy parameter: rebuilt as well
    ga.maxCoarseLevels = 3
    ls3 = ls_from(m2)
    ga.solve(ls3)
    assert ga.get_tuning("setups_reused") == 2 and ga.num_levels() == 4


def test_dataflow_sweep_bounded_wait_fails_loudly_then_recovers():
    """A reference-order dataflow sweep whose wavefronts may wait only once
    (exact_flow_spin_limit 1) gives up through its own bounded wait: the call
    fails with the sweep's error, every kernel exits normally, and the next call
    without the bound gives the oracle's bits again."""
    from fvm_amd._lib import FVM_E_NUMERIC, FvmError
    m = matgen.grid3d(10, seed=30)
    ga = gpu_amg(gsOrdering=0, nMaxIterations=5, relativeTolerance=1e-30)
    ga.set_tuning("exact_flow_spin_limit", 1)
    for call in ("smooth", "solve"):
        with pytest.raises(FvmError, match="stopped waiting") as e:
            getattr(ga, call)(ls_from(m))
        assert e.value.code == FVM_E_NUMERIC
    ga.set_tuning("exact_flow_spin_limit", 0)
    ls = ls_from(m)
    ga.solve(ls)
    want = oracle_for(m, n_max_iterations=5, relative_tolerance=1e-30).solve(m["b"])
    assert np.array_equal(ls.delta, want["x"])


def _continuity():
    return O.load_mm(os.path.join(G, "cav32_continuity_matrix.mat"), os.path.join(G, "cav32_continuity_matrix.rhs"))


@pytest.mark.parametrize("psr", [1, 0])
def test_continuity_matrix_exact(psr):
    """F:test/FLOW_CONTINUITY_MATRIX/GOLDEN/matrix.mat, the cavity's pressure-
    correction matrix: negative diagonal, positive off-diagonals, two stored
    zeros and one pinned row (diagonal -1). Maps, coarse matrices, delta and the
    residual history equal the oracle's over 60 reference-order cycles, with
    either stop rule (plain AMG converges slowly on it: the oracle needs about
    500 cycles for a 1e-3 reduction)."""
    m = _continuity()
    kw = dict(nMaxIterations=60, relativeTolerance=1e-30, parallelStopRule=psr)
    okw = dict(n_max_iterations=60, relative_tolerance=1e-30, parallel_stop_rule=psr)
    ga, oa, ls, r0, ores = run_pair(m, kw, okw)
    check_hierarchy(ga, oa, m)
    assert ga.lastCycles == ores["cycles"] == 59
    assert np.array_equal(ls.delta, ores["x"])
    np.testing.assert_allclose(ga.history(), ores["history"], rtol=1e-12)


@pytest.mark.parametrize("prec,order", [(0, 0), (1, 1)])
def test_continuity_matrix_fgmres_matches_dense_solve(prec, order):
    """FGMRES(30) around the AMG reaches 1e-12 on the same matrix and agrees with
    numpy.linalg.solve of A x = -b (the delta form r = b + A x) within the bound
    that the condition number (1.4e6) puts on a 1e-12 residual reduction."""
    f = _fvm()
    m = _continuity()
    n = m["n_self"]
    A = np.diag(m["diag"][:n]).astype(np.float64)
    for i in range(n):
        for p in range(m["row"][i], m["row"][i + 1]):
            A[i, m["col"][p]] += m["off"][p]
    xs = np.linalg.solve(A, -m["b"][:n])
    kr = f.FGMRES(gpu_amg(precision=prec, gsOrdering=order))
    kr.relativeTolerance = 1e-12
    kr.nMaxIterations = 400
    ls = ls_from(m)
    r0 = kr.solve(ls)
    assert kr.history()[-1] / r0 < 1e-12
    np.testing.assert_allclose(ls.delta[:n], xs, rtol=0, atol=1e-5 * np.abs(xs).max())


def test_convergence_lines_match_the_reference_golden(tmp_path):
    """AMG.cpp:237-272 with MultiFieldReduction::print: the lines of the cavity
    solve equal AMG_MERGING_THERMAL/proc1/GOLDEN/convergence.dat (parallel-build
    style: the first line and the one where the loop ends); the serial style
    prints every cycle with the same first and last lines."""
    m = O.load_mm(os.path.join(G, "cav32_matrix_mesh0.mat"), os.path.join(G, "cav32_matrix.rhs"))
    golden = open(os.path.join(G, "cav32_convergence_proc1.dat")).read().splitlines()[:2]
    out = {}
    for style in (1, 0):
        ga = gpu_amg(relativeTolerance=1e-9, maxCoarseLevels=20, nMaxIterations=2000, verbosity=2)
        ga.fieldName = "therm.temperature"
        ga.printStyle = style
        path = str(tmp_path / ("conv%d.dat" % style))
        ga.redirectPrintToFile(path)
        ga.solve(ls_from(m))
        ga.redirectPrintToScreen()
        out[style] = open(path).read().splitlines()
    assert out[1] == golden
    assert len(out[0]) == 57 and out[0][0] == golden[0] and out[0][-1] == golden[1]


def test_serial_stop_rule_exact():
    """AMG.cpp:199-204 (serial build): no coarse level of at most 3 rows is kept;
    maps, cycles and delta equal the oracle's with the same rule."""
    m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    ga, oa, ls, r0, ores = run_pair(m, dict(parallelStopRule=0), dict(parallel_stop_rule=0))
    check_hierarchy(ga, oa, m)
    sizes = [ga.level(l)["n"] for l in range(1, ga.num_levels())]
    assert sizes == [108, 48, 20, 8]
    assert ga.lastCycles == ores["cycles"]
    assert np.array_equal(ls.delta, ores["x"])
