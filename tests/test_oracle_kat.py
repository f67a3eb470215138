"""The CPU oracle against the reference's own known-answer tests.

KAT-226: src/fvm/test/testLinearSolver.out:5-11 (TESTS:1, mtest.py + MMReader):
  coarse levels 108, 48, 20, 8, 3 (parallel-build stop rule), "0: 6981.57",
  "40: 5.32223e-05" with the AMG defaults.
KAT-cav32: THERMAL_MATRIX/GOLDEN/matrix_mesh0.mat + matrix.rhs solved as in
  AMG_MERGING_THERMAL/testThermalParallel.py:384-398 (relTol 1e-9,
  maxCoarseLevels 20): "0: 63200", "56: 5.75812e-05".
"""
import os
import re

import numpy as np
import pytest

import oracle_lib as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def parse_golden(path):
    sizes, norms = [], {}
    for line in open(path):
        m = re.search(r"Created coarse level \d+ of size (\d+)", line)
        if m:
            sizes.append(int(m.group(1)))
        m = re.match(r"\s*(\d+): \[[^:]+ : ([-0-9.e+]+)\]", line)
        if m:
            norms[int(m.group(1))] = float(m.group(2))
    return sizes, norms


def fmt6(v):
    return float("%.6g" % v)


def test_kat226_levels_and_history():
    sizes, norms = parse_golden(os.path.join(G, "testLinearSolver.out"))
    assert sizes == [108, 48, 20, 8, 3]
    m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    assert m["n_self"] == 226 and not m["sym"]
    a = O.OracleAMG()
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    got = [a.level(l)["n_self"] for l in range(1, a.num_levels())]
    assert got == sizes
    res = a.solve(m["b"])
    last = max(norms)
    assert res["cycles"] == last == 40
    assert fmt6(res["r0"]) == norms[0]
    assert fmt6(res["rfinal"]) == norms[last]


def test_kat226_serial_stop_rule_drops_tiny_level():
    # serial build (AMG.cpp:199-204) never pushes a coarse level of size <= 3
    m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    a = O.OracleAMG(parallel_stop_rule=0)
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    assert [a.level(l)["n_self"] for l in range(1, a.num_levels())] == [108, 48, 20, 8]


def test_kat_cav32_history():
    _, norms = parse_golden(os.path.join(G, "cav32_convergence_proc1.dat"))
    m = O.load_mm(os.path.join(G, "cav32_matrix_mesh0.mat"), os.path.join(G, "cav32_matrix.rhs"))
    assert m["n_self"] == 1024
    assert np.abs(m["b"]).sum() == 63200.0
    a = O.OracleAMG(relative_tolerance=1e-9, max_coarse_levels=20, n_max_iterations=200000)
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    res = a.solve(m["b"])
    last = max(norms)
    assert res["cycles"] == last == 56
    assert fmt6(res["r0"]) == norms[0] == 63200
    assert fmt6(res["rfinal"]) == norms[last]


def test_iteration_cap_is_nmax_minus_one():
    # AMG.cpp:250: at most nMaxIterations-1 cycles
    m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    a = O.OracleAMG(n_max_iterations=10)
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    assert a.solve(m["b"])["cycles"] == 9


def test_zero_rhs_returns_immediately():
    m = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    a = O.OracleAMG()
    a.setup(m["n_self"], m["n_total"], m["row"], m["col"], m["diag"], m["off"])
    res = a.solve(np.zeros(226))
    assert res["r0"] == 0.0 and res["cycles"] == 0
