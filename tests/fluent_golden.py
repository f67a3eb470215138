"""Shared set-up of the reference's mesh-based goldens (test helper).

The reference's thermal tests read a Fluent case, set conductivity 1, the
specified temperatures 400 on zone 3 and 0 on zones 4, 5, 6 (all other walls
zero flux), start from the default initial temperature 300 and run one
ThermalModel::advance with AMG (relativeTolerance 1e-9, nMaxIterations 2000,
maxCoarseLevels 20):
  src/fvm/test/THERMAL_MATRIX/testThermalParallel.py (dumpMatrix before advance)
  src/fvm/test/PARALLEL_TESTS/testThermalParallel.py (temperatures after advance)
The goldens are compared by the reference to 1e-5 absolute
(src/fvm/etc/numfile_compare.py).
"""
import os
import re

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FLUENT = os.path.join(GOLDEN, "fluent")
T_INIT = 300.0
BC_VALUES = {3: 400.0, 4: 0.0, 5: 0.0, 6: 0.0}
SOLVER_CASES = [("cav32_thermal.cas", "QUAD_1024"), ("tri_894.cas", "TRI_894"), ("cav_hexa.cas", "HEXA_10K"),
                ("cav_tetra.cas", "TETRA_8K")]


def read_mesh(case):
    from fvm_amd.fluent import FluentCase
    return FluentCase(os.path.join(FLUENT, case)).read().getMeshList()[0]


def thermal_bcs(mesh):
    """(bc_type, bc_value) per face group: specified temperature on zones 3-6,
    zero heat flux on the other walls (ThermalModel_impl.h:60-79 defaults)."""
    types, values = [0], [0.0]
    for zid, ztype in zip(mesh.group_ids[1:], mesh.group_types[1:]):
        if zid in BC_VALUES:
            types.append(0)
            values.append(BC_VALUES[zid])
        elif ztype in ("wall", "symmetry"):
            types.append(1)
            values.append(0.0)
        else:
            raise ValueError("no boundary condition for zone %d (%s)" % (zid, ztype))
    return types, values


def read_matrix_golden():
    lines = open(os.path.join(GOLDEN, "cav32_matrix_mesh0.mat")).read().split("\n")
    n, _, nz = map(int, lines[1].split())
    rows = [[] for _ in range(n)]
    for ln in lines[2:2 + nz]:
        i, j, v = ln.split()
        rows[int(i) - 1].append((int(j) - 1, float(v)))
    rhs = np.loadtxt(os.path.join(GOLDEN, "cav32_matrix.rhs"))
    return rows, rhs


def dumped_rows(S):
    """Rows as ThermalModel::dumpMatrix writes them (ThermalModel_impl.h:538-552):
    the diagonal, then the self-cell columns in CSR order."""
    nc = S["n_self"]
    out = []
    for i in range(nc):
        r = [(i, S["diag"][i])]
        for p in range(S["row"][i], S["row"][i + 1]):
            if S["col"][p] < nc:
                r.append((int(S["col"][p]), S["off"][p]))
        out.append(r)
    return out


def read_temperature_golden(name):
    """(T per cell, cell centroid y per cell) of a tecplot golden."""
    t = open(os.path.join(FLUENT, "temp_%s_proc1.dat" % name)).read().split("\n")
    N = int(re.search(r"N = (\d+)", t[2]).group(1))
    E = int(re.search(r"E = (\d+)", t[2]).group(1))
    vals = " ".join(t[3:]).split()
    T = np.array(vals[3 * N:3 * N + E], float)
    cy = np.array(vals[3 * N + E:3 * N + 2 * E], float)
    return T, cy
