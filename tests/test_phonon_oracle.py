"""Pins of the phonon-direction assembly oracle (ofv_phonon in
oracle/fv_oracle.c: PhononCollisionDiscretization.h:54-68,
PhononConvectionDiscretization.h:150-175, GenericPhononBCS.h:67-131,
PhononModel.h:498-523) by closed forms the mathematics fixes:
- uniform hex mesh, constant direction s: every interior face carries
  F = h^2 s_axis; the upwind neighbour's coefficient is |F|, the downwind one 0;
- a cell's outflow equals its inflow (constant s, closed cell), so every row
  sums to -V/tau;
- x = 0, e0 = 1: b = V/tau plus |F| T_b on each inflow face of the x0 wall;
- inflow ghosts become identity rows, outflow ghosts are marked boundary and
  eliminate to nothing;
- s = (1, 0, 0): the system decouples into 1-D upwind chains whose solution is
  the recurrence x_i = (V/tau + h^2 x_{i-1}) / (V/tau + h^2), x_ghost = 1.
A swapped upwind side, a wrong sign of the relaxation term or a lost inflow
value fails one of these."""
import numpy as np
import pytest

import hexmesh
import oracle_lib as O


def fib_direction(i, n=256):
    z = 1.0 - (2 * i + 1) / n
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    r = np.sqrt(1.0 - z * z)
    return np.array([r * np.cos(phi), r * np.sin(phi), z])


def inflow_field(mesh, value_x0=1.0):
    """Cells 0; ghosts of the x0 wall hold value_x0, other ghosts 0."""
    x = np.zeros(mesh["n_cells"] + mesh["n_ghost"])
    fc = mesh["face_cells"].reshape(-1, 2)
    gp = mesh["group_ptr"]
    x[fc[gp[1]:gp[2], 1]] = value_x0
    return x


def dense(s):
    n = s["n_total"]
    A = np.zeros((n, n))
    A[np.arange(n), np.arange(n)] = s["diag"]
    for i in range(n):
        for p in range(s["row"][i], s["row"][i + 1]):
            A[i, s["col"][p]] += s["off"][p]
    return A


@pytest.mark.parametrize("i", [0, 37, 128, 200])
def test_uniform_upwind_coefficients_and_row_sums(i):
    N = 4
    h = 1.0 / N
    tau = 0.7
    m = hexmesh.hex_mesh(N)
    s = fib_direction(i)
    sys_ = O.fv_phonon(m, s, tau, np.ones(N ** 3), inflow_field(m))
    A = dense(sys_)
    nc = N ** 3
    cid = lambda a, b, c: a + N * (b + N * c)
    V = h ** 3
    for k in range(N):
        for j in range(N):
            for ii in range(N):
                c = cid(ii, j, k)
                for ax, (di, dj, dk) in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
                    for sgn in (1, -1):
                        a, b_, cc = ii + sgn * di, j + sgn * dj, k + sgn * dk
                        if not (0 <= a < N and 0 <= b_ < N and 0 <= cc < N):
                            continue
                        nb = cid(a, b_, cc)
                        upwind = sgn * s[ax] < 0   # neighbour lies against the flow
                        want = h * h * abs(s[ax]) if upwind else 0.0
                        assert abs(A[c, nb] - want) < 1e-15
    # interior rows (no boundary face) sum to -V/tau
    for k in range(1, N - 1):
        for j in range(1, N - 1):
            for ii in range(1, N - 1):
                c = cid(ii, j, k)
                assert abs(A[c, :nc].sum() + V / tau) < 1e-15


@pytest.mark.parametrize("i", [5, 100, 250])
def test_uniform_rhs_and_ghost_rows(i):
    N = 3
    h = 1.0 / N
    tau = 2.0
    m = hexmesh.hex_mesh(N)
    s = fib_direction(i)
    x = inflow_field(m)
    sys_ = O.fv_phonon(m, s, tau, np.ones(N ** 3), x)
    nc = N ** 3
    V = h ** 3
    want = np.full(nc, V / tau)
    fc = m["face_cells"].reshape(-1, 2)
    gp = m["group_ptr"]
    normals = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    for g in range(6):
        vn = float(np.dot(s, normals[g]))
        for f in range(gp[g + 1], gp[g + 2]):
            c0, gh = fc[f]
            if vn < 0:
                want[c0] += h * h * abs(vn) * x[gh]
                assert sys_["diag"][gh] == -1.0 and sys_["b"][gh] == 0.0 and not sys_["is_boundary"][gh]
            else:
                assert sys_["is_boundary"][gh] == 1
            # no coupling to the ghost remains in the interior row
            r = sys_["row"]
            for p in range(r[c0], r[c0 + 1]):
                if sys_["col"][p] == gh:
                    assert sys_["off"][p] == 0.0
    np.testing.assert_allclose(sys_["b"][:nc], want, rtol=1e-14, atol=1e-17)


def test_axis_direction_matches_1d_upwind_recurrence():
    N = 5
    h = 1.0 / N
    tau = 0.3
    m = hexmesh.hex_mesh(N)
    sys_ = O.fv_phonon(m, np.array([1.0, 0.0, 0.0]), tau, np.ones(N ** 3), inflow_field(m))
    nc = N ** 3
    A = dense(sys_)[:nc, :nc]
    delta = np.linalg.solve(A, -sys_["b"][:nc])
    V = h ** 3
    a = V / tau
    chain = []
    prev = 1.0
    for _ in range(N):
        prev = (a + h * h * prev) / (a + h * h)
        chain.append(prev)
    want = np.tile(np.array(chain), N * N)   # x = 0 before the solve: delta = solution
    np.testing.assert_allclose(delta, want, rtol=1e-12)


def test_jittered_row_sums_follow_divergence_theorem():
    """On a jittered mesh the net face flux of a closed cell is still zero for a
    constant direction, so interior rows sum to -V/tau."""
    N = 5
    m = hexmesh.hex_mesh(N, jitter=0.2, seed=9)
    g = O.fv_metrics(m)
    tau = 1.3
    sys_ = O.fv_phonon(m, fib_direction(77), tau, np.ones(N ** 3), inflow_field(m))
    A = dense(sys_)
    nc = N ** 3
    vol = g["vol"][:nc]
    cid = lambda a, b, c: a + N * (b + N * c)
    for k in range(1, N - 1):
        for j in range(1, N - 1):
            for ii in range(1, N - 1):
                c = cid(ii, j, k)
                assert abs(A[c, :nc].sum() + vol[c] / tau) < 1e-15
