"""Gray phonon BTE model loop on the device (fvm_phonon_advance: all directions of
an iteration assembled on the device and solved by one batched AMG) against the
oracle's loop (oracle/model_ref.py phonon_advance: the directions one after
another with the C oracle's assembly and AMG). In the reference row order each
direction's delta is the oracle's bit for bit, so the fields and the lattice
temperature are bit-identical after every iteration; the iteration norms agree
within 1e-12 relative (parallel vs sequential 1-norm sums). With the coloured
Gauss-Seidel the loop converges to the same temperatures within 1e-9."""
import os
import sys

import numpy as np
import pytest

import hexmesh
import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import model_ref as MR  # noqa: E402
from test_model_ref import _octants, _phonon_parts  # noqa: E402

pytestmark = pytest.mark.gpu


def _fib(K):
    i = np.arange(K)
    z = 1.0 - (2 * i + 1) / K
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    r = np.sqrt(1.0 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def _setup(case):
    if case == "slab":
        mesh = hexmesh.hex_mesh(4)
        vg = _octants()
        K = len(vg)
        tau = np.full(K, 0.05)
        args = (4, 0.0, 0)
    else:
        mesh = hexmesh.hex_mesh(5, 0.2, 8)
        vg = 1.3 * _fib(16)
        K = len(vg)
        tau = np.linspace(0.2, 5.0, K)
        args = (5, 0.2, 8)
    dk3 = np.full(K, 4 * np.pi / K)
    cp = np.linspace(0.8, 1.2, K)
    tref = np.zeros(K)
    wall = [0.0, 310.0, 290.0, 300.0, 305.0, 300.0, 295.0]
    return mesh, args, vg, dk3, tau, cp, tref, wall


@pytest.mark.parametrize("case", ["slab", "jittered"])
def test_phonon_model_matches_oracle_loop(case):
    from fvm_amd.phonon import PhononModel
    from fvm_amd.thermal import Mesh
    mesh, args, vg, dk3, tau, cp, tref, wall = _setup(case)
    nct = mesh["n_cells"] + mesh["n_ghost"]
    K = len(vg)
    niter = 6
    st = MR.phonon_init(nct, K, cp, tref, 300.0)
    g = O.fv_metrics(mesh)
    MR.phonon_walls(mesh, g["fA"], g["fAm"], vg, cp, tref, wall, st["e"])
    assemble, solve = _phonon_parts(mesh)
    norms_o, conv_o = MR.phonon_advance(assemble, solve, mesh, g, vg, dk3, tau, cp, tref, wall, st, niter,
                                        1e-30, 0.0)
    dm = Mesh.hex(*args)
    model = PhononModel(dm, vg, dk3, tau, cp, tref, wall, T_init=300.0)
    model.set_solver(gsOrdering=0, relativeTolerance=1e-12, nMaxIterations=400)
    model.callBoundaryConditions()
    keep = model.advance(niter, relTolerance=1e-30)
    print(case, "norms", model.norms, flush=True)
    assert keep and len(model.norms) == niter == len(norms_o)
    np.testing.assert_allclose(model.norms, norms_o, rtol=1e-12)
    TL, e, e0 = model.state()
    assert np.array_equal(TL, st["TL"])
    assert np.array_equal(e, st["e"])
    assert np.array_equal(e0, st["e0"])


def test_phonon_model_coloured_converges_to_oracle_temperature():
    from fvm_amd.phonon import PhononModel
    from fvm_amd.thermal import Mesh
    mesh, args, vg, dk3, tau, cp, tref, wall = _setup("jittered")
    nct = mesh["n_cells"] + mesh["n_ghost"]
    K = len(vg)
    st = MR.phonon_init(nct, K, cp, tref, 300.0)
    g = O.fv_metrics(mesh)
    MR.phonon_walls(mesh, g["fA"], g["fAm"], vg, cp, tref, wall, st["e"])
    assemble, solve = _phonon_parts(mesh)
    norms_o, conv_o = MR.phonon_advance(assemble, solve, mesh, g, vg, dk3, tau, cp, tref, wall, st, 300,
                                        1e-11, 0.0)
    assert conv_o
    model = PhononModel(Mesh.hex(*args), vg, dk3, tau, cp, tref, wall, T_init=300.0)
    model.set_solver(gsOrdering=1, relativeTolerance=1e-13, nMaxIterations=800)
    model.callBoundaryConditions()
    keep = model.advance(300, relTolerance=1e-11)
    assert not keep
    TL, e, _ = model.state()
    assert abs(len(model.norms) - len(norms_o)) <= 1
    nc = mesh["n_cells"]
    np.testing.assert_allclose(TL[:nc], st["TL"][:nc], rtol=1e-9)
