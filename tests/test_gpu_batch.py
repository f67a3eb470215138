"""Batched solve of independent systems (fvm_batch_*) on the GPU.

Bar: for every system of a batch, the correction delta, the initial and final
residual 1-norms, the cycle count and the level sizes are bit-identical to
that system's own fvm_amg solve with the same parameters (FP64, reference or
coloured Gauss-Seidel order), with the captured-graph replay and with plain
launches; and each
system agrees with the CPU oracle as the unbatched solve does (same cycles,
same delta within 1e-12 relative). The batch mixes sizes, directions,
symmetric and non-symmetric patterns, hierarchies of different depths and a
system with a zero right-hand side.
"""
import os

import numpy as np
import pytest

import hexmesh
import matgen
import oracle_lib as O

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fib(i, n=256):
    z = 1.0 - (2 * i + 1) / n
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    r = np.sqrt(1.0 - z * z)
    return np.array([r * np.cos(phi), r * np.sin(phi), z])


def _phonon_system(m, i, tau, seed):
    import torch
    from fvm_amd.thermal import Mesh, assemble_phonon
    mesh = Mesh.hex(m, 0.2, seed)
    ref = hexmesh.hex_mesh(m, 0.2, seed)
    nc = ref["n_cells"]
    x = np.zeros(mesh.n_total)
    fc = ref["face_cells"].reshape(-1, 2)
    gp = ref["group_ptr"]
    x[fc[gp[1]:gp[2], 1]] = 1.0
    e0 = torch.ones(nc, dtype=torch.float64, device="cuda:0")
    tx = torch.tensor(x, dtype=torch.float64, device="cuda:0")
    S = assemble_phonon(mesh, fib(i), tau, e0, tx)
    return S.to_host()


def _host_system(m):
    return dict(n_self=m["n_self"], n_total=m["n_total"], row=m["row"], col=m["col"], diag=m["diag"],
                off=m["off"], b=m["b"], is_boundary=m.get("is_boundary"))


def _to_dev(h):
    import torch
    t = lambda a, dt: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda:0")
    d = dict(n_self=h["n_self"], n_total=h["n_total"], row=t(h["row"], torch.int32),
             col=t(h["col"] if len(h["col"]) else np.zeros(1), torch.int32), diag=t(h["diag"], torch.float64),
             off=t(h["off"] if len(h["off"]) else np.zeros(1), torch.float64), b=t(h["b"], torch.float64))
    d["is_boundary"] = None if h.get("is_boundary") is None else t(h["is_boundary"], torch.uint8)
    return d


def _batch_inputs():
    systems = []
    for j, (m, i, tau, seed) in enumerate([(5, 0, 0.5, 1), (7, 37, 3.0, 2), (4, 128, 9.0, 3),
                                           (8, 200, 0.2, 4), (6, 255, 1.7, 5)]):
        systems.append(_phonon_system(m, i, tau, seed))
    kat = O.load_mm(os.path.join(G, "MatrixMarket226.dat"), os.path.join(G, "rhs226.dat"))
    systems.append(_host_system(kat))
    systems.append(_host_system(matgen.grid3d(9, seed=4)))
    z = _host_system(matgen.grid3d(4, seed=6))
    z["b"] = np.zeros_like(z["b"])
    systems.append(z)
    return systems


def _params(obj, **kw):
    obj.relativeTolerance = 1e-9
    obj.nMaxIterations = 200
    obj.verbosity = 0
    for k, v in kw.items():
        setattr(obj, k, v)


def _single(h, dev, **kw):
    import torch
    import fvm_amd
    a = fvm_amd.AMG()
    _params(a, **kw)
    a.setup_device(dev["n_self"], dev["n_total"], dev["row"], dev["col"], dev["diag"], dev["off"],
                   dev["is_boundary"])
    x = torch.zeros(dev["n_total"], dtype=torch.float64, device="cuda:0")
    r0, rf, cyc = a.solve_device(dev["b"], x)
    sizes = [a.level(l)["n"] for l in range(a.num_levels())]
    return x.cpu().numpy(), r0, rf, cyc, sizes


def _batched(devs, graph, **kw):
    import torch
    import fvm_amd
    B = fvm_amd.BatchAMG()
    _params(B, **kw)
    B.set_tuning("graph", 1 if graph else 0)
    B.setup_device(devs)
    xs = [torch.zeros(d["n_total"], dtype=torch.float64, device="cuda:0") for d in devs]
    r0, rf, cyc = B.solve_device([d["b"] for d in devs], xs)
    sizes = [list(B.level_sizes(k)) for k in range(len(devs))]
    return [x.cpu().numpy() for x in xs], r0, rf, cyc, sizes


@pytest.mark.parametrize("kw", [dict(), dict(cycleType=1), dict(cycleType=2, nPreSweeps=1),
                                dict(smootherType=1, nMaxIterations=60), dict(gsOrdering=1),
                                dict(gsOrdering=1, cycleType=1), dict(gsOrdering=1, smootherType=1, nMaxIterations=60)])
def test_batch_equals_unbatched_bitwise(kw):
    hosts = _batch_inputs()
    devs = [_to_dev(h) for h in hosts]
    xb, r0b, rfb, cb, sb = _batched(devs, True, **kw)
    xg, r0g, rfg, cg, sg = _batched(devs, False, **kw)
    for k, (h, d) in enumerate(zip(hosts, devs)):
        print("system", k, "n", h["n_self"], flush=True)
        x1, r01, rf1, c1, s1 = _single(h, d, **kw)
        n = h["n_self"]
        assert sb[k] == s1, (k, sb[k], s1)
        assert cb[k] == c1 and r0b[k] == r01 and rfb[k] == rf1, (k, cb[k], c1, r0b[k], r01, rfb[k], rf1)
        assert np.array_equal(xb[k][:n], x1[:n]), k
        # ghost entries untouched
        assert not np.any(xb[k][n:])
        # graph replay and plain launches agree
        assert np.array_equal(xb[k], xg[k]) and cb[k] == cg[k] and rfb[k] == rfg[k]
    assert cb[-1] == 0 and r0b[-1] == 0.0       # zero right-hand side: no cycle


def test_batch_against_oracle():
    hosts = _batch_inputs()
    devs = [_to_dev(h) for h in hosts]
    xb, r0b, rfb, cb, sb = _batched(devs, True)
    for k, h in enumerate(hosts):
        a = O.OracleAMG(relative_tolerance=1e-9, n_max_iterations=200)
        a.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h.get("is_boundary"))
        r = a.solve(h["b"])
        n = h["n_self"]
        assert cb[k] == r["cycles"], (k, cb[k], r["cycles"])
        assert sb[k] == [a.level(l)["n_self"] for l in range(a.num_levels())]
        np.testing.assert_allclose(r0b[k], r["r0"], rtol=1e-12)
        scale = max(np.abs(r["x"][:n]).max(), 1e-300)
        assert np.abs(xb[k][:n] - r["x"][:n]).max() <= 1e-12 * scale, k
