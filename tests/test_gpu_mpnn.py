"""GPU kernels of the learned correction (fvm_mpnn_*) against the FP64
restatement in oracle/mpnn_ref.py (itself pinned by finite differences in
test_mpnn_ref.py).

Bar: with the same (bf16-quantized) weights, the GPU forward pass (bf16
activations, MFMA with FP32 accumulation) agrees with the FP64 restatement
within 3e-2 relative (max-norm), and the GPU gradient within 5e-2 relative per
parameter block (L2) with cosine similarity above 0.995; training lowers the
loss; FGMRES with the correction converges to the FP64 AMG solution; a network
that makes things worse is switched off and the solve still converges.
"""
import os
import sys

import numpy as np
import pytest

import matgen

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mpnn_ref as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a, dt):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda:0")


def _graph(m):
    import torch
    return (m["n_self"], _dev(m["row"], torch.int32), _dev(m["col"], torch.int32), _dev(m["off"], torch.float64),
            _dev(m["diag"], torch.float64))


def _sys(N, seed, contrast=2.0):
    m = matgen.grid3d(N, seed=seed, ghosts=False, contrast=contrast)
    return m, M.Graph(m["row"], m["col"], m["off"], m["diag"])


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_matches_restatement(seed):
    import torch
    from fvm_amd import LearnedCorrection
    m, G = _sys(7, seed)
    rng = np.random.default_rng(seed)
    th = (M.init_params(seed) + 0.05 * rng.normal(size=M.n_params())).astype(np.float32)
    net = LearnedCorrection(seed=seed)
    net.set_params(th)
    assert np.array_equal(net.params(), th)
    net.set_graph(*_graph(m))
    r = rng.normal(size=G.n)
    e = torch.zeros(G.n, dtype=torch.float64, device="cuda:0")
    net.forward(_dev(r, torch.float64), e)
    ref = M.forward(M.quantize(th.astype(np.float64)), G, r)
    got = e.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("forward rel err", err, flush=True)
    assert err < 3e-2
    # accumulate mode adds
    net.forward(_dev(r, torch.float64), e, accumulate=True)
    np.testing.assert_allclose(e.cpu().numpy(), 2 * got, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_restatement(seed):
    import torch
    from fvm_amd import LearnedCorrection
    m, G = _sys(6, 10 + seed)
    rng = np.random.default_rng(20 + seed)
    th = (M.init_params(seed) + 0.05 * rng.normal(size=M.n_params())).astype(np.float32)
    net = LearnedCorrection(seed=seed)
    net.set_params(th)
    net.set_graph(*_graph(m))
    r = rng.normal(size=G.n)
    L = net.loss_grad(_dev(r, torch.float64))
    g = net.grad(reset=True).astype(np.float64)
    Lr, gr = M.grad(M.quantize(th.astype(np.float64)), G, r)
    print("loss", L, Lr, flush=True)
    assert abs(L - Lr) <= 3e-2 * Lr
    cos = g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr))
    assert cos > 0.995, cos
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.layer_shapes()])
    for b, (name, _) in enumerate(M.layer_shapes()):
        gb, rb = g[offs[b]:offs[b + 1]], gr[offs[b]:offs[b + 1]]
        if np.linalg.norm(rb) > 1e-6 * np.linalg.norm(gr):
            rel = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
            print(name, "rel", rel, flush=True)
            assert rel < 5e-2, (name, rel)
    # the buffer was reset
    assert not np.any(net.grad())


def test_energy_gradient_matches_restatement():
    import torch
    from fvm_amd import LearnedCorrection
    m, G = _sys(6, 31)
    rng = np.random.default_rng(41)
    th = (M.init_params(2) + 0.05 * rng.normal(size=M.n_params())).astype(np.float32)
    net = LearnedCorrection(seed=2)
    net.set_params(th)
    net.set_graph(*_graph(m))
    et = rng.normal(size=G.n)
    r = -G.matvec(et)
    L = net.loss_grad_pair(_dev(r, torch.float64), _dev(et, torch.float64))
    g = net.grad(reset=True).astype(np.float64)
    Lr, gr = M.grad_energy(M.quantize(th.astype(np.float64)), G, r, et)
    print("energy loss", L, Lr, flush=True)
    assert abs(L - Lr) <= 3e-2 * Lr
    cos = g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr))
    assert cos > 0.995, cos
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 5e-2


def test_training_lowers_the_loss():
    import torch
    from fvm_amd import LearnedCorrection
    m, G = _sys(8, 3)
    rng = np.random.default_rng(5)
    net = LearnedCorrection(seed=7)
    g = _graph(m)
    samples = [(g, _dev(rng.normal(size=G.n), torch.float64)) for _ in range(8)]
    curve = net.train(samples, steps=60, lr=3e-3, batch=4)
    print("loss", curve[0], "->", curve[-1], flush=True)
    assert np.mean(curve[-5:]) < 0.8 * np.mean(curve[:5])


def _fgmres(amg, net=None, stall=0.9):
    import fvm_amd
    fg = fvm_amd.FGMRES(amg)
    fg.restart = 30
    fg.relativeTolerance = 1e-10
    fg.nMaxIterations = 300
    if net is not None:
        fg.set_learned(net, stall)
    return fg


@pytest.mark.parametrize("prec", [0, 1])
def test_fgmres_with_learned_correction(prec):
    import torch
    import fvm_amd
    from fvm_amd import LearnedCorrection
    m = matgen.grid3d(16, seed=9, ghosts=False, contrast=2.0)
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.gsOrdering = 1
    amg.precision = prec
    g = _graph(m)
    amg.setup_device(m["n_self"], m["n_total"], g[1], g[2], g[4], g[3])
    b = _dev(m["b"], torch.float64)
    x0 = torch.zeros(m["n_self"], dtype=torch.float64, device="cuda:0")
    _fgmres(amg).solve_device(b, x0)
    net = LearnedCorrection(seed=2)
    fg = _fgmres(amg, net)
    x1 = torch.zeros(m["n_self"], dtype=torch.float64, device="cuda:0")
    r0, rf, it = fg.solve_device(b, x1)
    st = fg.learned_stats()
    print("iterations", it, st, flush=True)
    assert rf / r0 < 1e-10
    assert st["iterations_used"] > 0
    ref = x0.cpu().numpy()
    assert np.abs(x1.cpu().numpy() - ref).max() <= 1e-8 * np.abs(ref).max()


def test_bad_network_falls_back():
    import torch
    import fvm_amd
    from fvm_amd import LearnedCorrection
    m = matgen.grid3d(14, seed=4, ghosts=False, contrast=2.0)
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.gsOrdering = 1
    g = _graph(m)
    amg.setup_device(m["n_self"], m["n_total"], g[1], g[2], g[4], g[3])
    net = LearnedCorrection(seed=3)
    th = net.params()
    th[M.n_params() - 65:] *= 200.0        # readout blown up: corrections far too large
    net.set_params(th)
    fg = _fgmres(amg, net, stall=0.95)
    b = _dev(m["b"], torch.float64)
    x = torch.zeros(m["n_self"], dtype=torch.float64, device="cuda:0")
    r0, rf, it = fg.solve_device(b, x)
    st = fg.learned_stats()
    print("iterations", it, st, flush=True)
    assert rf / r0 < 1e-10
    assert st["fallback_at"] >= 0


def test_learned_ablation():
    """Train on residual/error pairs that plain AMG leaves on small jittered
    lognormal systems, then solve a held-out system with FGMRES + mixed AMG with
    and without the correction: both reach the tolerance; the iteration counts and
    times of both are reported (the comparison itself is the result)."""
    import time
    import torch
    import fvm_amd
    from fvm_amd import LearnedCorrection
    from fvm_amd._lib import check, lib
    import ctypes
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    rng = np.random.default_rng(11)
    samples = []
    keep = []
    for j in range(3):
        mesh, S = bench.thermal_system(torch, 16, 1.5, 1, 500 + j)
        amg = bench.make_amg(1, 0, maxit=10)
        bench.setup_on(amg, S)
        g = bench.level0_graph(torch, amg)
        zero = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
        for q in range(24):
            xs = torch.tensor(rng.normal(size=S.n_total), dtype=torch.float64, device="cuda:0")
            xs[S.n_self:] = 0
            b = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
            b[:S.n_self] = -bench.residual_of(torch, g, xs, zero)
            x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
            for _ in range(1 + q % 2):
                check(lib().fvm_amg_smooth_device(amg._h, ctypes.c_void_p(b.data_ptr()),
                                                  ctypes.c_void_p(x.data_ptr())))
            samples.append((g, bench.residual_of(torch, g, x, b).contiguous(),
                            (xs[:S.n_self] - x[:S.n_self]).contiguous()))
        keep.append((mesh, S, amg))
    net = LearnedCorrection(seed=1)
    curve = net.train(samples, steps=120, lr=2e-3, batch=8)
    assert np.mean(curve[-10:]) < 0.5 * np.mean(curve[:10])
    mesh, S = bench.thermal_system(torch, 40, 2.0, 8, 900)
    amg = bench.make_amg(1, 1)
    bench.setup_on(amg, S)
    b = S.device_pointers()["b"]
    out = {}
    for label, use in (("without", None), ("with", net)):
        fg = fvm_amd.FGMRES(amg)
        fg.relativeTolerance = 1e-8
        fg.nMaxIterations = 500
        fg.set_learned(use)
        x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r0, rf, it = fg.solve_device(b, x)
        torch.cuda.synchronize()
        out[label] = (it, time.perf_counter() - t0, rf / r0)
        assert rf / r0 < 1e-8
    print("ablation (iterations, seconds, ratio): without", out["without"], "with", out["with"], flush=True)
