"""The learned-preconditioner configuration at full size: 160^3 jittered mesh
(4.1M cells), conductivity exp(2 z) constant in 8^3-cell blocks, FP64 FGMRES(30)
with the mixed-precision AMG, with and without the learned correction (trained
on the GPU here on residual/error pairs that plain AMG leaves on 16^3 systems of
the per-cell sigma 1.5 distribution). Property checks: both solves reach 1e-8
on the FP64 true residual recomputed by this test from the level-0 CSR, the
two solutions agree within the tolerance's reach, and the iterations, times
and the network's fallback state are reported (the comparison itself is the
result: the correction is not expected to pay on this system)."""
import ctypes
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261202


def _true_residual_norm(torch, amg, S, x):
    """||b + A x||_1 over the self rows, from the hierarchy's level-0 CSR (this
    test's own sum, not the solver's)."""
    lv = amg.level(0)
    n = lv["n"]
    t = lambda a, dt: torch.tensor(a, dtype=dt, device="cuda:0")
    row, col, off, diag = t(lv["row"], torch.int64), t(lv["col"], torch.int64), t(lv["off"], torch.float64), \
        t(lv["diag"], torch.float64)
    b = torch.as_tensor(S.to_host()["b"][:n], device="cuda:0")
    rid = torch.repeat_interleave(torch.arange(n, device="cuda:0"), row[1:] - row[:-1])
    r = diag * x[:n] + b
    r = r.index_add(0, rid, off * x[col])
    return float(r.abs().sum()), float(b.abs().sum())


def test_learned_preconditioner_on_the_4m_cell_system():
    import torch
    import fvm_amd
    from fvm_amd import LearnedCorrection
    from fvm_amd._lib import check, lib
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    rng = np.random.default_rng(5)
    samples = []
    keep = []
    print("training pairs from plain AMG on 16^3 systems", flush=True)
    for j in range(4):
        mesh, S = bench.thermal_system(torch, 16, 1.5, 1, SEED + 100 + j)
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
    curve = net.train(samples, steps=200, lr=2e-3, batch=8)
    assert np.mean(curve[-10:]) < 0.5 * np.mean(curve[:10])
    del samples, keep
    print("160^3, sigma 2.0 in 8^3 blocks", flush=True)
    mesh, S = bench.thermal_system(torch, 160, 2.0, 8, SEED + 7)
    assert S.n_self == 160 ** 3
    amg = bench.make_amg(1, 1)
    bench.setup_on(amg, S)
    b = S.device_pointers()["b"]
    out, xs = {}, {}
    for label, use in (("without", None), ("with", net)):
        fg = fvm_amd.FGMRES(amg)
        fg.relativeTolerance = 1e-8
        fg.nMaxIterations = 1000
        fg.set_learned(use, 0.9)
        x = torch.zeros(S.n_total, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r0, rf, it = fg.solve_device(b, x)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rn, r0_own = _true_residual_norm(torch, amg, S, x)
        st = fg.learned_stats() if use is not None else {}
        out[label] = dict(iterations=it, seconds=round(dt, 4), ratio=rn / r0_own, **st)
        xs[label] = x[: S.n_self].clone()
        assert abs(r0_own - r0) <= 1e-10 * r0
        assert rn / r0_own < 1e-8
    print("160^3 learned ablation:", out, flush=True)
    assert out["with"]["iterations_used"] >= 0
    d = (xs["with"] - xs["without"]).abs().max().item() / xs["without"].abs().max().item()
    print("solutions differ by %.3g relative (both to 1e-8)" % d, flush=True)
    assert d < 1e-4
