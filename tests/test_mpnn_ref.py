"""Pins of the FP64 restatement of the learned correction (oracle/mpnn_ref.py):
its analytic gradient against central finite differences (per coordinate and
along random directions), the adjoint pair C / C^T, the matrix product against
a dense one, and scale equivariance (the network sees r / rms, so e(2r) = 2 e(r)).
A dropped term of the backward pass, a transposed weight or a wrong adjoint
fails one of these."""
import os
import sys

import numpy as np
import pytest

import matgen

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mpnn_ref as M  # noqa: E402


def _graph(N=5, seed=3):
    m = matgen.grid3d(N, seed=seed, ghosts=False)
    return M.Graph(m["row"], m["col"], m["off"], m["diag"]), m


def test_adjoint_and_matvec():
    G, m = _graph()
    rng = np.random.default_rng(0)
    h = rng.normal(size=(G.n, 7)); g = rng.normal(size=(G.n, 7))
    assert abs(np.sum(G.agg(h) * g) - np.sum(h * G.agg_t(g))) < 1e-10
    A = np.diag(m["diag"]).astype(float)
    for i in range(G.n):
        for p in range(m["row"][i], m["row"][i + 1]):
            A[i, m["col"][p]] += m["off"][p]
    x = rng.normal(size=G.n)
    np.testing.assert_allclose(G.matvec(x), A @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(A, A.T)     # the training loss assumes a symmetric matrix


def test_scale_equivariance():
    G, _ = _graph()
    th = M.init_params(1)
    r = np.random.default_rng(2).normal(size=G.n)
    np.testing.assert_allclose(M.forward(th, G, 2.0 * r), 2.0 * M.forward(th, G, r), rtol=1e-12)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    G, _ = _graph(4, seed)
    rng = np.random.default_rng(10 + seed)
    th = M.init_params(seed) + 0.05 * rng.normal(size=M.n_params())
    r = rng.normal(size=G.n)
    L0, g = M.grad(th, G, r)
    assert abs(L0 - M.loss(th, G, r)) < 1e-14 * max(1.0, L0)
    # directional derivatives
    for k in range(4):
        v = rng.normal(size=th.size)
        eps = 1e-6
        fd = (M.loss(th + eps * v, G, r) - M.loss(th - eps * v, G, r)) / (2 * eps)
        assert abs(fd - g @ v) <= 1e-6 * max(1.0, abs(fd)), (fd, g @ v)
    # single coordinates of every parameter block
    names = [n for n, _ in M.layer_shapes()]
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.layer_shapes()])
    for b, name in enumerate(names):
        for j in rng.choice(np.arange(offs[b], offs[b + 1]), size=min(3, offs[b + 1] - offs[b]), replace=False):
            eps = 1e-6
            tp = th.copy(); tp[j] += eps
            tm = th.copy(); tm[j] -= eps
            fd = (M.loss(tp, G, r) - M.loss(tm, G, r)) / (2 * eps)
            assert abs(fd - g[j]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (name, j, fd, g[j])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30], np.float32)
    r = M.bf16_round(x)
    # exact bf16 values stay; others land on the nearest of the two neighbours (ties to even)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    for a, b in zip(x, r):
        assert abs(a - b) <= abs(a) * 2.0 ** -8


@pytest.mark.parametrize("seed", [0, 1])
def test_energy_gradient_matches_finite_differences(seed):
    G, _ = _graph(4, 20 + seed)
    rng = np.random.default_rng(30 + seed)
    th = M.init_params(seed) + 0.05 * rng.normal(size=M.n_params())
    et = rng.normal(size=G.n)
    r = -G.matvec(et)                      # the pair satisfies A e* = -r
    L0, g = M.grad_energy(th, G, r, et)
    assert abs(L0 - M.loss_energy(th, G, r, et)) < 1e-13 * max(1.0, L0)
    for k in range(4):
        v = rng.normal(size=th.size)
        eps = 1e-6
        fd = (M.loss_energy(th + eps * v, G, r, et) - M.loss_energy(th - eps * v, G, r, et)) / (2 * eps)
        assert abs(fd - g @ v) <= 1e-6 * max(1.0, abs(fd)), (fd, g @ v)


def test_energy_loss_of_exact_error_is_zero_and_positive_otherwise():
    G, _ = _graph(4, 5)
    rng = np.random.default_rng(3)
    et = rng.normal(size=G.n)
    # -A is positive definite: the energy of any nonzero vector is positive
    for _ in range(5):
        d = rng.normal(size=G.n)
        assert -(d @ G.matvec(d)) > 0
