"""Learned correction for the AMG preconditioner (fvm_mpnn_* in the C ABI).

A 3-layer message-passing network over the graph of a CRMatrix maps a residual
to a correction (delta form, A e ~ -r); it runs on the matrix cores in bf16 with
FP32 accumulation and is trained on the GPU with the library's own forward and
backward kernels. It has no counterpart in the reference: it extends the
preconditioner that the reference's Krylov loops call (BCGStab.cpp:89,113).

    net = LearnedCorrection(seed=1)
    pairs = training_pairs(...)                 # residual/error pairs from plain AMG
    net.train(pairs, steps=..., lr=...)
    fg = FGMRES(amg); fg.set_learned(net)       # z += NN(A z - v) after every cycle
"""
import ctypes

import numpy as np

from ._lib import check, lib


def _dp(t):
    return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())


class LearnedCorrection:
    def __init__(self, device=0, seed=0):
        h = ctypes.c_void_p()
        check(lib().fvm_mpnn_create(ctypes.byref(h), int(device), int(seed)))
        self._h = h
        self.device = device
        self._graph = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().fvm_mpnn_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @staticmethod
    def num_params():
        return int(lib().fvm_mpnn_num_params())

    def params(self):
        th = np.zeros(self.num_params(), np.float32)
        check(lib().fvm_mpnn_get_params(self._h, ctypes.c_void_p(th.ctypes.data)))
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        if th.size != self.num_params():
            raise ValueError("expected %d parameters" % self.num_params())
        check(lib().fvm_mpnn_set_params(self._h, ctypes.c_void_p(th.ctypes.data)))

    def set_graph(self, n, row, col, off, diag):
        """Device CSR of the self rows (torch tensors or addresses), kept referenced."""
        self._graph = (row, col, off, diag)
        check(lib().fvm_mpnn_set_graph(self._h, int(n), _dp(row), _dp(col), _dp(off), _dp(diag)))

    def forward(self, r, e, accumulate=False):
        check(lib().fvm_mpnn_forward(self._h, _dp(r), _dp(e), 1 if accumulate else 0))
        return e

    def loss_grad(self, r):
        L = ctypes.c_double()
        check(lib().fvm_mpnn_loss_grad(self._h, _dp(r), ctypes.byref(L)))
        return L.value

    def loss_grad_pair(self, r, e):
        """Energy-norm loss of a residual/error pair (A e = -r); gradient accumulated."""
        L = ctypes.c_double()
        check(lib().fvm_mpnn_loss_grad_pair(self._h, _dp(r), _dp(e), ctypes.byref(L)))
        return L.value

    def grad(self, reset=False):
        g = np.zeros(self.num_params(), np.float32)
        check(lib().fvm_mpnn_get_grad(self._h, ctypes.c_void_p(g.ctypes.data), 1 if reset else 0))
        return g

    def adam_step(self, lr):
        check(lib().fvm_mpnn_adam_step(self._h, float(lr)))

    def train(self, samples, steps, lr=1e-3, batch=4, seed=0, log=None):
        """samples: list of (graph, residual) or (graph, residual, error) with
        graph = (n, row, col, off, diag) device CSR and device tensors; pairs train
        on the energy-norm error loss, residuals alone on the residual loss.
        Minibatches of `batch` samples, Adam with learning rate lr (cosine decay to
        lr/20). Returns the loss curve."""
        rng = np.random.default_rng(seed)
        curve = []
        cur = None
        for it in range(steps):
            idx = rng.choice(len(samples), size=min(batch, len(samples)), replace=False)
            tot = 0.0
            for k in sorted(idx, key=lambda j: id(samples[j][0])):
                g, r = samples[k][0], samples[k][1]
                if g is not cur:
                    self.set_graph(*g)
                    cur = g
                tot += self.loss_grad_pair(r, samples[k][2]) if len(samples[k]) > 2 else self.loss_grad(r)
            lr_t = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * it / max(steps, 1))))
            self.adam_step(lr_t)
            curve.append(tot / len(idx))
            if log and (it % max(1, steps // 10) == 0 or it == steps - 1):
                log("train step %d: loss %.4f" % (it, curve[-1]))
        return curve
