"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

FP64 numpy restatement of the learned message-passing correction of fvm_amd
(fvm_amd/csrc/mpnn.hip). It has no counterpart in the reference; it is the
checker of the GPU forward and backward kernels, and its own gradients are
pinned by central finite differences (tests/test_mpnn_ref.py).

Network over the graph of a CRMatrix (delta form: A e = -r):
  c_p   = -off_p / d_i                    (edge weight of entry p in row i)
  l_i   = (log|d_i| - mean) / (std + 1e-12)
  u_i   = -r_i / d_i ;  s = sqrt(mean(u^2)) + 1e-300 ;  v = u / s
  h0_i  = [v_i, (C v)_i, l_i, 0 ... 0]    (F0 = 16 features, 3 used)
  layer k = 0, 1, 2:  a = C h ;  h' = relu(Ws_k h + Wn_k a + b_k)   (64 features)
  y_i   = wo . h3_i + bo ;   e_i = s y_i
Losses of one sample (A symmetric; -A positive definite for the FV matrices):
  residual loss  L = ||r + A e||^2 / ||r||^2
  energy loss    L = (e - e*)^T (-A) (e - e*) / (e*^T (-A) e*)  for a pair (r, e*), A e* = -r.
Weights are used as the GPU uses them: bf16-rounded copies of FP32 masters.
"""
import numpy as np

H = 64
F0 = 16
LAYERS = 3


def bf16_round(a):
    """Round FP32 values to bfloat16 (round to nearest even), returned as FP64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_shapes():
    """[(name, shape)] in the packed parameter order of the GPU library."""
    out = []
    for k in range(LAYERS):
        fin = F0 if k == 0 else H
        out += [("Ws%d" % k, (H, fin)), ("Wn%d" % k, (H, fin)), ("b%d" % k, (H,))]
    out += [("wo", (H,)), ("bo", (1,))]
    return out


def n_params():
    return sum(int(np.prod(s)) for _, s in layer_shapes())


def unpack(theta):
    P, o = {}, 0
    for name, shp in layer_shapes():
        k = int(np.prod(shp))
        P[name] = np.asarray(theta[o:o + k], np.float64).reshape(shp)
        o += k
    return P


def pack(P):
    return np.concatenate([np.asarray(P[name], np.float64).reshape(-1) for name, _ in layer_shapes()])


def init_params(seed=0):
    """He-style initialization (the GPU library's fvm_mpnn_init draws its own)."""
    rng = np.random.default_rng(seed)
    P = {}
    for k in range(LAYERS):
        fin = F0 if k == 0 else H
        P["Ws%d" % k] = rng.normal(0, np.sqrt(1.0 / fin), (H, fin))
        P["Wn%d" % k] = rng.normal(0, np.sqrt(1.0 / fin), (H, fin))
        P["b%d" % k] = np.zeros(H)
    P["wo"] = rng.normal(0, np.sqrt(1.0 / H), H)
    P["bo"] = np.zeros(1)
    return pack(P)


def quantize(theta):
    """Weights as the kernels see them: matrices in bf16, biases and readout FP32."""
    P = unpack(theta)
    for k in range(LAYERS):
        P["Ws%d" % k] = bf16_round(P["Ws%d" % k])
        P["Wn%d" % k] = bf16_round(P["Wn%d" % k])
        P["b%d" % k] = np.asarray(P["b%d" % k], np.float32).astype(np.float64)
    P["wo"] = np.asarray(P["wo"], np.float32).astype(np.float64)
    P["bo"] = np.asarray(P["bo"], np.float32).astype(np.float64)
    return pack(P)


class Graph:
    def __init__(self, row, col, off, diag):
        self.n = len(diag)
        self.row = np.asarray(row, np.int64)
        self.col = np.asarray(col, np.int64)
        self.off = np.asarray(off, np.float64)
        self.diag = np.asarray(diag, np.float64)
        rid = np.repeat(np.arange(self.n), np.diff(self.row))
        self.rid = rid
        self.c = -self.off / self.diag[rid]
        ld = np.log(np.abs(self.diag))
        self.l = (ld - ld.mean()) / (ld.std() + 1e-12)

    def agg(self, h):
        """(C h)_i = sum_p c_p h_col(p), rows in CSR order."""
        out = np.zeros((self.n,) + h.shape[1:])
        np.add.at(out, self.rid, self.c.reshape((-1,) + (1,) * (h.ndim - 1)) * h[self.col])
        return out

    def agg_t(self, g):
        """(C^T g)_j = sum over entries p = (i, j) of c_p g_i."""
        out = np.zeros((self.n,) + g.shape[1:])
        np.add.at(out, self.col, self.c.reshape((-1,) + (1,) * (g.ndim - 1)) * g[self.rid])
        return out

    def matvec(self, x):
        """A x = diag x + off-diagonal part."""
        y = self.diag * x
        np.add.at(y, self.rid, self.off * x[self.col])
        return y


def forward(theta, G, r, keep=False):
    P = unpack(theta)
    u = -r / G.diag
    s = np.sqrt(np.mean(u * u)) + 1e-300
    v = u / s
    h = np.zeros((G.n, F0))
    h[:, 0] = v
    h[:, 1] = G.agg(v)
    h[:, 2] = G.l
    cache = []
    for k in range(LAYERS):
        a = G.agg(h)
        z = h @ P["Ws%d" % k].T + a @ P["Wn%d" % k].T + P["b%d" % k]
        hn = np.maximum(z, 0.0)
        cache.append((h, a, hn))
        h = hn
    y = h @ P["wo"] + P["bo"][0]
    e = s * y
    if keep:
        return e, dict(cache=cache, h=h, s=s)
    return e


def loss(theta, G, r):
    e = forward(theta, G, r)
    t = r + G.matvec(e)
    return float(t @ t / (r @ r))


def grad(theta, G, r):
    """Analytic gradient of loss() (A symmetric)."""
    P = unpack(theta)
    e, st = forward(theta, G, r, keep=True)
    t = r + G.matvec(e)
    rr = r @ r
    ge = 2.0 * G.matvec(t) / rr          # A^T t, A symmetric
    dy = st["s"] * ge
    D = {name: np.zeros(shp) for name, shp in layer_shapes()}
    D["wo"] = st["h"].T @ dy
    D["bo"] = np.array([dy.sum()])
    dh = np.outer(dy, P["wo"])
    for k in reversed(range(LAYERS)):
        h, a, hn = st["cache"][k]
        delta = dh * (hn > 0)
        D["Ws%d" % k] = delta.T @ h
        D["Wn%d" % k] = delta.T @ a
        D["b%d" % k] = delta.sum(0)
        if k > 0:
            dh = delta @ P["Ws%d" % k] + G.agg_t(delta @ P["Wn%d" % k])
    return float(t @ t / rr), pack(D)


def loss_energy(theta, G, r, et):
    e = forward(theta, G, r)
    d = e - et
    return float(-(d @ G.matvec(d)) / -(et @ G.matvec(et)))


def grad_energy(theta, G, r, et):
    """Analytic gradient of loss_energy() (A symmetric)."""
    P = unpack(theta)
    e, st = forward(theta, G, r, keep=True)
    d = e - et
    den = -(et @ G.matvec(et))
    Md = -G.matvec(d)
    ge = 2.0 * Md / den
    dy = st["s"] * ge
    D = {name: np.zeros(shp) for name, shp in layer_shapes()}
    D["wo"] = st["h"].T @ dy
    D["bo"] = np.array([dy.sum()])
    dh = np.outer(dy, P["wo"])
    for k in reversed(range(LAYERS)):
        h, a, hn = st["cache"][k]
        delta = dh * (hn > 0)
        D["Ws%d" % k] = delta.T @ h
        D["Wn%d" % k] = delta.T @ a
        D["b%d" % k] = delta.sum(0)
        if k > 0:
            dh = delta @ P["Ws%d" % k] + G.agg_t(delta @ P["Wn%d" % k])
    return float(d @ Md / den), pack(D)
