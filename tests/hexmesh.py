"""Test-side numpy restatement of the jittered hex mesh layout built by
fvm_mesh_create_hex (same face order, node order, ghost numbering and jitter
hash), so that CPU tests can drive the assembly oracle and GPU tests can check
the device mesh bit for bit."""
import numpy as np

M64 = (1 << 64) - 1


def _mix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def u01(seed, idx, stream):
    h = _mix64(_mix64(seed ^ ((0xD1B54A32D192ED03 * (stream + 1)) & M64)) ^ idx)
    return (h >> 11) * 2.0 ** -53


def hex_mesh(N, jitter=0.0, seed=0):
    h = 1.0 / N
    nn = (N + 1) ** 3
    nodes = np.zeros((nn, 3))
    for nid in range(nn):
        idx = (nid % (N + 1), (nid // (N + 1)) % (N + 1), nid // ((N + 1) ** 2))
        for a in range(3):
            v = float(idx[a]) * h
            if 0 < idx[a] < N:
                v += (2.0 * u01(seed, nid, a) - 1.0) * jitter * h
            nodes[nid, a] = v
    nid = lambda i, j, k: i + (N + 1) * (j + (N + 1) * k)
    cid = lambda i, j, k: i + N * (j + N * k)
    fn, fc = [], []
    for k in range(N):
        for j in range(N):
            for i in range(N - 1):
                fc.append((cid(i, j, k), cid(i + 1, j, k)))
                fn.append((nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i + 1, j + 1, k + 1), nid(i + 1, j, k + 1)))
    for k in range(N):
        for j in range(N - 1):
            for i in range(N):
                fc.append((cid(i, j, k), cid(i, j + 1, k)))
                fn.append((nid(i, j + 1, k), nid(i, j + 1, k + 1), nid(i + 1, j + 1, k + 1), nid(i + 1, j + 1, k)))
    for k in range(N - 1):
        for j in range(N):
            for i in range(N):
                fc.append((cid(i, j, k), cid(i, j, k + 1)))
                fn.append((nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)))
    n = N ** 3
    fint = len(fc)
    g = 0
    for grp in range(6):
        for a1 in range(N):
            for a0 in range(N):
                ghost = n + g
                g += 1
                if grp == 0:
                    c0 = cid(0, a0, a1); q = (nid(0, a0, a1), nid(0, a0, a1 + 1), nid(0, a0 + 1, a1 + 1), nid(0, a0 + 1, a1))
                elif grp == 1:
                    c0 = cid(N - 1, a0, a1); q = (nid(N, a0, a1), nid(N, a0 + 1, a1), nid(N, a0 + 1, a1 + 1), nid(N, a0, a1 + 1))
                elif grp == 2:
                    c0 = cid(a0, 0, a1); q = (nid(a0, 0, a1), nid(a0 + 1, 0, a1), nid(a0 + 1, 0, a1 + 1), nid(a0, 0, a1 + 1))
                elif grp == 3:
                    c0 = cid(a0, N - 1, a1); q = (nid(a0, N, a1), nid(a0, N, a1 + 1), nid(a0 + 1, N, a1 + 1), nid(a0 + 1, N, a1))
                elif grp == 4:
                    c0 = cid(a0, a1, 0); q = (nid(a0, a1, 0), nid(a0, a1 + 1, 0), nid(a0 + 1, a1 + 1, 0), nid(a0 + 1, a1, 0))
                else:
                    c0 = cid(a0, a1, N - 1); q = (nid(a0, a1, N), nid(a0 + 1, a1, N), nid(a0 + 1, a1 + 1, N), nid(a0, a1 + 1, N))
                fc.append((c0, ghost))
                fn.append(q)
    gptr = [0, fint] + [fint + (t + 1) * N * N for t in range(6)]
    return dict(N=N, nodes=nodes.reshape(-1), face_nodes=np.array(fn, np.int32).reshape(-1),
                face_cells=np.array(fc, np.int32).reshape(-1), n_cells=n, n_ghost=6 * N * N,
                group_ptr=np.array(gptr, np.int32))
