"""FP64 restatement of the selector MLP -- TEST INFRASTRUCTURE ONLY.

The same network as the device kernels (spmv-dare_amd/csrc/hip/mlp.hip):
n_in -> h1 -> h2 -> n_out, ReLU hidden layers, linear output, one flat
parameter vector with, per layer, W[n_in][n_out] (row-major) then b[n_out];
loss = sum((Y - T)^2 * M) / sum(M); Adam with beta1 0.9, beta2 0.999,
eps 1e-8. No counterpart in the reference. Pinned by central finite
differences of the loss (tests/test_mlp.py), so an error in a transposed
weight, a dropped ReLU mask or a wrong bias gradient fails there.
"""
import numpy as np


def layout(dims):
    offs = []
    o = 0
    for l in range(3):
        w = (o, o + dims[l] * dims[l + 1])
        o = w[1]
        b = (o, o + dims[l + 1])
        o = b[1]
        offs.append((w, b))
    return offs, o


def unpack(p, dims):
    offs, _ = layout(dims)
    out = []
    for l, ((w0, w1), (b0, b1)) in enumerate(offs):
        out.append((p[w0:w1].reshape(dims[l], dims[l + 1]), p[b0:b1]))
    return out


def forward(p, dims, X):
    """Returns (pre-activations, activations) per layer."""
    zs, acts = [], []
    a = np.asarray(X, dtype=np.float64)
    for l, (W, b) in enumerate(unpack(np.asarray(p, dtype=np.float64), dims)):
        z = a @ W + b
        a = np.maximum(z, 0.0) if l < 2 else z
        zs.append(z)
        acts.append(a)
    return zs, acts


def loss(p, dims, X, T, M=None):
    _, acts = forward(p, dims, X)
    M = np.ones_like(T, dtype=np.float64) if M is None else np.asarray(M, dtype=np.float64)
    d = (acts[-1] - T) * M
    return float((d * d).sum() / M.sum())


def gradients(p, dims, X, T, M=None):
    p = np.asarray(p, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    M = np.ones_like(T, dtype=np.float64) if M is None else np.asarray(M, dtype=np.float64)
    zs, acts = forward(p, dims, X)
    params = unpack(p, dims)
    offs, n = layout(dims)
    g = np.zeros(n)
    dz = 2.0 * (acts[-1] - T) * M / M.sum()
    for l in (2, 1, 0):
        a_in = X if l == 0 else acts[l - 1]
        (w0, w1), (b0, b1) = offs[l]
        g[w0:w1] = (a_in.T @ dz).reshape(-1)
        g[b0:b1] = dz.sum(axis=0)
        if l > 0:
            dz = (dz @ params[l][0].T) * (zs[l - 1] > 0)
    return g


def adam(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    return p - lr * (m1 / c1) / (np.sqrt(m2 / c2) + eps), m1, m2
