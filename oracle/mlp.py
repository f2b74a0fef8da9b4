"""CPU restatement of the learned selector's network and features -- TEST
INFRASTRUCTURE ONLY (imported by tests/ as the checker, never by the library).

The selector has no counterpart in the reference (it replaces the exhaustive
(d, c) search of SB/test/dspmv_test.cu:314-332), so this restatement is pinned
by mathematics instead: the backward pass against central finite differences
of the forward pass, the features against direct counting.

Network: x[32] -> h1 = relu(W1 x + b1)[64] -> h2 = relu(W2 h1 + b2)[64] ->
z = W3 h2 + b3 [96]; loss = mean over the batch of -sum_o T[o] log softmax(z)[o].
Parameters are one flat vector W1 b1 W2 b2 W3 b3 with W[out][in].
"""
from __future__ import annotations

import numpy as np

SHAPES = [(64, 32), (64,), (64, 64), (64,), (96, 64), (96,)]
NPARAMS = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p):
    out, o = [], 0
    for s in SHAPES:
        k = int(np.prod(s))
        out.append(np.asarray(p[o:o + k], np.float64).reshape(s))
        o += k
    return out


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    H1 = np.maximum(X @ W1.T + b1, 0.0)
    H2 = np.maximum(H1 @ W2.T + b2, 0.0)
    Z = H2 @ W3.T + b3
    return H1, H2, Z


def loss(p, X, T):
    _, _, Z = forward(p, X)
    mx = Z.max(axis=1, keepdims=True)
    lse = mx + np.log(np.exp(Z - mx).sum(axis=1, keepdims=True))
    return float((-(np.asarray(T, np.float64) * (Z - lse)).sum(axis=1)).mean())


def loss_grad(p, X, T):
    """(mean loss, gradient vector) by backpropagation."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    B = X.shape[0]
    H1, H2, Z = forward(p, X)
    mx = Z.max(axis=1, keepdims=True)
    E = np.exp(Z - mx)
    P = E / E.sum(axis=1, keepdims=True)
    dZ = (P - T) / B
    dW3, db3 = dZ.T @ H2, dZ.sum(0)
    dH2 = (dZ @ W3) * (H2 > 0)
    dW2, db2 = dH2.T @ H1, dH2.sum(0)
    dH1 = (dH2 @ W2) * (H1 > 0)
    dW1, db1 = dH1.T @ X, dH1.sum(0)
    g = np.concatenate([a.ravel() for a in (dW1, db1, dW2, db2, dW3, db3)])
    return loss(p, X, T), g


def adam(p, m1, m2, g, step, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step with L2 weight decay folded into the gradient."""
    g = g + wd * p
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** step)
    vh = m2 / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2


def features(rowptr, col, m, n, long_th):
    """The 32 selector features, by direct counting (float64); the column
    locality terms use the same strided sample as the device (stride =
    ceil(nnz / 2^20) when nnz > 2^20, else 1)."""
    rp = np.asarray(rowptr, np.int64)
    L = np.diff(rp)
    nnz = int(rp[-1]) if m else 0
    M, Z = max(m, 1), max(nnz, 1)
    mean = L.sum() / M
    var = max(0.0, float((L.astype(np.float64) ** 2).sum()) / M - mean * mean)
    sd = np.sqrt(var)
    f = np.zeros(32)
    f[0] = np.log2(m + 1.0) / 24
    f[1] = np.log2(n + 1.0) / 24
    f[2] = np.log2(nnz + 1.0) / 32
    f[3] = np.log2(mean + 1.0) / 12
    f[4] = min(sd / max(mean, 1e-30), 64.0) / 8
    f[5] = np.log2((L.max() if m else 0) + 1.0) / 24
    f[6] = (L == 0).sum() / M
    f[7] = L[L >= long_th].sum() / Z
    nz = L[L > 0]
    b = np.minimum(np.floor(np.log2(nz)).astype(np.int64), 15)
    f[8:24] = np.bincount(b, minlength=16)[:16] / M
    cls = np.where(nz < 4, 0, np.where(nz < 32, 1, np.where(nz < 1024, 2, 3)))
    f[24:28] = np.bincount(cls, weights=nz, minlength=4)[:4] / Z
    if nnz:
        stride = (nnz + (1 << 20) - 1) // (1 << 20) if nnz > (1 << 20) else 1
        k = np.arange(0, nnz, stride, dtype=np.int64)
        row = np.searchsorted(rp, k, side="right") - 1
        diag = row * n // m
        d = np.abs(np.asarray(col, np.int64)[k] - diag)
        loc = np.floor(np.log2(d.astype(np.float32) + np.float32(1)) * np.float32(1024)).astype(np.int64)
        f[28] = loc.sum() / 1024.0 / len(k) / 24
        f[29] = (d < 1024).mean()
        f[30] = (d < 32768).mean()
    f[31] = np.log2(sd + 1.0) / 12
    return f
