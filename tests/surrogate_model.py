"""Numpy model of the learned surrogate (16 -> 64 -> 64 -> 16 tanh MLP whose first
four outputs weight four su(N) basis vectors) with its analytic gradient; the
MFMA kernels are checked against it with their operands rounded to bfloat16."""
import numpy as np


def bf16(x):
    """Round float32 to bfloat16 (nearest even) and back."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(P):
    W1 = P[0:1024].reshape(16, 64)
    b1 = P[1024:1088]
    W2 = P[1088:5184].reshape(64, 64)
    b2 = P[5184:5248]
    W3 = P[5248:6272].reshape(64, 16)
    b3 = P[6272:6288]
    return W1, b1, W2, b2, W3, b3


def model(P, X, B, F, rnd):
    """Forward, loss and gradient of L = mean_rows |sum_k c_k B_k - F|^2; `rnd` rounds MFMA operands."""
    W1, b1, W2, b2, W3, b3 = unpack(P.astype(np.float64))
    X = X.astype(np.float64)
    h1 = np.tanh(rnd(X) @ rnd(W1) + b1)
    h2 = np.tanh(rnd(h1) @ rnd(W2) + b2)
    c = rnd(h2) @ rnd(W3) + b3
    S = np.einsum("rk,rka->ra", c[:, :4], B)
    res = S - F
    n = X.shape[0]
    loss = np.sum(res ** 2) / n
    g3 = np.zeros_like(c)
    g3[:, :4] = 2.0 / n * np.einsum("ra,rka->rk", res, B)
    dW3 = rnd(h2).T @ rnd(g3)
    dz2 = (rnd(g3) @ rnd(W3).T) * (1 - h2 ** 2)
    dW2 = rnd(h1).T @ rnd(dz2)
    dz1 = (rnd(dz2) @ rnd(W2).T) * (1 - h1 ** 2)
    dW1 = rnd(X).T @ rnd(dz1)
    dW3[:, 4:] = 0.0
    db3 = g3.sum(0)
    grad = np.concatenate([dW1.ravel(), dz1.sum(0), dW2.ravel(), dz2.sum(0), dW3.ravel(), db3])
    return S, loss, grad
