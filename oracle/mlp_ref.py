"""Reference forward/backward of the block-proposal MLP (test
infrastructure only): float64 numpy on the same parameters, with the same
bf16 rounding points as the GPU kernels (GEMM operands are bf16, products
accumulate in wide precision)."""
from __future__ import annotations

import numpy as np

D_IN, HID = 80, 256


def _bf16(x):
    """round-to-nearest-even to bfloat16, returned as float64"""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(desc, fdesc, emb):
    n = desc.shape[0]
    X = np.zeros((n, D_IN))
    X[:, :32] = emb[desc[:, 0]]
    X[:, 32:64] = emb[desc[:, 1]]
    X[:, 64:67] = fdesc
    X[np.arange(n), 67 + desc[:, 2]] = 1.0
    X[np.arange(n), 70 + desc[:, 3]] = 1.0
    return X


def forward_backward(P, desc, fdesc, target, round_bf16=True):
    """P: dict W1 (80x256), W2, W3, b1, b2, b3, w4, b4, emb.
    Returns (y, loss, grads dict).  round_bf16=False drops the bf16 rounding
    points, leaving the exact float64 network (the form the tests check
    against automatic differentiation)."""
    bf16 = _bf16 if round_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    n = desc.shape[0]
    X = features(desc, fdesc, P["emb"].astype(np.float64))
    Xb = bf16(X)
    W1, W2, W3 = bf16(P["W1"]), bf16(P["W2"]), bf16(P["W3"])
    H1 = np.maximum(Xb @ W1 + P["b1"], 0.0)
    H2 = np.maximum(bf16(H1) @ W2 + P["b2"], 0.0)
    H3 = np.maximum(bf16(H2) @ W3 + P["b3"], 0.0)
    y = H3 @ P["w4"] + P["b4"]
    e = y - target
    loss = np.mean(e * e)
    dy = 2.0 * e / n
    g = {}
    g["w4"] = H3.T @ dy
    g["b4"] = np.array([dy.sum()])
    dZ3 = np.where(H3 > 0, np.outer(dy, P["w4"]), 0.0)
    g["W3"] = bf16(H2).T @ bf16(dZ3)
    g["b3"] = dZ3.sum(axis=0)
    dZ2 = np.where(H2 > 0, bf16(dZ3) @ W3.T, 0.0)
    g["W2"] = bf16(H1).T @ bf16(dZ2)
    g["b2"] = dZ2.sum(axis=0)
    dZ1 = np.where(H1 > 0, bf16(dZ2) @ W2.T, 0.0)
    g["W1"] = Xb.T @ bf16(dZ1)
    g["b1"] = dZ1.sum(axis=0)
    dX = bf16(dZ1) @ W1.T
    demb = np.zeros_like(P["emb"], dtype=np.float64)
    np.add.at(demb, desc[:, 0], dX[:, :32])
    np.add.at(demb, desc[:, 1], dX[:, 32:64])
    g["emb"] = demb
    return y, loss, g


bf16 = _bf16
