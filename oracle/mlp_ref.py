"""CPU oracle of the count-kernel cost model (float64 numpy).

TEST INFRASTRUCTURE ONLY: checks the engine's HIP forward / backward kernels
(ktruss-explorer_amd/csrc/kt_learn.hip).  The reference has no learned part;
this restates the network from its definition:
    h1 = relu(W1 x + b1),  h2 = relu(W2 h1 + b2),  y = W3 h2 + b3
    loss = sum_s w_s sum_c m_sc (y_sc - t_sc)^2 / sum_s w_s
with the parameter vector laid out as W1[64][16] b1[64] W2[64][64] b2[64]
W3[C][64] b3[C] (row-major).  The backward pass is pinned independently by
central finite differences on this same function.
"""
import numpy as np

IN, HID = 16, 64


def unpack(p, C):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + HID * IN].reshape(HID, IN); o += HID * IN
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    W3 = p[o:o + C * HID].reshape(C, HID); o += C * HID
    b3 = p[o:o + C]; o += C
    assert o == len(p)
    return W1, b1, W2, b2, W3, b3


def forward(p, C, x):
    W1, b1, W2, b2, W3, b3 = unpack(p, C)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


def loss(p, C, x, t, m=None, w=None):
    y = forward(p, C, x)
    t = np.asarray(t, np.float64)
    m = np.ones_like(t) if m is None else np.asarray(m, np.float64)
    w = np.ones(len(t)) if w is None else np.asarray(w, np.float64)
    d = (y - t) * m
    return float((w[:, None] * d * d).sum() / w.sum())


def loss_grad(p, C, x, t, m=None, w=None):
    W1, b1, W2, b2, W3, b3 = unpack(p, C)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    m = np.ones_like(t) if m is None else np.asarray(m, np.float64)
    w = np.ones(len(t)) if w is None else np.asarray(w, np.float64)
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ W3.T + b3
    d = (y - t) * m
    L = float((w[:, None] * d * d).sum() / w.sum())
    dy = 2 * w[:, None] * d * m / w.sum()
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ W3) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


def finite_difference(p, C, x, t, m=None, w=None, idx=None, eps=1e-6):
    p = np.asarray(p, np.float64).copy()
    idx = range(len(p)) if idx is None else idx
    out = {}
    for i in idx:
        old = p[i]
        p[i] = old + eps
        lp = loss(p, C, x, t, m, w)
        p[i] = old - eps
        lm = loss(p, C, x, t, m, w)
        p[i] = old
        out[i] = (lp - lm) / (2 * eps)
    return out
