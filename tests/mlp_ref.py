"""float64 numpy restatement of the opening MLP (test infrastructure): forward pass, asymmetric
squared loss and its gradient by backpropagation, in the parameter layout of include/ptg.h
(W1[10][32], b1[32], W2[32][32] with rows = input units, b2[32], W3[32], b3)."""

import numpy as np

NIN, NH = 10, 32
NPARAM = NIN * NH + NH + NH * NH + NH + NH + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NIN * NH].reshape(NIN, NH); o += NIN * NH
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(x @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h2 @ W3 + b3


def loss(p, x, y, under):
    r = forward(p, x) - y
    w = np.where(r < 0, under, 1.0)
    return float(np.sum(w * r * r) / len(y))


def grad(p, x, y, under):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(x @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    r = h2 @ W3 + b3 - y
    w = np.where(r < 0, under, 1.0)
    g = 2 * w * r / len(y)
    dz2 = g[:, None] * W3[None, :] * (1 - h2 * h2)
    dz1 = (dz2 @ W2.T) * (1 - h1 * h1)
    return np.concatenate([(x.T @ dz1).ravel(), dz1.sum(0), (h1.T @ dz2).ravel(), dz2.sum(0), h2.T @ g, [g.sum()]])
