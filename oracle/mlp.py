"""FP64 numpy restatement of the learned delta proposal (features, MLP forward, backward, Adam) -- TEST
INFRASTRUCTURE ONLY: the gradient and feature checker of the GPU kernels.

The learned proposal has no counterpart in the reference; it is checked here against the mathematics itself
(finite differences of the loss for the backward pass) and used as the CPU restatement of the GPU kernels.
"""
from __future__ import annotations

import numpy as np

NBIN = 32
N_IN = 34
H = 128


def bins_of(s: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """Eigenvalue bin of each k: log-spaced edges (33), values outside clamped into the end bins."""
    return np.clip(np.searchsorted(edges, s, side="right") - 1, 0, NBIN - 1).astype(np.int32)


def features(uy: np.ndarray, bin_of_k: np.ndarray, n: float, c: float) -> np.ndarray:
    """(P, 34) features of the (neff, P) rotated columns uy: per eigenvalue bin, log(mean y^2 in the bin / mean
    y^2); a bin holding no eigenvalue takes the linear interpolation (in bin index) of its nearest occupied
    neighbours, and bins beyond the occupied range repeat the end value (all zero when no bin is occupied)."""
    y2 = uy * uy
    mean_all = np.maximum(y2.mean(axis=0), 1e-30)
    out = np.zeros((uy.shape[1], N_IN))
    occ = [b for b in range(NBIN) if np.any(bin_of_k == b)]
    for b in occ:
        out[:, b] = np.log(np.maximum(y2[bin_of_k == b].mean(axis=0), 1e-30) / mean_all)
    if occ:
        for b in range(NBIN):
            if b in occ:
                continue
            left = [o for o in occ if o < b]
            right = [o for o in occ if o > b]
            if left and right:
                lo, hi = left[-1], right[0]
                f = (b - lo) / (hi - lo)
                out[:, b] = (1 - f) * out[:, lo] + f * out[:, hi]
            else:
                out[:, b] = out[:, left[-1]] if left else out[:, right[0]]
    out[:, 32] = np.log(n) / 10
    out[:, 33] = c / 10
    return out


def unpack(theta: np.ndarray):
    o = 0
    W1 = theta[o:o + H * N_IN].reshape(H, N_IN); o += H * N_IN
    b1 = theta[o:o + H]; o += H
    W2 = theta[o:o + H * H].reshape(H, H); o += H * H
    b2 = theta[o:o + H]; o += H
    w3 = theta[o:o + H].reshape(1, H); o += H
    b3 = theta[o:o + 1]; o += 1
    return W1, b1, W2, b2, w3, b3


def n_params() -> int:
    return H * N_IN + H + H * H + H + H + 1


def forward(theta, x):
    W1, b1, W2, b2, w3, b3 = unpack(theta)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return (h2 @ w3.T + b3)[:, 0], (h1, h2)


def mse_and_grad(theta, x, target):
    """Mean squared error and its gradient with respect to the flat parameters (backpropagation)."""
    W1, b1, W2, b2, w3, b3 = unpack(theta)
    pred, (h1, h2) = forward(theta, x)
    B = x.shape[0]
    loss = np.mean((pred - target) ** 2)
    d = 2 * (pred - target) / B  # (B,)
    gw3 = d @ h2
    gb3 = d.sum(keepdims=True)
    dz2 = np.outer(d, w3[0]) * (1 - h2 ** 2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (1 - h1 ** 2)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3.ravel(), gb3])


def adam_step(theta, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
