"""CPU oracle of the learned warm start (numpy).  TEST INFRASTRUCTURE ONLY:
the checker of the device kernels in dachshund_amd/csrc/warmstart.hip.

The warm start has no counterpart in the reference; what the oracle fixes
is the network's definition: features (32 nearest neighbours under the
covariance metric, nearest first, ties by index), a 162 -> 64 -> 64 -> 64 ->
1 ReLU network with weights stored [in][out] in one flat vector, optional
BF16 rounding (nearest even) of the hidden activations, and the gradient of
the mean squared error.  Its backward pass is itself pinned by finite
differences (tests/test_warmstart.py).
"""
from __future__ import annotations

import numpy as np

K = 32
FEAT = 5 * K + 2
H = 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + FEAT * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
NPARAM = OFF_B4 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[OFF_W1:OFF_B1].reshape(FEAT, H), P[OFF_B1:OFF_W2],
            P[OFF_W2:OFF_B2].reshape(H, H), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3].reshape(H, H), P[OFF_B3:OFF_W4],
            P[OFF_W4:OFF_B4], P[OFF_B4])


def features(pixels_raw, var_s, l_perp, l_para):
    """Features of raw Pixel records (n, 5) {x, y, z, n, d}: brute-force
    32-NN under the metric (dx^2+dy^2)/(2 l_perp^2) + dz^2/(2 l_para^2),
    restricted (like the device search) to neighbours inside the table
    cutoff (metric < 8); absent neighbours are zeros."""
    p = np.asarray(pixels_raw, dtype=np.float64)
    n = len(p)
    cp, cl = 1 / (2 * l_perp ** 2), 1 / (2 * l_para ** 2)
    s = 1 / np.sqrt(var_s)
    out = np.zeros((n, FEAT))
    for i in range(n):
        d = p[:, :3] - p[i, :3]
        g = cp * (d[:, 0] ** 2 + d[:, 1] ** 2) + cl * d[:, 2] ** 2
        g[i] = np.inf
        order = np.lexsort((np.arange(n), g))
        nb = [j for j in order[:K] if g[j] * 24.875 < 199.0]
        for k, j in enumerate(nb):
            out[i, 5 * k:5 * k + 5] = [d[j, 0] / l_perp, d[j, 1] / l_perp,
                                       d[j, 2] / l_para, p[j, 3] * s,
                                       p[j, 4] * s]
        out[i, 5 * K] = p[i, 3] * s
        out[i, 5 * K + 1] = p[i, 4] * s
    return bf16(out.astype(np.float32))


def forward(P, feat, quant=True):
    """Returns (y, (x, h1, h2, h3))."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(P)
    x = np.asarray(feat, dtype=np.float64)

    def act(z):
        r = np.maximum(z, 0.0)
        return bf16(r.astype(np.float32)).astype(np.float64) if quant else r

    h1 = act(x @ W1 + b1)
    h2 = act(h1 @ W2 + b2)
    h3 = act(h2 @ W3 + b3)
    y = h3 @ W4 + b4
    return y, (x, h1, h2, h3)


def loss_and_grad(P, feat, target, quant=True):
    """Mean squared error and its gradient (straight through the BF16
    rounding), in the flat parameter layout."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(P)
    y, (x, h1, h2, h3) = forward(P, feat, quant)
    B = len(y)
    e = y - np.asarray(target, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 / B * e
    d3 = np.outer(dy, W4) * (h3 > 0)
    d2 = (d3 @ W3.T) * (h2 > 0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (x.T @ d1).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ d2).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = (h2.T @ d3).ravel()
    g[OFF_B3:OFF_W4] = d3.sum(0)
    g[OFF_W4:OFF_B4] = h3.T @ dy
    g[OFF_B4] = dy.sum()
    return loss, g
