"""CPU oracle of the learned distance bound's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the GPU kernels in meshlib_amd/csrc/mi_mlp.hip; only
tests/ import it.  It evaluates the same network with the same BF16 rounding points as the GPU
(features, weights, hidden activations and, in the backward pass, dZ are rounded to BF16 with
round-to-nearest-even; every product is accumulated in FP32 or better), so forward outputs and
gradients agree up to FP32 summation order and transcendental ulps.  The oracle's own gradient is
pinned by central finite differences in float64 (``exact=True`` turns all rounding off).

Network: x = (p - centre) / scale; features [x, sin(pi f_k x_d), cos(pi f_k x_d)], f_k = 2^(k/3),
k = 0..15 (99 dims, padded to 128); 4 x (Linear 128 + SiLU); Linear 1; softplus -> distance / scale.
Loss over a batch of n points: 1/(2n) sum (softplus(u) - label/scale)^2.
"""
from __future__ import annotations

import numpy as np

HID, LAYERS, BANDS = 128, 4, 16
NFEAT = 3 + 6 * BANDS
OFF_W, OFF_B = 0, LAYERS * HID * HID
OFF_W5 = OFF_B + LAYERS * HID
OFF_B5 = OFF_W5 + HID
NPARAMS = OFF_B5 + 1


def bf16(x, exact=False):
    """round float32 values to the nearest BF16 (ties to even), returned as float32"""
    x = np.asarray(x, dtype=np.float32)
    if exact:
        return x
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def pi_freq():
    return (np.pi * 2.0 ** (np.arange(BANDS) / 3.0)).astype(np.float32)


def features(pts, center, scale, exact=False):
    """[n, 128] features of the points (float32 arithmetic as on the GPU, then BF16)"""
    pts = np.asarray(pts, np.float32)
    inv = np.float32(1.0) / np.float32(scale)
    x = (pts - np.asarray(center, np.float32)) * inv
    n = x.shape[0]
    f = np.zeros((n, HID), np.float32)
    f[:, :3] = x
    pf = pi_freq()
    for k in range(BANDS):
        a = pf[k] * x
        f[:, 3 + 6 * k:6 + 6 * k] = np.sin(a.astype(np.float64)).astype(np.float32)
        f[:, 6 + 6 * k:9 + 6 * k] = np.cos(a.astype(np.float64)).astype(np.float32)
    return bf16(f, exact)


def unpack(theta):
    th = np.asarray(theta, np.float32)
    W = th[OFF_W:OFF_B].reshape(LAYERS, HID, HID)
    b = th[OFF_B:OFF_W5].reshape(LAYERS, HID)
    w5 = th[OFF_W5:OFF_B5]
    b5 = th[OFF_B5]
    return W, b, w5, b5


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(theta, pts, center, scale, exact=False):
    """returns (pred distance, cache) with the GPU's rounding points"""
    W, b, w5, b5 = unpack(theta)
    dt = np.float64
    X = features(pts, center, scale, exact).astype(dt)
    H = [X]
    Z = []
    for l in range(LAYERS):
        z = H[-1] @ bf16(W[l], exact).astype(dt).T + b[l].astype(dt)
        Z.append(z)
        h = z * _sig(z)
        H.append(bf16(h.astype(np.float32), exact).astype(dt) if not exact else h)
    u = H[-1] @ bf16(w5, exact).astype(dt) + dt(b5)
    pred = np.where(u > 20, u, np.log1p(np.exp(np.minimum(u, 20))))
    return pred * scale, (X, H, Z, u)


def loss(theta, pts, labels, center, scale, exact=False):
    pred, _ = forward(theta, pts, center, scale, exact)
    d = pred / scale - np.asarray(labels, np.float64) / scale
    return 0.5 * np.mean(d * d)


def gradient(theta, pts, labels, center, scale, exact=False):
    """gradient of loss() with the GPU's rounding of dZ to BF16 before the two GEMMs"""
    W, b, w5, b5 = unpack(theta)
    dt = np.float64
    pred, (X, H, Z, u) = forward(theta, pts, center, scale, exact)
    n = X.shape[0]
    y = np.asarray(labels, dt) / scale
    sp = np.where(u > 20, u, np.log1p(np.exp(np.minimum(u, 20))))
    du = (sp - y) / n * _sig(u)
    g = np.zeros(NPARAMS, dt)
    g[OFF_W5:OFF_B5] = du @ H[LAYERS]
    g[OFF_B5] = du.sum()
    dH = np.outer(du, bf16(w5, exact).astype(dt))
    for l in range(LAYERS - 1, -1, -1):
        s = _sig(Z[l])
        dz = dH * (s * (1 + Z[l] * (1 - s)))
        g[OFF_B + l * HID:OFF_B + (l + 1) * HID] = dz.sum(0)
        dzb = bf16(dz.astype(np.float32), exact).astype(dt) if not exact else dz
        g[OFF_W + l * HID * HID:OFF_W + (l + 1) * HID * HID] = (dzb.T @ H[l]).reshape(-1)
        dH = dzb @ bf16(W[l], exact).astype(dt)
    return g
