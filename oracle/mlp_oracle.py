"""CPU oracle of the learned router's network (test infrastructure only).

Restates, in numpy, the 64 -> 64 -> 64 -> 1 ReLU MLP of the learned
heavy-hitter form (include/qsketch_learned.h): input = the 64 bits of a key
(bit k -> input k, 0.0 or 1.0), BF16 weights and activations, FP32
accumulation, loss = mean((y - target)^2) / 2.  The reference has no learned
component, so there is no reference file to cite; the pins are mathematical:

  * bf16_round is round-to-nearest-even on the top 16 bits of an IEEE binary32
    (checked against hand-computed bit patterns, ties included);
  * grad(..., exact=True) evaluates the same formulas in float64 without
    rounding, and the tests compare it with central finite differences of
    loss(), which pins every term of the backward pass (a dropped ReLU mask, a
    transposed W2, a wrong bias index fail it);
  * grad(..., exact=False) is the same code with the rounding switched on, the
    configuration the GPU kernels compute.

Parameter layout (float32, 8385): W1[64][64] (unit, key bit), b1[64],
W2[64][64] (layer-2 unit, layer-1 unit), b2[64], w3[64], b3.
"""
from __future__ import annotations

import numpy as np

NPAR = 8385
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 4096, 4160, 8256, 8320, 8384


def bf16_round(a) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def bf16_bits(a) -> np.ndarray:
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def unpack(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(64, 64), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(64, 64),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def key_bits(keys) -> np.ndarray:
    k = np.asarray(keys, dtype=np.uint64)
    return ((k[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float32)


def forward(p, keys, exact: bool = False):
    """Scores and intermediates.  exact=False: BF16 weights/activations, FP32
    sums; exact=True: float64 throughout, no rounding."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: np.asarray(a, dt)) if exact else bf16_round
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, dt))
    X = key_bits(keys).astype(dt)
    W1r, W2r, w3r = rnd(W1).astype(dt), rnd(W2).astype(dt), rnd(w3).astype(dt)
    a1 = X @ W1r.T + b1
    h1 = rnd(np.maximum(a1, 0)).astype(dt)
    a2 = h1 @ W2r.T + b2
    h2 = rnd(np.maximum(a2, 0)).astype(dt)
    y = h2 @ w3r + b3
    return y, (X, a1, h1, a2, h2, W2r, w3r)


def loss(p, keys, targets, exact: bool = False) -> float:
    y, _ = forward(p, keys, exact)
    t = np.asarray(targets, np.float64 if exact else np.float32)
    return float(0.5 * np.mean((y - t) ** 2))


def grad(p, keys, targets, exact: bool = False):
    """(loss, gradient vector in the parameter layout)."""
    y, (X, a1, h1, a2, h2, W2r, w3r) = forward(p, keys, exact)
    dt = y.dtype
    t = np.asarray(targets, dt)
    n = y.shape[0]
    dy = (y - t) / n
    g = np.zeros(NPAR, dt)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dh2 = (dy[:, None] * w3r[None, :]) * (a2 > 0)
    g[OFF_W2:OFF_B2] = (dh2.T @ h1).reshape(-1)
    g[OFF_B2:OFF_W3] = dh2.sum(0)
    dh1 = (dh2 @ W2r) * (a1 > 0)
    g[OFF_W1:OFF_B1] = (dh1.T @ X).reshape(-1)
    g[OFF_B1:OFF_W2] = dh1.sum(0)
    return float(0.5 * np.mean((y - t) ** 2)), g


def init_params(seed: int, scale: float = 0.125) -> np.ndarray:
    rs = np.random.default_rng(seed)
    p = np.zeros(NPAR, np.float32)
    p[OFF_W1:OFF_B1] = rs.uniform(-scale, scale, 4096)
    p[OFF_W2:OFF_B2] = rs.uniform(-scale, scale, 4096)
    p[OFF_W3:OFF_B3] = rs.uniform(-scale, scale, 64)
    return p
