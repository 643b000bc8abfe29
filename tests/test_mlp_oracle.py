"""Pins of the CPU oracle of the learned router's network (oracle/mlp_oracle.py)."""
from __future__ import annotations

import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_round_known_values():
    vals = np.array([1.0, -2.0, 0.0, 3.140625, 1.00390625, 1.01171875, 1.0029296875, 65504.0, 1e-40],
                    np.float32)
    bits = mo.bf16_bits(vals)
    # 1 + 2^-8 (0x3f808000) is a tie with an even upper half -> 0x3f80;
    # 1 + 2^-7 + 2^-8 (0x3f818000) is a tie with an odd upper half -> 0x3f82;
    # 1 + 3 * 2^-10 (0x3f806000) rounds down; 65504 (0x477fe000) rounds up;
    # 1e-40 (0x000116c2, subnormal) keeps its upper half 0x0001
    assert list(bits) == [0x3F80, 0xC000, 0x0000, 0x4049, 0x3F80, 0x3F82, 0x3F80, 0x4780, 0x0001]


def test_bf16_round_is_idempotent_and_monotone():
    x = np.sort(np.random.default_rng(1).normal(0, 100, 10000).astype(np.float32))
    r = mo.bf16_round(x)
    assert np.array_equal(mo.bf16_round(r), r)
    assert np.all(np.diff(r) >= 0)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-30)


def _data(n, seed):
    rs = np.random.default_rng(seed)
    keys = rs.integers(0, 2 ** 63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
        rs.integers(0, 2, n, dtype=np.int64).astype(np.uint64)
    return keys, rs.normal(0, 1, n)


def test_exact_gradient_matches_finite_differences():
    keys, t = _data(24, 3)
    p = mo.init_params(5, 0.3).astype(np.float64)
    p[mo.OFF_B1:mo.OFF_W2] = 0.05
    p[mo.OFF_B2:mo.OFF_W3] = 0.05
    _, g = mo.grad(p, keys, t, exact=True)
    rs = np.random.default_rng(7)
    idx = np.concatenate([rs.choice(4096, 40, replace=False) + mo.OFF_W1,
                          rs.choice(64, 10, replace=False) + mo.OFF_B1,
                          rs.choice(4096, 40, replace=False) + mo.OFF_W2,
                          rs.choice(64, 10, replace=False) + mo.OFF_B2,
                          rs.choice(64, 10, replace=False) + mo.OFF_W3, [mo.OFF_B3]])
    h = 1e-6
    for e in idx:
        q = p.copy()
        q[e] += h
        lp = mo.loss(q, keys, t, exact=True)
        q[e] -= 2 * h
        lm = mo.loss(q, keys, t, exact=True)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[e]) <= 1e-6 + 1e-4 * abs(num), (e, num, g[e])


def test_rounded_gradient_close_to_exact():
    keys, t = _data(256, 4)
    p = mo.init_params(6)
    _, ge = mo.grad(p.astype(np.float64), keys, t, exact=True)
    _, gr = mo.grad(p, keys, t, exact=False)
    assert np.abs(gr - ge).max() <= 0.05 * np.abs(ge).max()


def test_gradient_descent_lowers_loss():
    keys, _ = _data(512, 8)
    t = ((keys >> np.uint64(60)).astype(np.float64))  # learnable from the top bits
    p = mo.init_params(9)
    l0 = None
    for _ in range(200):
        l, g = mo.grad(p, keys, t)
        l0 = l if l0 is None else l0
        p = (p - 0.05 * g).astype(np.float32)
    assert mo.loss(p, keys, t) < 0.5 * l0
