"""GPU checks of the learned heavy-hitter form: the network's forward passes
and gradients against the CPU oracle (oracle/mlp_oracle.py), training, and the
routed insert/query against the sketch oracle plus exact counts."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from oracle import pyoracle as po
from tests.helpers import to_dev, to_host_u32

pytestmark = pytest.mark.gpu


def _print_case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _learned(n_low=4093, n_high=4096, table_log2=12, key_bytes=8, rng_seed=3, init_seed=1):
    import qsketch_amd as q
    p = q.make_params(n_low, n_high, key_bytes=key_bytes, rng_seed=rng_seed)
    return q.LearnedSketch(p, table_log2=table_log2, init_seed=init_seed)


def _keys(n, seed):
    rs = np.random.default_rng(seed)
    return rs.integers(0, 2 ** 63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
        rs.integers(0, 2, n, dtype=np.int64).astype(np.uint64)


@pytest.mark.parametrize("mfma", [False, True])
def test_forward_matches_oracle(gpu, mfma):
    _print_case(test="mlp_forward", mfma=mfma)
    L = _learned()
    p = mo.init_params(11, 0.3)
    p[mo.OFF_B1:mo.OFF_W2] = np.linspace(-0.2, 0.2, 64)
    p[mo.OFF_B2:mo.OFF_W3] = np.linspace(0.1, -0.1, 64)
    p[mo.OFF_B3] = 0.7
    L.set_params(p)
    keys = _keys(5000, 2)
    y = L.score(to_dev(keys), mfma=mfma).cpu().numpy()
    ref, _ = mo.forward(p, keys)
    # same BF16/FP32 network; only the FP32 summation order differs (and with it,
    # rarely, a BF16 rounding of an activation)
    err = np.abs(y - ref)
    assert np.median(err) < 1e-5 * (1 + np.abs(ref).max())
    assert err.max() < 2e-2 * (1 + np.abs(ref).max()), err.max()


def test_mfma_forward_equals_valu_forward(gpu):
    _print_case(test="mlp_mfma_vs_valu")
    L = _learned(init_seed=5)
    keys = to_dev(_keys(1 << 16, 3))
    a = L.score(keys, mfma=False).cpu().numpy()
    b = L.score(keys, mfma=True).cpu().numpy()
    assert np.abs(a - b).max() < 2e-2 * (1 + np.abs(a).max())
    assert np.mean(np.abs(a - b) < 1e-5 * (1 + np.abs(a).max())) > 0.95


def test_gradients_match_oracle(gpu):
    _print_case(test="mlp_grad")
    import torch
    L = _learned()
    p = mo.init_params(12, 0.3)
    p[mo.OFF_B1:mo.OFF_W2] = 0.05
    p[mo.OFF_B2:mo.OFF_W3] = 0.05
    L.set_params(p)
    keys = _keys(1000, 5)
    t = np.random.default_rng(6).normal(0, 2, keys.size).astype(np.float32)
    loss, g = L.grad(to_dev(keys), torch.from_numpy(t).cuda())
    rl, rg = mo.grad(p, keys, t)
    assert abs(loss - rl) <= 1e-4 * (1 + abs(rl))
    scale = np.abs(rg).max()
    assert np.abs(g - rg).max() <= 1e-2 * scale, np.abs(g - rg).max() / scale
    # per block of the layout, direction agrees
    for a, b in ((mo.OFF_W1, mo.OFF_B1), (mo.OFF_W2, mo.OFF_B2), (mo.OFF_W3, mo.OFF_B3)):
        c = np.dot(g[a:b], rg[a:b]) / (np.linalg.norm(g[a:b]) * np.linalg.norm(rg[a:b]) + 1e-30)
        assert c > 0.999, (a, c)
    # gradient of an unchanged network twice: same result (no state left behind)
    loss2, g2 = L.grad(to_dev(keys), torch.from_numpy(t).cuda())
    assert np.allclose(g, g2, rtol=1e-5, atol=1e-7 * scale)


def test_training_fits_a_learnable_target(gpu):
    _print_case(test="mlp_train")
    import torch
    L = _learned()
    keys = _keys(1 << 14, 7)
    t = ((keys >> np.uint64(59)).astype(np.float32))  # a function of the top 5 bits
    dk, dt = to_dev(keys), torch.from_numpy(t).cuda()
    losses = []
    for it in range(300):
        sl = slice((it * 1024) % keys.size, (it * 1024) % keys.size + 1024)
        losses.append(L.step(dk[sl], dt[sl], lr=3e-3))
    l_end, _ = L.grad(dk, dt)
    assert l_end < 0.1 * losses[0], (losses[0], l_end)
    # Adam step equals the oracle's update on the same gradient (one step, fresh state)
    L2 = _learned(init_seed=9)
    p0 = L2.get_params()
    _, g = L2.grad(dk[:512], dt[:512])
    L2.step(dk[:512], dt[:512], lr=1e-3)
    p1 = L2.get_params()
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p0 - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.allclose(p1, ref, rtol=1e-5, atol=1e-6)


def test_routed_insert_and_query(gpu):
    """Heavy-routed keys are counted exactly; the others reach the sketch in
    stream order, so the sketch equals the oracle fed only them."""
    import qsketch_amd as q
    _print_case(test="learned_insert")
    z = q.Zipf(1.1, 1 << 14, device=False)
    keys = z.keys_u64_noisy_host(5, 1 << 18, sigma=1.5)
    L = _learned(n_low=4093, n_high=8192, table_log2=13)
    # route by the top bits: keys whose upper field is small are the heavy ones
    p = np.zeros(mo.NPAR, np.float32)
    W1 = np.zeros((64, 64), np.float32)
    W1[0, 48:53] = -1.0 * (2.0 ** np.arange(5))  # unit 0 = -(field & 31)
    p[mo.OFF_W1:mo.OFF_B1] = W1.reshape(-1)
    p[mo.OFF_B1] = 40.0
    W2 = np.zeros((64, 64), np.float32)
    W2[0, 0] = 1.0
    p[mo.OFF_W2:mo.OFF_B2] = W2.reshape(-1)
    p[mo.OFF_W3] = 1.0
    L.set_params(p)
    L.threshold = 36.5  # field <= 3
    dk = to_dev(keys)
    scores = L.score(dk).cpu().numpy()
    heavy = scores >= 36.5
    assert 0.2 < heavy.mean() < 0.95
    nh = L.insert(dk)
    assert nh == int(heavy.sum())
    # table == exact counts of the heavy keys
    tk, tc = L.table()
    hk, hc = np.unique(keys[heavy], return_counts=True)
    order = np.argsort(tk)
    assert np.array_equal(tk[order], hk) and np.array_equal(tc[order], hc)
    # sketch == oracle fed the light keys in order
    o = po.OracleSketch(4093, 8192, 1021, key_bytes=8, rng_seed=3)
    assert o.insert(keys[~heavy]) == 0
    st = L.export_sketch_state()
    assert np.array_equal(st["low"], o.low()) and np.array_equal(st["high"], o.high())
    assert st["next"] == o.next
    # query: exact for heavy-routed keys (0 if never inserted), sketch otherwise
    qk = np.concatenate([np.unique(keys), keys[:10] ^ np.uint64(0x5555)])
    est = to_host_u32(L.search(to_dev(qk)))
    qh = L.score(to_dev(qk)).cpu().numpy() >= 36.5
    uk, uc = np.unique(keys, return_counts=True)
    pos = np.clip(np.searchsorted(uk, qk), 0, uk.size - 1)
    true = np.where(uk[pos] == qk, uc[pos], 0)
    assert np.array_equal(est[qh], true[qh])
    assert np.array_equal(est[~qh], o.query(qk[~qh]))
