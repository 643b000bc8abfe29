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


def _exact_params(seed):
    """A network whose every activation and score is exactly representable:
    W1 and b1 on the grid 1/32 with |a1| < 2, two entries of +-1 per row of
    W2, b2, w3 and b3 on the grid 1/32.  Every partial sum of a layer is then a
    multiple of 2^-10 below 2^8, exact in FP32 in any order, and every BF16
    rounding of a weight or activation is the identity.  Any correct forward,
    whatever its summation order, gives the oracle's scores bit for bit."""
    rs = np.random.default_rng(seed)
    W1 = np.zeros((64, 64))
    W2 = np.zeros((64, 64))
    for u in range(64):
        W1[u, rs.choice(64, 8, replace=False)] = rs.integers(-3, 4, 8) / 32
        W2[u, rs.choice(64, 2, replace=False)] = rs.choice([-1.0, 1.0], 2)
    p = np.zeros(mo.NPAR, np.float64)
    p[mo.OFF_W1:mo.OFF_B1] = W1.reshape(-1)
    p[mo.OFF_B1:mo.OFF_W2] = rs.integers(-16, 17, 64) / 32
    p[mo.OFF_W2:mo.OFF_B2] = W2.reshape(-1)
    p[mo.OFF_B2:mo.OFF_W3] = rs.integers(-16, 17, 64) / 32
    p[mo.OFF_W3:mo.OFF_B3] = rs.integers(-32, 33, 64) / 32
    p[mo.OFF_B3] = rs.integers(-32, 33) / 32
    p32 = p.astype(np.float32)
    assert np.array_equal(mo.bf16_round(p32), p32)
    return p32


@pytest.mark.parametrize("mfma", [False, True])
def test_forward_bit_exact_on_an_exact_network(gpu, mfma):
    """Every score equals the oracle's bit for bit when no sum or rounding
    depends on the order (the network of _exact_params): a wrong weight index,
    a transposed layer, a dropped bias or ReLU changes some score."""
    _print_case(test="mlp_forward_exact", mfma=mfma)
    L = _learned()
    p = _exact_params(21)
    L.set_params(p)
    keys = _keys(20000, 2)
    y = L.score(to_dev(keys), mfma=mfma).cpu().numpy()
    ref32, (_, a1, _, a2, _, _, _) = mo.forward(p, keys)
    ref64, _ = mo.forward(p, keys, exact=True)
    assert np.array_equal(ref32, ref64.astype(np.float32))  # the network is exact
    assert (a1 > 0).mean() > 0.2 and (a2 > 0).mean() > 0.2  # both ReLUs matter
    bad = np.nonzero(y != ref32)[0]
    assert bad.size == 0, (bad.size, bad[:5], y[bad[:5]], ref32[bad[:5]])


@pytest.mark.parametrize("mfma", [False, True])
def test_forward_matches_oracle(gpu, mfma):
    """Random weights: the same BF16/FP32 network, but FP32 sums in another
    order can move an activation across a BF16 rounding boundary (about one
    key in a hundred here).  All other keys agree to the output layer's own
    summation error; a flipped rounding moves a score by a few BF16 steps."""
    _print_case(test="mlp_forward", mfma=mfma)
    L = _learned()
    p = mo.init_params(11, 0.3)
    p[mo.OFF_B1:mo.OFF_W2] = np.linspace(-0.2, 0.2, 64)
    p[mo.OFF_B2:mo.OFF_W3] = np.linspace(0.1, -0.1, 64)
    p[mo.OFF_B3] = 0.7
    L.set_params(p)
    keys = _keys(5000, 2)
    y = L.score(to_dev(keys), mfma=mfma).cpu().numpy().astype(np.float64)
    ref, (_, _, _, _, h2, _, w3r) = mo.forward(p, keys)
    ey = 70 * 2.0 ** -24 * (np.abs(h2.astype(np.float64)) @ np.abs(w3r.astype(np.float64)) + abs(p[mo.OFF_B3]))
    err = np.abs(y - ref)
    off = err > ey
    assert off.mean() < 0.05, off.mean()
    assert err.max() < 2e-2 * (1 + np.abs(ref).max()), err.max()


def test_mfma_forward_equals_valu_forward(gpu):
    _print_case(test="mlp_mfma_vs_valu")
    L = _learned(init_seed=5)
    keys = to_dev(_keys(1 << 16, 3))
    a = L.score(keys, mfma=False).cpu().numpy()
    b = L.score(keys, mfma=True).cpu().numpy()
    assert np.abs(a - b).max() < 2e-2 * (1 + np.abs(a).max())
    assert np.mean(np.abs(a - b) < 1e-5 * (1 + np.abs(a).max())) > 0.98


def _grad_abs_sums(p, keys, t):
    """Per parameter, the sum of the absolute values of the terms its gradient
    adds up (the scale of its rounding error)."""
    y, (X, a1, h1, a2, h2, W2r, w3r) = mo.forward(p, keys, exact=True)
    n = y.size
    ady = np.abs(y - t) / n
    s = np.zeros(mo.NPAR)
    s[mo.OFF_W3:mo.OFF_B3] = np.abs(h2).T @ ady
    s[mo.OFF_B3] = ady.sum()
    adh2 = (ady[:, None] * np.abs(w3r)[None, :]) * (a2 > 0)
    s[mo.OFF_W2:mo.OFF_B2] = (adh2.T @ np.abs(h1)).reshape(-1)
    s[mo.OFF_B2:mo.OFF_W3] = adh2.sum(0)
    adh1 = (adh2 @ np.abs(W2r)) * (a1 > 0)
    s[mo.OFF_W1:mo.OFF_B1] = (adh1.T @ X).reshape(-1)
    s[mo.OFF_B1:mo.OFF_W2] = adh1.sum(0)
    return s


def test_gradients_match_oracle(gpu):
    """On the exact network the forward is exact, so every gradient element
    must equal the float64 oracle within the rounding error of its own FP32
    sum: (2n + 256) u times the sum of the magnitudes of its terms."""
    _print_case(test="mlp_grad")
    import torch
    L = _learned()
    p = _exact_params(12)
    L.set_params(p)
    n = 1024
    keys = _keys(n, 5)
    t = (np.random.default_rng(6).integers(-96, 97, n) / 32).astype(np.float32)
    loss, g = L.grad(to_dev(keys), torch.from_numpy(t).cuda())
    rl, rg = mo.grad(p, keys, t, exact=True)
    absum = _grad_abs_sums(p, keys, t.astype(np.float64))
    u = 2.0 ** -24
    bound = (2 * n + 256) * u * absum
    err = np.abs(g.astype(np.float64) - rg)
    worst = int(np.argmax(err - bound))
    assert np.all(err <= bound), (worst, g[worst], rg[worst], bound[worst])
    assert (np.abs(rg) > 0).mean() > 0.3  # the check is not vacuous
    y, _ = mo.forward(p, keys, exact=True)
    assert abs(loss - rl) <= (n + 8) * u * 0.5 * np.mean((y - t) ** 2) + 1e-30
    # gradient of an unchanged network twice: same result (no state left behind)
    loss2, g2 = L.grad(to_dev(keys), torch.from_numpy(t).cuda())
    assert np.all(np.abs(g2.astype(np.float64) - rg) <= bound)


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
    # query (the table-first kernel: the table is far from full): the exact
    # count of a key held in the table, the sketch estimate of any other key,
    # never-inserted heavy-scored keys included
    qk = np.concatenate([np.unique(keys), keys[:10] ^ np.uint64(0x5555)])
    est = to_host_u32(L.search(to_dev(qk)))
    in_tab = np.isin(qk, tk)
    assert in_tab.sum() == hk.size
    assert np.array_equal(est[in_tab], hc[np.searchsorted(hk, qk[in_tab])])
    assert np.array_equal(est[~in_tab], o.query(qk[~in_tab]))
    never = ~np.isin(qk, keys)
    assert never.sum() > 0 and (L.score(to_dev(qk[never])).cpu().numpy() >= 36.5).any()


def _field_router(L, field_max):
    """Hand-set network: score = 40 - (bits 48..52 of the key) exactly (the
    noisy log-rank field of the cfg5 keys); threshold admits field <= field_max."""
    p = np.zeros(mo.NPAR, np.float32)
    W1 = np.zeros((64, 64), np.float32)
    W1[0, 48:53] = -1.0 * (2.0 ** np.arange(5))
    p[mo.OFF_W1:mo.OFF_B1] = W1.reshape(-1)
    p[mo.OFF_B1] = 40.0
    W2 = np.zeros((64, 64), np.float32)
    W2[0, 0] = 1.0
    p[mo.OFF_W2:mo.OFF_B2] = W2.reshape(-1)
    p[mo.OFF_W3] = 1.0
    L.set_params(p)
    thr = 40.0 - field_max - 0.5
    L.threshold = thr
    return p, thr


def test_network_and_threshold_fixed_while_holding_data(gpu):
    """Changing the routing with keys in would let a query look in the other
    part than the inserts went to: refused until clear."""
    import torch
    import qsketch_amd as q
    _print_case(test="learned_state_guard")
    z = q.Zipf(1.1, 1 << 12, device=False)
    keys = z.keys_u64_noisy_host(3, 1 << 16, sigma=1.5)
    L = _learned(n_low=4093, n_high=8192, table_log2=12,
                 key_bytes=8)
    p, thr = _field_router(L, 3)
    dk = to_dev(keys)
    t = torch.zeros(1024, dtype=torch.float32, device="cuda")
    L.insert(dk)
    with pytest.raises(q.QSketchError) as e:
        L.set_params(p)
    assert e.value.code == q.QS_ERR_STATE
    with pytest.raises(q.QSketchError) as e:
        L.step(dk[:1024], t)
    assert e.value.code == q.QS_ERR_STATE
    with pytest.raises(q.QSketchError) as e:
        L.threshold = thr + 1.0
    assert e.value.code == q.QS_ERR_STATE
    L.threshold = thr            # the same value is no change
    L.grad(dk[:1024], t)         # gradients leave the network as it is
    L.clear()
    L.set_params(p)
    L.threshold = thr + 1.0


def test_no_underestimate_across_inserts_with_refused_retraining(gpu):
    """Two inserts with a (refused) retraining between them: under the paper
    semantics no estimate falls below the exact count."""
    import torch
    import qsketch_amd as q
    _print_case(test="learned_two_inserts")
    z = q.Zipf(1.1, 1 << 14, device=False)
    k1 = z.keys_u64_noisy_host(5, 1 << 18, sigma=1.5)
    k2 = z.keys_u64_noisy_host(6, 1 << 18, sigma=1.5)
    pp = q.make_params(4093, 8192, key_bytes=8, rng_seed=3, semantics=q.QS_PAPER)
    L = q.LearnedSketch(pp, table_log2=13)
    p, thr = _field_router(L, 4)
    L.insert(to_dev(k1))
    p2 = p.copy()
    p2[mo.OFF_B1] = 30.0  # a retrained network would route other keys
    with pytest.raises(q.QSketchError):
        L.set_params(p2)
    L.insert(to_dev(k2))
    allk = np.concatenate([k1, k2])
    uk, uc = np.unique(allk, return_counts=True)
    est = to_host_u32(L.search(to_dev(uk))).astype(np.int64)
    assert np.all(est >= uc), int((est < uc).sum())


def test_full_table_sends_keys_to_the_sketch(gpu):
    """A table too small for the heavy keys: the keys it has no slot for are
    counted in the sketch, all their occurrences in stream order, so every key
    is counted entirely in one part and nothing is lost."""
    import qsketch_amd as q
    _print_case(test="learned_table_full")
    z = q.Zipf(1.1, 1 << 16, device=False)
    keys = z.keys_u64_noisy_host(9, 1 << 19, sigma=1.5)
    L = _learned(n_low=65537, n_high=1 << 14, table_log2=10)
    _, thr = _field_router(L, 11)
    heavy = (40.0 - ((keys >> np.uint64(48)) & np.uint64(31)).astype(np.float64)) >= thr
    assert np.unique(keys[heavy]).size > 3 * 1024  # far more heavy keys than slots
    dk = to_dev(keys)
    nh = L.insert(dk)
    assert L.table_full
    tk, tc = L.table()
    assert tk.size > 1000  # the table filled up
    in_table = heavy & np.isin(keys, tk)
    assert nh == int(in_table.sum())
    hk, hc = np.unique(keys[in_table], return_counts=True)
    order = np.argsort(tk)
    assert np.array_equal(tk[order], hk) and np.array_equal(tc[order], hc)
    o = po.OracleSketch(65537, 1 << 14, 1021, key_bytes=8, rng_seed=3)
    assert o.insert(keys[~in_table]) == 0
    st = L.export_sketch_state()
    assert np.array_equal(st["low"], o.low()) and np.array_equal(st["high"], o.high())
    assert st["next"] == o.next
    uk = np.unique(keys)
    est = to_host_u32(L.search(to_dev(uk)))
    tab = np.isin(uk, tk)
    pos = np.searchsorted(hk, uk[tab])
    assert np.array_equal(est[tab], hc[pos])
    assert np.array_equal(est[~tab], o.query(uk[~tab]))
    L.clear()
    assert not L.table_full


def test_learned_form_lowers_are_at_equal_memory(gpu):
    """What the trained network gains: the learned form, with its sketch
    shrunk by the table's and the network's bytes, against the plain sketch
    on the same stream (2^22 u64 keys, Zipf 1.1 over 2^18, upper 16 bits the
    noisy log-rank field; the network trained on the plain sketch's estimates
    of the first 2^20 items).  ARE over every distinct key against exact
    counts (the oracle's sort + count) must drop by at least 30%; the table
    holds exact counts and the sketch equals the oracle fed the other keys."""
    import torch
    import qsketch_amd as q
    from qsketch_amd import accuracy as acc
    from qsketch_amd import learned as lf
    _print_case(test="learned_gain", n=1 << 22)
    N, U, W, TOP, TL = 1 << 22, 1 << 18, 1 << 20, 1 << 10, 13
    keys = q.Zipf(1.1, U).keys_u64_noisy(seed=1, n=N, sigma=1.5)
    n_high = 1 << 14
    n_low = q.number_of_buckets(N, 8, 0.5)
    wk, tgt = lf.window_targets(keys[:W], n_low, n_high)
    L = q.LearnedSketch(q.make_params(lf.equal_memory_n_low(n_low, n_high, TL), n_high, key_bytes=8,
                                      rng_seed=1), table_log2=TL, init_seed=7)
    losses = lf.train_router(L, wk, tgt, steps=1500)
    assert losses[-1] < 0.5 * losses[0], losses
    L.threshold = lf.pick_threshold(L, wk, TOP)
    plain = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, n_high, 32, key_bytes=8, rng_seed=1)
    assert L.memory()["total"] <= plain.memory_bytes()
    plain.insert(keys)
    L.insert(keys)
    hk = keys.cpu().numpy().view(np.uint64)
    uk, uc = po.exact_counts(hk, key_bytes=8)
    # the GPU harness agrees with the oracle's exact counts
    gk, gc = acc.exact_counts(keys)
    assert np.array_equal(np.sort(gk.cpu().numpy().view(np.uint64)), uk)
    du = to_dev(uk)
    true = torch.from_numpy(uc.astype(np.int64))
    mp = acc.metrics(plain.search(du).cpu(), true)
    ml = acc.metrics(L.search(du).cpu(), true)
    assert ml["are"] <= 0.7 * mp["are"], (mp, ml)
    # the parts: exact table, sketch = oracle of the light keys in stream order
    tk, tc = L.table()
    assert not L.table_full and tk.size > TOP // 2
    heavy = np.isin(hk, tk)
    pos = np.searchsorted(uk, tk)
    assert np.array_equal(tc, uc[pos])
    o = po.OracleSketch(L.params.n_low, n_high, 1021, key_bytes=8, rng_seed=1)
    assert o.insert(hk[~heavy], nthreads=8) == 0
    st = L.export_sketch_state()
    assert np.array_equal(st["low"], o.low()) and np.array_equal(st["high"], o.high())
