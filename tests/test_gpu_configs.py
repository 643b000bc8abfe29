"""The remaining measured workloads on the GPU against the oracle, at their
full sizes, and the device key generators against their host twins:

  cfg1  2^22 u32 keys, Zipf 1.0 over 2^18, seed 1, sizing rule at WLF 0.5 and
        1.0 (src/main.cu:113, :390-391; lib/default.cu:39-54): whole state and
        every estimate; with the reference's own 1024-bucket high pool the
        insert runs out of high buckets
  cfg4  4096 sketches in one captured launch, every sketch's state and every
        distinct key's estimate against a per-sketch oracle
  cfg5  the learned form at 2^28 u64 keys: exact table, sketch = oracle fed
        the light keys, and at least 30% lower ARE than the plain sketch at
        equal memory
"""
from __future__ import annotations

import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import assert_same_state, make_pair, to_dev, to_host_u32

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


# ---------------------------------------------------------------------------
# generators: the timed streams (device) equal the parity streams (host)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("alpha,log2u", [(1.1, 26), (0.8, 26), (0.0, 26), (1.0, 18), (1.5, 24)])
def test_device_u32_generators_equal_host(gpu, alpha, log2u):
    import torch
    import qsketch_amd as q
    from qsketch_amd import workloads as wl
    _case(test="gen_u32", alpha=alpha, log2u=log2u)
    n = 1 << 24
    dev = wl.plain_keys(alpha, log2u, 28, seed=1, device=True, n=n).cpu().numpy().view(np.uint32)
    host = wl.plain_keys(alpha, log2u, 28, seed=1, device=False, n=n)
    assert np.array_equal(dev, host)
    # a window further into the stream, by offset
    off = (1 << 28) - 5000
    if alpha > 0:
        z = q.Zipf(alpha, 1 << log2u)
        d2 = z.keys_u32(1, 3000, offset=off).cpu().numpy().view(np.uint32)
        h2 = z.keys_u32_host(1, 3000, offset=off)
    else:
        d2 = q.uniform_u32(1, 3000, offset=off).cpu().numpy().view(np.uint32)
        h2 = q.uniform_u32_host(1, 3000, offset=off)
    assert np.array_equal(d2, h2)
    assert np.unique(host).size > 1000
    torch.cuda.synchronize()


def test_device_u64_noisy_generator_equals_host(gpu):
    from qsketch_amd import workloads as wl
    _case(test="gen_u64")
    n = 1 << 24
    dev = wl.learned_keys(device=True, n=n).cpu().numpy().view(np.uint64)
    host = wl.learned_keys(device=False, n=n)
    assert np.array_equal(dev, host)


def test_device_multi_streams_equal_host(gpu):
    """Every 37th stream of the 4096 (all lengths and exponents occur)."""
    import qsketch_amd as q
    from qsketch_amd import workloads as wl
    _case(test="gen_multi")
    keys, offs, lengths, alphas, seeds = wl.multi_keys(4096)
    hk = keys.cpu().numpy().view(np.uint32)
    for i in range(0, 4096, 37):
        a, b = int(offs[i]), int(offs[i + 1])
        h = q.Zipf(float(alphas[i]), 1 << 16, device=False).keys_u32_host(int(seeds[i]), b - a)
        assert np.array_equal(hk[a:b], h), i


# ---------------------------------------------------------------------------
# cfg1
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("wlf,n_low", [(0.5, 3145721), (1.0, 1572853)])
def test_cfg1_state_and_estimates(gpu, wlf, n_low):
    import qsketch_amd as q
    from qsketch_amd import workloads as wl
    _case(test="cfg1", wlf=wlf)
    keys = wl.plain_keys(**wl.CFG1, device=False)
    assert q.number_of_buckets(keys.size, 8, wlf) == n_low
    g, o = make_pair(n_low, n_high=1 << 13, rng_seed=1)
    g.insert(to_dev(keys))
    assert o.insert(keys, nthreads=_threads()) == 0
    assert_same_state(g, o, f"cfg1 wlf {wlf}")
    uq = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(uq))), o.query(uq))
    linked = o.next - 1025
    print(f"  linked buckets {linked}, distinct keys {uq.size}", flush=True)
    assert linked > 1023  # more than the reference's pool holds


def test_cfg1_reference_pool_runs_out(gpu):
    """The reference's n_high = 1024 (1,023 usable high buckets,
    src/main.cu:390-391) cannot hold cfg1's links: the oracle and the library
    both report the exhausted pool (the reference writes past its table,
    v3:119)."""
    import qsketch_amd as q
    from qsketch_amd import workloads as wl
    _case(test="cfg1_pool")
    keys = wl.plain_keys(**wl.CFG1, device=False)
    g, o = make_pair(3145721, n_high=1024, rng_seed=1)
    assert o.insert(keys) == -3  # ORQ_ERR_POOL
    with pytest.raises(q.QSketchError) as e:
        g.insert(to_dev(keys))
    assert e.value.code == q.QS_ERR_POOL
    assert g.stats()["pool_error"] == 1


# ---------------------------------------------------------------------------
# cfg4
# ---------------------------------------------------------------------------
def _oracle_sketch_check(p, keys, st, uq, est):
    """Oracle of one sketch of the set; returns a message on a difference."""
    o = po.OracleSketch(int(p.n_low), int(p.n_high), int(p.n_masks), rng_seed=int(p.rng_seed))
    rc = o.insert(keys)
    if rc != 0:
        return f"oracle insert {rc}"
    if st["seed"] != o.seed or st["next"] != o.next:
        return f"seed/next {st['seed']}/{st['next']} vs {o.seed}/{o.next}"
    if not np.array_equal(st["masks"], o.masks()):
        return "masks"
    if not np.array_equal(st["low"], o.low(copy=False)):
        return "low table"
    if not np.array_equal(st["high"], o.high(copy=False)):
        return "high table"
    if not np.array_equal(est, o.query(uq)):
        return "estimates"
    return None


def test_cfg4_4096_sketches_against_per_sketch_oracles(gpu):
    import torch
    import qsketch_amd as q
    from qsketch_amd import workloads as wl
    _case(test="cfg4", sketches=4096)
    K = 4096
    keys, offs, lengths, alphas, seeds = wl.multi_keys(K)
    params = wl.multi_params(lengths, seeds)
    S = q.SketchSet(params)
    d_offs = torch.from_numpy(offs.view(np.int64).copy()).cuda()
    g = S.capture_insert(keys, offs, d_offs)  # the measured path: one captured launch
    g.launch(0)
    S.sync()
    # distinct keys of every stream: one sort of (sketch, key)
    sid = torch.repeat_interleave(torch.arange(K, device="cuda", dtype=torch.int64),
                                  torch.from_numpy(lengths).cuda())
    comp = torch.unique((sid << 32) | (keys.to(torch.int64) & 0xFFFFFFFF))
    del sid
    qoffs = np.zeros(K + 1, np.uint64)
    qoffs[1:] = np.cumsum(torch.bincount(comp >> 32, minlength=K).cpu().numpy())
    qkeys = (comp & 0xFFFFFFFF).to(torch.int32)
    del comp
    est = S.search(qkeys, qoffs).cpu().numpy().view(np.uint32)
    hq = qkeys.cpu().numpy().view(np.uint32)
    hk = keys.cpu().numpy().view(np.uint32)
    print(f"  keys {hk.size}, distinct (sketch, key) pairs {hq.size}, linked "
          f"{S.stats()['linked_buckets']}", flush=True)
    bad = []
    with cf.ThreadPoolExecutor(_threads()) as ex:
        pending = {}
        for i in range(K):
            a, b = int(offs[i]), int(offs[i + 1])
            qa, qb = int(qoffs[i]), int(qoffs[i + 1])
            st = S.export_state(i)  # the library is called from this thread only
            pending[ex.submit(_oracle_sketch_check, params[i], hk[a:b], st, hq[qa:qb], est[qa:qb])] = i
            if len(pending) >= 2 * _threads():  # bound the states in flight
                done, _ = cf.wait(pending, return_when=cf.FIRST_COMPLETED)
                for f in done:
                    if f.result():
                        bad.append((pending[f], f.result()))
                    del pending[f]
        for f in cf.as_completed(pending):
            if f.result():
                bad.append((pending[f], f.result()))
    g.destroy()
    assert not bad, bad[:10]


# ---------------------------------------------------------------------------
# cfg5
# ---------------------------------------------------------------------------
def test_cfg5_learned_form_full_size(gpu):
    import torch
    import qsketch_amd as q
    from qsketch_amd import accuracy as acc
    from qsketch_amd import learned as lf
    from qsketch_amd import workloads as wl
    _case(test="cfg5", n=1 << 28)
    N, W, TOP, TL = 1 << 28, 1 << 26, 1 << 16, 19
    n_high = 1 << 18
    keys = wl.learned_keys(**wl.CFG5)
    n_low = q.number_of_buckets(N, 8, 0.5)
    wk, tgt = lf.window_targets(keys[:W], n_low, n_high)
    L = q.LearnedSketch(q.make_params(lf.equal_memory_n_low(n_low, n_high, TL), n_high, key_bytes=8,
                                      rng_seed=1), table_log2=TL, init_seed=7)
    losses = lf.train_router(L, wk, tgt, steps=3000)
    L.threshold = lf.pick_threshold(L, wk, TOP)
    del wk, tgt
    plain = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, n_high, 32, key_bytes=8, rng_seed=1)
    assert L.memory()["total"] <= plain.memory_bytes()
    plain.insert(keys)
    n_heavy = L.insert(keys)
    hk = keys.cpu().numpy().view(np.uint64)
    del keys
    uk, uc = po.exact_counts(hk, key_bytes=8)
    du = to_dev(uk)
    est_p = plain.search(du).cpu()
    est_l = L.search(du).cpu()
    true = torch.from_numpy(uc.astype(np.int64))
    mp, ml = acc.metrics(est_p, true), acc.metrics(est_l, true)
    print(f"  distinct {uk.size}, heavy share {n_heavy / N:.3f}, ARE plain {mp['are']:.3e} "
          f"learned {ml['are']:.3e}, loss {losses[0]:.3f} -> {losses[-1]:.3f}", flush=True)
    assert ml["are"] <= 0.7 * mp["are"], (mp, ml)
    # exact table
    tk, tc = L.table()
    assert not L.table_full
    heavy = np.isin(hk, tk)
    assert int(heavy.sum()) == n_heavy
    assert np.array_equal(tc, uc[np.searchsorted(uk, tk)])
    # the sketch part equals the oracle fed the light keys in stream order
    o = po.OracleSketch(int(L.params.n_low), n_high, 1021, key_bytes=8, rng_seed=1)
    assert o.insert(hk[~heavy], nthreads=_threads()) == 0
    st = L.export_sketch_state()
    assert st["next"] == o.next and np.array_equal(st["high"], o.high(copy=False))
    assert np.array_equal(st["low"], o.low(copy=False))
    del st
    light_q = ~np.isin(uk, tk)
    assert np.array_equal(est_l.numpy().view(np.uint32)[light_q], o.query(uk[light_q], nthreads=_threads()))
    del o
    # and the plain sketch equals its oracle on every distinct key
    op = po.OracleSketch(n_low, n_high, 1021, key_bytes=8, rng_seed=1)
    assert op.insert(hk, nthreads=_threads()) == 0
    assert np.array_equal(est_p.numpy().view(np.uint32), op.query(uk, nthreads=_threads()))
