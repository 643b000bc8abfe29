"""Parity at the full workload sizes: 2^28 u32 keys (Zipf 1.1 and 0.8 over
2^26, uniform) with the sizing rule's 201,326,557 buckets, and the overflow
stress case (2^27 keys, Zipf 1.5 over 2^24).  The oracle inserts the same
stream on the host's cores (bucket-range partitioned, byte-identical to the
sequential restatement); the whole exported state and the estimates of every
distinct key must be equal."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.mark.parametrize("alpha,log2n,log2u,n_high", [
    (1.1, 28, 26, 1 << 18),
    (0.8, 28, 26, 1 << 18),
    (0.0, 28, 26, 1 << 10),
    (1.5, 27, 24, 1 << 16),
])
def test_full_size_state_and_estimates(gpu, alpha, log2n, log2u, n_high):
    import torch
    import qsketch_amd as q
    print(f"CASE fullsize alpha={alpha} n=2^{log2n} u=2^{log2u}", flush=True)
    N = 1 << log2n
    if alpha > 0:
        keys = q.Zipf(alpha, 1 << log2u, device=False).keys_u32_host(1, N)
    else:
        keys = q.uniform_u32_host(1, N)
    n_low = q.number_of_buckets(N, 8, 0.5)
    g = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, n_high, 32, 1, 1021, rng_seed=1)
    d = torch.from_numpy(keys.view(np.int32)).cuda()
    g.insert(d)
    o = po.OracleSketch(n_low, n_high, 1021, rng_seed=1)
    assert o.insert(keys, nthreads=_threads()) == 0
    st = g.export_state()
    assert st["next"] == o.next
    assert np.array_equal(st["high"], o.high(copy=False))
    lo = st["low"]
    ol = o.low(copy=False)
    step = 1 << 26
    for a in range(0, lo.size, step):  # compare in slices to bound temporaries
        assert np.array_equal(lo[a:a + step], ol[a:a + step]), f"low table differs in words [{a}, {a + step})"
    del lo, st
    uq = torch.unique(d)
    est = g.search(uq).cpu().numpy().view(np.uint32)
    ref = o.query(uq.cpu().numpy().view(np.uint32), nthreads=_threads())
    assert np.array_equal(est, ref)
    print(f"  linked buckets {o.next - 1025}, distinct keys {uq.numel()}, "
          f"hot-filtered inserts {g.stats()['hot_keys']}", flush=True)
    if alpha >= 1.1:
        assert g.stats()["hot_keys"] > N // 2  # the skewed streams took the hot-key filter


@pytest.mark.parametrize("alpha,log2u", [(1.1, 22), (1.5, 20)])
def test_hot_filter_same_state_as_ordered_insert(gpu, alpha, log2u):
    """The hot-key filter (2^26 keys: the shortest call it takes) against the
    ordinary insert of the same keys and against the oracle; then a second
    call into the same sketch, whose hot buckets are linked from the start."""
    import torch
    import qsketch_amd as q
    print(f"CASE hot_filter alpha={alpha} u=2^{log2u}", flush=True)
    N = 1 << 26
    z = q.Zipf(alpha, 1 << log2u, device=False)
    k1, k2 = z.keys_u32_host(3, N), z.keys_u32_host(4, N)
    n_low = q.number_of_buckets(N, 8, 0.5)
    a = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 16, 32, rng_seed=2)
    b = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 16, 32, rng_seed=2)
    b.set_hot_filter(False)
    o = po.OracleSketch(n_low, 1 << 16, 1021, rng_seed=2)
    for k in (k1, k2):
        d = torch.from_numpy(k.view(np.int32)).cuda()
        a.insert(d)
        b.insert(d)
        assert o.insert(k, nthreads=_threads()) == 0
    assert a.stats()["hot_keys"] > N // 2, a.stats()["hot_keys"]
    assert b.stats()["hot_keys"] == 0
    assert a.stats()["keys_inserted"] == 2 * N
    sa, sb = a.export_state(), b.export_state()
    for f in ("low", "high", "masks", "seed", "next"):
        assert np.array_equal(sa[f], sb[f]), f
    assert sa["next"] == o.next
    assert np.array_equal(sa["high"], o.high(copy=False))
    assert np.array_equal(sa["low"], o.low(copy=False))
    uq = np.unique(np.concatenate([k1, k2]))
    est = a.search(torch.from_numpy(uq.view(np.int32)).cuda()).cpu().numpy().view(np.uint32)
    assert np.array_equal(est, o.query(uq, nthreads=_threads()))


def test_hot_filter_when_the_skew_changes_between_calls(gpu):
    """One sketch, 2^26-key calls: uniform, Zipf 1.1, uniform, uniform, Zipf
    1.1.  After a call without skew the next call finds its candidates before
    it inserts anything (and inserts a call without skew as one chunk); after
    any other call it finds them beside the prefix (and inserts the rest of a
    call without skew after the prefix).  The sequence takes each of the four
    cases; after every call the whole state equals the ordinary insert's."""
    import torch
    import qsketch_amd as q
    print("CASE hot_filter_skew_changes", flush=True)
    N = 1 << 26
    z = q.Zipf(1.1, 1 << 22, device=False).keys_u32_host(5, N)
    u = q.uniform_u32_host(6, N)
    n_low = q.number_of_buckets(N, 8, 0.5)
    a = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 18, 32, rng_seed=4)
    b = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 18, 32, rng_seed=4)
    b.set_hot_filter(False)
    dz = torch.from_numpy(z.view(np.int32)).cuda()
    du = torch.from_numpy(u.view(np.int32)).cuda()
    hot_before = 0
    for i, (name, d) in enumerate([("uniform", du), ("zipf", dz), ("uniform", du), ("uniform", du), ("zipf", dz)]):
        print(f"  call {i}: {name}", flush=True)
        a.insert(d)
        b.insert(d)
        sa, sb = a.export_state(), b.export_state()
        for f in ("low", "high", "next"):
            assert np.array_equal(sa[f], sb[f]), (i, name, f)
        hot = a.stats()["hot_keys"]
        if name == "zipf":
            assert hot - hot_before > N // 2, (i, hot - hot_before)  # the skewed calls took the filter
        else:
            assert hot == hot_before, i
        hot_before = hot
    assert a.stats()["keys_inserted"] == 5 * N
    assert b.stats()["hot_keys"] == 0


def test_autotune_of_a_filtered_insert_keeps_one_state(gpu):
    """The autotuner of a single sketch inserts a 2^26-key Zipf batch through
    the hot-key filter (as the sketch's own insert would) with each of the 16
    segment-pass shapes on a twin sketch: every shape must leave the same
    state (a checksum per shape; QS_ERR_STATE otherwise), and the sketch's
    own insert afterwards equals the unfiltered one."""
    import torch
    import qsketch_amd as q
    print("CASE autotune_filtered", flush=True)
    N = 1 << 26
    z = q.Zipf(1.1, 1 << 22, device=False).keys_u32_host(7, N)
    d = torch.from_numpy(z.view(np.int32)).cuda()
    n_low = q.number_of_buckets(N, 8, 0.5)
    a = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 18, 32, rng_seed=5)
    b = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 18, 32, rng_seed=5)
    b.set_hot_filter(False)
    r = a.autotune(d, reps=1)
    assert r["consistent"] and len(r["table"]) == 16
    assert a.stats()["keys_inserted"] == 0  # the twin took the inserts
    a.insert(d)
    b.insert(d)
    assert a.stats()["hot_keys"] > N // 2
    sa, sb = a.export_state(), b.export_state()
    for f in ("low", "high", "next"):
        assert np.array_equal(sa[f], sb[f]), f
