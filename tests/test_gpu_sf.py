"""GPU parity of the SF sketch (include/qsketch_sf.h, device_sketch_v42.h):
seeds, masks and their draw-order index, the split table, the table of
largest splits and every estimate equal the oracle's (oracle/qs_oracle.c
orq_sf_*), over several insert calls; with one split per counter it is the
single-level sketch."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_dev, to_host_u32

pytestmark = pytest.mark.gpu


def _case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _same(g, o, what):
    st = g.export_state()
    assert st["seeds"] == o.seeds(), what
    assert np.array_equal(st["masks"], o.masks()), what
    assert np.array_equal(st["index"], o.index()), what
    assert np.array_equal(st["split"], o.split()), f"{what}: split table"
    assert np.array_equal(st["search"], o.search()), f"{what}: search table"


@pytest.mark.parametrize("m,times,key_bytes,seed_sz,alpha", [
    (8, 3, 4, 1, 1.1),    # the reference driver's choice (src/main.cu:484-485)
    (16, 8, 4, 2, 0.8),
    (32, 2, 4, 1, None),  # uniform keys
    (8, 5, 8, 1, 1.1),    # 64-bit keys
])
def test_sf_state_and_estimates(gpu, m, times, key_bytes, seed_sz, alpha):
    import qsketch_amd as q
    _case(test="sf", m=m, times=times, key_bytes=key_bytes, alpha=alpha)
    n = 3067
    rs = np.random.default_rng(m + times)
    if key_bytes == 4:
        keys = (q.Zipf(alpha, 1 << 14, device=False).keys_u32_host(3, 1 << 19) if alpha
                else q.uniform_u32_host(3, 1 << 19))
    else:
        keys = q.Zipf(alpha, 1 << 14, device=False).keys_u64_noisy_host(3, 1 << 19, sigma=1.5)
    g = q.Sketch_GPU_SF_Host_Sub_Warp(n, m, times, seed_sz, 1021, key_bytes=key_bytes, rng_seed=11)
    o = po.OracleSF(n, m, times, 1021, seed_sz=seed_sz, key_bytes=key_bytes, rng_seed=11)
    st = g.export_state()
    assert not st["split"].any() and not st["search"].any()
    cut = int(rs.integers(1000, keys.size - 1000))
    for part in (keys[:cut], keys[cut:]):  # two calls: the update runs after each
        g.insert(to_dev(part))
        o.insert(part)
        _same(g, o, f"after {part.size} keys")
    uq = np.unique(keys)
    qk = np.concatenate([uq, uq[:100] ^ np.asarray(0x55, dtype=uq.dtype)])
    assert np.array_equal(to_host_u32(g.search(to_dev(qk))), o.query(qk))
    g.clear()
    st = g.export_state()
    assert not st["split"].any() and not st["search"].any()


def test_sf_with_one_split_equals_single_level_sketch(gpu):
    """times = 1 on the GPU: the SF sketch's counters are the single-level
    sketch's (same seeds and masks)."""
    import qsketch_amd as q
    _case(test="sf_times1")
    keys = q.Zipf(1.0, 1 << 15, device=False).keys_u32_host(5, 1 << 20)
    d = to_dev(keys)
    sf = q.Sketch_GPU_SF_Host_Sub_Warp(12281, 16, 1, 2, 1021, rng_seed=4)
    sl = q.Sketch_GPU_Sub_Warp(12281, 16, 2, 1021, rng_seed=4)
    sf.insert(d)
    sl.insert(d)
    assert np.array_equal(sf.export_state()["search"], sl.export_state()["table"])
    uq = to_dev(np.unique(keys))
    assert np.array_equal(to_host_u32(sf.search(uq)), to_host_u32(sl.search(uq)))


def test_sf_large_skewed_stream(gpu):
    """2^26 keys Zipf 1.1 over 2^22, the sizing rule's buckets for m = 8 at
    WLF 0.5, three splits: every table word and every estimate."""
    import qsketch_amd as q
    _case(test="sf_large")
    N = 1 << 26
    keys = q.Zipf(1.1, 1 << 22, device=False).keys_u32_host(7, N)
    n = q.number_of_buckets(N, 8, 0.5)
    g = q.Sketch_GPU_SF_Host_Sub_Warp(n, 8, 3, 1, 1021, rng_seed=2)
    o = po.OracleSF(n, 8, 3, 1021, rng_seed=2)
    g.insert(to_dev(keys))
    o.insert(keys)
    _same(g, o, "2^26 keys")
    uq = np.unique(keys)
    est = to_host_u32(g.search(to_dev(uq)))
    assert np.array_equal(est, o.query(uq))
    u, c = np.unique(keys, return_counts=True)
    assert np.all(est >= c)


def test_sf_short_and_threshold_calls(gpu):
    """Calls on both sides of the length where the insert starts grouping the
    keys by bucket first (2^16), and a 3-key call, into one sketch: the
    unsorted and the grouped paths leave the oracle's state."""
    import qsketch_amd as q
    _case(test="sf_paths")
    keys = q.Zipf(1.2, 1 << 16, device=False).keys_u32_host(9, 3 + (1 << 16) - 1 + (1 << 16) + 70001)
    g = q.Sketch_GPU_SF_Host_Sub_Warp(40009, 8, 4, 1, 1021, rng_seed=5)
    o = po.OracleSF(40009, 8, 4, 1021, rng_seed=5)
    a = 0
    for ln in (3, (1 << 16) - 1, 1 << 16, 70001):
        part = keys[a:a + ln]
        a += ln
        g.insert(to_dev(part))
        o.insert(part)
        _same(g, o, f"call of {ln} keys")
    uq = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(uq))), o.query(uq))


@pytest.mark.parametrize("n", [7, 300, 70001, (1 << 24) + 43])
def test_sf_grouping_for_every_bucket_width(gpu, n):
    """Tables whose bucket index has 3, 9, 17 and 25 bits: one or two
    grouping passes, over all of the bits or the top 16."""
    import qsketch_amd as q
    _case(test="sf_width", n=n)
    keys = q.uniform_u32_host(13, 1 << 18)
    keys[::3] = keys[7]  # a hot key
    g = q.Sketch_GPU_SF_Host_Sub_Warp(n, 16, 3, 1, 1021, rng_seed=8)
    o = po.OracleSF(n, 16, 3, 1021, rng_seed=8)
    g.insert(to_dev(keys))
    o.insert(keys)
    _same(g, o, f"n {n}")
    uq = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(uq))), o.query(uq))


@pytest.mark.slow
def test_sf_call_longer_than_one_chunk(gpu):
    """2^28 + 54321 keys in one call: two radix chunks (2^28 keys each at
    most), into a small table, against the oracle."""
    import torch
    import qsketch_amd as q
    _case(test="sf_chunks", n=(1 << 28) + 54321)
    N = (1 << 28) + 54321
    d = q.Zipf(1.1, 1 << 24).keys_u32(3, N)
    g = q.Sketch_GPU_SF_Host_Sub_Warp(1 << 20, 8, 3, 1, 1021, rng_seed=3)
    g.insert(d)
    keys = d.cpu().numpy().view(np.uint32)
    del d
    torch.cuda.empty_cache()
    o = po.OracleSF(1 << 20, 8, 3, 1021, rng_seed=3)
    o.insert(keys)
    _same(g, o, "two chunks")
    uq = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(uq))), o.query(uq))
