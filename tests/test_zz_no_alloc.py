"""The allocation-free insert (qs_insert_noalloc) against the insert it
varies (qs_insert_async): same whole state after each of two calls in a row,
the second smaller, and no device memory taken by the second call."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.helpers import assert_same_state, make_pair, to_dev, zipf_keys

pytestmark = pytest.mark.gpu


def _case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _same(a, b, what):
    sa, sb = a.export_state(), b.export_state()
    for k in ("low", "high", "masks"):
        assert np.array_equal(sa[k], sb[k]), f"{what}: {k} differs"
    assert sa["seed"] == sb["seed"] and sa["next"] == sb["next"], what


def test_noalloc_insert_equals_async_insert(gpu):
    import torch
    import qsketch_amd as q
    n_low, n_high = 12281, 4096
    a, o = make_pair(n_low, n_high=n_high, rng_seed=5)
    b = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, n_high, 32, rng_seed=5)
    k1 = zipf_keys(1.1, 1 << 14, q.NOALLOC_MAX_KEYS, seed=11)   # the largest batch
    k2 = zipf_keys(1.1, 1 << 14, 300_001, seed=12)              # then a smaller one
    d1, d2 = to_dev(k1), to_dev(k2)

    _case(test="noalloc", call=1, n=k1.size)
    a.insert_noalloc(d1)
    b.insert_async(d1, d1.numel(), 0)
    a.sync()
    b.sync()
    assert o.insert(k1) == 0
    _same(a, b, "after call 1")
    assert_same_state(a, o, "after call 1 (oracle)")
    assert a.stats()["linked_buckets"] > 100  # the allocation path ran

    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    _case(test="noalloc", call=2, n=k2.size)
    a.insert_noalloc(d2)
    a.sync()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 >= free0, f"the allocation-free insert took {free0 - free1} bytes of device memory"
    b.insert_async(d2, d2.numel(), 0)
    b.sync()
    assert o.insert(k2) == 0
    _same(a, b, "after call 2")
    assert_same_state(a, o, "after call 2 (oracle)")

    uq = np.unique(np.concatenate([k1, k2]))
    ea = a.search(to_dev(uq)).cpu().numpy().view(np.uint32)
    eb = b.search(to_dev(uq)).cpu().numpy().view(np.uint32)
    assert np.array_equal(ea, eb)
    assert np.array_equal(ea, o.query(uq))


def test_noalloc_limits_rejected_on_host(gpu):
    """Sizes beyond the static array's limits are refused before any GPU work."""
    import qsketch_amd as q
    lib = q.lib()
    keys = to_dev(np.arange(q.NOALLOC_MAX_KEYS + 1, dtype=np.uint32))
    s = q.Sketch_GPU_Sub_Warp_Level(12281, 8, 4096, 32, rng_seed=5)
    _case(test="noalloc_limits")
    rc = lib.qs_insert_noalloc(s.handle, keys.data_ptr(), q.NOALLOC_MAX_KEYS + 1, None)
    assert rc == q.QS_ERR_PARAM
    big_pool = q.Sketch_GPU_Sub_Warp_Level(12281, 8, q.NOALLOC_MAX_ALLOCS + 2, 32, rng_seed=5)
    rc = lib.qs_insert_noalloc(big_pool.handle, keys.data_ptr(), q.NOALLOC_MAX_ALLOCS + 1, None)
    assert rc == q.QS_ERR_PARAM
    # within the limits with that pool: a batch no larger than the allocation cap
    rc = lib.qs_insert_noalloc(big_pool.handle, keys.data_ptr(), q.NOALLOC_MAX_ALLOCS, None)
    assert rc == q.QS_OK
    big_pool.sync()
    assert big_pool.stats()["keys_inserted"] == q.NOALLOC_MAX_ALLOCS
