"""GPU parity: every state word and every estimate of the HIP path equals the
sequential oracle (the restated v3 semantics) on the same seeded inputs."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import (assert_same_state, make_pair, to_dev, to_host_u32, uniform_keys,
                           zipf_keys)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _print_case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("rng_seed", [1, 2, 12345])
def test_create_seed_and_masks_match_oracle_and_libstdcxx(gpu, rng_seed):
    _print_case(test="create", rng_seed=rng_seed)
    g, o = make_pair(1021, rng_seed=rng_seed)
    st = g.export_state()
    assert st["seed"] == o.seed
    assert np.array_equal(st["masks"], o.masks())
    gold = json.load(open(os.path.join(GOLD, "stdlib_rng.json")))
    rec = [c for c in gold["create"] if c["rng_seed"] == rng_seed]
    if rec:
        assert st["seed"] == rec[0]["seeds"][0]
        assert list(map(int, st["masks"])) == rec[0]["masks"]
    assert st["next"] == 1025
    assert not st["low"].any() and not st["high"].any()


def test_query_on_oracle_built_state(gpu):
    """Query kernel alone: import the oracle's state (with linked buckets)."""
    _print_case(test="query_import")
    keys = zipf_keys(1.1, 1 << 14, 1 << 18, seed=7)
    g, o = make_pair(3067, n_high=4096)
    assert o.insert(keys) == 0
    assert o.next > 1025 + 10  # some buckets linked
    g.import_state(o.low(), o.high(), o.masks(), o.seed, o.next)
    q = np.concatenate([np.unique(keys), uniform_keys(1 << 14, seed=3)])
    est = to_host_u32(g.search(to_dev(q)))
    assert np.array_equal(est, o.query(q))


@pytest.mark.parametrize("semantics", [0, 1])
@pytest.mark.parametrize("n_low,alpha,universe,n", [
    (1572853, None, None, 1 << 20),   # uniform, nothing links
    (12281, 1.1, 1 << 14, 1 << 20),   # skewed, many links
    (1021, 0.8, 1 << 12, 1 << 18),    # small table: most buckets link
    (13, 1.0, 64, 5000),              # tiny table
])
def test_insert_state_bit_exact(gpu, semantics, n_low, alpha, universe, n):
    _print_case(test="insert_state", semantics=semantics, n_low=n_low, alpha=alpha, n=n)
    keys = uniform_keys(n, seed=11) if alpha is None else zipf_keys(alpha, universe, n, seed=11)
    g, o = make_pair(n_low, n_high=1 << 14, semantics=semantics)
    assert o.insert(keys) == 0
    g.insert(to_dev(keys))
    assert_same_state(g, o, "insert")
    q = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(q))), o.query(q))


@pytest.mark.parametrize("semantics", [0, 1])
def test_quotient_item_form(gpu, semantics):
    """4093 masks need a 12-bit mask field; with 2,097,169 buckets the hash
    quotient h / n_low needs 11, so the insert items carry the quotient and the
    mask index is recomputed from it (the form large tables use).  Several
    inserts, chunked, with buckets linked across them."""
    _print_case(test="quotient_form", semantics=semantics)
    keys = zipf_keys(1.1, 1 << 16, 3 << 19, seed=21)
    g, o = make_pair(2097169, n_high=1 << 14, n_masks=4093, semantics=semantics, chunk_keys=300000)
    for part in np.array_split(keys, 3):
        assert o.insert(part) == 0
        g.insert(to_dev(part))
    assert o.next > 1025 + 100
    assert_same_state(g, o, "quotient form")
    q = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(q))), o.query(q))


def test_chunking_does_not_change_state(gpu):
    _print_case(test="chunking")
    keys = zipf_keys(1.2, 1 << 13, 300001, seed=5)
    states = []
    for chunk in (1000, 65536, 0):
        g, o = make_pair(4093, n_high=8192, chunk_keys=chunk)
        for part in np.array_split(keys, 7):
            g.insert(to_dev(part))
        states.append(g.export_state())
    o.insert(keys)
    for st in states:
        assert np.array_equal(st["low"], o.low())
        assert np.array_equal(st["high"], o.high())
        assert st["next"] == o.next


def test_u64_keys(gpu):
    _print_case(test="u64")
    import qsketch_amd as q
    z = q.Zipf(1.1, 1 << 14, device=False)
    keys = z.keys_u64_noisy_host(9, 1 << 18)
    g, o = make_pair(3067, n_high=8192, key_bytes=8)
    o.insert(keys)
    g.insert(to_dev(keys))
    assert_same_state(g, o, "u64")
    qk = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(qk))), o.query(qk))


def test_pool_exhaustion_is_an_error(gpu):
    _print_case(test="pool")
    import qsketch_amd as q
    keys = zipf_keys(1.0, 1 << 12, 1 << 17, seed=2)
    g = q.Sketch_GPU_Sub_Warp_Level(1021, 8, 16, 32)
    with pytest.raises(q.QSketchError) as ei:
        g.insert(to_dev(keys))
    assert ei.value.code == q.QS_ERR_POOL
    g.clear()
    st = g.export_state()
    assert st["next"] == 1025 and not st["low"].any()


def test_empty_and_query_before_insert(gpu):
    _print_case(test="empty")
    import torch
    g, o = make_pair(1021)
    g.insert(torch.empty(0, dtype=torch.int32, device="cuda"))
    q = uniform_keys(1000, seed=1)
    assert not to_host_u32(g.search(to_dev(q))).any()


def test_import_rejects_unreachable_states(gpu):
    """qs_import_state accepts the oracle's state and refuses, with
    QS_ERR_STATE and no change to the sketch, every state the insert cannot
    reach."""
    import qsketch_amd as q
    _print_case(test="import_reject")
    keys = zipf_keys(1.1, 1 << 12, 1 << 17, seed=4)
    g, o = make_pair(3067, n_high=4096)
    assert o.insert(keys) == 0
    low, high, masks, seed, nxt = o.low(), o.high(), o.masks(), o.seed, o.next
    assert nxt > 1030
    g.import_state(low, high, masks, seed, nxt)
    assert_same_state(g, o, "valid import")
    linked = np.nonzero(low[7::8])[0]
    free = np.nonzero(low[7::8] == 0)[0]

    def refused(what, lo=None, hi=None, ma=None, nx=None):
        with pytest.raises(q.QSketchError) as e:
            g.import_state(low if lo is None else lo, high if hi is None else hi,
                           masks if ma is None else ma, seed, nxt if nx is None else nx)
        assert e.value.code == q.QS_ERR_STATE, what
        assert_same_state(g, o, f"after refusing {what}")

    lo = low.copy(); lo[free[0] * 8 + 7] = 1024
    refused("link word 1024", lo=lo)
    lo = low.copy(); lo[free[0] * 8 + 7] = nxt
    refused("link word = next", lo=lo)
    lo = low.copy(); lo[free[0] * 8 + 7] = low[linked[0] * 8 + 7]
    refused("an id held by two buckets", lo=lo)
    refused("an id held by no bucket", nx=nxt + 1)
    lo = low.copy(); lo[free[0] * 8 + 2] = (lo[free[0] * 8 + 2] & ~np.uint32(0xff00)) | np.uint32(130 << 8)
    refused("byte counter 130", lo=lo)
    hi = high.copy(); hi[5] = 1
    refused("counts in high row 0", hi=hi)
    hi = high.copy(); hi[(nxt - 1024) * 32 + 3] = 1
    refused("counts in the row of an unallocated id", hi=hi)
    ma = masks.copy(); ma[7] |= np.uint32(1 << 28)
    refused("mask bit 28", ma=ma)
    ma = masks.copy(); ma[9] = np.uint32(0b1111)
    refused("mask of 4 bits", ma=ma)
    refused("next below 1025", nx=1024)
    refused("next beyond the pool", nx=1024 + 4096 + 1)
    # a byte of 129 is reachable
    lo = low.copy(); lo[free[0] * 8] = (lo[free[0] * 8] & ~np.uint32(0xff)) | np.uint32(129)
    g.import_state(lo, high, masks, seed, nxt)
