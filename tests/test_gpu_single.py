"""GPU parity of the single-level sub-warp sketch (include/qsketch_single.h)
against the oracle restatement of device_sketch_v2.h."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_dev, to_host_u32, uniform_keys, zipf_keys

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m", [8, 16, 32])
@pytest.mark.parametrize("alpha", [None, 1.1])
def test_single_level_state_and_estimates(gpu, m, alpha):
    import qsketch_amd as q
    print(f"CASE single m={m} alpha={alpha}", flush=True)
    n = q.number_of_buckets(1 << 20, m, 0.5)
    keys = uniform_keys(1 << 20, seed=4) if alpha is None else zipf_keys(alpha, 1 << 14, 1 << 20, seed=4)
    g = q.Sketch_GPU_Sub_Warp(n, m, rng_seed=3)
    o = po.OracleSingleLevel(n, m, 1021, rng_seed=3)
    st = g.export_state()
    assert st["seed"] == o.seed and np.array_equal(st["masks"], o.masks())
    g.insert(to_dev(keys[: 1 << 19]))
    g.insert(to_dev(keys[1 << 19:]))
    o.insert(keys)
    assert np.array_equal(g.export_state()["table"], o.table())
    uq = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(uq))), o.query(uq))
    g.clear()
    assert not g.export_state()["table"].any()


def test_single_level_u64(gpu):
    import qsketch_amd as q
    print("CASE single u64", flush=True)
    z = q.Zipf(1.2, 1 << 12, device=False)
    keys = z.keys_u64_noisy_host(2, 1 << 18)
    g = q.Sketch_GPU_Sub_Warp(4093, 8, key_bytes=8, rng_seed=6)
    o = po.OracleSingleLevel(4093, 8, 1021, key_bytes=8, rng_seed=6)
    g.insert(to_dev(keys))
    o.insert(keys)
    assert np.array_equal(g.export_state()["table"], o.table())
    uq = np.unique(keys)
    assert np.array_equal(to_host_u32(g.search(to_dev(uq))), o.query(uq))
