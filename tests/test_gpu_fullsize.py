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
    print(f"  linked buckets {o.next - 1025}, distinct keys {uq.numel()}", flush=True)
