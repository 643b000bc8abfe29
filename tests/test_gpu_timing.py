"""The timing hooks (qs_set_timing / qs_get_timing): a timed insert leaves the
same state as an untimed one, and its kernel-group times are consistent.

Each checked insert:
- every kernel group that ran has launches and positive time;
- the groups sum to the insert's kernel time;
- that time fits inside the call's wall time;
- the keys counted equal the keys inserted.

The radix passes and the steps after the segment pass share one event
between adjacent groups, so a shared event mis-assigned to the wrong pair
would break the sums.  Checked for the ordinary pipeline (2^20 keys) and for
the hot-key filter (2^26 Zipf keys: the sample, the prefix and cold chunks,
and the filter's own group)."""
from __future__ import annotations

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("log2n,filtered", [(20, False), (26, True)])
def test_timed_insert_same_state_and_consistent_times(gpu, log2n, filtered):
    import torch
    import qsketch_amd as q
    _case(test="timing", log2n=log2n)
    n = 1 << log2n
    keys = q.Zipf(1.1, 1 << 22).keys_u32(5, n)
    n_low = q.number_of_buckets(n, 8, 0.5)
    plain = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 18, 32, rng_seed=3)
    timed = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, 1 << 18, 32, rng_seed=3)
    plain.insert(keys)
    timed.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed.insert(keys)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    tim = timed.timing()
    timed.set_timing(False)

    a, b = plain.export_state(), timed.export_state()
    assert a["next"] == b["next"]
    assert np.array_equal(a["low"], b["low"]) and np.array_equal(a["high"], b["high"])

    ins = {k: v for k, v in tim["kernels"].items() if k != "query"}
    ran = {k for k, v in ins.items() if v["launches"]}
    assert {"sort_up", "sort_scan", "sort_down", "seg", "ordered", "assign", "tail"} <= ran, ran
    assert ("hot" in ran) == filtered, ran
    for k in ran:
        assert ins[k]["ms"] > 0.0, k
    total = sum(v["ms"] for v in ins.values())
    assert abs(total - tim["insert_kernel_ms"]) <= 1e-6 * max(1.0, total)
    assert tim["insert_kernel_launches"] == sum(v["launches"] for v in ins.values())
    assert 0.0 < tim["insert_kernel_ms"] <= wall_ms
    assert tim["insert_keys"] == n
    # four radix passes per pipeline chunk (two chunks when filtered)
    chunks = ins["seg"]["launches"]
    assert chunks == (2 if filtered else 1)
    assert ins["sort_up"]["launches"] == ins["sort_scan"]["launches"] == ins["sort_down"]["launches"]
    assert ins["sort_down"]["launches"] % chunks == 0
