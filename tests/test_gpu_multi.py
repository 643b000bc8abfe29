"""GPU parity of the batched multi-sketch form and of captured inserts.

Every sketch of a set must end in exactly the state (and give exactly the
estimates) of the oracle sketch with the same parameters fed only its own
stream, whatever the batch layout, the chunking and the capture."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_dev, to_host_u32, zipf_keys

pytestmark = pytest.mark.gpu


def _print_case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


LOWS = [13, 61, 1021, 4093, 12281, 65521]


def _specs(k, seed, n_high=1 << 12):
    rs = np.random.default_rng(seed)
    out = []
    for i in range(k):
        out.append({"n_low": LOWS[i % len(LOWS)], "n_high": int(n_high), "sem": i % 2,
                    "seed": int(rs.integers(1, 1 << 40)), "alpha": float(rs.uniform(0.7, 1.3)),
                    "universe": int(rs.choice([64, 1 << 12, 1 << 16])),
                    "n_masks": int(rs.choice([1021, 509, 61]))})
    return out


def _make(specs, key_bytes=4, chunk_keys=0):
    import qsketch_amd as q
    params = [q.make_params(s["n_low"], s["n_high"], n_masks=s["n_masks"], key_bytes=key_bytes,
                            semantics=s["sem"], rng_seed=s["seed"],
                            chunk_keys=chunk_keys if i == 0 else 0)
              for i, s in enumerate(specs)]
    S = q.SketchSet(params)
    O = [po.OracleSketch(s["n_low"], s["n_high"], s["n_masks"], key_bytes=key_bytes,
                         semantics=s["sem"], rng_seed=s["seed"]) for s in specs]
    return S, O


def _streams(specs, lengths, seed, key_bytes=4):
    out = []
    for i, (s, n) in enumerate(zip(specs, lengths)):
        k = zipf_keys(s["alpha"], s["universe"], int(n), seed=seed * 1000003 + i) if n else \
            np.zeros(0, np.uint32)
        if key_bytes == 8:
            k = (k.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64(i << 40)
        out.append(k)
    offs = np.zeros(len(out) + 1, np.uint64)
    offs[1:] = np.cumsum([a.size for a in out])
    dt = np.uint64 if key_bytes == 8 else np.uint32
    flat = np.concatenate(out).astype(dt) if offs[-1] else np.zeros(0, dt)
    return out, flat, offs


def _check_set(S, O, what):
    for i, o in enumerate(O):
        st = S.export_state(i)
        assert st["seed"] == o.seed, (what, i)
        assert np.array_equal(st["masks"], o.masks()), (what, i)
        if not np.array_equal(st["low"], o.low()):
            bad = np.nonzero(st["low"] != o.low())[0]
            raise AssertionError(f"{what}: sketch {i} low differs at {bad.size} words, first {bad[:6]}")
        assert np.array_equal(st["high"], o.high()), (what, i)
        assert st["next"] == o.next, (what, i, st["next"], o.next)


def _query_set(S, O, streams, key_bytes=4):
    qs = [np.unique(a) for a in streams]
    offs = np.zeros(len(qs) + 1, np.uint64)
    offs[1:] = np.cumsum([a.size for a in qs])
    if offs[-1] == 0:
        return
    flat = np.concatenate(qs)
    est = to_host_u32(S.search(to_dev(flat), offs))
    for i, o in enumerate(O):
        a, b = int(offs[i]), int(offs[i + 1])
        if b > a:
            assert np.array_equal(est[a:b], o.query(qs[i])), f"sketch {i} estimates"


@pytest.mark.parametrize("chunk_keys", [0, 5000])
def test_multi_insert_matches_each_oracle(gpu, chunk_keys):
    _print_case(test="multi_insert", chunk_keys=chunk_keys)
    specs = _specs(24, seed=3)
    S, O = _make(specs, chunk_keys=chunk_keys)
    rs = np.random.default_rng(5)
    for batch in range(2):
        lengths = [0 if (i + batch) % 7 == 3 else int(rs.integers(1, 1 << 15)) for i in range(len(specs))]
        lengths[batch] = 1 << 16
        streams, flat, offs = _streams(specs, lengths, seed=batch + 1)
        for o, a in zip(O, streams):
            if a.size:
                assert o.insert(a) == 0
        S.insert(to_dev(flat), offs)
        _check_set(S, O, f"batch {batch}")
        _query_set(S, O, streams)
    st = S.stats()
    assert st["n_sketches"] == len(specs) and st["pool_error"] == 0
    assert st["linked_buckets"] == sum(o.next - 1025 for o in O)


def test_multi_quotient_item_form(gpu):
    """A set whose every table has over 2^21 buckets and 4093 masks: the items
    carry the hash quotient and each sketch recomputes its mask index from its
    own n_low and n_masks."""
    _print_case(test="multi_quotient_form")
    specs = [{"n_low": nl, "n_high": 1 << 12, "sem": i % 2, "seed": 100 + i, "alpha": a, "universe": u,
              "n_masks": 4093}
             for i, (nl, a, u) in enumerate([(2097169, 1.1, 1 << 12), (2200009, 0.9, 1 << 14),
                                             (3000017, 1.2, 1 << 10)])]
    S, O = _make(specs)
    streams, flat, offs = _streams(specs, [200000, 150000, 250000], seed=8)
    S.insert(to_dev(flat), offs)
    for o, st in zip(O, streams):
        assert o.insert(st) == 0
    _check_set(S, O, "multi quotient form")
    _query_set(S, O, streams)


def test_multi_u64_keys(gpu):
    _print_case(test="multi_u64")
    specs = _specs(9, seed=8)
    S, O = _make(specs, key_bytes=8)
    lengths = [int(x) for x in np.random.default_rng(2).integers(100, 1 << 14, len(specs))]
    streams, flat, offs = _streams(specs, lengths, seed=4, key_bytes=8)
    for o, a in zip(O, streams):
        assert o.insert(a) == 0
    S.insert(to_dev(flat), offs)
    _check_set(S, O, "u64")
    _query_set(S, O, streams, key_bytes=8)


def test_multi_pool_error(gpu):
    import qsketch_amd as q
    _print_case(test="multi_pool_error")
    specs = _specs(4, seed=9)
    specs[2].update(n_low=13, n_high=3, alpha=1.0, universe=64)
    S, O = _make(specs)
    streams, flat, offs = _streams(specs, [1000, 1000, 20000, 1000], seed=6)
    assert O[2].insert(streams[2]) == q.QS_ERR_POOL
    with pytest.raises(q.QSketchError) as e:
        S.insert(to_dev(flat), offs)
    assert e.value.code == q.QS_ERR_POOL
    assert S.stats()["pool_error"] == 1
    S.clear()
    assert S.stats()["pool_error"] == 0


def test_multi_async_with_device_offsets(gpu):
    import torch
    _print_case(test="multi_async")
    specs = _specs(11, seed=12)
    S, O = _make(specs)
    lengths = [int(x) for x in np.random.default_rng(4).integers(0, 1 << 14, len(specs))]
    streams, flat, offs = _streams(specs, lengths, seed=9)
    for o, a in zip(O, streams):
        if a.size:
            assert o.insert(a) == 0
    d_offs = torch.from_numpy(offs.view(np.int64).copy()).cuda()
    S.insert_async(to_dev(flat), offs, d_offs, 0)
    S.sync()
    _check_set(S, O, "async")
    qk = np.concatenate([np.unique(a) for a in streams])
    qo = np.zeros(len(streams) + 1, np.uint64)
    qo[1:] = np.cumsum([np.unique(a).size for a in streams])
    d_q = to_dev(qk)
    cnt = torch.empty(qk.size, dtype=torch.int32, device="cuda")
    S.search_async(d_q, torch.from_numpy(qo.view(np.int64).copy()).cuda(), qk.size, cnt, 0)
    torch.cuda.synchronize()
    est = to_host_u32(cnt)
    for i, o in enumerate(O):
        a, b = int(qo[i]), int(qo[i + 1])
        assert np.array_equal(est[a:b], o.query(qk[a:b]))


def test_graph_replays_match_sequential_oracle(gpu):
    """A captured multi-sketch insert, replayed three times with new keys
    written into the captured buffers; eager work and synchronisations between
    the replays; every replay checked against the oracle."""
    import torch
    _print_case(test="graph_multi")
    specs = _specs(16, seed=21)
    S, O = _make(specs)
    lengths = [int(x) for x in np.random.default_rng(7).integers(1, 1 << 14, len(specs))]
    _, flat, offs = _streams(specs, lengths, seed=100)
    d_keys = to_dev(flat)
    d_offs = torch.from_numpy(offs.view(np.int64).copy()).cuda()
    g = S.capture_insert(d_keys, offs, d_offs)
    for rep in range(3):
        streams, flat, offs2 = _streams(specs, lengths, seed=200 + rep)
        assert np.array_equal(offs2, offs)
        d_keys.copy_(to_dev(flat))
        for o, a in zip(O, streams):
            assert o.insert(a) == 0
        g.launch(0)
        S.sync()
        _ = (d_keys.to(torch.int64) * 3).sum().item()  # unrelated eager work between replays
        _check_set(S, O, f"replay {rep}")
    assert S.stats()["keys_inserted"] == 3 * int(offs[-1])
    g.destroy()


def test_graph_single_sketch(gpu):
    import qsketch_amd as q
    _print_case(test="graph_single")
    g_sk = q.Sketch_GPU_Sub_Warp_Level(4093, 8, 1 << 13, 32, rng_seed=5)
    o = po.OracleSketch(4093, 1 << 13, 1021, rng_seed=5)
    n = 100003
    d_keys = to_dev(zipf_keys(1.1, 1 << 13, n, seed=1))
    gr = g_sk.capture_insert(d_keys, n)
    for rep in range(3):
        keys = zipf_keys(1.1, 1 << 13, n, seed=10 + rep)
        d_keys.copy_(to_dev(keys))
        assert o.insert(keys) == 0
        gr.launch(0)
        g_sk.sync()
        st = g_sk.export_state()
        assert np.array_equal(st["low"], o.low()) and np.array_equal(st["high"], o.high()), rep
        assert st["next"] == o.next
    gr.destroy()


SHAPES = [(0, 0)] + [(w, i) for w in (8, 16, 32, 64) for i in (1, 2, 4, 8)]


def test_every_insert_shape_matches_oracle(gpu):
    """The segment pass's shape (sub-wavefront width x items per lane) changes
    the schedule only: every shape leaves the oracle's state, on a stream with
    runs of every length (long hot runs, runs crossing windows, single keys)."""
    import qsketch_amd as q
    keys = np.concatenate([zipf_keys(1.1, 1 << 12, 1 << 17, seed=3),
                           zipf_keys(0.6, 1 << 20, 1 << 17, seed=4)])
    o = po.OracleSketch(12281, 1 << 13, 1021, rng_seed=9)
    assert o.insert(keys[: 1 << 17]) == 0
    assert o.insert(keys[1 << 17:]) == 0
    d1, d2 = to_dev(keys[: 1 << 17]), to_dev(keys[1 << 17:])
    for w, ipl in SHAPES:
        _print_case(test="shape", width=w, ipl=ipl)
        g = q.Sketch_GPU_Sub_Warp_Level(12281, 8, 1 << 13, 32, rng_seed=9)
        lib = q.lib()
        q.check(lib.qs_set_insert_shape(g.handle, w, ipl), "shape")
        g.insert(d1)
        g.insert(d2)
        st = g.export_state()
        assert np.array_equal(st["low"], o.low()), (w, ipl)
        assert np.array_equal(st["high"], o.high()), (w, ipl)
        assert st["next"] == o.next, (w, ipl)
        g.destroy()


def test_autotune_keeps_results_and_picks_a_shape(gpu):
    import torch
    _print_case(test="autotune")
    specs = _specs(12, seed=31)
    S, O = _make(specs)
    lengths = [int(x) for x in np.random.default_rng(9).integers(1, 1 << 14, len(specs))]
    streams, flat, offs = _streams(specs, lengths, seed=50)
    d_keys = to_dev(flat)
    d_offs = torch.from_numpy(offs.view(np.int64).copy()).cuda()
    r = S.autotune(d_keys, offs, d_offs, reps=1)
    assert r["consistent"] and len(r["table"]) == 16
    assert S.shape() == (r["best"]["width"], r["best"]["items_per_lane"])
    # the caller's set was not touched by the tuning runs
    assert S.stats()["keys_inserted"] == 0 and S.stats()["linked_buckets"] == 0
    S.insert(d_keys, offs)
    for o, a in zip(O, streams):
        assert o.insert(a) == 0
    _check_set(S, O, "tuned")
