"""Shared helpers for the parity tests (GPU path vs the CPU oracle)."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as po


def to_dev(a: np.ndarray):
    import torch
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32).copy()).cuda()
    if a.dtype == np.uint64:
        return torch.from_numpy(a.view(np.int64).copy()).cuda()
    raise TypeError(a.dtype)


def to_host_u32(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


def make_pair(n_low, n_high=1024, key_bytes=4, semantics=0, rng_seed=1, chunk_keys=0,
              n_masks=1021, seed_sz=1):
    """A GPU sketch and an oracle sketch built from the same parameters."""
    import qsketch_amd as q
    g = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, n_high, 32, seed_sz, n_masks, key_bytes=key_bytes,
                                    semantics=semantics, rng_seed=rng_seed, chunk_keys=chunk_keys)
    o = po.OracleSketch(n_low, n_high, n_masks, seed_sz=seed_sz, key_bytes=key_bytes,
                        semantics=semantics, rng_seed=rng_seed)
    return g, o


def assert_same_state(g, o, what=""):
    st = g.export_state()
    assert st["seed"] == o.seed, what
    assert np.array_equal(st["masks"], o.masks()), what
    lo, og = st["low"], o.low()
    if not np.array_equal(lo, og):
        bad = np.nonzero(lo != og)[0]
        raise AssertionError(f"{what}: low table differs at {bad.size} words, first {bad[:8]} "
                             f"gpu={lo[bad[:8]]} oracle={og[bad[:8]]}")
    hi, oh = st["high"], o.high()
    if not np.array_equal(hi, oh):
        bad = np.nonzero(hi != oh)[0]
        raise AssertionError(f"{what}: high table differs at {bad.size} words, first {bad[:8]} "
                             f"gpu={hi[bad[:8]]} oracle={oh[bad[:8]]}")
    assert st["next"] == o.next, f"{what}: next {st['next']} vs {o.next}"


def zipf_keys(alpha, universe, n, seed):
    import qsketch_amd as q
    z = q.Zipf(alpha, universe, device=False)
    return z.keys_u32_host(seed, n)


def uniform_keys(n, seed):
    import qsketch_amd as q
    return q.uniform_u32_host(seed, n)
