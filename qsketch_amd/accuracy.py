"""Accuracy harness: exact counts and error metrics on the GPU.

Mirrors the reference's accuracy test (include/test/test.h:186-291) and its
metrics (include/analyze/analyze.h:7-64):
  * exact counts by sort + reduce-by-key of the inserted stream (the reference
    keeps an std::unordered_map, include/sketch/host_sketch/host_sketch.h:15-63);
  * mean(est / true) * 100  -- the reference's reported accuracy
    (Percentage_Functor, analyze.h:15-29, averaged in test.h:248-273);
  * the number of under-estimates -- the reference's Count_Greater tripwire
    (analyze.h:49-64 counts est - true that wrapped below zero);
  * ARE = mean |est - true| / true and AAE = mean |est - true| over the queried
    distinct keys (the metrics of the learned-oracle comparison).  Queried keys
    that were never inserted (true = 0) count in AAE and the under-estimate
    tally but not in ARE or the mean ratio, where they would divide by zero.
"""
from __future__ import annotations


def exact_counts(keys):
    """Distinct keys (sorted) and their exact counts, on the keys' device."""
    import torch
    k = keys.view(torch.int64) if keys.dtype == torch.int64 else keys
    uniq, counts = torch.unique(k, sorted=True, return_counts=True)
    return uniq, counts


def metrics(est, true) -> dict:
    """est, true: tensors of the same length (any integer dtype)."""
    import torch
    e = est.to(torch.int64) & 0xFFFFFFFF
    t = true.to(torch.int64).to(e.device)
    d = (e - t).abs().to(torch.float64)
    pos = t > 0
    tf = t[pos].to(torch.float64)
    return {
        "keys": int(t.numel()),
        "are": float((d[pos] / tf).mean()) if tf.numel() else 0.0,
        "aae": float(d.mean()) if d.numel() else 0.0,
        "mean_est_over_true_pct": float((e[pos].to(torch.float64) / tf).mean() * 100.0) if tf.numel() else 0.0,
        "max_abs_error": int(d.max()) if d.numel() else 0,
        "under_estimates": int((e < t).sum()),
        "absent_keys": int((~pos).sum()),
    }


def evaluate(sketch, keys, query=None) -> dict:
    """Insert-side exact counts of `keys` vs the sketch's estimates.

    `query`: which keys to evaluate (default: every distinct key of `keys`).
    The sketch must already hold `keys`."""
    import torch
    uniq, counts = exact_counts(keys)
    if query is None:
        q, true = uniq, counts
    else:
        pos = torch.searchsorted(uniq, query)
        pos = pos.clamp(max=uniq.numel() - 1)
        true = torch.where(uniq[pos] == query, counts[pos], torch.zeros_like(counts[pos]))
        q = query
    est = sketch.search(q.contiguous(), q.numel())
    return metrics(est, true)
