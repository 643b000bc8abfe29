"""Host-side pipeline of the learned heavy-hitter form (include/qsketch_learned.h).

No counterpart in the reference.  The network is trained on what a plain
sketch estimated for the distinct keys of an earlier window of the stream
(target log2(1 + estimate)), with the library's own forward, backward and Adam
kernels; the threshold admits at most `top` of the window's keys; the learned
form's sketch gives up the bytes of the exact table and the network, so that
the whole form takes no more memory than the plain sketch it is compared with.
"""
from __future__ import annotations

from . import LearnedSketch, Sketch_GPU_Sub_Warp_Level, largest_prime_at_most, make_params


def window_targets(window_keys, n_low: int, n_high: int, rng_seed: int = 1):
    """Distinct keys of a window and log2(1 + the plain sketch's estimate) of each."""
    import torch
    sk = Sketch_GPU_Sub_Warp_Level(n_low, 8, n_high, 32, key_bytes=window_keys.element_size(),
                                   rng_seed=rng_seed)
    sk.insert(window_keys)
    wk = torch.unique(window_keys)
    est = sk.search(wk).to(torch.int64) & 0xFFFFFFFF
    sk.destroy()
    return wk, torch.log2(1.0 + est.to(torch.float32)).contiguous()


def train_router(L: LearnedSketch, wk, targets, steps: int, batch: int = 4096, lr: float = 2e-3,
                 seed: int = 3, loss_every: int = 100):
    """Adam steps on random batches of (key, target) pairs; returns the losses
    sampled every `loss_every` steps (and at the last)."""
    import torch
    gen = torch.Generator(device=wk.device)
    gen.manual_seed(seed)
    losses = []
    for it in range(steps):
        idx = torch.randint(0, wk.numel(), (batch,), device=wk.device, generator=gen)
        want = it % loss_every == 0 or it == steps - 1
        loss = L.step(wk[idx].contiguous(), targets[idx].contiguous(), lr=lr, want_loss=want)
        if want:
            losses.append(loss)
    return losses


def pick_threshold(L: LearnedSketch, wk, top: int) -> float:
    """The lowest score threshold that admits at most `top` of the window's
    keys (scores tie wherever the network ignores the random low key bits)."""
    import torch
    sc = L.score(wk)
    srt = torch.sort(sc, descending=True).values
    thr = float(srt[min(top, srt.numel()) - 1])
    if int((sc >= thr).sum()) > top:
        above = sc[sc > thr]
        thr = float(above.min()) if above.numel() else float("inf")
    return thr


def equal_memory_n_low(n_low_plain: int, n_high: int, table_log2: int, key_bytes: int = 8) -> int:
    """Bucket count of the learned form's sketch: the plain sketch's minus the
    32-byte buckets the table and the network take, kept prime (with the
    multiplicative hash an even bucket count lets low key bits steer buckets)."""
    probe = LearnedSketch(make_params(1021, n_high, key_bytes=key_bytes), table_log2=table_log2)
    m = probe.memory()
    probe.destroy()
    extra = m["table"] + m["mlp"]
    return largest_prime_at_most(n_low_plain - (extra + 31) // 32)
