"""Key streams of the measured workloads (include/qsketch_gen.h): the device
generators and their host twins give identical keys for the same seed.

  cfg1  2^22 u32 keys, Zipf 1.0 over 2^18, seed 1 (the CPU reference-path
        configuration; src/main.cu:113 sizes it at WLF 0.5)
  cfg2  2^28 u32 keys, Zipf 1.1 or 0.8 over 2^26, or uniform u32, seed 1
  cfg3  2^27 u32 keys, Zipf 1.5 over 2^24, seed 1
  cfg4  4096 streams, lengths log-uniform in [2^10, 2^20], Zipf exponent
        uniform in [0.7, 1.3], own seed and 2^16-key universe each
  cfg5  2^28 u64 keys, Zipf 1.1 over 2^24, upper 16 bits a noisy log-rank
"""
from __future__ import annotations

from . import Zipf, make_params, number_of_buckets, uniform_u32, uniform_u32_host

CFG1 = dict(alpha=1.0, log2u=18, log2n=22, seed=1)
CFG2 = dict(alpha=1.1, log2u=26, log2n=28, seed=1)
CFG3 = dict(alpha=1.5, log2u=24, log2n=27, seed=1)
CFG5 = dict(alpha=1.1, log2u=24, log2n=28, seed=1, sigma=1.5)


def plain_keys(alpha: float, log2u: int, log2n: int, seed: int = 1, device: bool = True, n: int | None = None):
    """u32 stream: Zipf(alpha) over 2^log2u ranks (rank -> key by a seeded
    permutation), or uniform u32 when alpha == 0.  The first n keys (default
    2^log2n): a torch tensor on the GPU, or a numpy array on the host."""
    n = (1 << log2n) if n is None else int(n)
    if alpha > 0:
        z = Zipf(alpha, 1 << log2u, device=device)
        return z.keys_u32(seed=seed, n=n) if device else z.keys_u32_host(seed, n)
    return uniform_u32(seed=seed, n=n) if device else uniform_u32_host(seed, n)


def learned_keys(alpha: float = 1.1, log2u: int = 24, log2n: int = 28, seed: int = 1, sigma: float = 1.5,
                 device: bool = True, n: int | None = None):
    """u64 stream of the learned-form workload: Zipf rank r -> (clip(round(
    log2(r + 1)) + N(0, sigma), 0, 31) << 48) | a 48-bit permutation of r."""
    n = (1 << log2n) if n is None else int(n)
    z = Zipf(alpha, 1 << log2u, device=device)
    return z.keys_u64_noisy(seed=seed, n=n, sigma=sigma) if device else z.keys_u64_noisy_host(seed, n, sigma)


def multi_workload(n_sk: int = 4096, seed: int = 1):
    """Per stream: length log-uniform in [2^10, 2^20], Zipf exponent uniform in
    [0.7, 1.3] (on a 0.01 grid, so that one sampler table serves many
    streams), and its own seed (which also keys its rank -> key permutation
    over its own 2^16-key universe)."""
    import numpy as np
    rs = np.random.default_rng(seed)
    lengths = np.rint(2.0 ** rs.uniform(10, 20, n_sk)).astype(np.int64)
    alphas = np.round(rs.uniform(0.7, 1.3, n_sk), 2)
    seeds = np.arange(1, n_sk + 1, dtype=np.int64) * 7919 + seed
    return lengths, alphas, seeds


def multi_keys(n_sk: int = 4096, seed: int = 1, device: bool = True):
    """(keys, offs, lengths, alphas, seeds): every stream's keys concatenated
    (stream i owns keys[offs[i]:offs[i+1]]), on the GPU or the host."""
    import numpy as np
    lengths, alphas, seeds = multi_workload(n_sk, seed)
    offs = np.zeros(n_sk + 1, np.uint64)
    offs[1:] = np.cumsum(lengths)
    total = int(offs[-1])
    samplers = {}
    if device:
        import torch
        keys = torch.empty(total, dtype=torch.int32, device="cuda")
    else:
        keys = np.empty(total, dtype=np.uint32)
    for i in range(n_sk):
        al = float(alphas[i])
        if al not in samplers:
            samplers[al] = Zipf(al, 1 << 16, device=device)
        a, b = int(offs[i]), int(offs[i + 1])
        if device:
            samplers[al].keys_u32(int(seeds[i]), b - a, out=keys[a:b])
        else:
            keys[a:b] = samplers[al].keys_u32_host(int(seeds[i]), b - a)
    return keys, offs, lengths, alphas, seeds


def multi_params(lengths, seeds):
    """One parameter set per stream: n_low by the reference sizing rule at
    WLF 0.5, a high pool of max(1024, n / 256) buckets, the stream's seed."""
    return [make_params(number_of_buckets(int(n), 8, 0.5), max(1024, int(n) // 256), rng_seed=int(s))
            for n, s in zip(lengths, seeds)]
