"""CPU-side checks of the C ABI library (no GPU compute): it loads, exports
every entry point the public headers declare, and its host-only functions
(sizing, primes, error strings, workload generators) behave."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def _declared(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(qs_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import qsketch_amd as q
    L = q.lib()
    declared = [n for h in HEADERS for n in _declared(h)]
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert not [n for n in declared if n not in q._SIGS], [n for n in declared if n not in q._SIGS]


def test_default_params_follow_reference_defaults():
    import qsketch_amd as q
    p = q.default_params(1 << 22, 0.5)
    assert (p.m_low, p.m_high, p.n_high, p.seed_sz, p.n_masks) == (8, 32, 1024, 1, 1021)
    assert (p.mask_ones, p.link_start, p.low_limit, p.key_bytes) == (3, 1024, 128, 4)
    assert p.n_low == 3145721
    assert q.default_params(1 << 28, 1.0).n_low == 100663291
    assert q.default_params(1 << 28, 0.5).n_low == 201326557


def test_primes_match_oracle():
    import qsketch_amd as q
    from oracle import pyoracle as po
    for x in [0, 1, 2, 3, 100, 1021, 1022, 4095, 4096, 123456789, 1 << 28, 1 << 33]:
        assert q.find_greatest_prime(x) == po.find_greatest_prime(x)
    for k in range(10, 30):
        assert q.number_of_buckets(1 << k, 8, 0.5) == po.number_of_buckets(1 << k, 8, 0.5)
        assert q.number_of_buckets(1 << k, 8, 1.0) == po.number_of_buckets(1 << k, 8, 1.0)


def test_error_strings_and_bad_params():
    import qsketch_amd as q
    L = q.lib()
    for code in (0, -1, -2, -3, -4, -5, -6, -7, -8, -99):
        assert L.qs_strerror(code)
    # every code has its own message; unknown codes share one
    msgs = [L.qs_strerror(c) for c in (0, -1, -2, -3, -4, -5, -6, -7, -8)]
    assert len(set(msgs)) == len(msgs) and L.qs_strerror(-99) not in msgs
    assert q.QS_ERR_INTERNAL == -8 and b"consistency" in L.qs_strerror(q.QS_ERR_INTERNAL)
    p = q.default_params(1 << 20)
    h = C.c_void_p()
    for field, bad in (("m_low", 4), ("m_high", 16), ("mask_ones", 2), ("link_start", 7),
                       ("low_limit", 100), ("key_bytes", 2), ("n_high", 1), ("n_masks", 0),
                       ("n_masks", 5000), ("n_low", 0), ("semantics", 7), ("seed_sz", 0)):
        pp = q.qs_params.from_buffer_copy(p)
        setattr(pp, field, bad)
        assert L.qs_create(C.byref(pp), C.byref(h)) == q.QS_ERR_PARAM, field
    pp = q.qs_params.from_buffer_copy(p)
    pp.n_low = (1 << 29) + 1  # 32-bit keys: the word offset must fit 32 bits (v3:42)
    assert L.qs_create(C.byref(pp), C.byref(h)) == q.QS_ERR_PARAM
    assert L.qs_default_params(C.byref(pp), 1 << 20, 0.0) == q.QS_ERR_PARAM
    assert L.qs_insert(None, None, 0) == q.QS_ERR_PARAM
    assert L.qs_query(None, None, 0, None) == q.QS_ERR_PARAM


def test_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import qsketch_amd as q
    p = q.default_params(1 << 20)
    h = C.c_void_p()
    assert q.lib().qs_create(C.byref(p), C.byref(h)) == q.QS_ERR_NODEVICE
    assert not h.value


# ---------------------------------------------------------------------------
# workload generators (host versions; the GPU versions share the sampler)
# ---------------------------------------------------------------------------
def test_zipf_sampler_matches_distribution():
    import qsketch_amd as q
    U = 1 << 12
    for alpha in (0.8, 1.1, 1.5):
        z = q.Zipf(alpha, U, device=False)
        r = z.ranks_host(7, 1 << 20)
        assert r.max() < U
        p = (np.arange(1, U + 1, dtype=np.float64) ** -alpha)
        p /= p.sum()
        obs = np.bincount(r, minlength=U)
        exp = p * r.size
        top = 64
        chi = ((obs[:top] - exp[:top]) ** 2 / exp[:top]).sum()
        assert chi < 3 * top, (alpha, chi)
        assert abs(obs[top:].sum() - exp[top:].sum()) < 6 * np.sqrt(exp[top:].sum())


def test_generators_are_deterministic_and_offsettable():
    import qsketch_amd as q
    z = q.Zipf(1.1, 1 << 16, device=False)
    a = z.keys_u32_host(3, 5000)
    b = z.keys_u32_host(3, 5000)
    assert np.array_equal(a, b)
    assert np.array_equal(z.keys_u32_host(3, 1000, offset=4000), a[4000:])
    assert not np.array_equal(z.keys_u32_host(4, 5000), a)
    u = q.uniform_u32_host(9, 1 << 16)
    assert np.array_equal(q.uniform_u32_host(9, 100, offset=50), u[50:150])
    assert abs(u.astype(np.float64).mean() / 2 ** 32 - 0.5) < 0.01


def test_rank_to_key_is_a_bijection_on_a_range():
    import qsketch_amd as q
    ks = np.array([q.rank_to_key32(5, r) for r in range(1 << 14)], dtype=np.uint64)
    assert np.unique(ks).size == ks.size


def test_noisy_u64_keys_encode_log_rank():
    import qsketch_amd as q
    z = q.Zipf(1.1, 1 << 16, device=False)
    k = z.keys_u64_noisy_host(11, 1 << 16, sigma=1.5)
    hi = (k >> np.uint64(48)).astype(np.int64)
    assert hi.min() >= 0 and hi.max() <= 31
    r = z.ranks_host(11, 1 << 16).astype(np.float64)
    base = np.clip(np.round(np.log2(r + 1)), 0, 31)
    # noise has mean ~0 and sd ~1.5 before clipping
    d = hi - base
    assert abs(d.mean()) < 0.5 and 0.8 < d[(base > 4) & (base < 26)].std() < 2.2
    # the same rank always maps to the same key
    assert np.unique(k).size == np.unique(z.ranks_host(11, 1 << 16)).size


def test_largest_prime_at_most():
    import qsketch_amd as q
    from oracle import pyoracle as po
    assert [q.largest_prime_at_most(x) for x in (2, 3, 4, 10, 1021, 1022, 201129416)] == \
        [2, 3, 3, 7, 1021, 1021, 201129389]
    # agrees with the reference's prime table below its largest entry (3 * 2^27)
    for k in range(2, 29):
        assert q.largest_prime_at_most(1 << k) == po.find_greatest_prime(1 << k)


# ---------------------------------------------------------------------------
# argument checks of the Python wrappers (no GPU: they fail before any call)
# ---------------------------------------------------------------------------
def test_wrapper_rejects_buffers_that_do_not_fit():
    import torch
    import qsketch_amd as q
    k32 = torch.zeros(10, dtype=torch.int32)
    k64 = torch.zeros(10, dtype=torch.int64)
    with pytest.raises(ValueError, match="holds 10"):
        q._keys_arg(k32, 11, 4, "x")            # n beyond the tensor
    with pytest.raises(ValueError, match="4-byte keys"):
        q._keys_arg(k64, None, 4, "x")          # u64 tensor on a u32 sketch
    with pytest.raises(ValueError, match="8-byte keys"):
        q._keys_arg(k32, None, 8, "x")          # u32 tensor on a u64 sketch
    with pytest.raises(ValueError, match="integer"):
        q._keys_arg(torch.zeros(10, dtype=torch.float32), None, 4, "x")
    with pytest.raises(ValueError, match="raw device address"):
        q._keys_arg(12345, None, 4, "x")
    with pytest.raises(ValueError, match="32-bit integer"):
        q._counts_arg(torch.zeros(10, dtype=torch.int64), 10, "x")
    with pytest.raises(ValueError, match="32-bit integer"):
        q._counts_arg(torch.zeros(10, dtype=torch.float32), 10, "x")
    with pytest.raises(ValueError, match="9 elements, 10 needed"):
        q._counts_arg(torch.zeros(9, dtype=torch.int32), 10, "x")
    with pytest.raises(ValueError, match="float32"):
        q._counts_arg(torch.zeros(10, dtype=torch.int32), 10, "x", floating=True)
    o = q._offs_array([0, 4, 12])
    with pytest.raises(ValueError, match="n = 12"):
        q._offs_keys(o, k32, 4, "x")            # offs[-1] beyond the keys
    with pytest.raises(ValueError, match="non-decreasing"):
        q._offs_array([0, 5, 3])
    # the checks pass for fitting buffers; the device check comes after them
    with pytest.raises(ValueError, match="device tensors"):
        q._keys_arg(k32, 10, 4, "x")


def test_u64_keys_take_one_seed():
    """hash_mul_add sums key * seed_i in 64 bits for 64-bit keys
    (device_hash_function.h:6-17); the device hash multiplies by one 32-bit
    seed, so the library refuses seed_sz > 1 with 64-bit keys."""
    import qsketch_amd as q
    L = q.lib()
    p = q.make_params(1021, 1024, key_bytes=8, seed_sz=2)
    h = C.c_void_p()
    assert L.qs_create(C.byref(p), C.byref(h)) == q.QS_ERR_PARAM
    assert L.qs_sl_create(1021, 8, 2, 1021, 8, 1, C.byref(h)) == q.QS_ERR_PARAM


def test_oracle_u64_hash_is_hash_mul_add_in_64_bits():
    """The oracle's 64-bit hash with two seeds equals sum_i key * seed_i mod
    2^64, the functor's arithmetic restated with Python integers; a 32-bit sum
    of the seeds (the plausible mistake) gives other buckets."""
    from oracle import pyoracle as po
    eng = po.Minstd(seed_bytes=po.seed_buffer(2, 0))
    seeds = [eng.uniform(1, 0xFFFFFFFF) for _ in range(2)]
    assert sum(seeds) >= 1 << 32  # rng_seed 2: the sum needs 33 bits
    o = po.OracleSketch(1000003, 1024, 1021, seed_sz=2, key_bytes=8, rng_seed=2)
    assert o.s.seed64 == sum(seeds)
    rs = np.random.default_rng(4)
    keys = rs.integers(0, 1 << 63, 2000, dtype=np.int64).view(np.uint64) * np.uint64(2) + np.uint64(1)
    o.insert(keys)
    want = np.zeros(o.s.n_low * 8, np.uint32)
    for k in keys.tolist():
        h = sum(k * s for s in seeds) % (1 << 64)
        b, mu = h % o.s.n_low, h % 1021
        m = int(o.masks()[mu])
        for j in range(28):
            if m >> j & 1:
                want[b * 8 + j // 4] += 1 << (8 * (j % 4))
    assert np.array_equal(o.low(), want)


# ---------------------------------------------------------------------------
# accuracy harness (qsketch_amd/accuracy.py runs on any torch device)
# ---------------------------------------------------------------------------
def test_accuracy_exact_counts_equal_the_oracle():
    """torch's sort + count (the GPU harness) against the oracle's radix sort +
    reduce-by-key, for u32 keys on both sides of 2^31 and for u64 keys."""
    import torch
    from oracle import pyoracle as po
    from qsketch_amd import accuracy as acc
    rs = np.random.default_rng(8)
    k32 = rs.integers(0, 1 << 32, 50000, dtype=np.uint64).astype(np.uint32) % np.uint32(3000) * np.uint32(1431655)
    gk, gc = acc.exact_counts(torch.from_numpy(k32.view(np.int32).copy()))
    uk, uc = po.exact_counts(k32, key_bytes=4)
    g = gk.numpy().view(np.uint32)
    order = np.argsort(g)  # torch sorts the signed view
    assert np.array_equal(g[order], uk) and np.array_equal(gc.numpy()[order], uc)
    k64 = rs.integers(0, 5000, 40000, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    gk, gc = acc.exact_counts(torch.from_numpy(k64.view(np.int64).copy()))
    uk, uc = po.exact_counts(k64, key_bytes=8)
    g = gk.numpy().view(np.uint64)
    order = np.argsort(g)
    assert np.array_equal(g[order], uk) and np.array_equal(gc.numpy()[order], uc)


def test_accuracy_metrics_on_hand_computed_cases():
    import torch
    from qsketch_amd import accuracy as acc
    est = torch.tensor([3, 5, 10, 0], dtype=torch.int32)
    true = torch.tensor([2, 5, 8, 1])
    m = acc.metrics(est, true)
    assert m["are"] == pytest.approx((1 / 2 + 0 + 2 / 8 + 1) / 4)
    assert m["aae"] == pytest.approx((1 + 0 + 2 + 1) / 4)
    assert m["mean_est_over_true_pct"] == pytest.approx((150 + 100 + 125 + 0) / 4)
    assert m["under_estimates"] == 1 and m["max_abs_error"] == 2
    # estimates are u32 bit patterns in an int32 tensor
    m = acc.metrics(torch.tensor([-1], dtype=torch.int32), torch.tensor([7]))
    assert m["aae"] == 4294967295 - 7 and m["under_estimates"] == 0
    # a key that was never inserted counts in AAE, not in ARE
    m = acc.metrics(torch.tensor([4, 2], dtype=torch.int32), torch.tensor([2, 0]))
    assert m["are"] == pytest.approx(1.0) and m["aae"] == pytest.approx(2.0) and m["absent_keys"] == 1


def test_sf_create_rejects_bad_parameters():
    import qsketch_amd as q
    L = q.lib()
    h = C.c_void_p()
    for args in ((1021, 7, 3, 1, 1021, 4), (1021, 8, 0, 1, 1021, 4), (1021, 8, 3, 0, 1021, 4),
                 (1021, 8, 3, 1, 0, 4), (1021, 8, 3, 1, 5000, 4), (1021, 8, 3, 1, 1021, 2),
                 (1021, 8, 3, 2, 1021, 8), (0, 8, 3, 1, 1021, 4), (1 << 32, 8, 1, 1, 1021, 8)):
        assert L.qs_sf_create(*args, 1, C.byref(h)) == q.QS_ERR_PARAM, args


def test_calibration_kernels_reject_bad_parameters():
    """The calibration entry points check their arguments on the host before
    anything is launched (a null buffer, an access size other than 16, 32, 64
    or 128 bytes, sizes not a multiple of 16, more accesses than lines)."""
    import qsketch_amd as q
    L = q.lib()
    fake = 1 << 20  # never dereferenced: every call below is refused first
    assert L.qs_calib_copy(None, fake, 1 << 20, None) == q.QS_ERR_PARAM
    assert L.qs_calib_copy(fake, fake, 17, None) == q.QS_ERR_PARAM
    assert L.qs_calib_fill(None, 1 << 20, 0, 0, None) == q.QS_ERR_PARAM
    assert L.qs_calib_fill(fake, 24, 0, 0, None) == q.QS_ERR_PARAM
    for size in (0, 8, 48, 256):
        assert L.qs_calib_gather(fake, 1 << 20, size, 1, 1, fake, None) == q.QS_ERR_PARAM, size
        assert L.qs_calib_rmw(fake, 1 << 20, size, 1, 1, 0, None) == q.QS_ERR_PARAM, size
    assert L.qs_calib_rmw(None, 1 << 20, 32, 1, 1, 0, None) == q.QS_ERR_PARAM
    assert L.qs_calib_rmw(fake, 1 << 10, 32, 33, 1, 1, None) == q.QS_ERR_PARAM  # 32 lines, 33 accesses
    assert L.qs_calib_rmw(fake, 1 << 10, 32, 0, 1, 1, None) == q.QS_OK  # nothing to do
