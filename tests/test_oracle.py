"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement fails it:
  * RNG: equality with this toolchain's libstdc++ (compiled here) and with the
    committed fixture tests/golden/stdlib_rng.json;
  * prime table: brute-force primality and "largest prime <= 2^k / 3*2^(k-1)";
  * sizing: the table sizes the reference itself logged
    (doc/data/qsketch/accuracy_0.5.dat memory column);
  * state machine: hand-derived scenarios from v3 (threshold > 128, first id
    1025, v3:187 j mod 8), an independent dictionary model, the textbook
    count-min special case, the 3N conservation law, and MT == sequential.
"""
from __future__ import annotations

import json
import os
import subprocess
from collections import Counter

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# RNG
# ---------------------------------------------------------------------------
def _check_rng_fixture(d):
    for s, v in d["minstd"].items():
        e = po.Minstd(int(s))
        assert [e.next() for _ in v] == v
    for c in d["seed_seq"]:
        inp = np.asarray(c["in"], dtype=np.uint32)
        assert list(map(int, po.seed_seq_generate(inp, c["n"]))) == c["out"]
    for c in d["uniform"]:
        e = po.Minstd(c["seed"])
        assert [e.uniform(c["a"], c["b"]) for _ in c["out"]] == c["out"]
    for c in d["create"]:
        e = po.Minstd(seed_bytes=po.seed_buffer(c["rng_seed"], 0))
        assert [e.next() for _ in range(4)] == c["seed_engine_raw"]
        e = po.Minstd(seed_bytes=po.seed_buffer(c["rng_seed"], 0))
        assert [e.uniform(1, 0xFFFFFFFF) for _ in range(6)] == c["seeds"]
        m = po.Minstd(seed_bytes=po.seed_buffer(c["rng_seed"], 1)).hashmask(1021).view(np.uint32)
        assert list(map(int, m)) == c["masks"]
        s = po.OracleSketch(1021, rng_seed=c["rng_seed"])
        assert s.seed == c["seeds"][0]
        assert np.array_equal(s.masks(), m)
        s3 = po.OracleSketch(1021, rng_seed=c["rng_seed"], seed_sz=3)
        assert s3.seed == sum(c["seeds"][:3]) & 0xFFFFFFFF


def test_rng_matches_committed_libstdcxx_fixture():
    _check_rng_fixture(json.load(open(os.path.join(GOLD, "stdlib_rng.json"))))


def test_rng_matches_libstdcxx_built_here(tmp_path):
    exe = tmp_path / "stdlib_rng"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(GOLD, "stdlib_rng.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout
    d = json.loads(out)
    assert d == json.load(open(os.path.join(GOLD, "stdlib_rng.json")))
    _check_rng_fixture(d)


def test_masks_invariants():
    for rs in (1, 5, 99):
        m = po.OracleSketch(1021, rng_seed=rs).masks()
        assert m.size == 1021
        assert all(bin(int(x)).count("1") == 3 for x in m)
        assert (m >> 28 == 0).all()
        # all 28 positions are reachable and roughly uniform
        pos = np.array([[(int(x) >> j) & 1 for j in range(28)] for x in m]).sum(0)
        assert pos.min() > 60 and pos.max() < 170


# ---------------------------------------------------------------------------
# primes and sizing
# ---------------------------------------------------------------------------
def _is_prime(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    i = 3
    while i * i <= n:
        if n % i == 0:
            return False
        i += 2
    return True


def _largest_prime_le(x):
    while not _is_prime(x):
        x -= 1
    return x


def test_prime_table_closed_form():
    table = []
    x = 1 << 40
    while True:
        p = po.find_greatest_prime(x)
        if p == 0:
            break
        if not table or p != table[-1]:
            table.append(p)
        x = p - 1
    assert len(table) == 55
    expected = set()
    for k in range(1, 29):
        expected.add(_largest_prime_le(1 << k))
        if k >= 2:
            expected.add(_largest_prime_le(3 << (k - 1)))
    expected.add(_largest_prime_le(3 << 27))  # 402653171 = largest prime <= 1.5 * 2^28
    assert sorted(table) == sorted(expected - {3} | {3})
    assert all(_is_prime(p) for p in table)
    for x in [2, 3, 10, 1000, 1021, 1022, 123456, 3145721, 3145720, 1 << 28, 10 ** 9]:
        p = po.find_greatest_prime(x)
        assert p <= x and p in table
        assert all(q > x or q <= p for q in table)


def test_sizing_matches_reference_logged_memory():
    # doc/data/qsketch/accuracy_0.5.dat, column Sub_Warp_LevelMemoryUsage (MiB)
    # = (n_low * 32 + 1024 * 128) / 2^20 for N = 2^15 .. 2^25 at WLF 0.5
    logged_05 = [0.874847, 1.624603, 3.124847, 6.124664, 12.124786, 24.124969, 48.124664,
                 96.124786, 192.124786, 384.124420, 768.124664]
    for k, mib in zip(range(15, 26), logged_05):
        n_low = po.number_of_buckets(1 << k, 8, 0.5)
        assert abs((n_low * 32 + 1024 * 128) / 2 ** 20 - mib) < 1e-5, (k, n_low)
    # accuracy_1.0.dat starts at N = 2^15 too
    logged_10 = [0.499786, 0.874847, 1.624603, 3.124847]
    for k, mib in zip(range(15, 19), logged_10):
        n_low = po.number_of_buckets(1 << k, 8, 1.0)
        assert abs((n_low * 32 + 1024 * 128) / 2 ** 20 - mib) < 1e-5, (k, n_low)
    assert po.number_of_buckets(1 << 22, 8, 0.5) == 3145721
    assert po.number_of_buckets(1 << 28, 8, 0.5) == 201326557
    assert po.number_of_buckets(1 << 28, 8, 1.0) == 100663291


# ---------------------------------------------------------------------------
# state machine
# ---------------------------------------------------------------------------
def _mask_bits(m):
    return [j for j in range(32) if (int(m) >> j) & 1]


def _sketch_with_mask(mask_bits, semantics=po.LITERAL, n_low=1, n_high=64):
    """Every key hashes to bucket 0 with the given mask (n_low = 1, 1 mask)."""
    m = sum(1 << j for j in mask_bits)
    return po.OracleSketch(n_low, n_high, 1, seed=1, masks=[m], semantics=semantics)


def test_hand_scenario_threshold_and_first_id():
    # one key, mask {1, 9, 20}: bytes 1, 9, 20 count to 129; insert 130 allocates
    s = _sketch_with_mask([1, 9, 20])
    keys = np.full(129, 77, dtype=np.uint32)
    s.insert(keys)
    low = s.low()
    b = low[:8].view(np.uint8)
    assert b[1] == b[9] == b[20] == 129 and b.sum() == 3 * 129
    assert low[7] == 0 and s.next == 1025
    assert s.query([77])[0] == 129
    s.insert(np.array([77], dtype=np.uint32))          # max selected = 129 > 128
    low = s.low()
    assert low[7] == 1025 and s.next == 1026            # v3:76 ids start at 1025
    g = s.high()[32:64]                                 # row 1025 - 1024 = 1
    # v3:187: the allocating insert adds to g[j mod 8]: {1, 1, 4}
    assert g[1] == 2 and g[4] == 1 and g.sum() == 3
    assert low[:7].view(np.uint8).sum() == 3 * 129      # low counters frozen
    # query: 129 + min(g[1], g[9], g[20]) = 129 + 0  (literal under-estimates by 1)
    assert s.query([77])[0] == 129
    s.insert(np.full(10, 77, dtype=np.uint32))          # linked: g[j] += 1
    g = s.high()[32:64]
    assert g[1] == 12 and g[9] == 10 and g[20] == 10 and g[4] == 1
    assert s.query([77])[0] == 129 + 10


def test_hand_scenario_paper_semantics():
    s = _sketch_with_mask([1, 9, 20], semantics=po.PAPER)
    s.insert(np.full(140, 5, dtype=np.uint32))
    g = s.high()[32:64]
    assert g[1] == g[9] == g[20] == 11 and g.sum() == 33
    assert s.query([5])[0] == 140                        # exact for a lone key


def test_hand_scenario_threshold_is_strict():
    # byte at exactly 128 still takes a low add (v3:160 "> 128")
    s = _sketch_with_mask([0, 5, 27])
    s.insert(np.full(128, 1, dtype=np.uint32))
    assert s.low()[7] == 0
    s.insert(np.array([1], dtype=np.uint32))
    assert s.low()[7] == 0 and s.low()[:8].view(np.uint8)[0] == 129
    s.insert(np.array([1], dtype=np.uint32))
    assert s.low()[7] == 1025


def _naive_model(keys, seed, masks, n_low, semantics, key_bytes=4):
    """Independent restatement: dictionary of 28-byte lists, explicit bit loops."""
    low = {}
    link = {}
    high = {}
    nxt = 1025
    for k in keys:
        k = int(k)
        h = (k * seed) & (0xFFFFFFFF if key_bytes == 4 else 0xFFFFFFFFFFFFFFFF)
        b = h % n_low
        m = int(masks[h % len(masks)])
        sel = [j for j in range(28) if m & (1 << j)]
        cnt = low.setdefault(b, [0] * 28)
        w = link.get(b, 0)
        if w > 1024:
            for j in sel:
                high[(w, j)] = high.get((w, j), 0) + 1
        elif max(cnt[j] for j in sel) > 128:
            if w == 0:
                w = nxt
                nxt += 1
                link[b] = w
            for j in sel:
                jj = j % 8 if semantics == po.LITERAL else j
                high[(w, jj)] = high.get((w, jj), 0) + 1
        else:
            for j in sel:
                cnt[j] += 1

    def query(k):
        k = int(k)
        h = (k * seed) & (0xFFFFFFFF if key_bytes == 4 else 0xFFFFFFFFFFFFFFFF)
        b = h % n_low
        m = int(masks[h % len(masks)])
        sel = [j for j in range(28) if m & (1 << j)]
        cnt = low.get(b, [0] * 28)
        lo = min(cnt[j] for j in sel)
        w = link.get(b, 0)
        hi = min(high.get((w, j), 0) for j in sel) if w > 1024 else 0
        return lo + hi

    return low, link, high, nxt, query


@pytest.mark.parametrize("semantics", [po.LITERAL, po.PAPER])
@pytest.mark.parametrize("n_low,universe,n", [(3, 5, 900), (7, 40, 3000), (13, 200, 6000),
                                               (61, 1000, 20000)])
def test_oracle_equals_naive_model(semantics, n_low, universe, n):
    rng = np.random.default_rng(n_low * 7 + semantics)
    # skewed small-universe stream forcing many allocations
    ranks = np.minimum(rng.zipf(1.3, n), universe) - 1
    keys = (ranks.astype(np.uint64) * 2654435761 % (1 << 32)).astype(np.uint32)
    s = po.OracleSketch(n_low, 4096, 1021, rng_seed=3, semantics=semantics)
    assert s.insert(keys) == 0
    low_d, link_d, high_d, nxt, query = _naive_model(keys, s.seed, s.masks(), n_low, semantics)
    low = s.low().reshape(n_low, 8)
    high = s.high().reshape(4096, 32)
    for b in range(n_low):
        got = list(low[b, :7].view(np.uint8)[:28])
        assert got == low_d.get(b, [0] * 28)
        assert low[b, 7] == link_d.get(b, 0)
    for w in range(1025, nxt):
        for j in range(32):
            assert high[w - 1024, j] == high_d.get((w, j), 0)
    assert s.next == nxt
    assert nxt > 1025  # the case exercised allocation
    uq = np.unique(keys)
    assert list(s.query(uq)) == [query(k) for k in uq]


def test_no_upgrade_case_is_textbook_count_min():
    """Without allocations the estimate is min over the selected byte counters,
    each = number of stream items hashing to (bucket, byte): computed by numpy
    sort/reduce independently of the oracle."""
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 1 << 32, 1 << 16, dtype=np.uint64).astype(np.uint32)
    n_low = 98299
    s = po.OracleSketch(n_low, 1024, 1021, rng_seed=9)
    s.insert(keys)
    assert s.next == 1025
    seed, masks = s.seed, s.masks()
    h = (keys.astype(np.uint64) * seed) & 0xFFFFFFFF
    b = (h % n_low).astype(np.int64)
    m = masks[(h % 1021).astype(np.int64)]
    cells = []
    for j in range(28):
        sel = ((m >> j) & 1).astype(bool)
        cells.append(b[sel] * 28 + j)
    cells = np.concatenate(cells)
    cnt = np.bincount(cells, minlength=n_low * 28).reshape(n_low, 28)
    uq = np.unique(keys)
    hq = (uq.astype(np.uint64) * seed) & 0xFFFFFFFF
    bq = (hq % n_low).astype(np.int64)
    mq = masks[(hq % 1021).astype(np.int64)]
    exp = np.array([min(cnt[bb, j] for j in range(28) if (int(mm) >> j) & 1) for bb, mm in zip(bq, mq)])
    assert np.array_equal(s.query(uq), exp.astype(np.uint32))
    # the stored bytes are the same counts
    assert np.array_equal(s.low().reshape(n_low, 8)[:, :7].copy().view(np.uint8).reshape(n_low, 28), cnt)


@pytest.mark.parametrize("semantics", [po.LITERAL, po.PAPER])
def test_conservation_and_bounds(semantics):
    from tests.helpers import zipf_keys
    keys = zipf_keys(1.1, 1 << 12, 1 << 18, seed=4)
    s = po.OracleSketch(1021, 8192, rng_seed=2, semantics=semantics)
    s.insert(keys)
    low = s.low().reshape(-1, 8)
    bytes_ = low[:, :7].copy().view(np.uint8)
    assert int(bytes_.sum(dtype=np.int64)) + int(s.high().sum(dtype=np.int64)) == 3 * keys.size
    unl = low[:, 7] == 0
    assert bytes_[unl].max() <= 129
    links = low[:, 7][~unl]
    assert np.array_equal(np.sort(links), np.arange(1025, s.next))
    uq, cnt = po.exact_counts(keys)
    est = s.query(uq).astype(np.int64)
    if semantics == po.PAPER:
        assert (est >= cnt).all()
    else:
        assert (est >= cnt.astype(np.int64) - 1).all()


@pytest.mark.parametrize("nthreads", [2, 3, 8])
def test_mt_insert_equals_sequential(nthreads):
    from tests.helpers import zipf_keys
    keys = zipf_keys(1.0, 1 << 13, 1 << 18, seed=8)
    a = po.OracleSketch(4093, 8192, rng_seed=5)
    b = po.OracleSketch(4093, 8192, rng_seed=5)
    for part in np.array_split(keys, 3):
        assert a.insert(part) == 0
        assert b.insert(part, nthreads=nthreads) == 0
    assert np.array_equal(a.low(), b.low())
    assert np.array_equal(a.high(), b.high())
    assert a.next == b.next
    uq = np.unique(keys)
    assert np.array_equal(a.query(uq), b.query(uq, nthreads=nthreads))


def test_pool_exhaustion_reports_error():
    from tests.helpers import zipf_keys
    keys = zipf_keys(1.0, 1 << 12, 1 << 17, seed=2)
    s = po.OracleSketch(1021, 16, rng_seed=1)
    assert s.insert(keys) == -3  # ORQ_ERR_POOL


def test_exact_counts_matches_dict():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 5000, 100000).astype(np.uint32)
    uq, cnt = po.exact_counts(keys)
    c = Counter(keys.tolist())
    assert list(uq) == sorted(c)
    assert list(cnt) == [c[k] for k in sorted(c)]
    k64 = (rng.integers(0, 300, 5000).astype(np.uint64) << np.uint64(40)) | np.uint64(7)
    uq, cnt = po.exact_counts(k64, key_bytes=8)
    c = Counter(k64.tolist())
    assert list(uq) == sorted(c) and list(cnt) == [c[k] for k in sorted(c)]


@pytest.mark.parametrize("wlf,lo,hi", [(0.5, 100.35, 100.95), (1.0, 102.0, 103.2)])
def test_accuracy_matches_reference_published(wlf, lo, hi):
    """Uniform u32 keys, N inserts, query the first N/4 keys of the stream:
    the reference logged mean(est/true)*100 of 100.55-100.73 (WLF 0.5) and
    102.41-102.81 (WLF 1.0) over N = 2^15..2^25
    (doc/data/qsketch/accuracy_{0.5,1.0}.dat, Sub_Warp_Level column)."""
    from tests.helpers import uniform_keys
    for k, seed in ((16, 1), (18, 2), (20, 3)):
        n = 1 << k
        keys = uniform_keys(n, seed)
        n_low = po.number_of_buckets(n, 8, wlf)
        s = po.OracleSketch(n_low, 1024, rng_seed=seed)
        s.insert(keys)
        q = keys[: n // 4]
        uq, cnt = po.exact_counts(keys)
        true = cnt[np.searchsorted(uq, q)]
        pct = (s.query(q) / true * 100.0).mean()
        assert lo < pct < hi, (wlf, k, pct)


# ---------------------------------------------------------------------------
# single-level sub-warp sketch (device_sketch_v2.h)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m", [8, 16, 32])
def test_single_level_matches_naive_model(m):
    rs = np.random.default_rng(m)
    n = 13
    o = po.OracleSingleLevel(n, m, 1021, rng_seed=5)
    masks = o.masks()
    assert all(bin(int(x)).count("1") == 3 and int(x) >> m == 0 for x in masks)
    keys = rs.integers(0, 200, 3000).astype(np.uint32)
    o.insert(keys)
    seed = o.seed
    table = np.zeros(n * m, dtype=np.int64)
    for k in keys.tolist():
        h = (k * seed) & 0xFFFFFFFF
        b, M = h % n, int(masks[h % 1021])
        for j in range(m):
            if M >> j & 1:
                table[b * m + j] += 1
    assert np.array_equal(o.table(), table)
    assert o.table().sum() == 3 * keys.size
    uq, cnt = np.unique(keys, return_counts=True)
    est = o.query(uq)
    assert np.all(est >= cnt)
    for k, e in zip(uq.tolist(), est.tolist()):
        h = (k * seed) & 0xFFFFFFFF
        b, M = h % n, int(masks[h % 1021])
        assert e == min(table[b * m + j] for j in range(m) if M >> j & 1)


def test_single_level_seed_and_masks_follow_the_two_level_engines():
    # same engines and call pattern as the two-level sketch: identical seed; the
    # masks draw from [0, m) instead of [0, 28)
    o2 = po.OracleSketch(1021, 1024, 1021, rng_seed=9)
    o1 = po.OracleSingleLevel(1021, 8, 1021, rng_seed=9)
    assert o1.seed == o2.seed
    m = po.Minstd(seed_bytes=po.seed_buffer(9, 1))
    ref = m.hashmask(1021, 8, 3, 0)
    assert np.array_equal(o1.masks(), ref.astype(np.uint32).reshape(-1))


# ---------------------------------------------------------------------------
# SF sketch (device_sketch_v42.h:207-480)
# ---------------------------------------------------------------------------
def _naive_sf(keys, n, m, times, seeds, masks, index, key_bytes=4):
    """The SF sketch written from v42's insert loop and update(), with Python
    integers and dicts (independent of the C restatement)."""
    mod = (1 << 32) if key_bytes == 4 else (1 << 64)
    split = {}
    for k in (int(x) for x in keys):
        h = (k * (seeds[0] % mod)) % mod
        b, mu = h % n, h % len(masks)
        for j in range(3):
            t = ((k * (seeds[j] % mod)) % mod) % times
            c = int(index[mu * 3 + j])
            split[(b, c, t)] = split.get((b, c, t), 0) + 1
    search = {}
    for (b, c, t), v in split.items():
        search[(b, c)] = max(search.get((b, c), 0), v)

    def query(k):
        h = (int(k) * (seeds[0] % mod)) % mod
        b, mu = h % n, h % len(masks)
        return min(search.get((b, c), 0) for c in range(m) if int(masks[mu]) >> c & 1)
    return split, search, query


@pytest.mark.parametrize("m,times,key_bytes", [(8, 3, 4), (16, 8, 4), (32, 2, 8), (8, 5, 8)])
def test_sf_matches_naive_model(m, times, key_bytes):
    rs = np.random.default_rng(m * times + key_bytes)
    dt = np.uint32 if key_bytes == 4 else np.uint64
    keys = rs.integers(0, 1 << 20, 3000).astype(dt) * dt(2654435761 if key_bytes == 4 else 11400714819323198485)
    keys = np.concatenate([keys, np.repeat(keys[:5], 40)])  # a few frequent keys
    rs.shuffle(keys)
    n = 61
    o = po.OracleSF(n, m, times, 1021, key_bytes=key_bytes, rng_seed=4)
    o.insert(keys[:2000])
    o.insert(keys[2000:])  # two calls: the update runs after each
    split, search, query = _naive_sf(keys, n, m, times, o.seeds(), o.masks(), o.index(), key_bytes)
    want = np.zeros(n * m * times, np.uint32)
    for (b, c, t), v in split.items():
        want[(b * m + c) * times + t] = v
    assert np.array_equal(o.split(), want)
    ws = np.zeros(n * m, np.uint32)
    for (b, c), v in search.items():
        ws[b * m + c] = v
    assert np.array_equal(o.search(), ws)
    q = np.concatenate([keys[:200], keys[:20] ^ dt(77)])
    assert [int(x) for x in o.query(q)] == [query(k) for k in q]
    # invariants: 3 increments per key; the estimate never falls below the count
    assert int(o.split().sum()) == 3 * keys.size
    u, c = np.unique(keys, return_counts=True)
    assert np.all(o.query(u) >= c)


def test_sf_mask_index_is_the_draw_order():
    """The index table lists each mask's bits in the order generate_hashmask
    drew them: replaying the draws with the engine restatement gives it."""
    o = po.OracleSF(101, 16, 3, 1021, rng_seed=9)
    eng = po.Minstd(seed_bytes=po.seed_buffer(9, 1))
    idx, masks = o.index(), o.masks()
    for i in range(1021):
        got = []
        while len(got) < 3:
            d = eng.uniform(0, 15)
            if d not in got:
                got.append(d)
        assert got == [int(x) for x in idx[i * 3:i * 3 + 3]], i
        assert int(masks[i]) == sum(1 << d for d in got)


def test_sf_with_one_split_is_the_single_level_sketch():
    """times = 1: every counter has one split, so the SF sketch is the
    single-level sketch with the same seed draws and masks."""
    rs = np.random.default_rng(5)
    keys = rs.integers(0, 1 << 16, 40000).astype(np.uint32)
    for m in (8, 16, 32):
        sf = po.OracleSF(4093, m, 1, 1021, seed_sz=2, rng_seed=6)
        sl = po.OracleSingleLevel(4093, m, 1021, seed_sz=2, rng_seed=6)
        sf.insert(keys)
        sl.insert(keys)
        assert np.array_equal(sf.search(), sl.table())
        assert np.array_equal(sf.query(keys[:500]), sl.query(keys[:500]))
