/*
 * qs_oracle.c -- CPU restatement of QSketch MSWS semantics (TEST INFRASTRUCTURE).
 *
 * This file is a checker.  The product library never links or loads it; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg call it.
 *
 * Function -> reference lines -> how it is pinned (tests/test_oracle*.py):
 *   orq_minstd_*           tool.h:189 (std::default_random_engine = minstd_rand0)
 *                          pinned: equality with libstdc++ std::minstd_rand0
 *                          (tests/golden/stdlib_rng.cpp, compiled with g++ here)
 *   orq_seed_seq_generate  tool.h:200-201 (std::seed_seq over 128 chars)
 *                          pinned: equality with libstdc++ std::seed_seq
 *   orq_uniform_u32        tool.h:212 and lib/hashmask.cu:61
 *                          (std::uniform_int_distribution<uint>)
 *                          pinned: equality with libstdc++, both the
 *                          down-scaling (0..27) and up-scaling (1..UINT_MAX) paths
 *   orq_find_greatest_prime  lib/qsketch.cu:50-115
 *                          pinned: every entry prime and the largest prime
 *                          <= 2^k or <= 3*2^(k-1) (brute-force primality)
 *   orq_number_of_buckets  device_sketch.h:371-376
 *                          pinned: the reference's own logged table sizes
 *                          (doc/data/qsketch/accuracy_*.dat memory column)
 *   orq_generate_hashmask  lib/hashmask.cu:32-82
 *                          pinned: invariants (3 distinct bits < 28) and
 *                          equality with the same call sequence run on libstdc++
 *   orq_insert / orq_query device_sketch_v3.h:87-191, 310-408 (canonical
 *                          sequential semantics, see below)
 *                          pinned: hand-derived scenarios (threshold 128,
 *                          first id 1025, v3:187 j mod 8 quirk), brute force
 *                          against an independent naive model, the textbook
 *                          count-min special case (no upgrade), the invariant
 *                          sum(low)+sum(high) = 3N, and the reference's own
 *                          published accuracy (mean est/true on uniform keys)
 *   orq_insert_mt          bucket-range partition of orq_insert
 *                          pinned: byte-identical state to orq_insert
 *   orq_sl_*               device_sketch_v2.h:67-283 (single-level sub-warp
 *                          sketch), construction device_sketch_v1.h:254-273
 *                          pinned: brute force against a naive Python model,
 *                          the conservation invariant sum = 3N, and the
 *                          textbook count-min bound (estimate >= true count)
 *   orq_sf_*               device_sketch_v42.h:207-480 (SF sketch: split
 *                          counters, max over the splits, min over the mask)
 *                          pinned: brute force against a naive Python model,
 *                          the invariants sum(split) = 3N and search = max of
 *                          the splits, estimate >= true count, and at
 *                          times = 1 equality with orq_sl (the same seeds and
 *                          masks: it is then the single-level sketch)
 *   orq_exact_counts       host_sketch.h:15-63 (std::unordered_map counter)
 *                          pinned: equality with a Python dict counter
 *
 * Canonical sequential semantics (per key, in stream order):
 *   h = hash(k); b = h % n_low; M = masks[h % n_masks]; w = link word of b
 *   if w > 1024:                          g[w][j] += 1 for j in M     (v3:111-129)
 *   elif max_{j in M} c_b[j] > 128:                                   (v3:160-161)
 *       if w == 0: w = next; next += 1                                (v3:68-85)
 *       g[w][q(j)] += 1 for j in M, q(j) = j mod 8 (v3:187) or j (paper)
 *   else:                                 c_b[j] += 1 for j in M      (v3:163-170)
 *   query: min_{j in M} c_b[j] + (w > 1024 ? min_{j in M} g[w][j] : 0)
 */
#define _GNU_SOURCE
#include "qs_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define LINK_START 1024u  /* lib/default.cu:45 NEXT_LEVEL_ID_START */
#define LOW_LIMIT  128u   /* v3:160 "max_count > 128" */
#define PROV_BIT   0x80000000u

/* ------------------------------------------------------------------------ */
/* std::minstd_rand0: x <- 16807 x mod (2^31 - 1)                           */
/* ------------------------------------------------------------------------ */
static const uint64_t MINSTD_M = 2147483647ull;

void orq_minstd_seed(orq_minstd *e, uint64_t s)
{
    uint64_t x = s % MINSTD_M;
    e->x = x == 0 ? 1 : x;
}

uint64_t orq_minstd_next(orq_minstd *e)
{
    e->x = (e->x * 16807ull) % MINSTD_M;
    return e->x;
}

/* std::seed_seq::generate ([rand.util.seedseq]); all arithmetic mod 2^32 */
void orq_seed_seq_generate(const uint32_t *v, size_t s, uint32_t *out, size_t n)
{
    if (n == 0)
        return;
    for (size_t i = 0; i < n; ++i)
        out[i] = 0x8b8b8b8bu;
    size_t t = (n >= 623) ? 11 : (n >= 68) ? 7 : (n >= 39) ? 5 : (n >= 7) ? 3 : (n - 1) / 2;
    size_t p = (n - t) / 2;
    size_t q = p + t;
    size_t m = (s + 1 > n) ? s + 1 : n;
    for (size_t k = 0; k < m; ++k) {
        uint32_t arg = out[k % n] ^ out[(k + p) % n] ^ out[(k + n - 1) % n];
        uint32_t r1 = 1664525u * (arg ^ (arg >> 27));
        uint32_t r2 = r1;
        if (k == 0)
            r2 += (uint32_t)s;
        else if (k <= s)
            r2 += (uint32_t)(k % n) + v[k - 1];
        else
            r2 += (uint32_t)(k % n);
        out[(k + p) % n] += r1;
        out[(k + q) % n] += r2;
        out[k % n] = r2;
    }
    for (size_t k = m; k < m + n; ++k) {
        uint32_t arg = out[k % n] + out[(k + p) % n] + out[(k + n - 1) % n];
        uint32_t r3 = 1566083941u * (arg ^ (arg >> 27));
        uint32_t r4 = r3 - (uint32_t)(k % n);
        out[(k + p) % n] ^= r3;
        out[(k + q) % n] ^= r4;
        out[k % n] = r4;
    }
}

/* eng.seed(std::seed_seq(buf, buf + n)) with buf a (signed) char array:
 * each char is sign-extended into the seed sequence (tool.h:200-201).  For
 * minstd_rand0, k = (lg(m) + 31) / 32 = 1, so the engine takes word 3 of a
 * 4-word generate() result. */
void orq_minstd_seed_bytes(orq_minstd *e, const int8_t *buf, size_t n)
{
    uint32_t *v = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i)
        v[i] = (uint32_t)(int32_t)buf[i];
    uint32_t arr[4];
    orq_seed_seq_generate(v, n, arr, 4);
    free(v);
    orq_minstd_seed(e, (uint64_t)arr[3]);
}

/* libstdc++ uniform_int_distribution<unsigned>::operator() for an engine with
 * range [1, 2^31-2]; __uctype is the 64-bit common type. */
static uint64_t uid_draw(orq_minstd *e, uint64_t a, uint64_t b)
{
    const uint64_t urngmin = 1, urngrange = MINSTD_M - 2; /* 2147483645 */
    const uint64_t urange = b - a;
    uint64_t ret;
    if (urngrange > urange) {
        const uint64_t uerange = urange + 1;
        const uint64_t scaling = urngrange / uerange;
        const uint64_t past = uerange * scaling;
        do
            ret = orq_minstd_next(e) - urngmin;
        while (ret >= past);
        ret /= scaling;
    } else if (urngrange < urange) {
        uint64_t tmp;
        do {
            const uint64_t uerngrange = urngrange + 1;
            tmp = uerngrange * uid_draw(e, 0, urange / uerngrange);
            ret = tmp + (orq_minstd_next(e) - urngmin);
        } while (ret > urange || ret < tmp);
    } else {
        ret = orq_minstd_next(e) - urngmin;
    }
    return ret + a;
}

uint32_t orq_uniform_u32(orq_minstd *e, uint32_t a, uint32_t b)
{
    return (uint32_t)uid_draw(e, a, b);
}

/* The reference seeds each engine from 128 bytes of /dev/urandom
 * (tool.h:192-201).  A deterministic build replaces those bytes with a
 * SplitMix64 stream of (rng_seed, stream); stream 0 = seed engine
 * (gpu_tool<uint>), stream 1 = mask engine (cpu_tool<uint>). */
static uint64_t splitmix64(uint64_t *st)
{
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orq_seed_buffer(uint64_t rng_seed, uint64_t stream, int8_t out[128])
{
    uint64_t st = rng_seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 16; ++i) {
        uint64_t z = splitmix64(&st);
        for (int k = 0; k < 8; ++k)
            out[i * 8 + k] = (int8_t)(uint8_t)(z >> (8 * k));
    }
}

/* ------------------------------------------------------------------------ */
/* sizing: lib/qsketch.cu:50-115, device_sketch.h:371-376                   */
/* ------------------------------------------------------------------------ */
static const uint64_t PRIMES[] = {
    402653171ull, 268435399ull, 201326557ull, 134217689ull, 100663291ull,
    67108859ull, 50331599ull, 33554393ull, 25165813ull, 16777213ull,
    12582893ull, 8388593ull, 6291449ull, 4194301ull, 3145721ull,
    2097143ull, 1572853ull, 1048573ull, 786431ull, 524287ull,
    393209ull, 262139ull, 196597ull, 131071ull, 98299ull,
    65521ull, 49139ull, 32749ull, 24571ull, 16381ull,
    12281ull, 8191ull, 6143ull, 4093ull, 3067ull,
    2039ull, 1531ull, 1021ull, 761ull, 509ull,
    383ull, 251ull, 191ull, 127ull, 89ull,
    61ull, 47ull, 31ull, 23ull, 13ull,
    11ull, 7ull, 5ull, 3ull, 2ull};

uint64_t orq_find_greatest_prime(uint64_t upper_bound)
{
    for (size_t i = 0; i < sizeof(PRIMES) / sizeof(PRIMES[0]); ++i)
        if (PRIMES[i] <= upper_bound)
            return PRIMES[i];
    return 0;
}

uint64_t orq_number_of_buckets(uint64_t insert_keys, uint32_t m, double factor)
{
    /* size_t table_sz_row = ((insert_keys_sz / factor) * HASH_MASK_ONES) / m; */
    double row = (((double)insert_keys / factor) * 3.0) / (double)m;
    return orq_find_greatest_prime((uint64_t)row);
}

/* ------------------------------------------------------------------------ */
/* masks: lib/hashmask.cu:5-82                                              */
/* ------------------------------------------------------------------------ */
void orq_generate_hashmask_idx(orq_minstd *eng, uint32_t n, uint32_t m_bits, uint32_t ones, uint32_t padding,
                               uint8_t *out, uint32_t *index)
{
    uint32_t sz = (m_bits + 7) / 8;
    for (uint32_t i = 0; i < n; ++i) {
        uint8_t *hm = out + (size_t)i * sz;
        uint32_t j = 0;
        while (j < ones) {
            uint32_t id = orq_uniform_u32(eng, 0, m_bits - padding - 1);
            uint8_t bit = (uint8_t)(1u << (id % 8));
            if ((hm[id / 8] & bit) == 0) {
                hm[id / 8] |= bit;
                if (index)
                    index[(size_t)i * ones + j] = id; /* lib/hashmask.cu:69-71 */
                ++j;
            }
        }
    }
}

void orq_generate_hashmask(orq_minstd *eng, uint32_t n, uint32_t m_bits,
                           uint32_t ones, uint32_t padding, uint8_t *out)
{
    uint32_t sz = (m_bits + 7) / 8;
    for (uint32_t i = 0; i < n; ++i) {
        uint8_t *hm = out + (size_t)i * sz;
        uint32_t j = 0;
        while (j < ones) {
            uint32_t id = orq_uniform_u32(eng, 0, m_bits - padding - 1);
            uint8_t bit = (uint8_t)(1u << (id % 8));
            if ((hm[id / 8] & bit) == 0) {
                hm[id / 8] |= bit;
                ++j;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* sketch                                                                   */
/* ------------------------------------------------------------------------ */
static int alloc_tables(orq_sketch *s)
{
    s->low = (uint32_t *)calloc((size_t)s->n_low * 8, sizeof(uint32_t));
    s->high = (uint32_t *)calloc((size_t)s->n_high * 32, sizeof(uint32_t));
    if (!s->low || !s->high) {
        free(s->low);
        free(s->high);
        s->low = s->high = NULL;
        return ORQ_ERR_NOMEM;
    }
    s->next = LINK_START + 1;
    s->n_triggers = 0;
    return ORQ_OK;
}

int orq_create_explicit(orq_sketch *s, uint64_t n_low, uint64_t n_high,
                        uint32_t n_masks, uint32_t seed, const uint32_t *masks,
                        uint32_t key_bytes, uint32_t semantics)
{
    memset(s, 0, sizeof(*s));
    if (n_low == 0 || n_high < 2 || n_masks == 0 || (key_bytes != 4 && key_bytes != 8) ||
        n_high > (uint64_t)0x7fffffffu - LINK_START)
        return ORQ_ERR_ARG;
    s->n_low = n_low;
    s->n_high = n_high;
    s->n_masks = n_masks;
    s->key_bytes = key_bytes;
    s->semantics = semantics;
    s->seed = seed;
    s->seed64 = seed;
    s->masks = (uint32_t *)malloc(sizeof(uint32_t) * n_masks);
    if (!s->masks)
        return ORQ_ERR_NOMEM;
    memcpy(s->masks, masks, sizeof(uint32_t) * n_masks);
    return alloc_tables(s);
}

int orq_create(orq_sketch *s, uint64_t n_low, uint64_t n_high, uint32_t n_masks,
               uint32_t seed_sz, uint32_t key_bytes, uint32_t semantics,
               uint64_t rng_seed)
{
    int8_t buf[128];
    orq_minstd seed_eng, mask_eng;
    orq_seed_buffer(rng_seed, 0, buf);
    orq_minstd_seed_bytes(&seed_eng, buf, 128);
    orq_seed_buffer(rng_seed, 1, buf);
    orq_minstd_seed_bytes(&mask_eng, buf, 128);

    /* Device_Seed::resize -> random(seed, seed_sz, 1, UINT_MAX): hash_mul_add
     * (device_hash_function.h:6-17) sums key * seed[i] in Hashed_T, so the
     * seeds act through their sum: mod 2^32 for 32-bit keys (Hashed_T = uint,
     * src/main.cu:5), mod 2^64 for 64-bit keys (the default Hashed_T =
     * size_t, where key * seed[i] is a 64-bit product). */
    uint64_t seed64 = 0;
    for (uint32_t i = 0; i < seed_sz; ++i)
        seed64 += orq_uniform_u32(&seed_eng, 1u, 0xffffffffu);
    const uint32_t seed = (uint32_t)seed64;

    uint8_t *mb = (uint8_t *)calloc((size_t)n_masks * 4, 1);
    uint32_t *masks = (uint32_t *)malloc(sizeof(uint32_t) * (n_masks ? n_masks : 1));
    if (!mb || !masks) {
        free(mb);
        free(masks);
        return ORQ_ERR_NOMEM;
    }
    orq_generate_hashmask(&mask_eng, n_masks, 32, 3, 4, mb);
    for (uint32_t i = 0; i < n_masks; ++i)
        masks[i] = (uint32_t)mb[4 * i] | ((uint32_t)mb[4 * i + 1] << 8) |
                   ((uint32_t)mb[4 * i + 2] << 16) | ((uint32_t)mb[4 * i + 3] << 24);
    int rc = orq_create_explicit(s, n_low, n_high, n_masks, seed, masks, key_bytes, semantics);
    s->seed64 = seed64;
    free(mb);
    free(masks);
    return rc;
}

void orq_destroy(orq_sketch *s)
{
    free(s->masks);
    free(s->low);
    free(s->high);
    memset(s, 0, sizeof(*s));
}

void orq_clear(orq_sketch *s)
{
    memset(s->low, 0, (size_t)s->n_low * 32);
    memset(s->high, 0, (size_t)s->n_high * 128);
    s->next = LINK_START + 1;
    s->n_triggers = 0;
}

uint64_t orq_hash(const orq_sketch *s, uint64_t key)
{
    if (s->key_bytes == 4)
        return (uint32_t)((uint32_t)key * s->seed);
    return key * s->seed64;
}

static inline uint64_t key_at(const orq_sketch *s, const void *keys, size_t i)
{
    return s->key_bytes == 4 ? ((const uint32_t *)keys)[i] : ((const uint64_t *)keys)[i];
}

static inline uint32_t byte_of(const uint32_t *bk, uint32_t j)
{
    return (bk[j >> 2] >> (8 * (j & 3))) & 0xffu;
}

/* High-level counters for link value w.  Provisional links (PROV_BIT) point
 * into a per-thread side array during orq_insert_mt. */
typedef struct {
    uint32_t *cnt;   /* cap * 32 */
    uint64_t *idx;   /* trigger stream index */
    uint64_t *bkt;   /* bucket */
    size_t    n, cap;
} prov_list;

static uint32_t *prov_push(prov_list *p, uint64_t idx, uint64_t b)
{
    if (p->n == p->cap) {
        size_t nc = p->cap ? p->cap * 2 : 64;
        p->cnt = (uint32_t *)realloc(p->cnt, nc * 32 * sizeof(uint32_t));
        p->idx = (uint64_t *)realloc(p->idx, nc * sizeof(uint64_t));
        p->bkt = (uint64_t *)realloc(p->bkt, nc * sizeof(uint64_t));
        p->cap = nc;
    }
    memset(p->cnt + p->n * 32, 0, 32 * sizeof(uint32_t));
    p->idx[p->n] = idx;
    p->bkt[p->n] = b;
    return p->cnt + 32 * p->n++;
}

/* One insert.  prov == NULL: real allocation from s->next. */
static inline int insert_one(orq_sketch *s, uint64_t key, uint64_t stream_idx,
                             prov_list *prov)
{
    uint64_t h = orq_hash(s, key);
    uint64_t b = h % s->n_low;
    uint32_t M = s->masks[h % s->n_masks];
    uint32_t *bk = s->low + b * 8;
    uint32_t w = bk[7];
    uint32_t *g;

    if (w > LINK_START) {
        g = (w & PROV_BIT) ? prov->cnt + 32 * (size_t)(w & ~PROV_BIT)
                           : s->high + (size_t)(w - LINK_START) * 32;
        for (uint32_t m = M; m; m &= m - 1)
            g[__builtin_ctz(m)] += 1;
        return ORQ_OK;
    }
    uint32_t mx = 0;
    for (uint32_t m = M; m; m &= m - 1) {
        uint32_t c = byte_of(bk, (uint32_t)__builtin_ctz(m));
        mx = c > mx ? c : mx;
    }
    if (mx > LOW_LIMIT) {
        if (w != 0)
            return ORQ_OK; /* reserved link values 1..1024: v3:183 skips the high add */
        if (prov) {
            g = prov_push(prov, stream_idx, b);
            bk[7] = PROV_BIT | (uint32_t)(prov->n - 1);
        } else {
            if ((uint64_t)(s->next - LINK_START) >= s->n_high)
                return ORQ_ERR_POOL;
            w = s->next++;
            s->n_triggers++;
            bk[7] = w;
            g = s->high + (size_t)(w - LINK_START) * 32;
        }
        for (uint32_t m = M; m; m &= m - 1) {
            uint32_t j = (uint32_t)__builtin_ctz(m);
            g[s->semantics == ORQ_LITERAL ? (j & 7u) : j] += 1;
        }
        return ORQ_OK;
    }
    for (uint32_t m = M; m; m &= m - 1) {
        uint32_t j = (uint32_t)__builtin_ctz(m);
        bk[j >> 2] += 1u << (8 * (j & 3));
    }
    return ORQ_OK;
}

int orq_insert(orq_sketch *s, const void *keys, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        int rc = insert_one(s, key_at(s, keys, i), i, NULL);
        if (rc)
            return rc;
    }
    return ORQ_OK;
}

typedef struct {
    orq_sketch *s;
    const void *keys;
    size_t n;
    uint64_t lo, hi;
    prov_list prov;
    uint32_t *out;
    const orq_sketch *cs;
} mt_job;

static void *insert_worker(void *arg)
{
    mt_job *j = (mt_job *)arg;
    orq_sketch *s = j->s;
    for (size_t i = 0; i < j->n; ++i) {
        uint64_t key = key_at(s, j->keys, i);
        uint64_t b = orq_hash(s, key) % s->n_low;
        if (b >= j->lo && b < j->hi)
            insert_one(s, key, i, &j->prov);
    }
    return NULL;
}

typedef struct { uint64_t idx; int t; size_t k; } trig_ref;

static int cmp_trig(const void *a, const void *b)
{
    const trig_ref *x = (const trig_ref *)a, *y = (const trig_ref *)b;
    return x->idx < y->idx ? -1 : x->idx > y->idx ? 1 : 0;
}

int orq_insert_mt(orq_sketch *s, const void *keys, size_t n, int nthreads)
{
    if (nthreads <= 1)
        return orq_insert(s, keys, n);
    mt_job *jobs = (mt_job *)calloc((size_t)nthreads, sizeof(mt_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].s = s;
        jobs[t].keys = keys;
        jobs[t].n = n;
        jobs[t].lo = s->n_low * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].hi = s->n_low * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, insert_worker, &jobs[t]);
    }
    size_t total = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        total += jobs[t].prov.n;
    }
    /* assign real ids in stream order of the triggering inserts */
    int rc = ORQ_OK;
    trig_ref *refs = (trig_ref *)malloc(sizeof(trig_ref) * (total ? total : 1));
    size_t k = 0;
    for (int t = 0; t < nthreads; ++t)
        for (size_t i = 0; i < jobs[t].prov.n; ++i) {
            refs[k].idx = jobs[t].prov.idx[i];
            refs[k].t = t;
            refs[k].k = i;
            ++k;
        }
    qsort(refs, total, sizeof(trig_ref), cmp_trig);
    if ((uint64_t)(s->next - LINK_START) + total > s->n_high)
        rc = ORQ_ERR_POOL;
    for (size_t i = 0; i < total; ++i) {
        prov_list *p = &jobs[refs[i].t].prov;
        uint64_t b = p->bkt[refs[i].k];
        if (rc == ORQ_OK) {
            uint32_t w = s->next++;
            s->n_triggers++;
            s->low[b * 8 + 7] = w;
            memcpy(s->high + (size_t)(w - LINK_START) * 32, p->cnt + 32 * refs[i].k,
                   32 * sizeof(uint32_t));
        } else {
            s->low[b * 8 + 7] = 0; /* state unspecified after a pool error; unlink */
        }
    }
    free(refs);
    for (int t = 0; t < nthreads; ++t) {
        free(jobs[t].prov.cnt);
        free(jobs[t].prov.idx);
        free(jobs[t].prov.bkt);
    }
    free(jobs);
    free(th);
    return rc;
}

static inline uint32_t query_one(const orq_sketch *s, uint64_t key)
{
    uint64_t h = orq_hash(s, key);
    uint64_t b = h % s->n_low;
    uint32_t M = s->masks[h % s->n_masks];
    const uint32_t *bk = s->low + b * 8;
    uint32_t w = bk[7];
    uint32_t lo = 0xffffffffu, hi = 0xffffffffu;
    for (uint32_t m = M; m; m &= m - 1) {
        uint32_t c = byte_of(bk, (uint32_t)__builtin_ctz(m));
        lo = c < lo ? c : lo;
    }
    if (w > LINK_START) {
        const uint32_t *g = s->high + (size_t)(w - LINK_START) * 32;
        for (uint32_t m = M; m; m &= m - 1) {
            uint32_t c = g[__builtin_ctz(m)];
            hi = c < hi ? c : hi;
        }
    } else {
        hi = 0;
    }
    return lo + hi;
}

void orq_query(const orq_sketch *s, const void *keys, size_t n, uint32_t *out)
{
    for (size_t i = 0; i < n; ++i)
        out[i] = query_one(s, key_at(s, keys, i));
}

static void *query_worker(void *arg)
{
    mt_job *j = (mt_job *)arg;
    for (size_t i = j->lo; i < j->hi; ++i)
        j->out[i] = query_one(j->cs, key_at(j->cs, j->keys, i));
    return NULL;
}

void orq_query_mt(const orq_sketch *s, const void *keys, size_t n, uint32_t *out,
                  int nthreads)
{
    if (nthreads <= 1) {
        orq_query(s, keys, n, out);
        return;
    }
    mt_job *jobs = (mt_job *)calloc((size_t)nthreads, sizeof(mt_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].cs = s;
        jobs[t].keys = keys;
        jobs[t].out = out;
        jobs[t].lo = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, query_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t)
        pthread_join(th[t], NULL);
    free(jobs);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* exact counting: LSD radix sort + reduce-by-key                           */
/* ------------------------------------------------------------------------ */
static void radix_sort_u64(uint64_t *a, size_t n, int nbytes)
{
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (int pass = 0; pass < nbytes; ++pass) {
        size_t cnt[257] = {0};
        int sh = 8 * pass;
        for (size_t i = 0; i < n; ++i)
            cnt[((a[i] >> sh) & 0xff) + 1]++;
        for (int d = 0; d < 256; ++d)
            cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; ++i)
            tmp[cnt[(a[i] >> sh) & 0xff]++] = a[i];
        memcpy(a, tmp, sizeof(uint64_t) * n);
    }
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* SF sketch (device_sketch_v42.h:207-480)                                   */
/* ------------------------------------------------------------------------ */
int orq_sf_create(orq_sf *s, uint64_t n, uint32_t m, uint32_t times, uint32_t n_masks, uint32_t seed_sz,
                  uint32_t key_bytes, uint64_t rng_seed)
{
    memset(s, 0, sizeof(*s));
    if (n == 0 || (m != 8 && m != 16 && m != 32) || times == 0 || n_masks == 0 || seed_sz == 0 ||
        (key_bytes != 4 && key_bytes != 8))
        return ORQ_ERR_ARG;
    int8_t buf[128];
    orq_minstd seed_eng, mask_eng;
    orq_seed_buffer(rng_seed, 0, buf);
    orq_minstd_seed_bytes(&seed_eng, buf, 128);
    orq_seed_buffer(rng_seed, 1, buf);
    orq_minstd_seed_bytes(&mask_eng, buf, 128);
    /* ds.resize(1 + HASH_MASK_ONES, ss) (v42:263): (1 + 3) * ss draws in a row;
     * insert uses groups 0..2 (seed + j * seed_sz, v42:346), group 0 also for
     * the bucket (v42:334) */
    for (uint32_t g = 0; g < 4; ++g) {
        uint64_t sum = 0;
        for (uint32_t i = 0; i < seed_sz; ++i)
            sum += orq_uniform_u32(&seed_eng, 1u, 0xffffffffu);
        if (g < 3)
            s->seed[g] = sum;
    }
    const uint32_t sz = (m + 7) / 8;
    uint8_t *mb = (uint8_t *)calloc((size_t)n_masks * sz, 1);
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * n_masks * 3);
    s->masks = (uint32_t *)malloc(sizeof(uint32_t) * n_masks);
    s->index = (uint8_t *)malloc((size_t)n_masks * 3);
    s->split = (uint32_t *)calloc((size_t)n * m * times, sizeof(uint32_t));
    s->search = (uint32_t *)calloc((size_t)n * m, sizeof(uint32_t));
    if (!mb || !idx || !s->masks || !s->index || !s->split || !s->search) {
        free(mb);
        free(idx);
        orq_sf_destroy(s);
        return ORQ_ERR_NOMEM;
    }
    orq_generate_hashmask_idx(&mask_eng, n_masks, m, 3, 0, mb, idx); /* v42:268-269 */
    for (uint32_t i = 0; i < n_masks; ++i) {
        uint32_t v = 0;
        for (uint32_t b = 0; b < sz; ++b)
            v |= (uint32_t)mb[(size_t)i * sz + b] << (8 * b);
        s->masks[i] = v;
        for (uint32_t j = 0; j < 3; ++j)
            s->index[(size_t)i * 3 + j] = (uint8_t)idx[(size_t)i * 3 + j];
    }
    free(mb);
    free(idx);
    s->n = n;
    s->m = m;
    s->times = times;
    s->n_masks = n_masks;
    s->key_bytes = key_bytes;
    return ORQ_OK;
}

void orq_sf_destroy(orq_sf *s)
{
    free(s->masks);
    free(s->index);
    free(s->split);
    free(s->search);
    memset(s, 0, sizeof(*s));
}

/* hash_mul_add of seed group g: k * s_g in Hashed_T (32 bits for 32-bit keys) */
static inline uint64_t sf_hash(const orq_sf *s, uint64_t k, uint32_t g)
{
    return s->key_bytes == 4 ? (uint32_t)((uint32_t)k * (uint32_t)s->seed[g]) : k * s->seed[g];
}

void orq_sf_insert(orq_sf *s, const void *keys, size_t n)
{
    for (size_t i = 0; i < n; ++i) { /* v42:330-355 */
        const uint64_t k = s->key_bytes == 4 ? ((const uint32_t *)keys)[i] : ((const uint64_t *)keys)[i];
        const uint64_t h = sf_hash(s, k, 0);
        const uint64_t b = h % s->n;
        const uint8_t *ix = s->index + (size_t)(h % s->n_masks) * 3;
        for (uint32_t j = 0; j < 3; ++j) {
            const uint64_t t = sf_hash(s, k, j) % s->times;
            s->split[(b * s->m + ix[j]) * s->times + t] += 1;
        }
    }
    /* update() (v42:302-327, update_table v42:72-95): every counter becomes
     * the largest of its splits */
    const size_t cells = (size_t)s->n * s->m;
    for (size_t c = 0; c < cells; ++c) {
        uint32_t v = 0;
        for (uint32_t t = 0; t < s->times; ++t)
            v = s->split[c * s->times + t] > v ? s->split[c * s->times + t] : v;
        s->search[c] = v;
    }
}

void orq_sf_query(const orq_sf *s, const void *keys, size_t n, uint32_t *out)
{
    for (size_t i = 0; i < n; ++i) { /* v42 search_kernel, calculation_s2_search v42:98-118 */
        const uint64_t k = s->key_bytes == 4 ? ((const uint32_t *)keys)[i] : ((const uint64_t *)keys)[i];
        const uint64_t h = sf_hash(s, k, 0);
        const uint32_t *row = s->search + (h % s->n) * s->m;
        uint32_t v = 0xffffffffu;
        for (uint32_t M = s->masks[h % s->n_masks]; M; M &= M - 1u) {
            const uint32_t c = row[__builtin_ctz(M)];
            v = c < v ? c : v;
        }
        out[i] = v;
    }
}

size_t orq_exact_counts(const void *keys, size_t n, uint32_t key_bytes, void *uniq,
                        uint32_t *counts)
{
    uint64_t *a = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i)
        a[i] = key_bytes == 4 ? ((const uint32_t *)keys)[i] : ((const uint64_t *)keys)[i];
    radix_sort_u64(a, n, key_bytes == 4 ? 4 : 8);
    size_t u = 0;
    for (size_t i = 0; i < n;) {
        size_t j = i;
        while (j < n && a[j] == a[i])
            ++j;
        if (key_bytes == 4)
            ((uint32_t *)uniq)[u] = (uint32_t)a[i];
        else
            ((uint64_t *)uniq)[u] = a[i];
        counts[u] = (uint32_t)(j - i);
        ++u;
        i = j;
    }
    free(a);
    return u;
}


/* ------------------------------------------------------------------------ */
/* single-level sub-warp sketch                                              */
/* ------------------------------------------------------------------------ */
int orq_sl_create(orq_sl *s, uint64_t n, uint32_t m, uint32_t n_masks, uint32_t seed_sz,
                  uint32_t key_bytes, uint64_t rng_seed)
{
    if (!s || n == 0 || (m != 8 && m != 16 && m != 32) || n_masks == 0 ||
        (key_bytes != 4 && key_bytes != 8))
        return ORQ_ERR_ARG;
    int8_t buf[128];
    orq_minstd seed_eng, mask_eng;
    orq_seed_buffer(rng_seed, 0, buf);
    orq_minstd_seed_bytes(&seed_eng, buf, 128);
    orq_seed_buffer(rng_seed, 1, buf);
    orq_minstd_seed_bytes(&mask_eng, buf, 128);
    uint64_t seed64 = 0; /* as in orq_create: 64-bit sum for 64-bit keys */
    for (uint32_t i = 0; i < seed_sz; ++i)
        seed64 += orq_uniform_u32(&seed_eng, 1u, 0xffffffffu);
    const uint32_t seed = (uint32_t)seed64;
    const uint32_t sz = (m + 7) / 8;
    uint8_t *mb = (uint8_t *)calloc((size_t)n_masks * sz, 1);
    s->masks = (uint32_t *)malloc(sizeof(uint32_t) * n_masks);
    s->table = (uint32_t *)calloc((size_t)n * m, sizeof(uint32_t));
    if (!mb || !s->masks || !s->table) {
        free(mb);
        free(s->masks);
        free(s->table);
        return ORQ_ERR_NOMEM;
    }
    orq_generate_hashmask(&mask_eng, n_masks, m, 3, 0, mb);
    for (uint32_t i = 0; i < n_masks; ++i) {
        uint32_t v = 0;
        for (uint32_t b = 0; b < sz; ++b)
            v |= (uint32_t)mb[(size_t)i * sz + b] << (8 * b);
        s->masks[i] = v;
    }
    free(mb);
    s->n = n;
    s->m = m;
    s->n_masks = n_masks;
    s->key_bytes = key_bytes;
    s->seed = seed;
    s->seed64 = seed64;
    return ORQ_OK;
}

void orq_sl_destroy(orq_sl *s)
{
    free(s->masks);
    free(s->table);
    s->masks = NULL;
    s->table = NULL;
}

static inline uint64_t sl_hash(const orq_sl *s, const void *keys, size_t i)
{
    if (s->key_bytes == 4)
        return (uint32_t)(((const uint32_t *)keys)[i] * s->seed);
    return ((const uint64_t *)keys)[i] * s->seed64;
}

void orq_sl_insert(orq_sl *s, const void *keys, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        const uint64_t h = sl_hash(s, keys, i);
        uint32_t *b = s->table + (h % s->n) * s->m;
        for (uint32_t M = s->masks[h % s->n_masks]; M; M &= M - 1)
            b[__builtin_ctz(M)] += 1u;
    }
}

void orq_sl_query(const orq_sl *s, const void *keys, size_t n, uint32_t *out)
{
    for (size_t i = 0; i < n; ++i) {
        const uint64_t h = sl_hash(s, keys, i);
        const uint32_t *b = s->table + (h % s->n) * s->m;
        uint32_t v = 0xffffffffu;
        for (uint32_t M = s->masks[h % s->n_masks]; M; M &= M - 1) {
            const uint32_t c = b[__builtin_ctz(M)];
            v = c < v ? c : v;
        }
        out[i] = v;
    }
}
