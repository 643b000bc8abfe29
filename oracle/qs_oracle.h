/*
 * qs_oracle.h -- CPU restatement of the QSketch two-level "Multi-level Sub-Warp
 * Sketch" (MSWS) semantics, used ONLY as a test checker and as the CPU baseline.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product library (qsketch_amd) links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it, and only to check or to time the
 * CPU path.
 *
 * Reference files restated here (paths relative to the reference checkout):
 *   include/sketch/device_sketch/device_sketch_v3.h      insert / query state machine
 *   include/sketch/device_sketch/device_hash_function.h:4-18   hash_mul_add
 *   lib/hashmask.cu:5-82                                  mask generation
 *   include/tool/tool.h:186-220                           RNG call pattern
 *   lib/qsketch.cu:50-115                                 prime table
 *   include/sketch/device_sketch/device_sketch.h:67-150, 268-319, 352-376
 *   include/sketch/host_sketch/host_sketch.h:15-63         exact counter
 *
 * Pinning status of each function is documented in the header comment of
 * qs_oracle.c.
 */
#ifndef QS_ORACLE_H
#define QS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- libstdc++ <random> restatement ------------------------------------ */
typedef struct { uint64_t x; } orq_minstd; /* std::minstd_rand0 */

void     orq_minstd_seed(orq_minstd *e, uint64_t s);
uint64_t orq_minstd_next(orq_minstd *e);
void     orq_seed_seq_generate(const uint32_t *v, size_t s, uint32_t *out, size_t n);
void     orq_minstd_seed_bytes(orq_minstd *e, const int8_t *buf, size_t n);
uint32_t orq_uniform_u32(orq_minstd *e, uint32_t a, uint32_t b);
void     orq_seed_buffer(uint64_t rng_seed, uint64_t stream, int8_t out[128]);

/* ---- sizing ------------------------------------------------------------- */
uint64_t orq_find_greatest_prime(uint64_t upper_bound);
uint64_t orq_number_of_buckets(uint64_t insert_keys, uint32_t m, double factor);

/* ---- masks -------------------------------------------------------------- */
/* out: n * ceil(m_bits/8) bytes, zero-filled by the caller */
void orq_generate_hashmask(orq_minstd *eng, uint32_t n, uint32_t m_bits,
                           uint32_t ones, uint32_t padding, uint8_t *out);

/* ---- the sketch --------------------------------------------------------- */
enum { ORQ_LITERAL = 0, ORQ_PAPER = 1 };
enum { ORQ_OK = 0, ORQ_ERR_POOL = -3, ORQ_ERR_ARG = -1, ORQ_ERR_NOMEM = -2 };

typedef struct {
    uint64_t  n_low;      /* low buckets, 32 B each: 28 byte counters + link word */
    uint64_t  n_high;     /* high buckets, 128 B each: 32 u32 counters */
    uint32_t  n_masks;    /* mask table entries (1021) */
    uint32_t  key_bytes;  /* 4 or 8 */
    uint32_t  semantics;  /* ORQ_LITERAL (v3:187 as written) or ORQ_PAPER */
    uint32_t  seed;       /* effective hash multiplier (sum of the seed_sz seeds) */
    uint32_t *masks;      /* n_masks little-endian u32 masks, bits 0..27 */
    uint32_t *low;        /* n_low * 8 words */
    uint32_t *high;       /* n_high * 32 words */
    uint32_t  next;       /* next high id, starts at 1025 */
    uint64_t  n_triggers; /* number of allocations so far */
    uint64_t  seed64;     /* the seed sum in 64 bits: 64-bit keys hash as
                             sum_i key * seed_i mod 2^64 = key * seed64 */
} orq_sketch;

/* Create with the reference's constructor call pattern (v3:544-563):
 * seeds drawn by uniform_int_distribution<uint>(1, UINT_MAX) from one engine,
 * masks by generate_hashmask(n_masks, 32, 3, 4) from a second engine.  The two
 * engines are seeded from 128-byte buffers derived from rng_seed. */
int  orq_create(orq_sketch *s, uint64_t n_low, uint64_t n_high, uint32_t n_masks,
                uint32_t seed_sz, uint32_t key_bytes, uint32_t semantics,
                uint64_t rng_seed);
/* Create with an explicit hash seed and mask table (for injected parameters). */
int  orq_create_explicit(orq_sketch *s, uint64_t n_low, uint64_t n_high,
                         uint32_t n_masks, uint32_t seed, const uint32_t *masks,
                         uint32_t key_bytes, uint32_t semantics);
void orq_destroy(orq_sketch *s);
void orq_clear(orq_sketch *s);

uint64_t orq_hash(const orq_sketch *s, uint64_t key);

/* Sequential stream-order insert.  Returns ORQ_ERR_POOL when an allocation
 * would exceed the high pool (the reference has no bounds check, v3:119). */
int  orq_insert(orq_sketch *s, const void *keys, size_t n);
/* Bucket-range partitioned insert on nthreads threads; state byte-identical to
 * orq_insert. */
int  orq_insert_mt(orq_sketch *s, const void *keys, size_t n, int nthreads);
void orq_query(const orq_sketch *s, const void *keys, size_t n, uint32_t *out);
void orq_query_mt(const orq_sketch *s, const void *keys, size_t n, uint32_t *out,
                  int nthreads);

/* ---- single-level sub-warp sketch (device_sketch_v2.h:67-283) ----------- */
/* n buckets of m u32 counters (m in {8, 16, 32}); 1021 masks of 3 distinct
 * bits in [0, m) from generate_hashmask(n_masks, m, 3, 0) (the constructor
 * call of device_sketch_v1.h:267-273); insert adds 1 to each selected counter
 * (v2:54-65), query takes their minimum (v2:155-190). */
typedef struct {
    uint64_t  n;
    uint32_t  m;
    uint32_t  n_masks;
    uint32_t  key_bytes;
    uint32_t  seed;
    uint32_t *masks;   /* bits 0..m-1 */
    uint32_t *table;   /* n * m counters */
    uint64_t  seed64;  /* the seed sum in 64 bits (64-bit keys) */
} orq_sl;
int  orq_sl_create(orq_sl *s, uint64_t n, uint32_t m, uint32_t n_masks, uint32_t seed_sz,
                   uint32_t key_bytes, uint64_t rng_seed);
void orq_sl_destroy(orq_sl *s);
void orq_sl_insert(orq_sl *s, const void *keys, size_t n);
void orq_sl_query(const orq_sl *s, const void *keys, size_t n, uint32_t *out);

/* ---- SF sketch with split counters (device_sketch_v42.h:207-480) -------- */
/* n buckets of m counters (m in {8, 16, 32}), each counter split into `times`
 * sub-counters.  Seeds: 1 + 3 groups of seed_sz draws (v42 ds.resize(1 +
 * HASH_MASK_ONES, ss)); group g acts through its sum s_g.  Masks:
 * generate_hashmask(n_masks, m, 3, 0) with the index of each mask's bits in
 * draw order.  Insert of k (v42 insert, host loop): h = k*s_0; b = h % n;
 * mu = h % n_masks; for j = 0..2: split[b][idx[mu][j]][(k*s_j) % times] += 1.
 * After each insert call (v42 update, update_table): search[b][c] =
 * max_t split[b][c][t].  Query (v42 search_kernel): min over the mask's bits c
 * of search[b][c]. */
typedef struct {
    uint64_t  n;
    uint32_t  m;
    uint32_t  times;
    uint32_t  n_masks;
    uint32_t  key_bytes;
    uint64_t  seed[3];   /* group sums (64-bit, as hash_mul_add sums for 64-bit keys) */
    uint32_t *masks;     /* bits 0..m-1 */
    uint8_t  *index;     /* n_masks * 3: each mask's bits in draw order */
    uint32_t *split;     /* n * m * times */
    uint32_t *search;    /* n * m */
} orq_sf;
int  orq_sf_create(orq_sf *s, uint64_t n, uint32_t m, uint32_t times, uint32_t n_masks, uint32_t seed_sz,
                   uint32_t key_bytes, uint64_t rng_seed);
void orq_sf_destroy(orq_sf *s);
void orq_sf_insert(orq_sf *s, const void *keys, size_t n);
void orq_sf_query(const orq_sf *s, const void *keys, size_t n, uint32_t *out);
/* generate_hashmask with the index output (lib/hashmask.cu:32-82): each
 * mask's ones in the order they were drawn */
void orq_generate_hashmask_idx(orq_minstd *eng, uint32_t n, uint32_t m_bits, uint32_t ones, uint32_t padding,
                               uint8_t *out, uint32_t *index);

/* ---- exact counting (sort + reduce-by-key) ------------------------------ */
/* returns number of distinct keys; uniq/counts sized n by the caller */
size_t orq_exact_counts(const void *keys, size_t n, uint32_t key_bytes,
                        void *uniq, uint32_t *counts);

#ifdef __cplusplus
}
#endif
#endif
