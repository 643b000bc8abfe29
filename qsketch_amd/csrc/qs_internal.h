// qs_internal.h -- internal types and device helpers of the MI355X QSketch.
//
// Layout in HBM (identical to the reference so that state export is a copy):
//   low  : uint32[n_low][8]   bytes 0..27 = byte counters, word 7 = link
//          (0 = unlinked, >1024 = high id)    device_sketch_v3.h:263, :119
//   high : uint32[n_high][32] counters of high id w at row (w - 1024)
//   masks: uint32[n_masks]    bit j selects byte counter j (bits 28..31 zero)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <random>

#include "../../include/qsketch.h"

#define QS_LINK_START 1024u
#define QS_LOW_LIMIT 128u
#define QS_MAX_MASKS 4096u
#define QS_WAVE 64

// ---------------------------------------------------------------------------
// Exact division by a run-time 32-bit constant (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, Fig. 4.1):
//   l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
//   t = mulhi(m, n), q = (t + ((n - t) >> 1)) >> (l - 1)      for d >= 2.
// Exact for every 32-bit n; the tests compare it with % by brute force.
// ---------------------------------------------------------------------------
struct qs_fastdiv {
    uint32_t d;
    uint32_t m;
    uint32_t sh; // l - 1
    uint32_t one; // d == 1
};

static inline qs_fastdiv qs_fastdiv_make(uint32_t d)
{
    qs_fastdiv f;
    f.d = d;
    if (d <= 1) {
        f.m = 0;
        f.sh = 0;
        f.one = 1;
        return f;
    }
    uint32_t l = 32u - (uint32_t)__builtin_clz(d - 1u);
    uint64_t two_l = 1ull << l;
    f.m = (uint32_t)(((1ull << 32) * (two_l - d)) / d + 1ull);
    f.sh = l - 1u;
    f.one = 0;
    return f;
}

__host__ __device__ static inline uint32_t qs_mulhi32(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

__host__ __device__ static inline uint32_t qs_fastmod(uint32_t n, const qs_fastdiv &f)
{
    if (f.one)
        return 0u;
    uint32_t t = qs_mulhi32(f.m, n);
    uint32_t q = (t + ((n - t) >> 1)) >> f.sh;
    return n - q * f.d;
}

// quotient n / d of the same division
__host__ __device__ static inline uint32_t qs_fastquot(uint32_t n, const qs_fastdiv &f)
{
    if (f.one)
        return n;
    const uint32_t t = qs_mulhi32(f.m, n);
    return (t + ((n - t) >> 1)) >> f.sh;
}

template <int KB> struct qs_key;
template <> struct qs_key<4> { typedef uint32_t T; };
template <> struct qs_key<8> { typedef uint64_t T; };
template <int KB> using KeyT = typename qs_key<KB>::T;

// Device descriptor of one sketch of a set.  A plain sketch is a set of one;
// the batched multi-sketch form holds thousands.  Every table of every sketch
// lives in one pool per table kind, addressed through these offsets.
struct qs_sdesc {
    uint64_t low_off;   // words into the low pool (multiple of 8)
    uint64_t high_off;  // words into the high pool (multiple of 32)
    uint64_t bm_off;    // bits into the long-segment bitmap (multiple of 32)
    uint64_t tail_off;  // entries into the tail-start pool (= high_off / 32)
    uint32_t masks_off; // words into the mask pool
    uint32_t n_masks;
    uint32_t n_low;
    uint32_t n_high;
    uint32_t seed;      // effective hash multiplier
    uint32_t paper;     // QS_PAPER semantics
    qs_fastdiv fd_low;
    qs_fastdiv fd_masks;
};

// Bucket index and mask index of a key (hash_mul_add, device_hash_function.h:4-18,
// then v3:41-44).  32-bit keys hash in 32 bits (Hashed_T = uint, src/main.cu:5);
// 64-bit keys hash in 64 bits (the functor's default Hashed_T = size_t).
__device__ static inline uint32_t qs_bucket_d(uint32_t key, const qs_sdesc &d)
{
    return qs_fastmod(key * d.seed, d.fd_low);
}
__device__ static inline uint32_t qs_bucket_d(uint64_t key, const qs_sdesc &d)
{
    return (uint32_t)((key * (uint64_t)d.seed) % (uint64_t)d.n_low);
}
__device__ static inline uint32_t qs_maskidx_d(uint32_t key, const qs_sdesc &d)
{
    return qs_fastmod(key * d.seed, d.fd_masks);
}
__device__ static inline uint32_t qs_maskidx_d(uint64_t key, const qs_sdesc &d)
{
    return (uint32_t)((key * (uint64_t)d.seed) % (uint64_t)d.n_masks);
}

// One insert after hashing, in 64 bits, built once in the first radix pass:
//   [ bucket : bb | sketch : sb | mask field : bm | stream index : ib ]
// The sort key is the low bb + sb bits (bucket within sketch), so a bucket
// segment is a run of equal "cell" values.  The stream index is the position
// within the sketch's part of the chunk, so allocations of one sketch can be
// ordered by (sketch, index) in 32 bits.
// The mask field holds the mask index h % n_masks, or (quotient form, 32-bit
// keys only) the hash quotient q = h / n_low, from which h = q * n_low + b and
// the mask index follow.  The quotient form is used when it is narrower: on
// large tables (n_low = 2^27.6 needs 5 bits, 1021 masks 10), which leaves the
// stream index room for 2^28-key chunks.
struct qs_pack {
    uint32_t bb, sb, bm, ib;
    uint32_t mshift;   // bb + sb
    uint32_t ishift;   // bb + sb + bm
    uint32_t mmask;    // (1 << bm) - 1
    uint32_t unm;      // every sketch's mask table size when they are all equal (else 0):
                       // sketch sid's masks then start at sid * unm in the mask pool
    uint32_t quot;     // the mask field holds the hash quotient
    uint32_t q_low;    // a single sketch's n_low (quotient form)
    qs_fastdiv q_fd;   // a single sketch's n_masks divisor (quotient form)
    uint64_t bmask;    // (1 << bb) - 1
    uint64_t cmask;    // (1 << (bb + sb)) - 1: the cell (sketch, bucket)
};

static inline uint32_t qs_bits_for(uint64_t n) // bits to hold 0 .. n-1 (at least 1)
{
    return n > 1 ? 64u - (uint32_t)__builtin_clzll(n - 1) : 1u;
}

// Field widths for a set of nsk sketches whose tables have min_low .. max_low
// buckets and at most max_masks masks, for keys of key_bytes bytes.  The
// stream index gets what is left of 64 bits, but at most 32 - sb bits (the
// allocation order key (sketch, index) is 32 bits) and at most 30 bits (device
// code shifts 32-bit values by ib; a shift by 32 would wrap to a shift by 0).
// Returns false when fewer than 12 bits remain.  A single sketch's n_low and
// n_masks go into the pack for the quotient form (set q_low, q_fd after).
static inline bool qs_make_pack(uint64_t max_low, uint64_t min_low, uint32_t nsk, uint64_t max_masks,
                                uint32_t key_bytes, qs_pack *pk)
{
    qs_pack p;
    p.bb = qs_bits_for(max_low);
    p.sb = nsk > 1 ? qs_bits_for(nsk) : 0u;
    p.bm = qs_bits_for(max_masks);
    p.quot = 0;
    p.q_low = 0;
    p.q_fd = qs_fastdiv_make(1u);
    if (key_bytes == 4 && min_low >= 1) {
        const uint32_t qb = qs_bits_for(0xffffffffull / min_low + 1ull); // quotients 0 .. (2^32 - 1) / min_low
        if (qb < p.bm) {
            p.bm = qb;
            p.quot = 1;
        }
    }
    int ib = 64 - (int)(p.bb + p.sb + p.bm);
    ib = ib < 32 - (int)p.sb ? ib : 32 - (int)p.sb;
    ib = ib < 30 ? ib : 30;
    if (ib < 12)
        return false;
    p.ib = (uint32_t)ib;
    p.mshift = p.bb + p.sb;
    p.ishift = p.bb + p.sb + p.bm;
    p.mmask = (1u << p.bm) - 1u;
    p.unm = 0;
    p.bmask = (1ull << p.bb) - 1ull;
    p.cmask = (1ull << (p.bb + p.sb)) - 1ull;
    *pk = p;
    return true;
}

__host__ __device__ static inline uint64_t it_cell(uint64_t it, const qs_pack &pk)
{
    return it & pk.cmask;
}
__host__ __device__ static inline uint32_t it_b(uint64_t it, const qs_pack &pk)
{
    return (uint32_t)(it & pk.bmask);
}
__host__ __device__ static inline uint32_t it_sid(uint64_t it, const qs_pack &pk)
{
    return (uint32_t)((it & pk.cmask) >> pk.bb);
}
__host__ __device__ static inline uint32_t it_mu(uint64_t it, const qs_pack &pk)
{
    return (uint32_t)(it >> pk.mshift) & pk.mmask;
}
__host__ __device__ static inline uint32_t it_idx(uint64_t it, const qs_pack &pk)
{
    return (uint32_t)(it >> pk.ishift);
}

// mask index of an item: the mask field, or from the quotient form's
// h = q * n_low + b (exact in 32 bits: it is the hash itself)
template <bool MULTI>
__host__ __device__ static inline uint32_t item_mu(uint64_t it, const qs_pack &pk, const qs_sdesc *D)
{
    const uint32_t f = it_mu(it, pk);
    if (!pk.quot)
        return f;
    const uint32_t b = it_b(it, pk);
    if (!MULTI)
        return qs_fastmod(f * pk.q_low + b, pk.q_fd);
    const qs_sdesc &d = D[it_sid(it, pk)];
    return qs_fastmod(f * d.n_low + b, d.fd_masks);
}

// Device-resident control block shared by a set.
// sticky device error bits
#define QS_EF_POOL 1u // a high pool is exhausted
#define QS_EF_SORT 2u // a radix scatter computed a position outside the chunk

static inline int qs_err_code(uint32_t err)
{
    return (err & QS_EF_SORT) ? QS_ERR_INTERNAL : err ? QS_ERR_POOL : QS_OK;
}

struct qs_dev_state {
    uint32_t n_ord;     // bucket segments queued for the in-order replay
    uint32_t n_long;    // long bucket segments (handled by the windowed tail kernel)
    uint32_t n_trig;    // allocations triggered in the current chunk
    uint32_t err;       // sticky error bits (QS_EF_*)
    // running totals since create / clear (statistics only)
    unsigned long long replay_keys;   // inserts replayed in stream order
    unsigned long long ord_segments;  // segments sent to the replay
    unsigned long long long_segments; // segments handled by the tail kernel
    unsigned long long chunks_ord;    // chunks with at least one replayed segment
};
// k_segw reserves entries of both lists with one 64-bit atomic on (n_ord, n_long)
static_assert(offsetof(qs_dev_state, n_ord) % 8 == 0 && offsetof(qs_dev_state, n_long) == offsetof(qs_dev_state, n_ord) + 4,
              "n_ord and n_long form one aligned 64-bit word");

// One allocation found by the in-order replay (filled before ids are known).
struct qs_trig {
    uint32_t idx;        // stream index of the triggering insert (within its sketch)
    uint32_t bucket;     // low bucket
    uint32_t tail_start; // first sorted position handled by the tail kernel
    uint32_t sid;        // sketch
    uint32_t hc[32];     // high counters accumulated up to tail_start (incl. trigger)
};

// Device scratch of the insert pipeline for chunks of up to `chunk` keys.  A
// handle allocates its own (qs_ensure_scratch); the allocation-free insert
// carves one out of a static device array instead.  Nothing in it needs to
// start at zero: every buffer is written before it is read within a chunk.
struct qs_scratch {
    uint64_t chunk;         // capacity (keys)
    uint64_t *it[2];        // packed chunk items sorted by cell (double buffer)
    uint32_t *hist;         // radix histograms (both sorts), qs_scratch_hist_words(chunk) words
    uint8_t *dig;           // the next radix pass's digit of each sorted item (one byte per item)
    uint32_t *ord;          // heads of segments for the in-order replay
    uint32_t *lng;          // heads of long segments
    qs_trig *trig;          // allocations found by the replay
    uint32_t trig_cap;      // entries of trig, tk, tv: at most the pool (sum of n_high) and the chunk
    uint32_t *tk[2];        // trigger sort keys / values
    uint32_t *tv[2];
};

// Host-side handle: a set of nsk sketches (nsk == 1 for the plain API).
struct qs_sketch {
    uint32_t nsk;
    uint32_t key_bytes;
    int device;
    qs_params *ps;       // per sketch
    uint32_t *seeds;     // per sketch (effective multiplier)
    uint32_t *h_masks;   // concatenated mask tables (host copy)
    qs_sdesc *h_desc;
    qs_sdesc *d_desc;
    uint64_t low_words, high_words, mask_words, bm_words, tail_words;
    qs_pack pk;
    uint64_t max_stream; // longest per-sketch part of one chunk (2^ib)

    uint32_t *d_low;
    uint32_t *d_high;
    uint32_t *d_masks;
    uint32_t *d_next;       // per sketch: next high id (starts at 1025)
    qs_dev_state *d_st;
    uint32_t *d_longbm;     // bitmap over (sketch, bucket) cells
    uint32_t *d_tail_of;    // per high bucket: tail start (current chunk)
    uint64_t *d_offs;       // per sketch: start of its keys in the current batch (+ end)
    uint64_t *h_offs;       // pinned staging for d_offs

    qs_scratch sc;          // the insert scratch this handle allocates (qs_ensure_scratch)
    qs_dev_state *h_st;     // pinned host mirror

    int timing;             // record events around kernel launches
    void *tim;              // event lists (qs_sketch.hip)
    qs_timing tot;

    uint64_t keys_inserted;
    uint64_t chunks;
    int sticky_err;
    int graphs;             // live captured graphs (they hold the scratch pointers)
    uint32_t seg_w;         // segment kernel shape: group width (0 = the 64-lane k_seg)
    uint32_t seg_ipl;       //   and items per lane
    int hot_filter;         // blocking inserts may use the hot-key filter (qs_hot.hip)
    void *hot;              // its scratch
    uint64_t hot_keys;      // inserts counted by the filter since create / clear
};

// hot-key filter (qs_hot.hip): blocking insert of 32-bit keys into a plain
// sketch; *used = 0 and nothing inserted when the call does not qualify
int qs_insert_hot(qs_sketch *s, const void *d_keys, uint64_t n, hipStream_t stream, int *used);
void qs_hot_free(qs_sketch *s);
// the same filter for a sink whose inserts all commute: the keys split into
// hot keys, counted per table slot (tkey[slot]; part[b * nslots + slot] for
// persistent block b < nblk, summed by the caller) and cold keys (stream
// order), all in the filter's scratch *state until the next call
struct qs_hot_split {
    int split;             // 0: nothing split (short call or too little skew)
    const uint32_t *cold;
    uint64_t ncold;
    const uint32_t *part;
    uint32_t nblk;
    const uint32_t *tkey;
    uint32_t nslots;
};
int qs_hot_split_keys(void **state, const uint32_t *keys, uint64_t n, hipStream_t stream, qs_hot_split *out);
void qs_hot_state_free(void *state);

// Stable LSD radix sort of (key, value) uint32 pairs, 8 bits per pass, over the
// low key_bits bits.  Input in keys[0]/vals[0]; returns the index (0 or 1) of
// the buffer pair holding the sorted result.
int qs_sort_pairs(uint32_t *keys[2], uint32_t *vals[2], uint32_t *d_hist, uint32_t count,
                  int key_bits, hipStream_t stream);
// Same, with the element count read from device memory at run time (max_count
// sizes the grid); capture-safe: no host synchronisation.
int qs_sort_pairs_dev(uint32_t *keys[2], uint32_t *vals[2], uint32_t *d_hist,
                      const uint32_t *d_count, uint32_t max_count, int key_bits, hipStream_t stream);
size_t qs_sort_hist_words(uint32_t max_count);
// Exclusive scan of each row hist[d][0..ntiles) for d < nbins; totals[d] = row sum.
void qs_scan_digits(uint32_t *hist, uint32_t ntiles, uint32_t nbins, uint32_t *totals,
                    hipStream_t stream);

// insert pipeline (qs_insert.hip).  offs: host array of nsk + 1 batch offsets
// (sketch i owns keys[offs[i], offs[i+1])).  blocking: synchronise at the end
// and report a pool error; otherwise the call never waits on the device.
// d_offs: device copy of offs (read by the kernels when nsk > 1); may be null
// for a single sketch.  capture: the launch sequence is being recorded into a
// graph (no timing events, no synchronisation).
// fixed: run in this scratch (null: the handle's own, allocated on demand); the
// batch must then fit one chunk of it.
int qs_insert_set(qs_sketch *s, const void *d_keys, const uint64_t *offs, const uint64_t *d_offs,
                  hipStream_t stream, bool blocking, bool capture, const qs_scratch *fixed = nullptr);
int qs_ensure_scratch(qs_sketch *s, uint64_t chunk);
// allocation entries the insert scratch needs for a chunk of `chunk` keys
uint32_t qs_trig_cap(const qs_sketch *s, uint64_t chunk);
// Lay out a scratch for chunk keys and tcap allocations from base (256-byte
// aligned pieces); returns the bytes it spans.  base may be null to size it.
uint64_t qs_scratch_bytes(uint64_t chunk, uint32_t tcap, qs_scratch *sc, uint8_t *base);
// segment kernel shapes the insert supports (group width x items per lane)
bool qs_seg_shape_ok(uint32_t w, uint32_t ipl);
// keys per chunk for a batch with host offsets offs
uint64_t qs_chunk_keys(const qs_sketch *s, const uint64_t *offs);
void qs_free_scratch(qs_sketch *s);
uint64_t qs_scratch_hist_words(uint64_t chunk);

// the reference's seed and mask generation (qs_sketch.hip): the libstdc++
// engine seeded from a 128-byte buffer of (rng_seed, stream), and
// generate_hashmask's draw of 3 distinct bits in [0, m_bits - padding)
void qs_seed_engine(std::default_random_engine &eng, uint64_t rng_seed, uint64_t stream);
void qs_make_masks(std::default_random_engine &eng, uint32_t n, uint32_t m_bits, uint32_t padding,
                   uint32_t *out);

// timing hooks (qs_sketch.hip)
struct qs_evpair {
    hipEvent_t a, b;
    int kind; // QS_TK_* kernel group, or QS_TIMING_KERNELS for a chunk marker
    uint64_t keys;
    bool shared_a; // a is the previous pair's b (one record ends one group and begins the next)
};
void qs_tim_begin(qs_sketch *s, qs_evpair &e, int kind, uint64_t keys, hipStream_t st);
void qs_tim_end(qs_sketch *s, qs_evpair &e, hipStream_t st);
// end `prev` and begin `next` with one event: for adjacent groups, with
// nothing launched between (each record idles the GPU for microseconds)
void qs_tim_next(qs_sketch *s, qs_evpair &prev, qs_evpair &next, int kind, uint64_t keys, hipStream_t st);

