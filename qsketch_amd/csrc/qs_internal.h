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

// Bucket index and mask index of a key (hash_mul_add, device_hash_function.h:4-18,
// then v3:41-44).  32-bit keys hash in 32 bits (Hashed_T = uint, src/main.cu:5);
// 64-bit keys hash in 64 bits (the functor's default Hashed_T = size_t).
struct qs_hashpar {
    qs_fastdiv low;
    qs_fastdiv masks;
    uint32_t seed;
    uint32_t n_low;
    uint32_t n_masks;
};

template <int KB> struct qs_key;
template <> struct qs_key<4> { typedef uint32_t T; };
template <> struct qs_key<8> { typedef uint64_t T; };

__device__ static inline void qs_hash_index(uint32_t key, const qs_hashpar &hp, uint32_t &b,
                                            uint32_t &mu)
{
    uint32_t h = key * hp.seed;
    b = qs_fastmod(h, hp.low);
    mu = qs_fastmod(h, hp.masks);
}

__device__ static inline void qs_hash_index(uint64_t key, const qs_hashpar &hp, uint32_t &b,
                                            uint32_t &mu)
{
    uint64_t h = key * (uint64_t)hp.seed;
    b = (uint32_t)(h % (uint64_t)hp.n_low);
    mu = (uint32_t)(h % (uint64_t)hp.n_masks);
}

// One insert after hashing: bucket | mask index << bb | chunk-local stream
// index << (bb + bm), in 64 bits.  Built once in the first radix pass, so no
// later pass or kernel hashes again.
struct qs_pack {
    uint32_t bb;     // bucket bits
    uint32_t bm;     // mask-index bits
    uint32_t ishift; // bb + bm
    uint32_t mmask;  // (1 << bm) - 1
    uint64_t bmask;  // (1 << bb) - 1
};

__host__ __device__ static inline uint32_t it_b(uint64_t it, const qs_pack &pk)
{
    return (uint32_t)(it & pk.bmask);
}
__host__ __device__ static inline uint32_t it_mu(uint64_t it, const qs_pack &pk)
{
    return (uint32_t)(it >> pk.bb) & pk.mmask;
}
__host__ __device__ static inline uint32_t it_idx(uint64_t it, const qs_pack &pk)
{
    return (uint32_t)(it >> pk.ishift);
}

// Small device-resident control block of one sketch.
struct qs_dev_state {
    uint32_t next;      // next high id (starts at 1025)
    uint32_t n_ord;     // bucket segments queued for the in-order replay
    uint32_t n_long;    // long bucket segments (handled by the windowed tail kernel)
    uint32_t n_trig;    // allocations triggered in the current chunk
    uint32_t err;       // sticky error: 1 = pool exhausted
    uint32_t pad[3];
};

// One allocation found by the in-order replay (filled before ids are known).
struct qs_trig {
    uint32_t idx;        // chunk-local stream index of the triggering insert
    uint32_t bucket;     // low bucket
    uint32_t tail_start; // first sorted position handled by the tail kernel
    uint32_t pad;
    uint32_t hc[32];     // high counters accumulated up to tail_start (incl. trigger)
};

// Host-side handle.
struct qs_sketch {
    qs_params p;
    int device;
    uint32_t seed;
    uint32_t *h_masks;   // host copy
    qs_hashpar hp;
    uint64_t chunk;      // capacity (keys) of the chunk scratch
    int bucket_bits;     // bits of (n_low - 1)
    qs_pack pk;          // layout of a packed (bucket, mask index, stream index) item

    uint32_t *d_low;
    uint32_t *d_high;
    uint32_t *d_masks;
    qs_dev_state *d_st;
    uint32_t *d_longbm;     // bitmap over buckets: segment handled by the tail kernel
    uint64_t *d_it[2];      // packed chunk items sorted by bucket (double buffer)
    uint32_t *d_hist;       // radix histograms (shared by both sorts)
    uint32_t *d_ord;        // heads of segments for the in-order replay
    uint32_t *d_long;       // heads of long segments
    qs_trig *d_trig;
    uint32_t trig_cap;
    uint32_t *d_tk[2];      // trigger sort keys / values
    uint32_t *d_tv[2];
    uint32_t *d_tail_of;    // n_high: tail start per high bucket (current chunk)
    qs_dev_state *h_st;     // pinned host mirror

    int timing;             // record events around kernel launches
    void *tim;              // event lists (qs_sketch.hip)
    qs_timing tot;

    uint64_t keys_inserted;
    uint64_t chunks;
    uint64_t chunks_fix;
    uint64_t fix_keys;
    int sticky_err;
};

// Stable LSD radix sort of (key, value) uint32 pairs, 8 bits per pass, over the
// low key_bits bits.  Input in keys[0]/vals[0]; returns the index (0 or 1) of
// the buffer pair holding the sorted result.
int qs_sort_pairs(uint32_t *keys[2], uint32_t *vals[2], uint32_t *d_hist, uint32_t count,
                  int key_bits, hipStream_t stream);
size_t qs_sort_hist_words(uint32_t max_count);
// Exclusive scan of each row hist[d][0..ntiles) for d < nbins; totals[d] = row sum.
void qs_scan_digits(uint32_t *hist, uint32_t ntiles, uint32_t nbins, uint32_t *totals,
                    hipStream_t stream);

// insert pipeline (qs_insert.hip)
int qs_insert_chunks(qs_sketch *s, const void *d_keys, size_t n, hipStream_t stream);
int qs_ensure_scratch(qs_sketch *s, uint64_t chunk);
void qs_free_scratch(qs_sketch *s);
uint64_t qs_scratch_hist_words(uint64_t chunk);

// timing hooks (qs_sketch.hip)
struct qs_evpair {
    hipEvent_t a, b;
    int kind; // QS_TK_* kernel group, or QS_TIMING_KERNELS for a chunk marker
    uint64_t keys;
};
void qs_tim_begin(qs_sketch *s, qs_evpair &e, int kind, uint64_t keys, hipStream_t st);
void qs_tim_end(qs_sketch *s, qs_evpair &e, hipStream_t st);

// ---------------------------------------------------------------------------
// shared device helpers
// ---------------------------------------------------------------------------
__device__ static inline uint32_t qs_bucket(uint32_t key, const qs_hashpar &hp)
{
    return qs_fastmod(key * hp.seed, hp.low);
}

__device__ static inline uint32_t qs_bucket(uint64_t key, const qs_hashpar &hp)
{
    return (uint32_t)((key * (uint64_t)hp.seed) % (uint64_t)hp.n_low);
}

__device__ static inline uint32_t qs_maskidx(uint32_t key, const qs_hashpar &hp)
{
    return qs_fastmod(key * hp.seed, hp.masks);
}

__device__ static inline uint32_t qs_maskidx(uint64_t key, const qs_hashpar &hp)
{
    return (uint32_t)((key * (uint64_t)hp.seed) % (uint64_t)hp.n_masks);
}

template <int KB> using KeyT = typename qs_key<KB>::T;
