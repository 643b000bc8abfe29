/*
 * qsketch.h -- C ABI of the MI355X-native QSketch (two-level Multi-level
 * Sub-Warp Sketch, "MSWS").
 *
 * The reference has no library boundary: its host API is the C++ virtual
 * interface Sketch<Key_T, Count_T> (include/sketch/sketch.h:5-28) implemented
 * by Sketch_GPU_Sub_Warp_Level (include/sketch/device_sketch/device_sketch_v3.h:
 * 524-634), constructed with the default.h parameter set (lib/default.cu:29-63).
 * Each entry point below replaces one member of that interface; the comment on
 * each cites the member it replaces.
 *
 * Conventions (same as the reference unless stated):
 *   - keys and counts are DEVICE pointers (the reference passes device buffers,
 *     include/test/test.h:143-170);
 *   - qs_insert / qs_query are blocking, like the reference's
 *     cudaDeviceSynchronize after each launch (v3:592, v3:620);
 *   - a handle is used by one host thread at a time (no internal locks, like
 *     the reference);
 *   - errors are negative return codes instead of the reference's undefined
 *     behaviour (unchecked high pool, v3:119; release builds ignore errors,
 *     include/qsketch.h:104-108).
 *
 * Semantics: every estimate equals the sequential stream-order restatement of
 * v3 insert/search (v3:87-191, v3:310-408): the GPU insert is deterministic,
 * including the allocation order of high buckets.
 */
#ifndef QSKETCH_H
#define QSKETCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------- */
#define QS_OK                 0
#define QS_ERR_PARAM         -1  /* invalid parameter / argument             */
#define QS_ERR_NOMEM         -2  /* device or host allocation failed         */
#define QS_ERR_POOL          -3  /* high-bucket pool exhausted; state invalid until qs_clear */
#define QS_ERR_HIP           -4  /* a HIP runtime call failed                */
#define QS_ERR_STATE         -5  /* imported state is malformed              */
#define QS_ERR_NODEVICE      -6  /* no usable GPU                            */
#define QS_ERR_INTERNAL      -8  /* a device-side consistency check failed; state invalid until qs_clear */

/* ---- insert semantics of the allocation-triggering insert --------------- */
#define QS_LITERAL 0 /* v3:187 as written: trigger insert adds to g[j mod 8] */
#define QS_PAPER   1 /* doc/main.tex:433 intent: trigger insert adds to g[j]  */

/* Parameter set.  Defaults are lib/default.cu:39-54 and the reference
 * program's constructor call src/main.cu:390-391:
 *   Sketch_GPU_Sub_Warp_Level(n, 8, 1024, 32, 1, 1021).                     */
typedef struct qs_params {
    uint64_t n_low;      /* number of 32-byte low buckets (v3:544 _n_low)          */
    uint32_t m_low;      /* words per low bucket: must be 8 (v3:544 _m_low)        */
    uint32_t m_high;     /* words per high bucket: must be 32 (v3:545 _m_high)     */
    uint64_t n_high;     /* high buckets incl. the unused bucket 0 (v3:545 _n_high)*/
    uint32_t seed_sz;    /* hash seeds summed by hash_mul_add (lib/default.cu:52)  */
    uint32_t n_masks;    /* mask table size, 1021 (lib/default.cu:54)             */
    uint32_t mask_ones;  /* bits per mask: must be 3 (lib/default.cu:53)           */
    uint32_t link_start; /* NEXT_LEVEL_ID_START: must be 1024 (lib/default.cu:45)  */
    uint32_t low_limit;  /* promote when a selected byte exceeds it: must be 128 (v3:160) */
    uint32_t key_bytes;  /* 4 (uint32 keys, src/main.cu:3) or 8 (uint64 keys)      */
    int32_t  semantics;  /* QS_LITERAL (default) or QS_PAPER                      */
    uint32_t hash_seed;  /* 0: draw the seeds from rng_seed; else use this value   */
    uint64_t rng_seed;   /* replaces the reference's /dev/urandom engine seeding (tool.h:192-201) */
    uint64_t chunk_keys; /* keys per internal insert chunk (0 = default 2^28, capped by the packing) */
} qs_params;

typedef struct qs_sketch qs_sketch;

/* Fill *p with the default parameter set, n_low sized by the reference rule
 * Sketch_GPU::number_of_buckets(expected_keys, 8, wlf)
 * (device_sketch.h:371-376 + find_greatest_prime, lib/qsketch.cu:50-115).
 * Replaces the reference program's sizing call, src/main.cu:371-372. */
int qs_default_params(qs_params *p, uint64_t expected_keys, double wlf);

/* find_greatest_prime (lib/qsketch.cu:50-115). */
uint64_t qs_find_greatest_prime(uint64_t upper_bound);

/* Construct: allocate and zero both tables, draw the seeds and build the mask
 * table.  Replaces the constructor Sketch_GPU_Sub_Warp_Level(...) (v3:544-563)
 * and Sketch_GPU(...) (device_sketch.h:352-369).  Runs on the current HIP
 * device. */
int qs_create(const qs_params *p, qs_sketch **out);

/* Free all device memory.  Replaces ~Sketch_GPU (device_sketch.h:387-394). */
void qs_destroy(qs_sketch *s);

/* Insert n keys (device pointer to uint32 or uint64, per key_bytes).
 * Replaces virtual int Sketch::insert(Key_T*, size_t) (sketch.h:7; v3:570-599). */
int qs_insert(qs_sketch *s, const void *d_keys, size_t n);

/* Estimate the counts of n keys into d_counts (device, uint32[n]).
 * Replaces virtual int Sketch::search(Key_T*, size_t, Count_T*) (sketch.h:8;
 * v3:600-632). */
int qs_query(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts);

/* Zero both tables and reset the allocator.  Replaces virtual void
 * Sketch::clear() (sketch.h:14; device_sketch.h:115-124, 378-381). */
int qs_clear(qs_sketch *s);

/* Stream-ordered variants (no synchronisation; stream is a hipStream_t or
 * NULL for the null stream).  Build additions with no reference counterpart:
 * the reference only has blocking calls on the default stream.  qs_insert_async
 * still performs host-side chunking; errors are reported by the next blocking
 * call or qs_sync. */
int qs_insert_async(qs_sketch *s, const void *d_keys, size_t n, void *stream);
int qs_query_async(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts,
                   void *stream);
int qs_sync(qs_sketch *s);

/* Hot-key filter of the blocking qs_insert (build addition; on by default).
 * A call of at least 2^26 32-bit keys into a plain sketch whose first 2^20
 * keys show skew (keys occurring 8 or more times there covering 30% or more
 * of them) inserts its first 2^25 keys the ordinary way; after that, the
 * inserts of those frequent keys whose bucket is now linked are counted and
 * added to their high rows at once (such inserts commute), and only the other
 * keys go through the ordered insert.  The result is the same state, bit for
 * bit.  enable = 0 turns it off for this handle. */
int qs_set_hot_filter(qs_sketch *s, int enable);

/* Allocation-free insert.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * qs_insert_async(s, d_keys, n, stream) and stream-ordered like it (errors
 * surface at the next blocking call or qs_sync), but the insert scratch lives
 * in one static device array of 33 MiB per GPU, held from library load until
 * the process ends, so the call allocates and frees nothing.  Limits: a plain
 * sketch (a set of one), n <= QS_NOALLOC_MAX_KEYS, and at most
 * QS_NOALLOC_MAX_ALLOCS possible high-bucket allocations, i.e.
 * min(n_high, n) <= QS_NOALLOC_MAX_ALLOCS; otherwise QS_ERR_PARAM.  The
 * current device must be the sketch's.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (several streams, several
 * threads) race and corrupt each other's scratch; order them on one stream. */
#define QS_NOALLOC_MAX_KEYS   (1u << 20)
#define QS_NOALLOC_MAX_ALLOCS (1u << 16)
int qs_insert_noalloc(qs_sketch *s, const void *d_keys, size_t n, void *stream);

/* ---- state import / export ---------------------------------------------- */
/* Fills the reference layout: low  = uint32[n_low*8]  (word 7 = link),
 * high = uint32[n_high*32], masks = uint32[n_masks] (little-endian bit j = byte
 * counter j), next = the next high id.  Host pointers; any may be NULL.
 * Fills the gap left by the declared-but-undefined read_file / write_file
 * (include/qsketch.h:70-71). */
int qs_export_state(qs_sketch *s, uint32_t *low, uint32_t *high, uint32_t *masks,
                    uint32_t *seed, uint32_t *next);
/* Replace the whole state.  Host pointers, none may be NULL.  QS_ERR_STATE
 * unless the state is one the insert can reach: 1025 <= next <= 1024 + n_high;
 * masks of 3 bits below bit 28; link words 0 or in [1025, next), each id of
 * [1025, next) held by exactly one bucket; byte counters <= 129; the high
 * rows of row 0 and of ids >= next zero. */
int qs_import_state(qs_sketch *s, const uint32_t *low, const uint32_t *high,
                    const uint32_t *masks, uint32_t seed, uint32_t next);

/* ---- statistics (build addition) -----------------------------------------
 * For a set of sketches (qs_multi_create) the sizes and linked_buckets are
 * sums over the set; next_id and seed are sketch 0's.                       */
typedef struct qs_stats {
    uint64_t n_low, n_high, n_masks;
    uint64_t bytes_low, bytes_high;     /* table sizes, cf. show_memory_usage
                                           (device_sketch.h:403-407) */
    uint64_t keys_inserted;
    uint64_t chunks;                    /* internal insert chunks processed */
    uint64_t chunks_with_fixup;         /* chunks with at least one in-order replay */
    uint64_t fixup_keys;                /* inserts replayed in stream order */
    uint64_t linked_buckets;            /* high buckets in use (= next - 1025) */
    uint32_t next_id;
    uint32_t seed;
    uint64_t ordered_segments;          /* bucket segments sent to the in-order replay */
    uint64_t long_segments;             /* bucket segments handled by the tail pass */
    uint32_t n_sketches;
    uint32_t pool_error;                /* 1 once a high pool overflowed */
    uint64_t hot_keys;                  /* inserts counted by the hot-key filter */
} qs_stats;
int qs_get_stats(qs_sketch *s, qs_stats *st);

/* ---- kernel timing (build addition) --------------------------------------
 * When enabled, the library records HIP events on the launch stream around
 * every kernel launch of the insert path and of the query (the reference only
 * timed whole blocking calls on the host, include/tool/tool.h:26-50).
 * Enabling resets the totals; qs_get_timing synchronises and resolves them.
 * kernel_ms[k] / kernel_launches[k] per kernel group k:                      */
#define QS_TK_SORT_UP   0  /* radix sort by bucket: digit histograms         */
#define QS_TK_SORT_SCAN 1  /* radix sort: scan of the histograms             */
#define QS_TK_SORT_DOWN 2  /* radix sort: stable scatter                     */
#define QS_TK_SEG       3  /* per-bucket-segment apply (the table sweep)     */
#define QS_TK_ORDERED   4  /* in-order replay of segments near the limit     */
#define QS_TK_ASSIGN    5  /* allocation ids in stream order                 */
#define QS_TK_TAIL      6  /* long (hot) segments: windowed high adds        */
#define QS_TK_QUERY     7  /* query kernel                                   */
#define QS_TK_HOT       8  /* hot-key filter (qs_set_hot_filter)             */
#define QS_TIMING_KERNELS 9
typedef struct qs_timing {
    double   insert_kernel_ms;      /* sum of kernel_ms[0..6] */
    uint64_t insert_kernel_launches;
    uint64_t insert_keys;           /* keys inserted while timing */
    double   query_kernel_ms;       /* = kernel_ms[QS_TK_QUERY] */
    uint64_t query_kernel_launches;
    uint64_t query_keys;
    double   kernel_ms[QS_TIMING_KERNELS];
    uint64_t kernel_launches[QS_TIMING_KERNELS];
} qs_timing;
int qs_set_timing(qs_sketch *s, int enable);
int qs_get_timing(qs_sketch *s, qs_timing *t);

const char *qs_strerror(int code);

/* ---- batched multi-sketch (build addition) --------------------------------
 * The reference runs one sketch per object and one launch per insert call.
 * A "set" holds n independent sketches (each with its own parameters, seeds,
 * masks and high pool; all with the same key width) and updates all of them
 * with one launch sequence per batch: a batch is one key array in which
 * sketch i owns keys[offs[i] .. offs[i+1]) (offs has n + 1 entries, offs[0] =
 * 0).  Every sketch of a set ends in exactly the state the plain sketch with
 * the same parameters reaches on its own stream.  The plain entry points
 * (qs_insert, qs_query, qs_export_state, qs_import_state) take sets of one
 * only; qs_clear, qs_destroy, qs_sync, qs_get_stats and timing take any set.
 * Calls on one set must be ordered (one stream at a time).                  */
int qs_multi_create(const qs_params *ps, uint32_t n_sketches, qs_sketch **out);
uint32_t qs_multi_count(const qs_sketch *s);
int qs_multi_params(const qs_sketch *s, uint32_t sketch, qs_params *p);
/* blocking; h_offs is a host array */
int qs_multi_insert(qs_sketch *s, const void *d_keys, const uint64_t *h_offs);
int qs_multi_query(qs_sketch *s, const void *d_keys, const uint64_t *h_offs, uint32_t *d_counts);
/* stream-ordered: h_offs (host) plans the chunks, d_offs (device, n + 1
 * entries, same values) is read by the kernels and must stay valid until the
 * work has run; pool errors surface at qs_sync */
int qs_multi_insert_async(qs_sketch *s, const void *d_keys, const uint64_t *h_offs,
                          const uint64_t *d_offs, void *stream);
int qs_multi_query_async(qs_sketch *s, const void *d_keys, const uint64_t *d_offs, uint64_t n_keys,
                         uint32_t *d_counts, void *stream);
int qs_multi_export_state(qs_sketch *s, uint32_t sketch, uint32_t *low, uint32_t *high,
                          uint32_t *masks, uint32_t *seed, uint32_t *next);
int qs_multi_import_state(qs_sketch *s, uint32_t sketch, const uint32_t *low, const uint32_t *high,
                          const uint32_t *masks, uint32_t seed, uint32_t next);

/* Captured insert: the launch sequence of one insert batch (any set, a set of
 * one included) recorded once into a hipGraph and replayed with one call.  The
 * graph reads d_keys and d_offs at replay time, so the caller may refill them
 * between replays with the same layout h_offs.  Destroy graphs before their
 * set; while a graph lives, import is refused (QS_ERR_STATE: the graph holds
 * the hash parameters) and so is a later batch that needs larger scratch
 * buffers than the graph's.  Build addition with no reference counterpart. */
typedef struct qs_graph qs_graph;
int qs_graph_capture_insert(qs_sketch *s, const void *d_keys, const uint64_t *h_offs,
                            const uint64_t *d_offs, qs_graph **out);
int qs_graph_launch(qs_graph *g, void *stream);
void qs_graph_destroy(qs_graph *g);

/* ---- insert shape and autotuner (build addition) --------------------------
 * The per-bucket-segment pass of the insert runs in groups of `width` lanes
 * (a sub-wavefront: 8, 16, 32 or 64) over `items_per_lane` (1, 2, 4 or 8)
 * consecutive sorted inserts per lane; (0, 0) selects the original 64-lane,
 * one-item kernel.  A new sketch or set starts with (64, 4).  The shape never changes the result.  The autotuner runs
 * one insert batch with all 16 shapes on a twin set (the caller's set is not
 * modified; the twin needs as much device memory as the set), checks that all
 * of them leave the same state, and keeps the shape with the fastest segment
 * pass on the set (QS_ERR_STATE, shape unchanged, if two shapes disagree).
 * A single sketch's batch is inserted as qs_insert would (through the hot-key
 * filter when the sketch has it on and the batch takes it).                 */
typedef struct qs_tune_entry {
    uint32_t width, items_per_lane;
    double seg_ms;     /* segment pass, mean over reps */
    double insert_ms;  /* whole insert launch sequence, mean over reps */
    uint64_t checksum; /* state checksum after the insert */
} qs_tune_entry;
typedef struct qs_tune_result {
    uint32_t n;          /* entries measured */
    uint32_t best;       /* index of the chosen entry */
    uint32_t consistent; /* 1 if every shape left the same state */
    uint32_t pad;
    qs_tune_entry e[16];
} qs_tune_result;
int qs_set_insert_shape(qs_sketch *s, uint32_t width, uint32_t items_per_lane);
int qs_get_insert_shape(const qs_sketch *s, uint32_t *width, uint32_t *items_per_lane);
int qs_autotune_insert(qs_sketch *s, const void *d_keys, const uint64_t *h_offs, const uint64_t *d_offs,
                       int reps, qs_tune_result *out);

#ifdef __cplusplus
}
#endif
#endif /* QSKETCH_H */
