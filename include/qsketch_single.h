/*
 * qsketch_single.h -- the single-level sub-warp sketch on MI355X.
 *
 * Replaces Sketch_GPU_Sub_Warp (include/sketch/device_sketch/device_sketch_v2.h:
 * 290-410, built on Sketch_GPU_Warp, device_sketch_v1.h:254-273): n buckets of
 * m 32-bit counters (m = 8, 16 or 32; the reference driver uses 8,
 * src/main.cu:330), a table of n_masks masks of 3 distinct bits in [0, m).
 * A key selects bucket h % n and mask h % n_masks (h = key * seed, the
 * hash_mul_add of device_hash_function.h:4-18); insert adds 1 to each selected
 * counter (v2:54-65), query returns their minimum (v2:155-190).  It is the
 * plain equal-memory comparator of the two-level sketch.  Seeds and masks come
 * from the same engines and call pattern as qs_create (rng_seed).
 *
 * Keys and counts are device pointers; calls block, like the reference
 * (v2:345, v2:372).  Counters add commutatively, so the GPU result equals the
 * sequential one whatever the schedule.
 */
#ifndef QSKETCH_SINGLE_H
#define QSKETCH_SINGLE_H

#include <stddef.h>
#include <stdint.h>

#include "qsketch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qs_sl qs_sl;

/* Sketch_GPU_Sub_Warp(_n, _m, _ss, _n_hash_mask_table) (v2:307-309) */
int qs_sl_create(uint64_t n, uint32_t m, uint32_t seed_sz, uint32_t n_masks, uint32_t key_bytes,
                 uint64_t rng_seed, qs_sl **out);
void qs_sl_destroy(qs_sl *s);
/* insert / search (v2:326-377) */
int qs_sl_insert(qs_sl *s, const void *d_keys, size_t n);
int qs_sl_query(qs_sl *s, const void *d_keys, size_t n, uint32_t *d_counts);
/* zero the counters (Device_Hash_Table::clear, device_sketch.h:115-124) */
int qs_sl_clear(qs_sl *s);
/* host copies: table u32[n * m], masks u32[n_masks]; any pointer may be NULL */
int qs_sl_export(qs_sl *s, uint32_t *table, uint32_t *masks, uint32_t *seed);

#ifdef __cplusplus
}
#endif
#endif /* QSKETCH_SINGLE_H */
