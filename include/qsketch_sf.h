/*
 * qsketch_sf.h -- the SF sketch with split counters, Sketch_GPU_SF_Host_Sub_Warp
 * (include/sketch/device_sketch/device_sketch_v42.h:207-480).
 *
 * n buckets of m counters (m in {8, 16, 32}); every counter is split into
 * `times` sub-counters.  Seeds: (1 + 3) * seed_sz draws (v42:263), groups 0..2
 * used; masks: generate_hashmask(n_masks, m, 3, 0) with each mask's three
 * counters in draw order (v42:268-269).  Insert of k: h = k*s_0, bucket
 * b = h % n, mask mu = h % n_masks; for j = 0..2 the j-th counter of mask mu
 * gets +1 in split (k*s_j) % times (v42:329-355).  After every insert call each
 * counter becomes the largest of its splits (update, v42:302-327).  Query: the
 * smallest of the mask's counters (search_kernel, v42:120-199).
 *
 * The reference inserts on the host (std::atomic splits) and copies to the
 * GPU; here insert and update run on the GPU.  Keys and counts are device
 * pointers; calls are blocking.  64-bit keys take one seed (seed_sz = 1).
 */
#ifndef QSKETCH_SF_H
#define QSKETCH_SF_H

#include <stddef.h>
#include <stdint.h>

#include "qsketch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qs_sf qs_sf;

/* Sketch_GPU_SF_Host_Sub_Warp(_n, _m, _times, _ss, _n_hash_mask_table)
 * (v42:245-276); rng_seed replaces the reference's /dev/urandom engines.
 * QS_ERR_PARAM unless 1 <= n < 2^32, m in {8, 16, 32}, 1 <= times <= 256,
 * n * m * times <= 2^36, 1 <= n_masks <= 4096, key_bytes 4 or 8 (8: seed_sz
 * 1). */
int qs_sf_create(uint64_t n, uint32_t m, uint32_t times, uint32_t seed_sz, uint32_t n_masks, uint32_t key_bytes,
                 uint64_t rng_seed, qs_sf **out);
/* ~Sketch_GPU_SF_Host_Sub_Warp (v42:468-478) */
void qs_sf_destroy(qs_sf *s);
/* insert (v42:329-384, including the update of v42:358).  Calls of 2^16 keys
 * or more take scratch of 2 * key_bytes bytes per key (at most 2^28 4-byte or
 * 2^27 8-byte keys; longer calls run in chunks), kept until destroy.
 * QS_ERR_INTERNAL: a device-side consistency check failed (the state is
 * invalid until qs_sf_clear).  A split that passes 2^32 - 1 within one call
 * wraps, and its counter then need not be the largest split. */
int qs_sf_insert(qs_sf *s, const void *d_keys, size_t n);
/* search (v42:385-459) */
int qs_sf_query(qs_sf *s, const void *d_keys, size_t n, uint32_t *d_counts);
/* clear (v42:461-464) */
int qs_sf_clear(qs_sf *s);
/* Host copies (any may be NULL): split = uint32[n*m*times] ([bucket][counter]
 * [split]), search = uint32[n*m], masks = uint32[n_masks], index =
 * uint8[n_masks*3], seeds = uint64[3] (the group sums). */
int qs_sf_export(qs_sf *s, uint32_t *split, uint32_t *search, uint32_t *masks, uint8_t *index, uint64_t *seeds);

#ifdef __cplusplus
}
#endif
#endif /* QSKETCH_SF_H */
