/*
 * qsketch_calib.h -- kernels of known HBM traffic (build addition, no
 * counterpart in the reference).  They state the ceilings the sketch kernels
 * are measured against (a streaming copy; random reads of a given size, the
 * query's access pattern; read-modify-writes of buckets, random or sorted,
 * the segment pass's pattern) and calibrate the rocprofv3 FETCH_SIZE and
 * WRITE_SIZE counters.  Stream-ordered, no synchronisation.
 */
#ifndef QSKETCH_CALIB_H
#define QSKETCH_CALIB_H

#include <stddef.h>
#include <stdint.h>

#include "qsketch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Copy bytes (a multiple of 16) from d_src to d_dst, 16 bytes per lane. */
int qs_calib_copy(const void *d_src, void *d_dst, uint64_t bytes, void *stream);
/* Fill bytes (a multiple of 16) of d_dst with the 32-bit value, 16 bytes per
 * lane; nontemporal != 0: non-temporal stores. */
int qs_calib_fill(void *d_dst, uint64_t bytes, uint32_t value, int nontemporal, void *stream);
/* n random reads of access_bytes (16, 32, 64 or 128) from the
 * access_bytes-aligned lines of d_buf (buf_bytes long), read i at the line
 * picked by a hash of (seed, i); the xor of each read goes to d_out[i]
 * (uint32[n]). */
int qs_calib_gather(const void *d_buf, uint64_t buf_bytes, uint32_t access_bytes, uint64_t n, uint64_t seed,
                    uint32_t *d_out, void *stream);
/* n read-modify-writes of access_bytes (16, 32, 64 or 128; every 32-bit word
 * incremented) on the access_bytes-aligned lines of d_buf (buf_bytes long,
 * n <= buf_bytes / access_bytes): sorted == 0, access i at the line picked by
 * a hash of (seed, i); sorted != 0, in ascending order with random gaps
 * (access i in the i-th of n equal ranges of lines), the pattern of a pass
 * over distinct buckets in bucket order.  Accesses that meet race (the
 * values are not used). */
int qs_calib_rmw(void *d_buf, uint64_t buf_bytes, uint32_t access_bytes, uint64_t n, uint64_t seed, int sorted,
                 void *stream);

#ifdef __cplusplus
}
#endif
#endif /* QSKETCH_CALIB_H */
