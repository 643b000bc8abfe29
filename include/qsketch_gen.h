/*
 * qsketch_gen.h -- synthetic key streams for the QSketch workloads.
 *
 * Build addition.  The reference draws its keys with cuRAND
 * (include/tool/tool.h:390-418, uniform uint32) and a host "freq" mix
 * (tool.h:222-252); it has no Zipf generator.  These generators are
 * counter-based and integer-only at sampling time, so the device and host
 * versions return identical keys for the same (seed, offset):
 *   x_i   = SplitMix64 finaliser of (seed, i)
 *   Zipf  : Walker/Vose alias table of P(r) ~ r^-alpha, r = 1..U (built once on
 *           the host in double precision), column = hi32(x) * U >> 32,
 *           accept if lo32(x) < threshold[column] else alias[column]
 *   key   : 4-round Feistel permutation of the 0-based rank (distinct ranks ->
 *           distinct keys), keyed by seed
 * All device entry points take device output pointers and an optional
 * hipStream_t (NULL = null stream) and do not synchronise.
 */
#ifndef QSKETCH_GEN_H
#define QSKETCH_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qs_zipf qs_zipf;

/* Alias table of Zipf(alpha) over ranks 1..universe (universe <= 2^32).
 * with_device != 0 also uploads it to the current GPU. */
int  qs_zipf_create(double alpha, uint64_t universe, int with_device, qs_zipf **out);
void qs_zipf_destroy(qs_zipf *z);

/* Keys i = offset .. offset+n-1 of the stream (seed). */
int qs_gen_zipf_u32(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n, uint32_t *d_out,
                    void *stream);
int qs_gen_zipf_u32_host(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                         uint32_t *out);
/* 0-based ranks instead of keys (for tests of the sampler). */
int qs_gen_zipf_rank_host(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                          uint32_t *out);
int qs_gen_uniform_u32(uint64_t seed, uint64_t offset, uint64_t n, uint32_t *d_out, void *stream);
int qs_gen_uniform_u32_host(uint64_t seed, uint64_t offset, uint64_t n, uint32_t *out);

/* The rank -> key permutation used by the Zipf streams. */
uint32_t qs_rank_to_key32(uint64_t seed, uint32_t rank);

/* 64-bit keys of the learned-oracle workload: Zipf rank r (0-based) ->
 * key = (clip(round(log2(r+1)) + round(N(0, sigma)), 0, 31) << 48) | F48(r),
 * with F48 a 48-bit Feistel permutation and the noise drawn once per rank. */
int qs_gen_zipf_u64_noisy(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                          double sigma, uint64_t *d_out, void *stream);
int qs_gen_zipf_u64_noisy_host(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                               double sigma, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
