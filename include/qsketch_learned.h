/*
 * qsketch_learned.h -- learned heavy-hitter form of the MSWS sketch (build
 * addition; the reference has no learned component).
 *
 * A small MLP scores each key from its 64 bits (64 -> 64 -> 64 -> 1, ReLU,
 * BF16 weights and activations, FP32 accumulation).  Keys whose score reaches
 * a threshold are counted exactly in a GPU hash table; every other key goes to
 * an ordinary MSWS sketch (include/qsketch.h), in stream order, so that part
 * keeps the plain sketch's semantics bit for bit.  A query returns the exact
 * count for a key held in the table and the sketch estimate otherwise.
 *
 * Training is the caller's loop over minibatches: qs_mlp_step computes the
 * mean-squared-error gradients of a batch (hand-written forward and backward
 * kernels) and applies one Adam step to FP32 master weights, refreshing the
 * BF16 copies the kernels read.  Targets are what a plain sketch estimated
 * for keys of an earlier window (e.g. log2(1 + estimate)).
 *
 * Which part counts a key is a function of the network and the threshold, so
 * both are fixed while the form holds data: qs_mlp_set_params, qs_mlp_step
 * and qs_learned_set_threshold return QS_ERR_STATE after an insert until
 * qs_learned_clear.  (Changing them with data in would let a query look for a
 * key in the other part than its inserts went to.)  Train a fresh or cleared
 * form, or a separate one, and copy the weights with get/set_params.
 *
 * A full table does not lose keys: a key the table has no slot for (the
 * table is full, or its probe window is) is counted in the sketch instead,
 * with all its occurrences of that insert call and later ones, in stream
 * order.  Every key is thus counted entirely in one part, and a query
 * returns the table count of a key held in the table and the sketch estimate
 * of any other key.
 *
 * Parameter layout (FP32, 8385 values): W1[64][64] (row = hidden unit,
 * column = key bit), b1[64], W2[64][64] (row = layer-2 unit, column = layer-1
 * unit), b2[64], w3[64], b3.  Key bit k (k = 0 is the least significant) is
 * input k, as 0.0 or 1.0.
 */
#ifndef QSKETCH_LEARNED_H
#define QSKETCH_LEARNED_H

#include <stddef.h>
#include <stdint.h>

#include "qsketch.h"

#ifdef __cplusplus
extern "C" {
#endif

#define QS_MLP_PARAMS 8385
#define QS_ERR_TABLE  -7  /* reserved (an exact table that fills no longer fails an insert) */

typedef struct qs_learned qs_learned;

/* Sketch part with parameters *p (any key_bytes; keys of 8 bytes use the
 * whole key as MLP input, keys of 4 bytes zero-extend), an exact table of
 * 2^table_log2 slots (12 bytes each: u64 key, u32 count; 10 <= table_log2 <=
 * 28), and MLP weights initialised from init_seed (uniform +-1/sqrt(fan_in),
 * biases 0). */
int qs_learned_create(const qs_params *p, uint32_t table_log2, uint64_t init_seed, qs_learned **out);
void qs_learned_destroy(qs_learned *L);
/* The sketch handle of the non-heavy keys (for stats, export, timing). */
qs_sketch *qs_learned_sketch(qs_learned *L);

/* ---- the network ---------------------------------------------------------- */
/* FP32 master weights in the layout above (host arrays of QS_MLP_PARAMS). */
int qs_mlp_get_params(qs_learned *L, float *h_params);
int qs_mlp_set_params(qs_learned *L, const float *h_params);
/* One training step on n (key, target) pairs (device arrays; keys per the
 * sketch's key_bytes, targets FP32): forward, backward (loss = mean of
 * (y - target)^2 / 2), Adam (lr, beta1, beta2, eps; step counter kept
 * inside).  *loss (host, may be NULL) receives the batch loss before the
 * update; h_grad (host, may be NULL) receives the FP32 gradients. */
int qs_mlp_step(qs_learned *L, const void *d_keys, const float *d_targets, uint32_t n, float lr,
                float beta1, float beta2, float eps, float *loss, float *h_grad);
/* Gradients and loss only (no update). */
int qs_mlp_grad(qs_learned *L, const void *d_keys, const float *d_targets, uint32_t n, float *loss,
                float *h_grad);
/* Scores of n keys into d_scores (device FP32).  mfma = 0: the per-key VALU
 * forward of the training kernels; mfma = 1: the matrix-core forward the
 * router uses.  Both compute the same BF16/FP32 network. */
int qs_mlp_score(qs_learned *L, const void *d_keys, uint64_t n, float *d_scores, int mfma);

/* ---- routing -------------------------------------------------------------- */
/* Route keys with score >= threshold to the exact table. */
int qs_learned_set_threshold(qs_learned *L, float threshold);
float qs_learned_threshold(const qs_learned *L);
/* Insert n keys (device pointer): route, count the heavy ones exactly, insert
 * the others into the sketch in stream order.  Blocking.  *n_heavy (may be
 * NULL) receives how many of the n went to the table. */
int qs_learned_insert(qs_learned *L, const void *d_keys, uint64_t n, uint64_t *n_heavy);
int qs_learned_query(qs_learned *L, const void *d_keys, uint64_t n, uint32_t *d_counts);
/* Zero the table and the sketch (the network is kept). */
int qs_learned_clear(qs_learned *L);
/* Table contents: up to cap (key, count) pairs into host arrays; returns the
 * number of occupied slots in *n_used. */
int qs_learned_table(qs_learned *L, uint64_t *h_keys, uint32_t *h_counts, uint64_t cap, uint64_t *n_used);
/* 1 once the table has turned a key away to the sketch (until clear), else 0;
 * negative on a bad handle. */
int qs_learned_table_full(const qs_learned *L);
/* Bytes of the table, the MLP (BF16 weights + FP32 biases) and the sketch. */
int qs_learned_memory(const qs_learned *L, uint64_t *table_bytes, uint64_t *mlp_bytes, uint64_t *sketch_bytes);

#ifdef __cplusplus
}
#endif
#endif /* QSKETCH_LEARNED_H */
