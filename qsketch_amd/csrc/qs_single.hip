// qs_single.hip -- the single-level sub-warp sketch (device_sketch_v2.h) on gfx950.
//
// The reference gives each key a sub-warp of m lanes, one counter each, and
// adds with one atomic per selected counter (v2:54-65); under skew the hot
// keys' counters then serialise (same-line atomics cost ~11 ns each here).
// k_sl_insert instead counts a tile of keys per workgroup in an LDS table keyed
// by (bucket, mask index) and adds each distinct pair's count with three
// atomics; the counters end the same because the adds commute.  k_sl_query is
// one lane per key: the three selected counters lie in one 32-128 byte bucket.
#include "qs_internal.h"

#include "../../include/qsketch_single.h"

#include <cstdlib>
#include <cstring>
#include <vector>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess)                                                                   \
            return QS_ERR_HIP;                                                                   \
    } while (0)

#define SL_THREADS 256
#define SL_TILE 8192u
#define SL_BATCH 8u
#define SL_SLOTS 2048u
#define SL_PROBES 16u

struct qs_sl {
    uint64_t n;
    uint32_t m;
    uint32_t n_masks;
    uint32_t key_bytes;
    uint32_t seed;
    qs_fastdiv fd_n, fd_masks;
    uint32_t *h_masks;
    uint32_t *d_masks;
    uint32_t *d_table;
};

struct sl_par {
    qs_fastdiv fd_n, fd_masks;
    uint64_t n;
    uint32_t n_masks, seed, m;
};

template <int KB> __device__ static inline void sl_index(KeyT<KB> key, const sl_par &p, uint32_t &b, uint32_t &mu);

template <> __device__ inline void sl_index<4>(uint32_t key, const sl_par &p, uint32_t &b, uint32_t &mu)
{
    const uint32_t h = key * p.seed;
    b = qs_fastmod(h, p.fd_n);
    mu = qs_fastmod(h, p.fd_masks);
}

template <> __device__ inline void sl_index<8>(uint64_t key, const sl_par &p, uint32_t &b, uint32_t &mu)
{
    const uint64_t h = key * (uint64_t)p.seed;
    b = (uint32_t)(h % p.n);
    mu = (uint32_t)(h % (uint64_t)p.n_masks);
}

__device__ static inline void sl_add(uint32_t *__restrict__ table, uint32_t m, uint32_t b, uint32_t M, uint32_t c)
{
    uint32_t *row = table + (size_t)b * m;
    for (; M; M &= M - 1u)
        atomicAdd(row + __builtin_ctz(M), c);
}

template <int KB>
__global__ __launch_bounds__(SL_THREADS) void k_sl_insert(const KeyT<KB> *__restrict__ keys, uint64_t n, sl_par p,
                                                          const uint32_t *__restrict__ masks,
                                                          uint32_t *__restrict__ table)
{
    __shared__ unsigned long long s_key[SL_SLOTS];
    __shared__ uint32_t s_cnt[SL_SLOTS];
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    const uint32_t t = threadIdx.x;
    for (uint32_t e = t; e < SL_SLOTS; e += SL_THREADS) {
        s_key[e] = ~0ull;
        s_cnt[e] = 0;
    }
    for (uint32_t e = t; e < p.n_masks; e += SL_THREADS)
        s_masks[e] = masks[e];
    __syncthreads();
    const uint64_t tb = (uint64_t)blockIdx.x * SL_TILE;
    for (uint32_t r0 = 0; r0 < SL_TILE / SL_THREADS; r0 += SL_BATCH) {
        KeyT<KB> kk[SL_BATCH];
#pragma unroll
        for (uint32_t k = 0; k < SL_BATCH; ++k) {
            const uint64_t i = tb + (uint64_t)(r0 + k) * SL_THREADS + t;
            kk[k] = i < n ? keys[i] : (KeyT<KB>)0;
        }
#pragma unroll
        for (uint32_t k = 0; k < SL_BATCH; ++k) {
            const uint64_t i = tb + (uint64_t)(r0 + k) * SL_THREADS + t;
            if (i >= n)
                continue;
            uint32_t b, mu;
            sl_index<KB>(kk[k], p, b, mu);
            const unsigned long long key = ((unsigned long long)b << 12) | mu;
            uint32_t sl = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 53) & (SL_SLOTS - 1u);
            bool done = false;
            for (uint32_t pr = 0; pr < SL_PROBES && !done; ++pr, sl = (sl + 1u) & (SL_SLOTS - 1u)) {
                unsigned long long cur = s_key[sl];
                if (cur == ~0ull)
                    cur = atomicCAS(&s_key[sl], ~0ull, key);
                if (cur == ~0ull || cur == key) {
                    atomicAdd(&s_cnt[sl], 1u);
                    done = true;
                }
            }
            if (!done) // crowded table: add directly
                sl_add(table, p.m, b, s_masks[mu], 1u);
        }
    }
    __syncthreads();
    for (uint32_t e = t; e < SL_SLOTS; e += SL_THREADS) {
        const uint32_t c = s_cnt[e];
        if (c) {
            const unsigned long long key = s_key[e];
            sl_add(table, p.m, (uint32_t)(key >> 12), s_masks[key & 0xfffu], c);
        }
    }
}

template <int KB>
__global__ __launch_bounds__(SL_THREADS) void k_sl_query(const KeyT<KB> *__restrict__ keys, uint64_t n, sl_par p,
                                                         const uint32_t *__restrict__ masks,
                                                         const uint32_t *__restrict__ table,
                                                         uint32_t *__restrict__ out)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t e = threadIdx.x; e < p.n_masks; e += SL_THREADS)
        s_masks[e] = masks[e];
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * SL_THREADS + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * SL_THREADS) {
        uint32_t b, mu;
        sl_index<KB>(keys[i], p, b, mu);
        const uint32_t *row = table + (size_t)b * p.m;
        uint32_t v = 0xffffffffu;
        for (uint32_t M = s_masks[mu]; M; M &= M - 1u)
            v = min(v, row[__builtin_ctz(M)]);
        out[i] = v;
    }
}

static sl_par par_of(const qs_sl *s)
{
    sl_par p;
    p.fd_n = s->fd_n;
    p.fd_masks = s->fd_masks;
    p.n = s->n;
    p.n_masks = s->n_masks;
    p.seed = s->seed;
    p.m = s->m;
    return p;
}

extern "C" void qs_sl_destroy(qs_sl *s)
{
    if (!s)
        return;
    hipDeviceSynchronize();
    hipFree(s->d_masks);
    hipFree(s->d_table);
    free(s->h_masks);
    free(s);
}

extern "C" int qs_sl_create(uint64_t n, uint32_t m, uint32_t seed_sz, uint32_t n_masks, uint32_t key_bytes,
                            uint64_t rng_seed, qs_sl **out)
{
    if (!out)
        return QS_ERR_PARAM;
    *out = nullptr;
    if (n < 1 || (m != 8 && m != 16 && m != 32) || seed_sz < 1 || n_masks < 1 || n_masks > QS_MAX_MASKS ||
        (key_bytes != 4 && key_bytes != 8))
        return QS_ERR_PARAM;
    if (key_bytes == 4 && n > 0xffffffffull)
        return QS_ERR_PARAM;
    if (key_bytes == 8 && seed_sz > 1)
        return QS_ERR_PARAM; // 64-bit keys sum the seeds in 64 bits (device_hash_function.h:6-17)
    if (n * m > (1ull << 36))
        return QS_ERR_PARAM;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return QS_ERR_NODEVICE;
    qs_sl *s = (qs_sl *)calloc(1, sizeof(qs_sl));
    if (!s)
        return QS_ERR_NOMEM;
    s->n = n;
    s->m = m;
    s->n_masks = n_masks;
    s->key_bytes = key_bytes;
    // Device_Seed::resize then generate_hashmask(n_masks, m, 3, 0), the call
    // order of device_sketch_v1.h:267-270
    std::default_random_engine seed_eng, mask_eng;
    qs_seed_engine(seed_eng, rng_seed, 0);
    qs_seed_engine(mask_eng, rng_seed, 1);
    std::uniform_int_distribution<unsigned int> dis(1, 0xffffffffu);
    uint32_t seed = 0;
    for (uint32_t k = 0; k < seed_sz; ++k)
        seed += dis(seed_eng);
    s->seed = seed;
    s->h_masks = (uint32_t *)malloc(sizeof(uint32_t) * n_masks);
    if (!s->h_masks) {
        free(s);
        return QS_ERR_NOMEM;
    }
    qs_make_masks(mask_eng, n_masks, m, 0, s->h_masks);
    s->fd_n = qs_fastdiv_make(n <= 0xffffffffull ? (uint32_t)n : 1u);
    s->fd_masks = qs_fastdiv_make(n_masks);
    const size_t tbytes = (size_t)n * m * 4u;
    if (hipMalloc(&s->d_masks, n_masks * 4u) != hipSuccess || hipMalloc(&s->d_table, tbytes) != hipSuccess) {
        qs_sl_destroy(s);
        return QS_ERR_NOMEM;
    }
    if (hipMemcpy(s->d_masks, s->h_masks, n_masks * 4u, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->d_table, 0, tbytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        qs_sl_destroy(s);
        return QS_ERR_HIP;
    }
    *out = s;
    return QS_OK;
}

extern "C" int qs_sl_clear(qs_sl *s)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    CK(hipMemset(s->d_table, 0, (size_t)s->n * s->m * 4u));
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_sl_insert(qs_sl *s, const void *d_keys, size_t n)
{
    if (!s || (!d_keys && n))
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    const sl_par p = par_of(s);
    const dim3 grid((uint32_t)((n + SL_TILE - 1) / SL_TILE)), block(SL_THREADS);
    if (s->key_bytes == 4)
        hipLaunchKernelGGL(k_sl_insert<4>, grid, block, 0, 0, (const uint32_t *)d_keys, (uint64_t)n, p, s->d_masks,
                           s->d_table);
    else
        hipLaunchKernelGGL(k_sl_insert<8>, grid, block, 0, 0, (const uint64_t *)d_keys, (uint64_t)n, p, s->d_masks,
                           s->d_table);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(0));
    return QS_OK;
}

extern "C" int qs_sl_query(qs_sl *s, const void *d_keys, size_t n, uint32_t *d_counts)
{
    if (!s || ((!d_keys || !d_counts) && n))
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    const sl_par p = par_of(s);
    uint64_t blocks = (n + SL_THREADS - 1) / SL_THREADS;
    if (blocks > 8192)
        blocks = 8192;
    const dim3 grid((uint32_t)blocks), block(SL_THREADS);
    if (s->key_bytes == 4)
        hipLaunchKernelGGL(k_sl_query<4>, grid, block, 0, 0, (const uint32_t *)d_keys, (uint64_t)n, p, s->d_masks,
                           s->d_table, d_counts);
    else
        hipLaunchKernelGGL(k_sl_query<8>, grid, block, 0, 0, (const uint64_t *)d_keys, (uint64_t)n, p, s->d_masks,
                           s->d_table, d_counts);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(0));
    return QS_OK;
}

extern "C" int qs_sl_export(qs_sl *s, uint32_t *table, uint32_t *masks, uint32_t *seed)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    if (table)
        CK(hipMemcpy(table, s->d_table, (size_t)s->n * s->m * 4u, hipMemcpyDeviceToHost));
    if (masks)
        memcpy(masks, s->h_masks, s->n_masks * 4u);
    if (seed)
        *seed = s->seed;
    return QS_OK;
}
