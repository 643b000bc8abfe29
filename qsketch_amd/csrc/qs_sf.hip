// qs_sf.hip -- the SF sketch (device_sketch_v42.h:207-480) on gfx950.
//
// n buckets of m counters (m in {8, 16, 32}); each counter is split into
// `times` sub-counters.  An insert of k adds 1 to one split of each of the
// three counters its mask selects, the split chosen by k's hash under seed
// group j (v42:340-352); after every insert call each counter becomes the
// largest of its splits (update, v42:302-327); a query takes the smallest of
// the mask's counters (search_kernel, v42:120-199).
//
// The reference runs the insert loop on the host over std::atomic split
// counters, copies them to the GPU in batches and takes the maximum there.
// Here the split table lives in HBM: k_sf_insert counts a tile of keys per
// workgroup in an LDS table keyed by the key and adds each distinct key's
// count to its three splits with atomics (the adds commute, so the splits end
// as the reference's); k_sf_update streams the split table once and writes
// every counter's largest split; k_sf_query reads the key's bucket in 16-byte
// pieces, m / 4 lanes per key, and takes the minimum of the mask's counters.
#include "qs_internal.h"

#include "../../include/qsketch_sf.h"

#include <cstdlib>
#include <cstring>
#include <vector>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess)                                                                   \
            return QS_ERR_HIP;                                                                   \
    } while (0)

#define SF_THREADS 256
#define SF_TILE 8192u
#define SF_BATCH 8u
#define SF_SLOTS 2048u
#define SF_PROBES 16u
#define SF_EMPTY 0xffffffffffffffffull

struct qs_sf {
    uint64_t n;
    uint32_t m, times, n_masks, key_bytes;
    uint64_t seed[3];   // group sums of the seed draws (groups 0..2)
    uint32_t *h_masks;  // bit c selects counter c
    uint32_t *h_index;  // the three counters of each mask in draw order, 8 bits each
    uint32_t *d_masks;
    uint32_t *d_index;
    uint32_t *d_split;  // n * m * times
    uint32_t *d_search; // n * m
};

struct sf_par {
    qs_fastdiv fd_n, fd_masks, fd_times;
    uint64_t n;
    uint32_t n_masks, m, times;
    uint32_t s32[3];
    uint64_t s64[3];
};

// bucket, mask index and the split of each of the three counters
template <int KB> struct sf_idx;
template <> struct sf_idx<4> {
    __device__ static inline void get(uint64_t k, const sf_par &p, uint32_t &b, uint32_t &mu, uint32_t t[3])
    {
        const uint32_t key = (uint32_t)k;
        const uint32_t h0 = key * p.s32[0];
        b = qs_fastmod(h0, p.fd_n);
        mu = qs_fastmod(h0, p.fd_masks);
        t[0] = qs_fastmod(h0, p.fd_times);
        t[1] = qs_fastmod(key * p.s32[1], p.fd_times);
        t[2] = qs_fastmod(key * p.s32[2], p.fd_times);
    }
};
template <> struct sf_idx<8> {
    __device__ static inline void get(uint64_t key, const sf_par &p, uint32_t &b, uint32_t &mu, uint32_t t[3])
    {
        const uint64_t h0 = key * p.s64[0];
        b = (uint32_t)(h0 % p.n);
        mu = (uint32_t)(h0 % (uint64_t)p.n_masks);
        t[0] = (uint32_t)(h0 % (uint64_t)p.times);
        t[1] = (uint32_t)((key * p.s64[1]) % (uint64_t)p.times);
        t[2] = (uint32_t)((key * p.s64[2]) % (uint64_t)p.times);
    }
};

template <int KB>
__device__ static inline void sf_add(uint32_t *__restrict__ split, const sf_par &p, const uint32_t *s_index,
                                     uint64_t key, uint32_t c)
{
    uint32_t b, mu, t[3];
    sf_idx<KB>::get(key, p, b, mu, t);
    const uint32_t ix = s_index[mu];
    uint32_t *row = split + (size_t)b * p.m * p.times;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        atomicAdd(row + ((ix >> (8 * j)) & 0xffu) * p.times + t[j], c);
}

template <int KB>
__global__ __launch_bounds__(SF_THREADS) void k_sf_insert(const KeyT<KB> *__restrict__ keys, uint64_t n, sf_par p,
                                                          const uint32_t *__restrict__ index,
                                                          uint32_t *__restrict__ split)
{
    __shared__ unsigned long long s_key[SF_SLOTS];
    __shared__ uint32_t s_cnt[SF_SLOTS];
    __shared__ uint32_t s_index[QS_MAX_MASKS];
    const uint32_t t = threadIdx.x;
    for (uint32_t e = t; e < SF_SLOTS; e += SF_THREADS) {
        s_key[e] = SF_EMPTY;
        s_cnt[e] = 0;
    }
    for (uint32_t e = t; e < p.n_masks; e += SF_THREADS)
        s_index[e] = index[e];
    __syncthreads();
    const uint64_t tb = (uint64_t)blockIdx.x * SF_TILE;
    for (uint32_t r0 = 0; r0 < SF_TILE / SF_THREADS; r0 += SF_BATCH) {
        uint64_t kk[SF_BATCH];
#pragma unroll
        for (uint32_t k = 0; k < SF_BATCH; ++k) { // every load issued before any probe
            const uint64_t i = tb + (uint64_t)(r0 + k) * SF_THREADS + t;
            kk[k] = i < n ? (uint64_t)keys[i] : 0ull;
        }
#pragma unroll
        for (uint32_t k = 0; k < SF_BATCH; ++k) {
            const uint64_t i = tb + (uint64_t)(r0 + k) * SF_THREADS + t;
            if (i >= n)
                continue;
            const uint64_t key = kk[k];
            bool done = false;
            if (key != SF_EMPTY) { // the empty marker itself is added directly
                uint32_t sl = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 53) & (SF_SLOTS - 1u);
                for (uint32_t pr = 0; pr < SF_PROBES && !done; ++pr, sl = (sl + 1u) & (SF_SLOTS - 1u)) {
                    unsigned long long cur = s_key[sl];
                    if (cur == SF_EMPTY)
                        cur = atomicCAS(&s_key[sl], SF_EMPTY, (unsigned long long)key);
                    if (cur == SF_EMPTY || cur == key) {
                        atomicAdd(&s_cnt[sl], 1u);
                        done = true;
                    }
                }
            }
            if (!done) // crowded table: add directly
                sf_add<KB>(split, p, s_index, key, 1u);
        }
    }
    __syncthreads();
    for (uint32_t e = t; e < SF_SLOTS; e += SF_THREADS) {
        const uint32_t c = s_cnt[e];
        if (c)
            sf_add<KB>(split, p, s_index, s_key[e], c);
    }
}

// every counter: the largest of its splits (update_table, v42:72-95)
__global__ __launch_bounds__(256) void k_sf_update(const uint32_t *__restrict__ split, uint64_t cells, uint32_t times,
                                                   uint32_t *__restrict__ search)
{
    for (uint64_t c = (uint64_t)blockIdx.x * 256u + threadIdx.x; c < cells; c += (uint64_t)gridDim.x * 256u) {
        const uint32_t *s = split + c * times;
        uint32_t v = 0;
        for (uint32_t t = 0; t < times; ++t)
            v = max(v, __builtin_nontemporal_load(s + t));
        search[c] = v;
    }
}

// G = m / 4 lanes per key, one 16-byte quarter (m = 16), half (m = 8) or
// eighth (m = 32) of the bucket each, so a key costs one request per 64
// bytes of bucket (one 128-byte line at most); the lanes reduce their minima.
template <int KB, int G>
__global__ __launch_bounds__(SF_THREADS) void k_sf_query(const KeyT<KB> *__restrict__ keys, uint64_t n, sf_par p,
                                                         const uint32_t *__restrict__ masks,
                                                         const uint32_t *__restrict__ search,
                                                         uint32_t *__restrict__ out)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t e = threadIdx.x; e < p.n_masks; e += SF_THREADS)
        s_masks[e] = masks[e];
    __syncthreads();
    const uint32_t part = threadIdx.x % G;
    const uint64_t per_block = SF_THREADS / G;
    for (uint64_t i = (uint64_t)blockIdx.x * per_block + threadIdx.x / G; i < n; i += (uint64_t)gridDim.x * per_block) {
        uint32_t b, mu, t[3];
        sf_idx<KB>::get((uint64_t)keys[i], p, b, mu, t);
        const uint4 v = reinterpret_cast<const uint4 *>(search + (size_t)b * p.m)[part];
        const uint32_t M = (s_masks[mu] >> (4u * part)) & 0xfu;
        uint32_t r = 0xffffffffu;
        r = (M & 1u) ? min(r, v.x) : r;
        r = (M & 2u) ? min(r, v.y) : r;
        r = (M & 4u) ? min(r, v.z) : r;
        r = (M & 8u) ? min(r, v.w) : r;
#pragma unroll
        for (int d = 1; d < G; d <<= 1)
            r = min(r, (uint32_t)__shfl_xor((int)r, d, 64));
        if (part == 0)
            out[i] = r;
    }
}

// ===========================================================================
// host
// ===========================================================================
static sf_par sf_par_of(const qs_sf *s)
{
    sf_par p;
    p.fd_n = qs_fastdiv_make(s->key_bytes == 4 ? (uint32_t)s->n : 1u);
    p.fd_masks = qs_fastdiv_make(s->n_masks);
    p.fd_times = qs_fastdiv_make(s->times);
    p.n = s->n;
    p.n_masks = s->n_masks;
    p.m = s->m;
    p.times = s->times;
    for (int g = 0; g < 3; ++g) {
        p.s32[g] = (uint32_t)s->seed[g];
        p.s64[g] = s->seed[g];
    }
    return p;
}

extern "C" void qs_sf_destroy(qs_sf *s)
{
    if (!s)
        return;
    (void)hipDeviceSynchronize();
    (void)hipFree(s->d_masks);
    (void)hipFree(s->d_index);
    (void)hipFree(s->d_split);
    (void)hipFree(s->d_search);
    free(s->h_masks);
    free(s->h_index);
    free(s);
}

extern "C" int qs_sf_create(uint64_t n, uint32_t m, uint32_t times, uint32_t seed_sz, uint32_t n_masks,
                            uint32_t key_bytes, uint64_t rng_seed, qs_sf **out)
{
    if (!out)
        return QS_ERR_PARAM;
    *out = nullptr;
    if (n < 1 || (m != 8 && m != 16 && m != 32) || times < 1 || times > 256 || seed_sz < 1 || n_masks < 1 ||
        n_masks > QS_MAX_MASKS || (key_bytes != 4 && key_bytes != 8))
        return QS_ERR_PARAM;
    if (key_bytes == 4 && n > 0xffffffffull)
        return QS_ERR_PARAM;
    if (key_bytes == 8 && seed_sz > 1)
        return QS_ERR_PARAM; // 64-bit keys sum the seeds in 64 bits (device_hash_function.h:6-17)
    if (n * m * (uint64_t)times > (1ull << 36))
        return QS_ERR_PARAM;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return QS_ERR_NODEVICE;
    qs_sf *s = (qs_sf *)calloc(1, sizeof(qs_sf));
    if (!s)
        return QS_ERR_NOMEM;
    s->n = n;
    s->m = m;
    s->times = times;
    s->n_masks = n_masks;
    s->key_bytes = key_bytes;
    // ds.resize(1 + HASH_MASK_ONES, ss) (v42:263): (1 + 3) * ss seeds drawn in
    // a row; the insert uses groups 0..2 (seed + j * seed_sz, v42:346), group
    // 0 also for the bucket and the mask (v42:333-334)
    std::default_random_engine seed_eng, mask_eng;
    qs_seed_engine(seed_eng, rng_seed, 0);
    qs_seed_engine(mask_eng, rng_seed, 1);
    std::uniform_int_distribution<unsigned int> dis(1, std::numeric_limits<unsigned int>::max());
    for (uint32_t g = 0; g < 4; ++g) {
        uint64_t sum = 0;
        for (uint32_t k = 0; k < seed_sz; ++k)
            sum += dis(seed_eng);
        if (g < 3)
            s->seed[g] = sum;
    }
    // generate_hashmask(n_masks, m, 3, 0) with its index table (v42:268-269,
    // lib/hashmask.cu:32-82): each mask's three counters in the order drawn
    s->h_masks = (uint32_t *)calloc(n_masks, sizeof(uint32_t));
    s->h_index = (uint32_t *)calloc(n_masks, sizeof(uint32_t));
    if (!s->h_masks || !s->h_index) {
        qs_sf_destroy(s);
        return QS_ERR_NOMEM;
    }
    std::uniform_int_distribution<unsigned int> cdis(0, m - 1);
    for (uint32_t i = 0; i < n_masks; ++i) {
        uint32_t j = 0, mask = 0, ix = 0;
        while (j < 3) {
            const uint32_t id = cdis(mask_eng);
            if (!(mask >> id & 1u)) {
                mask |= 1u << id;
                ix |= id << (8 * j);
                ++j;
            }
        }
        s->h_masks[i] = mask;
        s->h_index[i] = ix;
    }
    const uint64_t cells = n * m;
    bool ok = hipMalloc(&s->d_masks, 4ull * n_masks) == hipSuccess &&
              hipMalloc(&s->d_index, 4ull * n_masks) == hipSuccess &&
              hipMalloc(&s->d_split, 4ull * cells * times) == hipSuccess &&
              hipMalloc(&s->d_search, 4ull * cells) == hipSuccess;
    if (!ok) {
        qs_sf_destroy(s);
        return QS_ERR_NOMEM;
    }
    if (hipMemcpy(s->d_masks, s->h_masks, 4ull * n_masks, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_index, s->h_index, 4ull * n_masks, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->d_split, 0, 4ull * cells * times) != hipSuccess ||
        hipMemset(s->d_search, 0, 4ull * cells) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        qs_sf_destroy(s);
        return QS_ERR_HIP;
    }
    *out = s;
    return QS_OK;
}

static uint32_t sf_grid(uint64_t want)
{
    return (uint32_t)(want < 1 ? 1 : (want > 8192 ? 8192 : want));
}

extern "C" int qs_sf_insert(qs_sf *s, const void *d_keys, size_t n)
{
    if (!s || (!d_keys && n))
        return QS_ERR_PARAM;
    if (n) {
        const sf_par p = sf_par_of(s);
        const dim3 grid((uint32_t)((n + SF_TILE - 1) / SF_TILE)), block(SF_THREADS);
        if (s->key_bytes == 4)
            hipLaunchKernelGGL(k_sf_insert<4>, grid, block, 0, 0, (const uint32_t *)d_keys, (uint64_t)n, p,
                               s->d_index, s->d_split);
        else
            hipLaunchKernelGGL(k_sf_insert<8>, grid, block, 0, 0, (const uint64_t *)d_keys, (uint64_t)n, p,
                               s->d_index, s->d_split);
        CK(hipGetLastError());
    }
    // update() after every insert call, as the reference does (v42:358)
    const uint64_t cells = s->n * s->m;
    hipLaunchKernelGGL(k_sf_update, dim3(sf_grid((cells + 255) / 256)), dim3(256), 0, 0, s->d_split, cells,
                       s->times, s->d_search);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_sf_query(qs_sf *s, const void *d_keys, size_t n, uint32_t *d_counts)
{
    if (!s || ((!d_keys || !d_counts) && n))
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    const sf_par p = sf_par_of(s);
    const uint32_t G = s->m / 4u; // lanes per key
    const dim3 grid(sf_grid(((uint64_t)n * G + SF_THREADS - 1) / SF_THREADS)), block(SF_THREADS);
#define SF_Q(KB, GG)                                                                                             \
    hipLaunchKernelGGL((k_sf_query<KB, GG>), grid, block, 0, 0, (const KeyT<KB> *)d_keys, (uint64_t)n, p,         \
                       s->d_masks, s->d_search, d_counts)
    if (s->key_bytes == 4) {
        if (G == 2)
            SF_Q(4, 2);
        else if (G == 4)
            SF_Q(4, 4);
        else
            SF_Q(4, 8);
    } else {
        if (G == 2)
            SF_Q(8, 2);
        else if (G == 4)
            SF_Q(8, 4);
        else
            SF_Q(8, 8);
    }
#undef SF_Q
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_sf_clear(qs_sf *s)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    CK(hipMemset(s->d_split, 0, 4ull * s->n * s->m * s->times));
    CK(hipMemset(s->d_search, 0, 4ull * s->n * s->m));
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_sf_export(qs_sf *s, uint32_t *split, uint32_t *search, uint32_t *masks, uint8_t *index,
                            uint64_t *seeds)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    if (split)
        CK(hipMemcpy(split, s->d_split, 4ull * s->n * s->m * s->times, hipMemcpyDeviceToHost));
    if (search)
        CK(hipMemcpy(search, s->d_search, 4ull * s->n * s->m, hipMemcpyDeviceToHost));
    if (masks)
        memcpy(masks, s->h_masks, 4ull * s->n_masks);
    if (index)
        for (uint32_t i = 0; i < s->n_masks; ++i)
            for (uint32_t j = 0; j < 3; ++j)
                index[(size_t)i * 3 + j] = (uint8_t)(s->h_index[i] >> (8 * j));
    if (seeds)
        for (int g = 0; g < 3; ++g)
            seeds[g] = s->seed[g];
    return QS_OK;
}
