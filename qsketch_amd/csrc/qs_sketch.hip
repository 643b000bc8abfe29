// qs_sketch.hip -- MI355X-native QSketch (two-level MSWS): host API + kernels.
//
// Reference path replaced: include/sketch/device_sketch/device_sketch_v3.h
// (insert_kernel v3:193-290, search_kernel v3:410-521, atomic_malloc v3:68-85,
// class Sketch_GPU_Sub_Warp_Level v3:524-634) and its parameter / mask / seed
// set-up (lib/default.cu, lib/hashmask.cu, include/tool/tool.h:186-220).
//
// The insert pipeline lives in qs_insert.hip; this file holds the host API
// (parameters, seeds and masks with the reference's libstdc++ call pattern,
// create / destroy / clear / import / export / stats / timing) and the query
// kernel.
#include "qs_internal.h"

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess) {                                                                 \
            return QS_ERR_HIP;                                                                   \
        }                                                                                        \
    } while (0)

#define INS_THREADS 256

// Pending event pairs of the timing hooks, resolved by qs_get_timing.
struct qs_timer {
    std::vector<qs_evpair> pend;
};

void qs_tim_begin(qs_sketch *s, qs_evpair &e, int kind, uint64_t keys, hipStream_t st)
{
    e.kind = kind;
    e.keys = keys;
    e.a = e.b = nullptr;
    if (!s->timing)
        return;
    hipEventCreate(&e.a);
    hipEventCreate(&e.b);
    hipEventRecord(e.a, st);
}

void qs_tim_end(qs_sketch *s, qs_evpair &e, hipStream_t st)
{
    if (!s->timing || !e.a)
        return;
    hipEventRecord(e.b, st);
    ((qs_timer *)s->tim)->pend.push_back(e);
}

static void tim_resolve(qs_sketch *s)
{
    if (!s->tim)
        return;
    qs_timer *t = (qs_timer *)s->tim;
    for (auto &e : t->pend) {
        float ms = 0.f;
        hipEventSynchronize(e.b);
        hipEventElapsedTime(&ms, e.a, e.b);
        if (e.kind >= 0 && e.kind < QS_TIMING_KERNELS) {
            s->tot.kernel_ms[e.kind] += ms;
            s->tot.kernel_launches[e.kind]++;
            if (e.kind == QS_TK_QUERY) {
                s->tot.query_kernel_ms += ms;
                s->tot.query_kernel_launches++;
                s->tot.query_keys += e.keys;
            } else {
                s->tot.insert_kernel_ms += ms;
                s->tot.insert_kernel_launches++;
            }
        } else if (e.kind == QS_TIMING_KERNELS) {
            s->tot.insert_keys += e.keys; // marker: keys of one insert chunk
        }
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    t->pend.clear();
}

// ===========================================================================
// Host: parameter set, prime table, seeds and masks
// ===========================================================================

// find_greatest_prime's table (lib/qsketch.cu:51-107): the largest primes not
// above 2^k and 3*2^(k-1).  Kept identical because it fixes n_low, and n_low
// fixes every estimate.
static const uint64_t k_prime_table[] = {
    402653171ull, 268435399ull, 201326557ull, 134217689ull, 100663291ull, 67108859ull,
    50331599ull,  33554393ull,  25165813ull,  16777213ull,  12582893ull,  8388593ull,
    6291449ull,   4194301ull,   3145721ull,   2097143ull,   1572853ull,   1048573ull,
    786431ull,    524287ull,    393209ull,    262139ull,    196597ull,    131071ull,
    98299ull,     65521ull,     49139ull,     32749ull,     24571ull,     16381ull,
    12281ull,     8191ull,      6143ull,      4093ull,      3067ull,      2039ull,
    1531ull,      1021ull,      761ull,       509ull,       383ull,       251ull,
    191ull,       127ull,       89ull,        61ull,        47ull,        31ull,
    23ull,        13ull,        11ull,        7ull,         5ull,         3ull,
    2ull};

extern "C" uint64_t qs_find_greatest_prime(uint64_t upper_bound)
{
    for (uint64_t p : k_prime_table)
        if (p <= upper_bound)
            return p;
    return 0;
}

extern "C" int qs_default_params(qs_params *p, uint64_t expected_keys, double wlf)
{
    if (!p || !(wlf > 0.0))
        return QS_ERR_PARAM;
    memset(p, 0, sizeof(*p));
    p->m_low = 8;
    p->m_high = 32;
    p->n_high = 1024;
    p->seed_sz = 1;
    p->n_masks = 1021;
    p->mask_ones = 3;
    p->link_start = QS_LINK_START;
    p->low_limit = QS_LOW_LIMIT;
    p->key_bytes = 4;
    p->semantics = QS_LITERAL;
    p->hash_seed = 0;
    p->rng_seed = 1;
    p->chunk_keys = 0;
    // Sketch_GPU::number_of_buckets (device_sketch.h:371-376), in double then
    // truncated, with m = 8 words and HASH_MASK_ONES = 3.
    double row = (((double)expected_keys / wlf) * 3.0) / 8.0;
    p->n_low = qs_find_greatest_prime((uint64_t)row);
    return p->n_low ? QS_OK : QS_ERR_PARAM;
}

extern "C" const char *qs_strerror(int code)
{
    switch (code) {
    case QS_OK: return "ok";
    case QS_ERR_PARAM: return "invalid parameter";
    case QS_ERR_NOMEM: return "out of memory";
    case QS_ERR_POOL: return "high-bucket pool exhausted (state invalid until qs_clear)";
    case QS_ERR_HIP: return "HIP runtime error";
    case QS_ERR_STATE: return "malformed state";
    case QS_ERR_NODEVICE: return "no usable GPU";
    default: return "unknown error";
    }
}

// The reference seeds std::default_random_engine from 128 bytes of
// /dev/urandom through std::seed_seq (tool.h:189-201).  Here the 128 bytes are
// a SplitMix64 stream of (rng_seed, stream) so that a sketch is reproducible;
// stream 0 feeds the seed engine (gpu_tool<uint>), stream 1 the mask engine
// (cpu_tool<uint>).
static void seed_engine(std::default_random_engine &eng, uint64_t rng_seed, uint64_t stream)
{
    char buf[128];
    uint64_t st = rng_seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 16; ++i) {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        for (int k = 0; k < 8; ++k)
            buf[i * 8 + k] = (char)(uint8_t)(z >> (8 * k));
    }
    std::seed_seq sq(buf, buf + 128);
    eng.seed(sq);
}

// generate_hashmask(nullptr, n, 32, 3, 4) (lib/hashmask.cu:32-82): draw bit
// positions uniformly in [0, 27] until three distinct ones are set.
static void make_masks(std::default_random_engine &eng, uint32_t n, uint32_t *out)
{
    std::uniform_int_distribution<unsigned int> dis(0, 32 - 4 - 1);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t m = 0;
        uint32_t ones = 0;
        while (ones < 3) {
            uint32_t id = dis(eng);
            if (!(m & (1u << id))) {
                m |= 1u << id;
                ++ones;
            }
        }
        out[i] = m;
    }
}

// ===========================================================================
// Device helpers
// ===========================================================================

__device__ static inline uint32_t sel_word(const uint4 &a, const uint4 &b, uint32_t k)
{
    uint32_t x01 = (k & 1u) ? a.y : a.x;
    uint32_t x23 = (k & 1u) ? a.w : a.z;
    uint32_t x45 = (k & 1u) ? b.y : b.x;
    uint32_t x67 = (k & 1u) ? b.w : b.z;
    uint32_t x03 = (k & 2u) ? x23 : x01;
    uint32_t x47 = (k & 2u) ? x67 : x45;
    return (k & 4u) ? x47 : x03;
}

__device__ static inline uint32_t byte_at(uint32_t w, uint32_t j)
{
    return (w >> ((j & 3u) * 8u)) & 0xffu;
}

// any byte selected by add (bytes of add are 0 or 1..3) has old value > 128
__device__ static inline bool over_limit(uint32_t old, uint32_t add)
{
    bool r = false;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r |= (((add >> (8 * k)) & 0xffu) != 0u) && (((old >> (8 * k)) & 0xffu) > QS_LOW_LIMIT);
    return r;
}

// Split a 3-bit mask into at most three (word, packed add) pairs.
struct low_adds {
    uint32_t k[3];
    uint32_t a[3];
};

__device__ static inline low_adds split_mask(uint32_t M)
{
    uint32_t j1 = (uint32_t)__builtin_ctz(M);
    uint32_t M1 = M & (M - 1u);
    uint32_t j2 = (uint32_t)__builtin_ctz(M1);
    uint32_t M2 = M1 & (M1 - 1u);
    uint32_t j3 = (uint32_t)__builtin_ctz(M2);
    uint32_t k1 = j1 >> 2, k2 = j2 >> 2, k3 = j3 >> 2;
    uint32_t a1 = 1u << ((j1 & 3u) * 8u), a2 = 1u << ((j2 & 3u) * 8u), a3 = 1u << ((j3 & 3u) * 8u);
    low_adds r;
    r.k[0] = k1;
    r.a[0] = a1;
    r.k[1] = k2;
    r.a[1] = 0;
    r.k[2] = k3;
    r.a[2] = 0;
    if (k2 == k1)
        r.a[0] += a2;
    else
        r.a[1] = a2;
    if (k3 == k1)
        r.a[0] += a3;
    else if (k3 == k2)
        r.a[1] += a3;
    else
        r.a[2] = a3;
    return r;
}

__device__ static inline bool bit_test(const uint32_t *bm, uint32_t b)
{
    return (bm[b >> 5] >> (b & 31u)) & 1u;
}

__device__ static inline uint32_t lds_mask(const uint32_t *s_masks, uint32_t mu)
{
    return s_masks[mu];
}


// ===========================================================================
// Query: two lanes per key, each loading one 16-byte half of the 32-byte
// bucket, so the pair issues a single coalesced request per key (a random
// access costs one DRAM activation whatever its size, so requests per key are
// what bounds this kernel).  Replaces search_kernel (v3:410-521):
// est = min over the three selected byte counters
//       + (linked ? min over the same three high counters : 0).
// ===========================================================================
__device__ static inline uint32_t sel4(const uint4 &v, uint32_t k)
{
    uint32_t lo = (k & 1u) ? v.y : v.x;
    uint32_t hi = (k & 1u) ? v.w : v.z;
    return (k & 2u) ? hi : lo;
}

template <int KB>
__global__ __launch_bounds__(INS_THREADS) void k_query(const KeyT<KB> *__restrict__ keys, uint32_t n,
                                                       const uint32_t *__restrict__ low,
                                                       const uint32_t *__restrict__ high,
                                                       const uint32_t *__restrict__ masks,
                                                       qs_hashpar hp, uint32_t *__restrict__ out)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t i = threadIdx.x; i < hp.n_masks; i += INS_THREADS)
        s_masks[i] = masks[i];
    __syncthreads();
    const uint32_t half = threadIdx.x & 1u;
    const uint64_t stride = (uint64_t)gridDim.x * (INS_THREADS / 2);
    for (uint64_t i = (uint64_t)blockIdx.x * (INS_THREADS / 2) + (threadIdx.x >> 1); i < n; i += stride) {
        const KeyT<KB> key = keys[i];
        const uint32_t M = s_masks[qs_maskidx(key, hp)];
        const uint4 v = reinterpret_cast<const uint4 *>(low + (size_t)qs_bucket(key, hp) * 8u)[half];
        const uint32_t Mh = half ? (M >> 16) : (M & 0xffffu);
        uint32_t lo = 0xffffffffu;
        for (uint32_t m = Mh; m; m &= m - 1u) {
            const uint32_t jj = (uint32_t)__builtin_ctz(m);
            lo = min(lo, (sel4(v, jj >> 2) >> ((jj & 3u) * 8u)) & 0xffu);
        }
        lo = min(lo, (uint32_t)__shfl_xor(lo, 1, 64));
        const uint32_t wo = (uint32_t)__shfl_xor(v.w, 1, 64);
        const uint32_t w = half ? v.w : wo;
        uint32_t hi = 0;
        if (w > QS_LINK_START) {
            const uint32_t *g = high + (size_t)(w - QS_LINK_START) * 32u + 16u * half;
            hi = 0xffffffffu;
            for (uint32_t m = Mh; m; m &= m - 1u)
                hi = min(hi, g[__builtin_ctz(m)]);
            hi = min(hi, (uint32_t)__shfl_xor(hi, 1, 64));
        }
        if (!half)
            out[i] = lo + hi;
    }
}

// ===========================================================================
// Host: create / destroy / scratch
// ===========================================================================

static int g_cu_count = 0;

static int reset_state(qs_sketch *s, uint32_t next)
{
    qs_dev_state h;
    memset(&h, 0, sizeof(h));
    h.next = next;
    CK(hipMemcpy(s->d_st, &h, sizeof(h), hipMemcpyHostToDevice));
    return QS_OK;
}

static int validate_params(const qs_params *p)
{
    if (!p)
        return QS_ERR_PARAM;
    if (p->m_low != 8 || p->m_high != 32 || p->mask_ones != 3 || p->link_start != QS_LINK_START ||
        p->low_limit != QS_LOW_LIMIT)
        return QS_ERR_PARAM;
    if (p->key_bytes != 4 && p->key_bytes != 8)
        return QS_ERR_PARAM;
    if (p->n_low < 1 || p->n_low >= (1ull << 31))
        return QS_ERR_PARAM;
    if (p->key_bytes == 4 && p->n_low > (1ull << 29)) // v3:42 computes the word offset in 32 bits
        return QS_ERR_PARAM;
    if (p->n_high < 2 || p->n_high > (uint64_t)0x7fffffffu - QS_LINK_START)
        return QS_ERR_PARAM;
    if (p->n_masks < 1 || p->n_masks > QS_MAX_MASKS)
        return QS_ERR_PARAM;
    if (p->seed_sz < 1 || (p->semantics != QS_LITERAL && p->semantics != QS_PAPER))
        return QS_ERR_PARAM;
    return QS_OK;
}

extern "C" void qs_destroy(qs_sketch *s)
{
    if (!s)
        return;
    if (s->tim) {
        hipDeviceSynchronize();
        tim_resolve(s);
        delete (qs_timer *)s->tim;
        s->tim = nullptr;
    }
    qs_free_scratch(s);
    hipFree(s->d_low);
    hipFree(s->d_high);
    hipFree(s->d_masks);
    hipFree(s->d_st);
    hipFree(s->d_longbm);
    hipFree(s->d_tail_of);
    if (s->h_st)
        hipHostFree(s->h_st);
    free(s->h_masks);
    free(s);
}

static void set_hashpar(qs_sketch *s)
{
    s->hp.low = qs_fastdiv_make((uint32_t)s->p.n_low);
    s->hp.masks = qs_fastdiv_make(s->p.n_masks);
    s->hp.seed = s->seed;
    s->hp.n_low = (uint32_t)s->p.n_low;
    s->hp.n_masks = s->p.n_masks;
}

extern "C" int qs_create(const qs_params *p, qs_sketch **out)
{
    if (!out)
        return QS_ERR_PARAM;
    *out = nullptr;
    int rc = validate_params(p);
    if (rc)
        return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return QS_ERR_NODEVICE;
    qs_sketch *s = (qs_sketch *)calloc(1, sizeof(qs_sketch));
    if (!s)
        return QS_ERR_NOMEM;
    s->p = *p;
    if (hipGetDevice(&s->device) != hipSuccess) {
        free(s);
        return QS_ERR_HIP;
    }
    if (g_cu_count == 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, s->device) == hipSuccess)
            g_cu_count = prop.multiProcessorCount;
        if (g_cu_count <= 0)
            g_cu_count = 256;
    }

    // seeds (Device_Seed::resize, device_sketch.h:280-302) and masks (v3:558-562)
    std::default_random_engine seed_eng, mask_eng;
    seed_engine(seed_eng, p->rng_seed, 0);
    seed_engine(mask_eng, p->rng_seed, 1);
    uint32_t seed = 0;
    {
        std::uniform_int_distribution<unsigned int> dis(1, std::numeric_limits<unsigned int>::max());
        for (uint32_t i = 0; i < p->seed_sz; ++i)
            seed += dis(seed_eng);
    }
    s->seed = p->hash_seed ? p->hash_seed : seed;
    s->h_masks = (uint32_t *)malloc(sizeof(uint32_t) * p->n_masks);
    if (!s->h_masks) {
        free(s);
        return QS_ERR_NOMEM;
    }
    make_masks(mask_eng, p->n_masks, s->h_masks);
    set_hashpar(s);

    size_t low_bytes = (size_t)p->n_low * 32u, high_bytes = (size_t)p->n_high * 128u;
    size_t viol_words = (size_t)((p->n_low + 31) / 32);
    s->bucket_bits = p->n_low > 1 ? 64 - __builtin_clzll(p->n_low - 1) : 0;
    s->pk.bb = s->bucket_bits > 0 ? (uint32_t)s->bucket_bits : 1u;
    s->pk.bm = p->n_masks > 1 ? 32u - (uint32_t)__builtin_clz(p->n_masks - 1u) : 1u;
    s->pk.ishift = s->pk.bb + s->pk.bm;
    s->pk.mmask = (1u << s->pk.bm) - 1u;
    s->pk.bmask = (1ull << s->pk.bb) - 1ull;
    bool ok = hipMalloc(&s->d_low, low_bytes) == hipSuccess &&
              hipMalloc(&s->d_high, high_bytes) == hipSuccess &&
              hipMalloc(&s->d_masks, sizeof(uint32_t) * p->n_masks) == hipSuccess &&
              hipMalloc(&s->d_st, sizeof(qs_dev_state)) == hipSuccess &&
              hipMalloc(&s->d_longbm, sizeof(uint32_t) * viol_words) == hipSuccess &&
              hipMalloc(&s->d_tail_of, sizeof(uint32_t) * p->n_high) == hipSuccess &&
              hipHostMalloc(&s->h_st, sizeof(qs_dev_state), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        qs_destroy(s);
        return QS_ERR_NOMEM;
    }
    if (hipMemset(s->d_low, 0, low_bytes) != hipSuccess ||
        hipMemset(s->d_high, 0, high_bytes) != hipSuccess ||
        hipMemset(s->d_longbm, 0, sizeof(uint32_t) * viol_words) != hipSuccess ||
        hipMemcpy(s->d_masks, s->h_masks, sizeof(uint32_t) * p->n_masks, hipMemcpyHostToDevice) !=
            hipSuccess ||
        reset_state(s, QS_LINK_START + 1) != QS_OK || hipDeviceSynchronize() != hipSuccess) {
        qs_destroy(s);
        return QS_ERR_HIP;
    }
    *out = s;
    return QS_OK;
}

extern "C" int qs_clear(qs_sketch *s)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    CK(hipMemset(s->d_low, 0, (size_t)s->p.n_low * 32u));
    CK(hipMemset(s->d_high, 0, (size_t)s->p.n_high * 128u));
    CK(hipMemset(s->d_longbm, 0, sizeof(uint32_t) * (size_t)((s->p.n_low + 31) / 32)));
    int rc = reset_state(s, QS_LINK_START + 1);
    if (rc)
        return rc;
    CK(hipDeviceSynchronize());
    s->keys_inserted = 0;
    s->sticky_err = 0;
    return QS_OK;
}

static uint32_t grid_for(uint64_t n, uint32_t threads, uint32_t per_cu)
{
    uint64_t blocks = (n + threads - 1) / threads;
    uint64_t cap = (uint64_t)g_cu_count * per_cu;
    if (blocks > cap)
        blocks = cap;
    return blocks ? (uint32_t)blocks : 1u;
}

static int insert_dispatch(qs_sketch *s, const void *d_keys, size_t n, hipStream_t stream)
{
    if (!s || (!d_keys && n))
        return QS_ERR_PARAM;
    if (s->sticky_err)
        return s->sticky_err;
    if (n == 0)
        return QS_OK;
    return qs_insert_chunks(s, d_keys, n, stream);
}

extern "C" int qs_insert(qs_sketch *s, const void *d_keys, size_t n)
{
    int rc = insert_dispatch(s, d_keys, n, 0);
    if (rc)
        return rc;
    CK(hipStreamSynchronize(0));
    return QS_OK;
}

extern "C" int qs_insert_async(qs_sketch *s, const void *d_keys, size_t n, void *stream)
{
    return insert_dispatch(s, d_keys, n, (hipStream_t)stream);
}

static int query_dispatch(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts,
                          hipStream_t stream)
{
    if (!s || ((!d_keys || !d_counts) && n))
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    if (n > 0xffffffffull)
        return QS_ERR_PARAM;
    uint32_t nn = (uint32_t)n;
    qs_evpair ev;
    qs_tim_begin(s, ev, QS_TK_QUERY, nn, stream);
    if (s->p.key_bytes == 4)
        hipLaunchKernelGGL(k_query<4>, dim3(grid_for((uint64_t)nn * 2, INS_THREADS, 8)), dim3(INS_THREADS), 0, stream,
                           (const uint32_t *)d_keys, nn, s->d_low, s->d_high, s->d_masks, s->hp,
                           d_counts);
    else
        hipLaunchKernelGGL(k_query<8>, dim3(grid_for((uint64_t)nn * 2, INS_THREADS, 8)), dim3(INS_THREADS), 0, stream,
                           (const uint64_t *)d_keys, nn, s->d_low, s->d_high, s->d_masks, s->hp,
                           d_counts);
    qs_tim_end(s, ev, stream);
    CK(hipGetLastError());
    return QS_OK;
}

extern "C" int qs_query(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts)
{
    int rc = query_dispatch(s, d_keys, n, d_counts, 0);
    if (rc)
        return rc;
    CK(hipStreamSynchronize(0));
    return QS_OK;
}

extern "C" int qs_query_async(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts,
                              void *stream)
{
    return query_dispatch(s, d_keys, n, d_counts, (hipStream_t)stream);
}

extern "C" int qs_sync(qs_sketch *s)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    return s->sticky_err;
}

// ===========================================================================
// State export / import (reference layout) and statistics
// ===========================================================================
extern "C" int qs_export_state(qs_sketch *s, uint32_t *low, uint32_t *high, uint32_t *masks,
                               uint32_t *seed, uint32_t *next)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    if (low)
        CK(hipMemcpy(low, s->d_low, (size_t)s->p.n_low * 32u, hipMemcpyDeviceToHost));
    if (high)
        CK(hipMemcpy(high, s->d_high, (size_t)s->p.n_high * 128u, hipMemcpyDeviceToHost));
    if (masks)
        memcpy(masks, s->h_masks, sizeof(uint32_t) * s->p.n_masks);
    if (seed)
        *seed = s->seed;
    if (next) {
        qs_dev_state h;
        CK(hipMemcpy(&h, s->d_st, sizeof(h), hipMemcpyDeviceToHost));
        *next = h.next;
    }
    return QS_OK;
}

extern "C" int qs_import_state(qs_sketch *s, const uint32_t *low, const uint32_t *high,
                               const uint32_t *masks, uint32_t seed, uint32_t next)
{
    if (!s || !low || !high || !masks)
        return QS_ERR_PARAM;
    const uint64_t n_high = s->p.n_high;
    if (next < QS_LINK_START + 1 || (uint64_t)(next - QS_LINK_START) > n_high)
        return QS_ERR_STATE;
    for (uint32_t i = 0; i < s->p.n_masks; ++i)
        if (__builtin_popcount(masks[i]) != 3 || (masks[i] >> 28) != 0)
            return QS_ERR_STATE;
    for (uint64_t b = 0; b < s->p.n_low; ++b) {
        uint32_t w = low[b * 8 + 7];
        if (w != 0 && (w <= QS_LINK_START || w >= next))
            return QS_ERR_STATE;
    }
    CK(hipDeviceSynchronize());
    memcpy(s->h_masks, masks, sizeof(uint32_t) * s->p.n_masks);
    s->seed = seed;
    set_hashpar(s);
    CK(hipMemcpy(s->d_low, low, (size_t)s->p.n_low * 32u, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_high, high, (size_t)n_high * 128u, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_masks, masks, sizeof(uint32_t) * s->p.n_masks, hipMemcpyHostToDevice));
    CK(hipMemset(s->d_longbm, 0, sizeof(uint32_t) * (size_t)((s->p.n_low + 31) / 32)));
    int rc = reset_state(s, next);
    if (rc)
        return rc;
    CK(hipDeviceSynchronize());
    s->sticky_err = 0;
    s->keys_inserted = 0;
    return QS_OK;
}

extern "C" int qs_set_timing(qs_sketch *s, int enable)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    if (!s->tim)
        s->tim = new qs_timer();
    tim_resolve(s);
    memset(&s->tot, 0, sizeof(s->tot));
    s->timing = enable ? 1 : 0;
    return QS_OK;
}

extern "C" int qs_get_timing(qs_sketch *s, qs_timing *t)
{
    if (!s || !t)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    tim_resolve(s);
    *t = s->tot;
    return QS_OK;
}

extern "C" int qs_get_stats(qs_sketch *s, qs_stats *st)
{
    if (!s || !st)
        return QS_ERR_PARAM;
    qs_dev_state h;
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(&h, s->d_st, sizeof(h), hipMemcpyDeviceToHost));
    memset(st, 0, sizeof(*st));
    st->n_low = s->p.n_low;
    st->n_high = s->p.n_high;
    st->n_masks = s->p.n_masks;
    st->bytes_low = s->p.n_low * 32u;
    st->bytes_high = s->p.n_high * 128u;
    st->keys_inserted = s->keys_inserted;
    st->chunks = s->chunks;
    st->chunks_with_fixup = s->chunks_fix;
    st->fixup_keys = s->fix_keys;
    st->next_id = h.next;
    st->linked_buckets = h.next - (QS_LINK_START + 1);
    st->seed = s->seed;
    return QS_OK;
}
