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
#include <mutex>
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
    e.shared_a = false;
    if (!s->timing)
        return;
    // timing only: no system-scope fence at the records (the host reads the
    // sketch's memory after its own synchronisation, never through these)
    (void)hipEventCreateWithFlags(&e.a, hipEventDisableSystemFence);
    (void)hipEventCreateWithFlags(&e.b, hipEventDisableSystemFence);
    (void)hipEventRecord(e.a, st);
}

void qs_tim_end(qs_sketch *s, qs_evpair &e, hipStream_t st)
{
    if (!s->timing || !e.a)
        return;
    (void)hipEventRecord(e.b, st);
    ((qs_timer *)s->tim)->pend.push_back(e);
}

void qs_tim_next(qs_sketch *s, qs_evpair &prev, qs_evpair &next, int kind, uint64_t keys, hipStream_t st)
{
    if (!s->timing || !prev.a) {
        qs_tim_end(s, prev, st);
        qs_tim_begin(s, next, kind, keys, st);
        return;
    }
    next.kind = kind;
    next.keys = keys;
    (void)hipEventRecord(prev.b, st); // ends prev and begins next
    ((qs_timer *)s->tim)->pend.push_back(prev);
    next.a = prev.b;
    next.shared_a = true;
    (void)hipEventCreateWithFlags(&next.b, hipEventDisableSystemFence);
}

static void tim_resolve(qs_sketch *s)
{
    if (!s->tim)
        return;
    qs_timer *t = (qs_timer *)s->tim;
    for (auto &e : t->pend) {
        float ms = 0.f;
        (void)hipEventSynchronize(e.b);
        (void)hipEventElapsedTime(&ms, e.a, e.b);
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
    }
    // destroyed after every elapsed time is read: a shared event is one
    // pair's b and the next pair's a, and goes with the first
    for (auto &e : t->pend) {
        if (!e.shared_a)
            (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
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
    case -7: return "exact table full (learned form: more routed keys than slots)";
    case QS_ERR_INTERNAL: return "device consistency check failed (state invalid until qs_clear)";
    default: return "unknown error";
    }
}

// The reference seeds std::default_random_engine from 128 bytes of
// /dev/urandom through std::seed_seq (tool.h:189-201).  Here the 128 bytes are
// a SplitMix64 stream of (rng_seed, stream) so that a sketch is reproducible;
// stream 0 feeds the seed engine (gpu_tool<uint>), stream 1 the mask engine
// (cpu_tool<uint>).
void qs_seed_engine(std::default_random_engine &eng, uint64_t rng_seed, uint64_t stream)
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

// generate_hashmask(nullptr, n, m_bits, 3, padding) (lib/hashmask.cu:32-82):
// draw bit positions uniformly in [0, m_bits - padding - 1] until three
// distinct ones are set.  The two-level sketch uses (32, 4), i.e. [0, 27].
void qs_make_masks(std::default_random_engine &eng, uint32_t n, uint32_t m_bits, uint32_t padding,
                   uint32_t *out)
{
    std::uniform_int_distribution<unsigned int> dis(0, m_bits - padding - 1);
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
// Query: two lanes per key, each loading one 16-byte half of the 32-byte
// bucket, so the pair issues a single coalesced request per key (a random
// access costs one DRAM activation whatever its size, so requests per key are
// what bounds this kernel).  Replaces search_kernel (v3:410-521):
// est = min over the three selected byte counters
//       + (linked ? min over the same three high counters : 0).
// A batch of several sketches finds each key's sketch from the batch offsets.
// ===========================================================================
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ static inline uint32_t sel4(const uint4 &v, uint32_t k)
{
    uint32_t lo = (k & 1u) ? v.y : v.x;
    uint32_t hi = (k & 1u) ? v.w : v.z;
    return (k & 2u) ? hi : lo;
}

__device__ static inline uint32_t query_sid(const uint64_t *__restrict__ offs, uint32_t nsk, uint64_t pos)
{
    uint32_t lo = 0, hi = nsk; // offs[lo] <= pos < offs[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (offs[mid] <= pos)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <int KB, bool MULTI>
__global__ __launch_bounds__(INS_THREADS) void k_query(const KeyT<KB> *__restrict__ keys, uint32_t n, qs_sdesc d0,
                                                       const qs_sdesc *__restrict__ D,
                                                       const uint64_t *__restrict__ offs, uint32_t nsk,
                                                       const uint32_t *__restrict__ low_pool,
                                                       const uint32_t *__restrict__ high_pool,
                                                       const uint32_t *__restrict__ mask_pool,
                                                       uint32_t *__restrict__ out)
{
    __shared__ uint32_t s_masks[MULTI ? 1 : QS_MAX_MASKS];
    if (!MULTI)
        for (uint32_t i = threadIdx.x; i < d0.n_masks; i += INS_THREADS)
            s_masks[i] = mask_pool[i];
    __syncthreads();
    const uint32_t half = threadIdx.x & 1u;
    const uint64_t stride = (uint64_t)gridDim.x * (INS_THREADS / 2);
    for (uint64_t i = (uint64_t)blockIdx.x * (INS_THREADS / 2) + (threadIdx.x >> 1); i < n; i += stride) {
        const KeyT<KB> key = keys[i];
        const uint32_t *low = low_pool;
        const uint32_t *high = high_pool;
        uint32_t M, b;
        if (MULTI) {
            const qs_sdesc &d = D[query_sid(offs, nsk, i)];
            M = mask_pool[d.masks_off + qs_maskidx_d(key, d)];
            b = qs_bucket_d(key, d);
            low += d.low_off;
            high += d.high_off;
        } else {
            M = s_masks[qs_maskidx_d(key, d0)];
            b = qs_bucket_d(key, d0);
        }
        // read once: a non-temporal load (measured 0.433 -> 0.405 ms for 16.6 M keys)
        const u32x4 vv = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(low + (size_t)b * 8u) + half);
        const uint4 v = make_uint4(vv[0], vv[1], vv[2], vv[3]);
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
// Host: create / destroy
// ===========================================================================

static int g_cu_count = 0;

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
    // hash_mul_add sums key * seed_i in 64 bits for 64-bit keys
    // (device_hash_function.h:6-17): the seed sum can exceed 32 bits, and the
    // device hash multiplies by a 32-bit seed, so 64-bit keys take one seed
    if (p->key_bytes == 8 && p->seed_sz > 1 && p->hash_seed == 0)
        return QS_ERR_PARAM;
    return QS_OK;
}

// zero the device control block and restart every sketch's allocator at 1025
static int reset_state(qs_sketch *s)
{
    qs_dev_state h;
    memset(&h, 0, sizeof(h));
    CK(hipMemcpy(s->d_st, &h, sizeof(h), hipMemcpyHostToDevice));
    std::vector<uint32_t> nx(s->nsk, QS_LINK_START + 1);
    CK(hipMemcpy(s->d_next, nx.data(), sizeof(uint32_t) * s->nsk, hipMemcpyHostToDevice));
    return QS_OK;
}

// zero the control block only (counts, statistics, pool error)
static int reset_counts(qs_sketch *s)
{
    qs_dev_state h;
    memset(&h, 0, sizeof(h));
    CK(hipMemcpy(s->d_st, &h, sizeof(h), hipMemcpyHostToDevice));
    return QS_OK;
}

static void set_hash(qs_sdesc &d, uint32_t seed)
{
    d.seed = seed;
    d.fd_low = qs_fastdiv_make(d.n_low);
    d.fd_masks = qs_fastdiv_make(d.n_masks);
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
    qs_hot_free(s);
    hipFree(s->d_low);
    hipFree(s->d_high);
    hipFree(s->d_masks);
    hipFree(s->d_next);
    hipFree(s->d_st);
    hipFree(s->d_longbm);
    hipFree(s->d_tail_of);
    hipFree(s->d_offs);
    hipFree(s->d_desc);
    if (s->h_st)
        hipHostFree(s->h_st);
    free(s->ps);
    free(s->seeds);
    free(s->h_masks);
    free(s->h_desc);
    free(s);
}

static int create_set(const qs_params *ps, uint32_t nsk, qs_sketch **out)
{
    if (!out || !ps || nsk < 1 || nsk > (1u << 20))
        return QS_ERR_PARAM;
    *out = nullptr;
    for (uint32_t i = 0; i < nsk; ++i) {
        int rc = validate_params(ps + i);
        if (rc)
            return rc;
        if (ps[i].key_bytes != ps[0].key_bytes)
            return QS_ERR_PARAM;
    }
    // item packing: [bucket | sketch | mask field | stream index] in 64 bits,
    // and (sketch, stream index) in 32 bits for the allocation order
    uint64_t max_low = 0, min_low = ~0ull, max_masks = 0;
    for (uint32_t i = 0; i < nsk; ++i) {
        max_low = ps[i].n_low > max_low ? ps[i].n_low : max_low;
        min_low = ps[i].n_low < min_low ? ps[i].n_low : min_low;
        max_masks = ps[i].n_masks > max_masks ? ps[i].n_masks : max_masks;
    }
    qs_pack pk;
    if (!qs_make_pack(max_low, min_low, nsk, max_masks, ps[0].key_bytes, &pk))
        return QS_ERR_PARAM; // too many sketches / too large tables for one set
    pk.q_low = (uint32_t)ps[0].n_low;
    pk.q_fd = qs_fastdiv_make(ps[0].n_masks);
    pk.unm = ps[0].n_masks;
    for (uint32_t i = 1; i < nsk; ++i)
        if (ps[i].n_masks != ps[0].n_masks)
            pk.unm = 0;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return QS_ERR_NODEVICE;
    qs_sketch *s = (qs_sketch *)calloc(1, sizeof(qs_sketch));
    if (!s)
        return QS_ERR_NOMEM;
    s->nsk = nsk;
    s->key_bytes = ps[0].key_bytes;
    s->pk = pk;
    // segment-pass shape: 64-lane groups, 4 items per lane (the autotuner's
    // pick for both the 2^28-key Zipf stream and the 4096-sketch batch)
    s->seg_w = 64;
    s->seg_ipl = 4;
    s->hot_filter = 1;
    s->max_stream = 1ull << pk.ib;
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
    s->ps = (qs_params *)malloc(sizeof(qs_params) * nsk);
    s->seeds = (uint32_t *)malloc(sizeof(uint32_t) * nsk);
    s->h_desc = (qs_sdesc *)calloc(nsk, sizeof(qs_sdesc));
    uint64_t mask_words = 0;
    for (uint32_t i = 0; i < nsk; ++i)
        mask_words += ps[i].n_masks;
    s->h_masks = (uint32_t *)malloc(sizeof(uint32_t) * mask_words);
    if (!s->ps || !s->seeds || !s->h_desc || !s->h_masks) {
        qs_destroy(s);
        return QS_ERR_NOMEM;
    }
    memcpy(s->ps, ps, sizeof(qs_params) * nsk);

    uint64_t low_w = 0, high_w = 0, bm_bits = 0, tail_e = 0, mask_w = 0;
    for (uint32_t i = 0; i < nsk; ++i) {
        const qs_params &p = ps[i];
        // seeds (Device_Seed::resize, device_sketch.h:280-302) and masks (v3:558-562)
        std::default_random_engine seed_eng, mask_eng;
        qs_seed_engine(seed_eng, p.rng_seed, 0);
        qs_seed_engine(mask_eng, p.rng_seed, 1);
        uint32_t seed = 0;
        std::uniform_int_distribution<unsigned int> dis(1, std::numeric_limits<unsigned int>::max());
        for (uint32_t k = 0; k < p.seed_sz; ++k)
            seed += dis(seed_eng);
        s->seeds[i] = p.hash_seed ? p.hash_seed : seed;
        qs_make_masks(mask_eng, p.n_masks, 32, 4, s->h_masks + mask_w);
        qs_sdesc &d = s->h_desc[i];
        d.low_off = low_w;
        d.high_off = high_w;
        d.bm_off = bm_bits;
        d.tail_off = tail_e;
        d.masks_off = (uint32_t)mask_w;
        d.n_masks = p.n_masks;
        d.n_low = (uint32_t)p.n_low;
        d.n_high = (uint32_t)p.n_high;
        d.paper = p.semantics == QS_PAPER;
        set_hash(d, s->seeds[i]);
        low_w += p.n_low * 8u;
        high_w += p.n_high * 32u;
        bm_bits += (p.n_low + 31) & ~31ull;
        tail_e += p.n_high;
        mask_w += p.n_masks;
    }
    s->low_words = low_w;
    s->high_words = high_w;
    s->bm_words = bm_bits / 32;
    s->tail_words = tail_e;
    s->mask_words = mask_w;
    bool ok = hipMalloc(&s->d_low, low_w * 4) == hipSuccess && hipMalloc(&s->d_high, high_w * 4) == hipSuccess &&
              hipMalloc(&s->d_masks, mask_w * 4) == hipSuccess &&
              hipMalloc(&s->d_next, sizeof(uint32_t) * nsk) == hipSuccess &&
              hipMalloc(&s->d_st, sizeof(qs_dev_state)) == hipSuccess &&
              hipMalloc(&s->d_longbm, s->bm_words * 4) == hipSuccess &&
              hipMalloc(&s->d_tail_of, tail_e * 4) == hipSuccess &&
              hipMalloc(&s->d_offs, sizeof(uint64_t) * (nsk + 1)) == hipSuccess &&
              hipMalloc(&s->d_desc, sizeof(qs_sdesc) * nsk) == hipSuccess &&
              hipHostMalloc(&s->h_st, sizeof(qs_dev_state), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        qs_destroy(s);
        return QS_ERR_NOMEM;
    }
    if (hipMemset(s->d_low, 0, low_w * 4) != hipSuccess || hipMemset(s->d_high, 0, high_w * 4) != hipSuccess ||
        hipMemset(s->d_longbm, 0, s->bm_words * 4) != hipSuccess ||
        hipMemcpy(s->d_masks, s->h_masks, mask_w * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_desc, s->h_desc, sizeof(qs_sdesc) * nsk, hipMemcpyHostToDevice) != hipSuccess ||
        reset_state(s) != QS_OK || hipDeviceSynchronize() != hipSuccess) {
        qs_destroy(s);
        return QS_ERR_HIP;
    }
    *out = s;
    return QS_OK;
}

extern "C" int qs_create(const qs_params *p, qs_sketch **out)
{
    if (!out)
        return QS_ERR_PARAM;
    *out = nullptr;
    int rc = validate_params(p);
    if (rc)
        return rc;
    return create_set(p, 1, out);
}

extern "C" int qs_multi_create(const qs_params *ps, uint32_t n_sketches, qs_sketch **out)
{
    return create_set(ps, n_sketches, out);
}

extern "C" uint32_t qs_multi_count(const qs_sketch *s)
{
    return s ? s->nsk : 0u;
}

extern "C" int qs_multi_params(const qs_sketch *s, uint32_t sketch, qs_params *p)
{
    if (!s || !p || sketch >= s->nsk)
        return QS_ERR_PARAM;
    *p = s->ps[sketch];
    return QS_OK;
}

extern "C" int qs_clear(qs_sketch *s)
{
    if (!s)
        return QS_ERR_PARAM;
    // Blocking, as the reference's clear: every earlier use of the tables has
    // finished before the fills start, and the fills and the control block's
    // reset have finished when the call returns.  In between they are queued
    // back to back on the null stream, with no host round trip between them.
    CK(hipDeviceSynchronize());
    CK(hipMemsetAsync(s->d_low, 0, s->low_words * 4, 0));
    CK(hipMemsetAsync(s->d_high, 0, s->high_words * 4, 0));
    CK(hipMemsetAsync(s->d_longbm, 0, s->bm_words * 4, 0));
    qs_dev_state h;
    memset(&h, 0, sizeof(h));
    std::vector<uint32_t> nx(s->nsk, QS_LINK_START + 1);
    CK(hipMemcpyAsync(s->d_st, &h, sizeof(h), hipMemcpyHostToDevice, 0));
    CK(hipMemcpyAsync(s->d_next, nx.data(), sizeof(uint32_t) * s->nsk, hipMemcpyHostToDevice, 0));
    CK(hipDeviceSynchronize()); // (h and nx are read before this returns)
    s->keys_inserted = 0;
    s->chunks = 0;
    s->sticky_err = 0;
    s->hot_keys = 0;
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

// ===========================================================================
// Host: insert / query entry points
// ===========================================================================
static int insert_single(qs_sketch *s, const void *d_keys, size_t n, hipStream_t stream, bool blocking)
{
    if (!s || (!d_keys && n))
        return QS_ERR_PARAM;
    if (s->nsk != 1)
        return QS_ERR_PARAM;
    if (s->sticky_err)
        return s->sticky_err;
    if (n == 0)
        return QS_OK;
    if (blocking && s->hot_filter) {
        int used = 0;
        const int rc = qs_insert_hot(s, d_keys, n, stream, &used);
        if (rc || used)
            return rc;
    }
    const uint64_t offs[2] = {0, (uint64_t)n};
    return qs_insert_set(s, d_keys, offs, nullptr, stream, blocking, false);
}

extern "C" int qs_set_hot_filter(qs_sketch *s, int enable)
{
    if (!s)
        return QS_ERR_PARAM;
    s->hot_filter = enable != 0;
    return QS_OK;
}

extern "C" int qs_insert(qs_sketch *s, const void *d_keys, size_t n)
{
    return insert_single(s, d_keys, n, 0, true);
}

extern "C" int qs_insert_async(qs_sketch *s, const void *d_keys, size_t n, void *stream)
{
    return insert_single(s, d_keys, n, (hipStream_t)stream, false);
}

// ---------------------------------------------------------------------------
// Allocation-free insert: the insert scratch (items, radix counts, digit
// bytes, replay and long-segment lists, allocation records and their sort
// buffers) is carved out of one static device array instead of the handle's
// own buffers.  Each GPU holds its own copy of the array from module load to
// process end; its address on a device comes from hipGetSymbolAddress, taken
// once per device with that device current.  The array's name is never used
// in host code beyond that call.
// ---------------------------------------------------------------------------
#define QS_NOALLOC_BYTES (33ull << 20)
#define QS_NOALLOC_DEVICES 64

static __device__ __align__(256) uint8_t scratch[QS_NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[QS_NOALLOC_DEVICES];

// the array's device address on device dev (the current device)
static int noalloc_scratch(int dev, uint8_t **out)
{
    if (dev < 0 || dev >= QS_NOALLOC_DEVICES)
        return QS_ERR_PARAM;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
            return QS_ERR_HIP;
        g_scratch_ptr[dev] = (uint8_t *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return QS_OK;
}

extern "C" int qs_insert_noalloc(qs_sketch *s, const void *d_keys, size_t n, void *stream)
{
    if (!s || (!d_keys && n))
        return QS_ERR_PARAM;
    if (s->nsk != 1 || n > QS_NOALLOC_MAX_KEYS)
        return QS_ERR_PARAM;
    if (s->sticky_err)
        return s->sticky_err;
    if (n == 0)
        return QS_OK;
    const uint32_t tcap = qs_trig_cap(s, n);
    if (tcap > QS_NOALLOC_MAX_ALLOCS)
        return QS_ERR_PARAM;
    int dev = -1;
    CK(hipGetDevice(&dev));
    if (dev != s->device)
        return QS_ERR_PARAM; // the sketch lives on another device
    uint8_t *base = nullptr;
    int rc = noalloc_scratch(dev, &base);
    if (rc)
        return rc;
    qs_scratch sc;
    if (qs_scratch_bytes(n, tcap, &sc, base) > QS_NOALLOC_BYTES)
        return QS_ERR_INTERNAL; // the limits above do not fit the array
    const uint64_t offs[2] = {0, (uint64_t)n};
    return qs_insert_set(s, d_keys, offs, nullptr, (hipStream_t)stream, false, false, &sc);
}

static int check_offs(const qs_sketch *s, const uint64_t *h_offs)
{
    if (!h_offs || h_offs[0] != 0)
        return QS_ERR_PARAM;
    for (uint32_t i = 0; i < s->nsk; ++i)
        if (h_offs[i + 1] < h_offs[i])
            return QS_ERR_PARAM;
    return QS_OK;
}

extern "C" int qs_multi_insert(qs_sketch *s, const void *d_keys, const uint64_t *h_offs)
{
    if (!s || check_offs(s, h_offs) || (!d_keys && h_offs[s->nsk]))
        return QS_ERR_PARAM;
    if (s->sticky_err)
        return s->sticky_err;
    if (h_offs[s->nsk] == 0)
        return QS_OK;
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(s->d_offs, h_offs, sizeof(uint64_t) * (s->nsk + 1), hipMemcpyHostToDevice));
    return qs_insert_set(s, d_keys, h_offs, s->d_offs, 0, true, false);
}

extern "C" int qs_multi_insert_async(qs_sketch *s, const void *d_keys, const uint64_t *h_offs,
                                     const uint64_t *d_offs, void *stream)
{
    if (!s || check_offs(s, h_offs) || (!d_keys && h_offs[s->nsk]) || (!d_offs && s->nsk > 1))
        return QS_ERR_PARAM;
    if (s->sticky_err)
        return s->sticky_err;
    if (h_offs[s->nsk] == 0)
        return QS_OK;
    return qs_insert_set(s, d_keys, h_offs, d_offs, (hipStream_t)stream, false, false);
}

static int query_launch(qs_sketch *s, const void *d_keys, uint64_t n, const uint64_t *d_offs, uint32_t *d_counts,
                        hipStream_t stream)
{
    if (n == 0)
        return QS_OK;
    if (n > 0xffffffffull)
        return QS_ERR_PARAM;
    const uint32_t nn = (uint32_t)n;
    const qs_sdesc d0 = s->h_desc[0];
    const bool multi = s->nsk > 1;
    const dim3 grid(grid_for((uint64_t)nn * 2, INS_THREADS, 8)), block(INS_THREADS);
    qs_evpair ev;
    qs_tim_begin(s, ev, QS_TK_QUERY, nn, stream);
    if (s->key_bytes == 4) {
        if (multi)
            hipLaunchKernelGGL((k_query<4, true>), grid, block, 0, stream, (const uint32_t *)d_keys, nn, d0,
                               s->d_desc, d_offs, s->nsk, s->d_low, s->d_high, s->d_masks, d_counts);
        else
            hipLaunchKernelGGL((k_query<4, false>), grid, block, 0, stream, (const uint32_t *)d_keys, nn, d0,
                               s->d_desc, d_offs, s->nsk, s->d_low, s->d_high, s->d_masks, d_counts);
    } else {
        if (multi)
            hipLaunchKernelGGL((k_query<8, true>), grid, block, 0, stream, (const uint64_t *)d_keys, nn, d0,
                               s->d_desc, d_offs, s->nsk, s->d_low, s->d_high, s->d_masks, d_counts);
        else
            hipLaunchKernelGGL((k_query<8, false>), grid, block, 0, stream, (const uint64_t *)d_keys, nn, d0,
                               s->d_desc, d_offs, s->nsk, s->d_low, s->d_high, s->d_masks, d_counts);
    }
    qs_tim_end(s, ev, stream);
    CK(hipGetLastError());
    return QS_OK;
}

static int query_single(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts, hipStream_t stream)
{
    if (!s || ((!d_keys || !d_counts) && n) || s->nsk != 1)
        return QS_ERR_PARAM;
    return query_launch(s, d_keys, n, nullptr, d_counts, stream);
}

extern "C" int qs_query(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts)
{
    int rc = query_single(s, d_keys, n, d_counts, 0);
    if (rc)
        return rc;
    CK(hipStreamSynchronize(0));
    return QS_OK;
}

extern "C" int qs_query_async(qs_sketch *s, const void *d_keys, size_t n, uint32_t *d_counts,
                              void *stream)
{
    return query_single(s, d_keys, n, d_counts, (hipStream_t)stream);
}

extern "C" int qs_multi_query(qs_sketch *s, const void *d_keys, const uint64_t *h_offs, uint32_t *d_counts)
{
    if (!s || check_offs(s, h_offs) || ((!d_keys || !d_counts) && h_offs[s->nsk]))
        return QS_ERR_PARAM;
    const uint64_t n = h_offs[s->nsk];
    if (n == 0)
        return QS_OK;
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(s->d_offs, h_offs, sizeof(uint64_t) * (s->nsk + 1), hipMemcpyHostToDevice));
    int rc = query_launch(s, d_keys, n, s->d_offs, d_counts, 0);
    if (rc)
        return rc;
    CK(hipStreamSynchronize(0));
    return QS_OK;
}

extern "C" int qs_multi_query_async(qs_sketch *s, const void *d_keys, const uint64_t *d_offs, uint64_t n,
                                    uint32_t *d_counts, void *stream)
{
    if (!s || ((!d_keys || !d_counts || (!d_offs && s->nsk > 1)) && n))
        return QS_ERR_PARAM;
    return query_launch(s, d_keys, n, d_offs, d_counts, (hipStream_t)stream);
}

extern "C" int qs_sync(qs_sketch *s)
{
    if (!s)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost));
    if (s->h_st->err)
        s->sticky_err = qs_err_code(s->h_st->err);
    return s->sticky_err;
}

// ===========================================================================
// Captured insert (hipGraph): the whole launch sequence of one batch layout,
// replayed with one call.  The graph reads the keys and the device offsets
// at replay time, so the caller may refill them between replays.
// ===========================================================================
struct qs_graph {
    qs_sketch *s;
    hipGraph_t g;
    hipGraphExec_t x;
    uint64_t keys;
    uint64_t chunks;
};

extern "C" int qs_graph_capture_insert(qs_sketch *s, const void *d_keys, const uint64_t *h_offs,
                                       const uint64_t *d_offs, qs_graph **out)
{
    if (!s || !out || check_offs(s, h_offs) || !d_keys || h_offs[s->nsk] == 0 || (!d_offs && s->nsk > 1))
        return QS_ERR_PARAM;
    *out = nullptr;
    if (s->sticky_err)
        return s->sticky_err;
    // size the scratch first: allocation cannot be captured
    const uint64_t n = h_offs[s->nsk];
    const uint64_t cap = qs_chunk_keys(s, h_offs);
    int rc = qs_ensure_scratch(s, n < cap ? n : cap);
    if (rc)
        return rc;
    CK(hipDeviceSynchronize());
    hipStream_t cs;
    CK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    qs_graph *g = (qs_graph *)calloc(1, sizeof(qs_graph));
    if (!g) {
        hipStreamDestroy(cs);
        return QS_ERR_NOMEM;
    }
    g->s = s;
    g->keys = n;
    g->chunks = (n + cap - 1) / cap;
    bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        rc = qs_insert_set(s, d_keys, h_offs, d_offs, cs, false, true);
        ok = hipStreamEndCapture(cs, &g->g) == hipSuccess && rc == QS_OK;
    }
    if (ok)
        ok = hipGraphInstantiate(&g->x, g->g, nullptr, nullptr, 0) == hipSuccess;
    hipStreamDestroy(cs);
    if (!ok) {
        if (g->g)
            hipGraphDestroy(g->g);
        free(g);
        return rc ? rc : QS_ERR_HIP;
    }
    s->graphs++;
    *out = g;
    return QS_OK;
}

extern "C" int qs_graph_launch(qs_graph *g, void *stream)
{
    if (!g)
        return QS_ERR_PARAM;
    if (g->s->sticky_err)
        return g->s->sticky_err;
    CK(hipGraphLaunch(g->x, (hipStream_t)stream));
    g->s->keys_inserted += g->keys;
    g->s->chunks += g->chunks;
    return QS_OK;
}

extern "C" void qs_graph_destroy(qs_graph *g)
{
    if (!g)
        return;
    hipDeviceSynchronize();
    hipGraphExecDestroy(g->x);
    hipGraphDestroy(g->g);
    g->s->graphs--;
    free(g);
}

// ===========================================================================
// Insert shape and its autotuner.  The segment kernel's group width and items
// per lane change only the schedule, never the result: every shape leaves the
// same state.  The tuner times each shape on a twin set (same parameters, so
// the caller's set is untouched), checks that all twins end in the same state
// (a checksum of both pools), and keeps the shape with the fastest segment pass.
// ===========================================================================
__global__ void k_checksum(const uint32_t *__restrict__ a, uint64_t n, uint64_t salt,
                           unsigned long long *__restrict__ out)
{
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc += (uint64_t)a[i] * (2u * (i + salt) + 1u);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        acc += (uint64_t)__shfl_xor((long long)acc, o, 64);
    if ((threadIdx.x & 63u) == 0)
        atomicAdd(out, (unsigned long long)acc);
}

static int set_checksum(qs_sketch *s, uint64_t *sum)
{
    unsigned long long *d = nullptr;
    CK(hipMalloc(&d, sizeof(unsigned long long)));
    hipMemset(d, 0, sizeof(unsigned long long));
    hipLaunchKernelGGL(k_checksum, dim3(grid_for(s->low_words, 256, 4)), dim3(256), 0, 0, s->d_low, s->low_words,
                       0ull, d);
    hipLaunchKernelGGL(k_checksum, dim3(grid_for(s->high_words, 256, 4)), dim3(256), 0, 0, s->d_high,
                       s->high_words, s->low_words, d);
    unsigned long long h = 0;
    hipError_t e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess)
        return QS_ERR_HIP;
    *sum = h;
    return QS_OK;
}

extern "C" int qs_set_insert_shape(qs_sketch *s, uint32_t width, uint32_t items_per_lane)
{
    if (!s || !qs_seg_shape_ok(width, items_per_lane))
        return QS_ERR_PARAM;
    if (s->graphs > 0)
        return QS_ERR_STATE; // captured graphs hold the current shape
    s->seg_w = width;
    s->seg_ipl = items_per_lane;
    return QS_OK;
}

extern "C" int qs_get_insert_shape(const qs_sketch *s, uint32_t *width, uint32_t *items_per_lane)
{
    if (!s || !width || !items_per_lane)
        return QS_ERR_PARAM;
    *width = s->seg_w;
    *items_per_lane = s->seg_ipl;
    return QS_OK;
}

extern "C" int qs_autotune_insert(qs_sketch *s, const void *d_keys, const uint64_t *h_offs, const uint64_t *d_offs,
                                  int reps, qs_tune_result *out)
{
    if (!s || !out || check_offs(s, h_offs) || !d_keys || h_offs[s->nsk] == 0 || (!d_offs && s->nsk > 1) ||
        reps < 1)
        return QS_ERR_PARAM;
    if (s->graphs > 0)
        return QS_ERR_STATE;
    memset(out, 0, sizeof(*out));
    qs_sketch *t = nullptr;
    int rc = create_set(s->ps, s->nsk, &t);
    if (rc)
        return rc;
    t->hot_filter = s->hot_filter;
    static const uint32_t widths[4] = {8, 16, 32, 64}, ipls[4] = {1, 2, 4, 8};
    double best = 1e300;
    uint64_t ref_sum = 0;
    out->consistent = 1;
    for (int wi = 0; wi < 4 && rc == QS_OK; ++wi)
        for (int ii = 0; ii < 4 && rc == QS_OK; ++ii) {
            qs_tune_entry &e = out->e[out->n];
            e.width = widths[wi];
            e.items_per_lane = ipls[ii];
            t->seg_w = e.width;
            t->seg_ipl = e.items_per_lane;
            double seg = 0, ins = 0;
            for (int r = 0; r <= reps && rc == QS_OK; ++r) { // the first run is a warm-up
                rc = qs_clear(t);
                if (rc)
                    break;
                rc = qs_set_timing(t, r > 0);
                if (rc)
                    break;
                // a single sketch's insert as the sketch's own would run it (the
                // hot-key filter sends a long skewed call through two chunks)
                rc = s->nsk == 1 ? insert_single(t, d_keys, h_offs[1], 0, true)
                                 : qs_insert_set(t, d_keys, h_offs, d_offs, 0, true, false);
                if (rc)
                    break;
                if (r > 0) {
                    qs_timing tm;
                    rc = qs_get_timing(t, &tm);
                    seg += tm.kernel_ms[QS_TK_SEG];
                    ins += tm.insert_kernel_ms;
                }
            }
            if (rc)
                break;
            qs_set_timing(t, 0);
            e.seg_ms = seg / reps;
            e.insert_ms = ins / reps;
            rc = set_checksum(t, &e.checksum);
            if (rc)
                break;
            if (out->n == 0)
                ref_sum = e.checksum;
            else if (e.checksum != ref_sum)
                out->consistent = 0;
            if (e.seg_ms < best) {
                best = e.seg_ms;
                out->best = out->n;
            }
            out->n++;
        }
    qs_destroy(t);
    if (rc)
        return rc;
    if (!out->consistent)
        return QS_ERR_STATE; // shapes disagree: keep the current shape
    s->seg_w = out->e[out->best].width;
    s->seg_ipl = out->e[out->best].items_per_lane;
    return QS_OK;
}

// ===========================================================================
// State export / import (reference layout) and statistics
// ===========================================================================
static int export_at(qs_sketch *s, uint32_t sid, uint32_t *low, uint32_t *high, uint32_t *masks, uint32_t *seed,
                     uint32_t *next)
{
    if (!s || sid >= s->nsk)
        return QS_ERR_PARAM;
    const qs_sdesc &d = s->h_desc[sid];
    CK(hipDeviceSynchronize());
    if (low)
        CK(hipMemcpy(low, s->d_low + d.low_off, (size_t)d.n_low * 32u, hipMemcpyDeviceToHost));
    if (high)
        CK(hipMemcpy(high, s->d_high + d.high_off, (size_t)d.n_high * 128u, hipMemcpyDeviceToHost));
    if (masks)
        memcpy(masks, s->h_masks + d.masks_off, sizeof(uint32_t) * d.n_masks);
    if (seed)
        *seed = d.seed;
    if (next)
        CK(hipMemcpy(next, s->d_next + sid, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return QS_OK;
}

static int import_at(qs_sketch *s, uint32_t sid, const uint32_t *low, const uint32_t *high, const uint32_t *masks,
                     uint32_t seed, uint32_t next)
{
    if (!s || sid >= s->nsk || !low || !high || !masks)
        return QS_ERR_PARAM;
    if (s->graphs > 0)
        return QS_ERR_STATE; // captured graphs hold the hash parameters
    qs_sdesc &d = s->h_desc[sid];
    const uint64_t n_high = d.n_high;
    if (next < QS_LINK_START + 1 || (uint64_t)(next - QS_LINK_START) > n_high)
        return QS_ERR_STATE;
    for (uint32_t i = 0; i < d.n_masks; ++i)
        if (__builtin_popcount(masks[i]) != 3 || (masks[i] >> 28) != 0)
            return QS_ERR_STATE;
    // a state the insert can reach: link words 0 or an id below next, every
    // id in [1025, next) held by exactly one bucket (the bump allocator of
    // v3:68-85 hands each out once), byte counters <= 129 (a byte is raised
    // only while every selected byte is <= 128, v3:160-170), and the high
    // rows of unallocated ids (and the unused row 0) zero
    std::vector<uint8_t> seen(next - QS_LINK_START - 1, 0);
    for (uint64_t b = 0; b < d.n_low; ++b) {
        const uint32_t *bk = low + b * 8;
        const uint32_t w = bk[7];
        if (w != 0) {
            if (w <= QS_LINK_START || w >= next || seen[w - QS_LINK_START - 1])
                return QS_ERR_STATE;
            seen[w - QS_LINK_START - 1] = 1;
        }
        for (int k = 0; k < 7; ++k)
            for (int j = 0; j < 4; ++j)
                if (((bk[k] >> (8 * j)) & 0xffu) > QS_LOW_LIMIT + 1u)
                    return QS_ERR_STATE;
    }
    for (uint8_t v : seen)
        if (!v)
            return QS_ERR_STATE;
    for (uint64_t r = 0; r < n_high; ++r) {
        if (r != 0 && r + QS_LINK_START < next)
            continue;
        for (int j = 0; j < 32; ++j)
            if (high[r * 32 + j] != 0)
                return QS_ERR_STATE;
    }
    CK(hipDeviceSynchronize());
    memcpy(s->h_masks + d.masks_off, masks, sizeof(uint32_t) * d.n_masks);
    s->seeds[sid] = seed;
    set_hash(d, seed);
    CK(hipMemcpy(s->d_low + d.low_off, low, (size_t)d.n_low * 32u, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_high + d.high_off, high, (size_t)n_high * 128u, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_masks + d.masks_off, masks, sizeof(uint32_t) * d.n_masks, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_desc + sid, &d, sizeof(qs_sdesc), hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_next + sid, &next, sizeof(uint32_t), hipMemcpyHostToDevice));
    CK(hipMemset(s->d_longbm, 0, s->bm_words * 4));
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_export_state(qs_sketch *s, uint32_t *low, uint32_t *high, uint32_t *masks,
                               uint32_t *seed, uint32_t *next)
{
    if (!s || s->nsk != 1)
        return QS_ERR_PARAM;
    return export_at(s, 0, low, high, masks, seed, next);
}

extern "C" int qs_import_state(qs_sketch *s, const uint32_t *low, const uint32_t *high,
                               const uint32_t *masks, uint32_t seed, uint32_t next)
{
    if (!s || s->nsk != 1)
        return QS_ERR_PARAM;
    int rc = import_at(s, 0, low, high, masks, seed, next);
    if (rc)
        return rc;
    s->sticky_err = 0;
    s->keys_inserted = 0;
    return reset_counts(s);
}

extern "C" int qs_multi_export_state(qs_sketch *s, uint32_t sketch, uint32_t *low, uint32_t *high,
                                     uint32_t *masks, uint32_t *seed, uint32_t *next)
{
    return export_at(s, sketch, low, high, masks, seed, next);
}

extern "C" int qs_multi_import_state(qs_sketch *s, uint32_t sketch, const uint32_t *low, const uint32_t *high,
                                     const uint32_t *masks, uint32_t seed, uint32_t next)
{
    return import_at(s, sketch, low, high, masks, seed, next);
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
    std::vector<uint32_t> nx(s->nsk);
    CK(hipMemcpy(nx.data(), s->d_next, sizeof(uint32_t) * s->nsk, hipMemcpyDeviceToHost));
    memset(st, 0, sizeof(*st));
    for (uint32_t i = 0; i < s->nsk; ++i) {
        st->n_low += s->ps[i].n_low;
        st->n_high += s->ps[i].n_high;
        st->n_masks += s->ps[i].n_masks;
        st->linked_buckets += nx[i] - (QS_LINK_START + 1);
    }
    st->bytes_low = st->n_low * 32u;
    st->bytes_high = st->n_high * 128u;
    st->keys_inserted = s->keys_inserted;
    st->chunks = s->chunks;
    st->chunks_with_fixup = h.chunks_ord;
    st->fixup_keys = h.replay_keys;
    st->ordered_segments = h.ord_segments;
    st->long_segments = h.long_segments;
    st->next_id = nx[0];
    st->seed = s->h_desc[0].seed;
    st->n_sketches = s->nsk;
    st->pool_error = (h.err & QS_EF_POOL) ? 1u : 0u;
    st->hot_keys = s->hot_keys;
    return QS_OK;
}
