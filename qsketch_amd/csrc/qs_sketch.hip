// qs_sketch.hip -- MI355X-native QSketch (two-level MSWS): host API + kernels.
//
// Reference path replaced: include/sketch/device_sketch/device_sketch_v3.h
// (insert_kernel v3:193-290, search_kernel v3:410-521, atomic_malloc v3:68-85,
// class Sketch_GPU_Sub_Warp_Level v3:524-634) and its parameter / mask / seed
// set-up (lib/default.cu, lib/hashmask.cu, include/tool/tool.h:186-220).
//
// Insert is deterministic: its result equals the sequential stream-order
// semantics of v3 (see the oracle) including the order in which high buckets
// are allocated.  Per chunk of keys:
//   1. k_insert (optimistic, one lane per key): linked buckets -> three u32
//      atomics on the high bucket (commutative); unlinked buckets -> returning
//      packed-byte atomics on the low words; a lane that sees a selected byte
//      >= 129 before its add flags the bucket in a bitmap.
//      No flag on a bucket  <=>  no selected byte passed 129 in any order
//      <=>  no allocation would happen for it in stream order, so the
//      commutative result is exact.
//   2. only if some bucket was flagged: every insert of a flagged bucket is
//      subtracted again (exact mod 2^32, so even a wrapped byte is restored),
//      compacted in stream order, radix-sorted by bucket (stable), and each
//      bucket's segment is replayed in stream order by one wavefront
//      (k_ordered) up to its allocation point; the rest of a long segment is
//      added to the high bucket in parallel (k_tail).  Allocations of a chunk
//      get ids in stream order of their triggering insert (k_assign).
#include "qs_internal.h"

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess) {                                                                 \
            return QS_ERR_HIP;                                                                   \
        }                                                                                        \
    } while (0)

#define INS_THREADS 256
#define RB_THREADS 256
#define RB_ITEMS 8
#define RB_TILE (RB_THREADS * RB_ITEMS)

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

template <int KB> using KeyT = typename qs_key<KB>::T;

// ===========================================================================
// Query: one lane per key, the 32-byte bucket as two 16-byte loads.
// Replaces search_kernel (v3:410-521): est = min over the three selected byte
// counters + (linked ? min over the same three high counters : 0).
// ===========================================================================
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
    const uint32_t stride = gridDim.x * INS_THREADS;
    for (uint32_t i = blockIdx.x * INS_THREADS + threadIdx.x; i < n; i += stride) {
        uint32_t b, mu;
        qs_hash_index(keys[i], hp, b, mu);
        uint32_t M = lds_mask(s_masks, mu);
        const uint4 *bp = reinterpret_cast<const uint4 *>(low + (size_t)b * 8u);
        uint4 a = bp[0];
        uint4 c = bp[1];
        uint32_t j1 = (uint32_t)__builtin_ctz(M);
        uint32_t M1 = M & (M - 1u);
        uint32_t j2 = (uint32_t)__builtin_ctz(M1);
        uint32_t M2 = M1 & (M1 - 1u);
        uint32_t j3 = (uint32_t)__builtin_ctz(M2);
        uint32_t c1 = byte_at(sel_word(a, c, j1 >> 2), j1);
        uint32_t c2 = byte_at(sel_word(a, c, j2 >> 2), j2);
        uint32_t c3 = byte_at(sel_word(a, c, j3 >> 2), j3);
        uint32_t est = min(c1, min(c2, c3));
        uint32_t w = c.w;
        if (w > QS_LINK_START) {
            const uint32_t *g = high + (size_t)(w - QS_LINK_START) * 32u;
            est += min(g[j1], min(g[j2], g[j3]));
        }
        out[i] = est;
    }
}

// ===========================================================================
// Insert step 1: optimistic commutative insert.
// ===========================================================================
template <int KB>
__global__ __launch_bounds__(INS_THREADS) void k_insert(const KeyT<KB> *__restrict__ keys, uint32_t n,
                                                        uint32_t *__restrict__ low,
                                                        uint32_t *__restrict__ high,
                                                        const uint32_t *__restrict__ masks,
                                                        qs_hashpar hp, uint32_t *__restrict__ viol,
                                                        qs_dev_state *__restrict__ st)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t i = threadIdx.x; i < hp.n_masks; i += INS_THREADS)
        s_masks[i] = masks[i];
    __syncthreads();
    const uint32_t stride = gridDim.x * INS_THREADS;
    for (uint32_t i = blockIdx.x * INS_THREADS + threadIdx.x; i < n; i += stride) {
        uint32_t b, mu;
        qs_hash_index(keys[i], hp, b, mu);
        uint32_t M = lds_mask(s_masks, mu);
        uint32_t *bk = low + (size_t)b * 8u;
        uint32_t w = bk[7];
        if (w > QS_LINK_START) {
            uint32_t *g = high + (size_t)(w - QS_LINK_START) * 32u;
            uint32_t m = M;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                atomicAdd(g + __builtin_ctz(m), 1u);
                m &= m - 1u;
            }
        } else {
            low_adds la = split_mask(M);
            bool v = false;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (la.a[r]) {
                    uint32_t old = atomicAdd(bk + la.k[r], la.a[r]);
                    v |= over_limit(old, la.a[r]);
                }
            }
            if (v) {
                atomicOr(viol + (b >> 5), 1u << (b & 31u));
                __hip_atomic_store(&st->any_viol, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// ===========================================================================
// Insert step 2a: undo the optimistic adds of every unlinked insert whose
// bucket is flagged, and count them per tile of RB_TILE consecutive keys.
// ===========================================================================
template <int KB>
__global__ __launch_bounds__(RB_THREADS) void k_rollback(const KeyT<KB> *__restrict__ keys, uint32_t n,
                                                         uint32_t *__restrict__ low,
                                                         const uint32_t *__restrict__ masks,
                                                         qs_hashpar hp,
                                                         const uint32_t *__restrict__ viol,
                                                         uint32_t *__restrict__ tile_cnt)
{
    __shared__ uint32_t s_w[RB_THREADS / QS_WAVE];
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * RB_TILE;
    uint32_t cnt = 0;
#pragma unroll
    for (int r = 0; r < RB_ITEMS; ++r) {
        uint32_t i = base + r * RB_THREADS + t;
        if (i < n) {
            uint32_t b, mu;
            qs_hash_index(keys[i], hp, b, mu);
            if (bit_test(viol, b)) {
                uint32_t *bk = low + (size_t)b * 8u;
                if (bk[7] <= QS_LINK_START) {
                    low_adds la = split_mask(masks[mu]);
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        if (la.a[q])
                            atomicSub(bk + la.k[q], la.a[q]);
                    ++cnt;
                }
            }
        }
    }
    // block sum
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        cnt += __shfl_xor(cnt, o, 64);
    if ((t & 63u) == 0)
        s_w[t >> 6] = cnt;
    __syncthreads();
    if (t == 0) {
        uint32_t s = 0;
        for (int k = 0; k < RB_THREADS / QS_WAVE; ++k)
            s += s_w[k];
        tile_cnt[blockIdx.x] = s;
    }
}

// Exclusive scan of n values (single workgroup); total to *total.
__global__ __launch_bounds__(1024) void k_scan_excl(const uint32_t *__restrict__ in,
                                                    uint32_t *__restrict__ out, uint32_t n,
                                                    uint32_t *__restrict__ total)
{
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    if (t == 0)
        s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += 1024) {
        uint32_t i = base + t;
        uint32_t v = i < n ? in[i] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o)
                x += y;
        }
        if (lane == 63)
            s_w[wv] = x;
        __syncthreads();
        if (t < 16) {
            uint32_t a = s_w[t], sc = a;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                uint32_t y = __shfl_up(sc, o, 16);
                if ((t & 15u) >= (uint32_t)o)
                    sc += y;
            }
            s_w[t] = sc - a;
        }
        __syncthreads();
        uint32_t carry = s_carry;
        if (i < n)
            out[i] = carry + s_w[wv] + x - v;
        __syncthreads();
        if (t == 1023)
            s_carry = carry + s_w[wv] + x;
        __syncthreads();
    }
    if (t == 0)
        *total = s_carry;
}

// Insert step 2b: write the flagged inserts (bucket, stream index) in stream
// order at their tile's offset.
template <int KB>
__global__ __launch_bounds__(RB_THREADS) void k_compact(const KeyT<KB> *__restrict__ keys, uint32_t n,
                                                        const uint32_t *__restrict__ low,
                                                        qs_hashpar hp,
                                                        const uint32_t *__restrict__ viol,
                                                        const uint32_t *__restrict__ tile_cnt,
                                                        const uint32_t *__restrict__ tile_off,
                                                        uint32_t *__restrict__ rb,
                                                        uint32_t *__restrict__ ri)
{
    __shared__ uint32_t s_w[RB_THREADS / QS_WAVE];
    if (tile_cnt[blockIdx.x] == 0)
        return;
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * RB_TILE;
    uint32_t run = tile_off[blockIdx.x];
    for (int r = 0; r < RB_ITEMS; ++r) {
        uint32_t i = base + r * RB_THREADS + t;
        bool f = false;
        uint32_t b = 0, mu;
        if (i < n) {
            qs_hash_index(keys[i], hp, b, mu);
            f = bit_test(viol, b) && low[(size_t)b * 8u + 7u] <= QS_LINK_START;
        }
        uint64_t bal = __ballot(f);
        if (lane == 0)
            s_w[wv] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t off = run, tot = 0;
        for (uint32_t k = 0; k < RB_THREADS / QS_WAVE; ++k) {
            uint32_t c = s_w[k];
            if (k < wv)
                off += c;
            tot += c;
        }
        if (f) {
            uint32_t pos = off + (uint32_t)__popcll(bal & lt);
            rb[pos] = b;
            ri[pos] = i;
        }
        run += tot;
        __syncthreads();
    }
}

__global__ void k_heads(const uint32_t *__restrict__ rb, uint32_t r_n, uint32_t *__restrict__ heads,
                        qs_dev_state *__restrict__ st)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < r_n && (p == 0 || rb[p] != rb[p - 1])) {
        uint32_t k = atomicAdd(&st->n_heads, 1u);
        heads[k] = p;
    }
}

__device__ static inline uint32_t kth_set_bit(uint64_t m, uint32_t k)
{
    // position of the k-th (0-based) set bit of m; caller guarantees popc(m) > k
    uint32_t pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        uint64_t lowm = (w == 64) ? ~0ull : ((1ull << w) - 1ull);
        uint32_t c = (uint32_t)__popcll(m & lowm);
        if (k >= c) {
            k -= c;
            m >>= w;
            pos += (uint32_t)w;
        }
    }
    return pos;
}

// Insert step 3: replay each flagged bucket's segment in stream order on one
// wavefront until its allocation point (at most ceil(28*129/3) inserts can
// precede it).  Lane j (< 28) holds byte counter j; a batch of 64 inserts is
// transposed into 28 ballots B_j, and the first insert that would see a
// selected byte > 128 is the (129 - c_j)-th set bit of B_j, minimised over j.
template <int KB>
__global__ __launch_bounds__(256) void k_ordered(const KeyT<KB> *__restrict__ keys,
                                                 const uint32_t *__restrict__ rb,
                                                 const uint32_t *__restrict__ ri, uint32_t r_n,
                                                 const uint32_t *__restrict__ heads,
                                                 uint32_t *__restrict__ low,
                                                 const uint32_t *__restrict__ masks, qs_hashpar hp,
                                                 uint32_t *__restrict__ viol,
                                                 qs_trig *__restrict__ trig, uint32_t trig_cap,
                                                 qs_dev_state *__restrict__ st, int paper)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (blockDim.x / 64u);
    const uint32_t n_heads = st->n_heads;
    for (uint32_t h = gw; h < n_heads; h += nw) {
        const uint32_t p0 = heads[h];
        const uint32_t b = rb[p0];
        uint32_t *bk = low + (size_t)b * 8u;
        uint32_t word = lane < 7u ? bk[lane] : 0u;
        uint32_t c = __shfl(word, (int)(lane >> 2), 64);
        c = lane < 28u ? ((c >> ((lane & 3u) * 8u)) & 0xffu) : 0u;
        uint32_t hc = 0;
        bool fired = false;
        uint32_t t_idx = 0, tail_start = 0;
        for (uint32_t p = p0;; p += 64u) {
            uint32_t q = p + lane;
            bool valid = q < r_n && rb[q] == b;
            uint64_t vm = __ballot(valid);
            if (vm == 0)
                break;
            uint32_t idx = 0, M = 0;
            if (valid) {
                uint32_t bb, mu;
                idx = ri[q];
                qs_hash_index(keys[idx], hp, bb, mu);
                M = masks[mu];
            }
            uint64_t Bj = 0;
#pragma unroll
            for (int j = 0; j < 28; ++j) {
                uint64_t bal = __ballot((M >> j) & 1u);
                Bj = (lane == (uint32_t)j) ? bal : Bj;
            }
            uint32_t need = c > QS_LOW_LIMIT ? 0u : (QS_LOW_LIMIT + 1u - c);
            uint32_t pos = 64u;
            if (lane < 28u && (uint32_t)__popcll(Bj) > need)
                pos = kth_set_bit(Bj, need);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
                pos = min(pos, (uint32_t)__shfl_xor(pos, o, 64));
            if (pos == 64u) {
                c += (uint32_t)__popcll(Bj);
            } else {
                uint64_t before = pos == 0u ? 0ull : ((1ull << pos) - 1ull);
                uint64_t after = pos >= 63u ? 0ull : (~0ull << (pos + 1u));
                c += (uint32_t)__popcll(Bj & before);
                uint32_t Mt = __shfl(M, (int)pos, 64);
                t_idx = __shfl(idx, (int)pos, 64);
                if (paper)
                    hc += lane < 32u ? ((Mt >> lane) & 1u) : 0u;
                else
                    hc += lane < 8u ? (uint32_t)__popc(Mt & (0x01010101u << lane)) : 0u;
                hc += (uint32_t)__popcll(Bj & after);
                fired = true;
                tail_start = p + 64u;
                break;
            }
            if (vm != ~0ull)
                break;
        }
        // write the low counters back (bucket is owned by this wavefront)
        uint32_t w4 = (uint32_t)__shfl(c, (int)(4u * lane) & 63, 64) |
                      ((uint32_t)__shfl(c, (int)(4u * lane + 1u) & 63, 64) << 8) |
                      ((uint32_t)__shfl(c, (int)(4u * lane + 2u) & 63, 64) << 16) |
                      ((uint32_t)__shfl(c, (int)(4u * lane + 3u) & 63, 64) << 24);
        if (lane < 7u)
            bk[lane] = w4;
        if (lane == 0)
            atomicAnd(viol + (b >> 5), ~(1u << (b & 31u)));
        if (fired) {
            uint32_t slot = 0;
            if (lane == 0)
                slot = atomicAdd(&st->n_trig, 1u);
            slot = __shfl(slot, 0, 64);
            if (slot < trig_cap) {
                qs_trig *tr = trig + slot;
                if (lane == 0) {
                    tr->idx = t_idx;
                    tr->bucket = b;
                    tr->tail_start = tail_start;
                    tr->pad = 0;
                }
                if (lane < 32u)
                    tr->hc[lane] = hc;
            } else if (lane == 0) {
                atomicOr(&st->err, 1u); // more allocations than the pool holds
            }
        }
    }
}

__global__ void k_trig_keys(const qs_trig *__restrict__ trig, uint32_t n, uint32_t *__restrict__ tk,
                            uint32_t *__restrict__ tv)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) {
        tk[r] = trig[r].idx;
        tv[r] = r;
    }
}

// Insert step 4: ids in stream order of the triggering inserts (the sequential
// bump allocator, v3:68-85, with next starting at 1025).
__global__ __launch_bounds__(256) void k_assign(const qs_trig *__restrict__ trig,
                                                const uint32_t *__restrict__ order, uint32_t n_trig,
                                                uint32_t *__restrict__ low, uint32_t *__restrict__ high,
                                                uint32_t *__restrict__ tail_of, uint64_t n_high,
                                                qs_dev_state *__restrict__ st)
{
    const uint32_t base = st->next;
    if ((uint64_t)(base - QS_LINK_START) + n_trig > n_high) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            atomicOr(&st->err, 1u);
        return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t r = gw; r < n_trig; r += nw) {
        const qs_trig *tr = trig + order[r];
        uint32_t id = base + r;
        uint32_t row = id - QS_LINK_START;
        if (lane < 32u)
            high[(size_t)row * 32u + lane] = tr->hc[lane];
        if (lane == 0) {
            low[(size_t)tr->bucket * 8u + 7u] = id;
            tail_of[row] = tr->tail_start;
        }
    }
}

// Insert step 5: fix-up inserts after the allocation's 64-insert batch are
// plain high adds; a wavefront whose active lanes share one bucket (the usual
// case: a long hot segment) adds popcount(B_j) once per counter.
template <int KB>
__global__ __launch_bounds__(256) void k_tail(const KeyT<KB> *__restrict__ keys,
                                              const uint32_t *__restrict__ rb,
                                              const uint32_t *__restrict__ ri, uint32_t r_n,
                                              const uint32_t *__restrict__ low,
                                              uint32_t *__restrict__ high,
                                              const uint32_t *__restrict__ masks, qs_hashpar hp,
                                              const uint32_t *__restrict__ tail_of)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < r_n; base += stride) {
        uint32_t p = base + threadIdx.x;
        bool act = false;
        uint32_t b = 0, w = 0, M = 0;
        if (p < r_n) {
            b = rb[p];
            w = low[(size_t)b * 8u + 7u];
            if (w > QS_LINK_START && p >= tail_of[w - QS_LINK_START]) {
                uint32_t bb, mu;
                qs_hash_index(keys[ri[p]], hp, bb, mu);
                M = masks[mu];
                act = true;
            }
        }
        uint64_t am = __ballot(act);
        if (am == 0)
            continue;
        uint32_t first = (uint32_t)__builtin_ctzll(am);
        uint32_t b0 = __shfl(b, (int)first, 64);
        uint32_t w0 = __shfl(w, (int)first, 64);
        uint64_t same = __ballot(act && b == b0);
        if (same == am) {
            uint64_t Bj = 0;
#pragma unroll
            for (int j = 0; j < 28; ++j) {
                uint64_t bal = __ballot(act && ((M >> j) & 1u));
                Bj = (lane == (uint32_t)j) ? bal : Bj;
            }
            uint32_t cnt = (uint32_t)__popcll(Bj);
            if (lane < 28u && cnt)
                atomicAdd(high + (size_t)(w0 - QS_LINK_START) * 32u + lane, cnt);
        } else if (act) {
            uint32_t *g = high + (size_t)(w - QS_LINK_START) * 32u;
            uint32_t m = M;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                atomicAdd(g + __builtin_ctz(m), 1u);
                m &= m - 1u;
            }
        }
    }
}

__global__ void k_finalize(qs_dev_state *st, uint32_t n_trig)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (st->err == 0)
            st->next += n_trig;
        st->any_viol = 0;
        st->r_n = 0;
        st->n_heads = 0;
        st->n_trig = 0;
    }
}

// ===========================================================================
// Host: create / destroy / scratch
// ===========================================================================

static int g_cu_count = 0;

static void free_scratch(qs_sketch *s)
{
    hipFree(s->d_tile_cnt);
    hipFree(s->d_tile_off);
    for (int k = 0; k < 2; ++k) {
        hipFree(s->d_rb[k]);
        hipFree(s->d_ri[k]);
        hipFree(s->d_tk[k]);
        hipFree(s->d_tv[k]);
        s->d_rb[k] = s->d_ri[k] = s->d_tk[k] = s->d_tv[k] = nullptr;
    }
    hipFree(s->d_hist);
    hipFree(s->d_heads);
    hipFree(s->d_trig);
    s->d_tile_cnt = s->d_tile_off = s->d_hist = s->d_heads = nullptr;
    s->d_trig = nullptr;
    s->chunk = 0;
}

static int ensure_scratch(qs_sketch *s, uint32_t chunk)
{
    if (s->chunk >= chunk)
        return QS_OK;
    free_scratch(s);
    uint32_t tiles = (chunk + RB_TILE - 1) / RB_TILE;
    uint64_t tcap = s->p.n_high < (uint64_t)chunk ? s->p.n_high : (uint64_t)chunk;
    s->trig_cap = (uint32_t)tcap;
    size_t hw = qs_sort_hist_words(chunk);
    CK(hipMalloc(&s->d_tile_cnt, sizeof(uint32_t) * tiles));
    CK(hipMalloc(&s->d_tile_off, sizeof(uint32_t) * tiles));
    for (int k = 0; k < 2; ++k) {
        CK(hipMalloc(&s->d_rb[k], sizeof(uint32_t) * chunk));
        CK(hipMalloc(&s->d_ri[k], sizeof(uint32_t) * chunk));
        CK(hipMalloc(&s->d_tk[k], sizeof(uint32_t) * s->trig_cap));
        CK(hipMalloc(&s->d_tv[k], sizeof(uint32_t) * s->trig_cap));
    }
    CK(hipMalloc(&s->d_hist, sizeof(uint32_t) * hw));
    CK(hipMalloc(&s->d_heads, sizeof(uint32_t) * chunk));
    CK(hipMalloc(&s->d_trig, sizeof(qs_trig) * s->trig_cap));
    s->chunk = chunk;
    s->tile_n = tiles;
    return QS_OK;
}

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
    free_scratch(s);
    hipFree(s->d_low);
    hipFree(s->d_high);
    hipFree(s->d_masks);
    hipFree(s->d_st);
    hipFree(s->d_viol);
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
    bool ok = hipMalloc(&s->d_low, low_bytes) == hipSuccess &&
              hipMalloc(&s->d_high, high_bytes) == hipSuccess &&
              hipMalloc(&s->d_masks, sizeof(uint32_t) * p->n_masks) == hipSuccess &&
              hipMalloc(&s->d_st, sizeof(qs_dev_state)) == hipSuccess &&
              hipMalloc(&s->d_viol, sizeof(uint32_t) * viol_words) == hipSuccess &&
              hipMalloc(&s->d_tail_of, sizeof(uint32_t) * p->n_high) == hipSuccess &&
              hipHostMalloc(&s->h_st, sizeof(qs_dev_state), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        qs_destroy(s);
        return QS_ERR_NOMEM;
    }
    if (hipMemset(s->d_low, 0, low_bytes) != hipSuccess ||
        hipMemset(s->d_high, 0, high_bytes) != hipSuccess ||
        hipMemset(s->d_viol, 0, sizeof(uint32_t) * viol_words) != hipSuccess ||
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
    CK(hipMemset(s->d_viol, 0, sizeof(uint32_t) * (size_t)((s->p.n_low + 31) / 32)));
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

static uint64_t default_chunk(const qs_sketch *s)
{
    return s->p.chunk_keys ? s->p.chunk_keys : (1ull << 24);
}

// Chunk schedule: a chunk is never larger than what the sketch has already
// absorbed (floor 2^20), so the first chunks of a fresh sketch -- where hot
// keys allocate and the fix-up is large -- stay small.  Any schedule gives the
// same state; only the cost changes.
static uint32_t next_chunk(const qs_sketch *s, size_t remaining)
{
    uint64_t cap = default_chunk(s);
    uint64_t c = s->keys_inserted;
    if (c < (1ull << 20))
        c = 1ull << 20;
    if (c > cap)
        c = cap;
    if (c > remaining)
        c = remaining;
    if (c > 0x7fffffffull)
        c = 0x7fffffffull;
    return (uint32_t)c;
}

template <int KB>
static int fixup_chunk(qs_sketch *s, const KeyT<KB> *keys, uint32_t cn, hipStream_t stream)
{
    const uint32_t tiles = (cn + RB_TILE - 1) / RB_TILE;
    hipLaunchKernelGGL(k_rollback<KB>, dim3(tiles), dim3(RB_THREADS), 0, stream, keys, cn, s->d_low,
                       s->d_masks, s->hp, s->d_viol, s->d_tile_cnt);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, stream, s->d_tile_cnt, s->d_tile_off,
                       tiles, &s->d_st->r_n);
    hipLaunchKernelGGL(k_compact<KB>, dim3(tiles), dim3(RB_THREADS), 0, stream, keys, cn, s->d_low,
                       s->hp, s->d_viol, s->d_tile_cnt, s->d_tile_off, s->d_rb[0], s->d_ri[0]);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    const uint32_t r_n = s->h_st->r_n;
    s->fix_keys += r_n;
    if (r_n == 0)
        return QS_OK;
    int bits = 32 - __builtin_clz((uint32_t)(s->p.n_low > 1 ? s->p.n_low - 1 : 1));
    int cur = qs_sort_pairs(s->d_rb, s->d_ri, s->d_hist, r_n, bits, stream);
    if (cur < 0)
        return QS_ERR_HIP;
    const uint32_t *rb = s->d_rb[cur], *ri = s->d_ri[cur];
    hipLaunchKernelGGL(k_heads, dim3((r_n + 255) / 256), dim3(256), 0, stream, rb, r_n, s->d_heads,
                       s->d_st);
    hipLaunchKernelGGL(k_ordered<KB>, dim3(grid_for((uint64_t)r_n * 64, 256, 8)), dim3(256), 0,
                       stream, keys, rb, ri, r_n, s->d_heads, s->d_low, s->d_masks, s->hp, s->d_viol,
                       s->d_trig, s->trig_cap, s->d_st, s->p.semantics == QS_PAPER ? 1 : 0);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    uint32_t n_trig = s->h_st->n_trig;
    if (n_trig > s->trig_cap)
        n_trig = s->trig_cap;
    if (n_trig) {
        hipLaunchKernelGGL(k_trig_keys, dim3((n_trig + 255) / 256), dim3(256), 0, stream, s->d_trig,
                           n_trig, s->d_tk[0], s->d_tv[0]);
        int tbits = 32 - __builtin_clz(cn > 1 ? cn - 1 : 1);
        int tc = qs_sort_pairs(s->d_tk, s->d_tv, s->d_hist, n_trig, tbits, stream);
        if (tc < 0)
            return QS_ERR_HIP;
        hipLaunchKernelGGL(k_assign, dim3(grid_for((uint64_t)n_trig * 64, 256, 8)), dim3(256), 0,
                           stream, s->d_trig, s->d_tv[tc], n_trig, s->d_low, s->d_high, s->d_tail_of,
                           (uint64_t)s->p.n_high, s->d_st);
        hipLaunchKernelGGL(k_tail<KB>, dim3(grid_for(r_n, 256, 8)), dim3(256), 0, stream, keys, rb, ri,
                           r_n, s->d_low, s->d_high, s->d_masks, s->hp, s->d_tail_of);
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, stream, s->d_st, n_trig);
    CK(hipGetLastError());
    return QS_OK;
}

template <int KB>
static int insert_impl(qs_sketch *s, const void *d_keys, size_t n, hipStream_t stream)
{
    const KeyT<KB> *keys = (const KeyT<KB> *)d_keys;
    size_t done = 0;
    while (done < n) {
        uint32_t cn = next_chunk(s, n - done);
        int rc = ensure_scratch(s, cn);
        if (rc)
            return rc;
        const KeyT<KB> *ck = keys + done;
        hipLaunchKernelGGL(k_insert<KB>, dim3(grid_for(cn, INS_THREADS, 8)), dim3(INS_THREADS), 0,
                           stream, ck, cn, s->d_low, s->d_high, s->d_masks, s->hp, s->d_viol, s->d_st);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        if (s->h_st->any_viol) {
            rc = fixup_chunk<KB>(s, ck, cn, stream);
            if (rc)
                return rc;
            s->chunks_fix++;
            CK(hipMemcpyAsync(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost, stream));
            CK(hipStreamSynchronize(stream));
        }
        s->chunks++;
        s->keys_inserted += cn;
        done += cn;
        if (s->h_st->err) {
            s->sticky_err = QS_ERR_POOL;
            return s->sticky_err;
        }
    }
    return QS_OK;
}

static int insert_dispatch(qs_sketch *s, const void *d_keys, size_t n, hipStream_t stream)
{
    if (!s || (!d_keys && n))
        return QS_ERR_PARAM;
    if (s->sticky_err)
        return s->sticky_err;
    if (n == 0)
        return QS_OK;
    return s->p.key_bytes == 4 ? insert_impl<4>(s, d_keys, n, stream)
                               : insert_impl<8>(s, d_keys, n, stream);
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
    if (s->p.key_bytes == 4)
        hipLaunchKernelGGL(k_query<4>, dim3(grid_for(nn, INS_THREADS, 8)), dim3(INS_THREADS), 0, stream,
                           (const uint32_t *)d_keys, nn, s->d_low, s->d_high, s->d_masks, s->hp,
                           d_counts);
    else
        hipLaunchKernelGGL(k_query<8>, dim3(grid_for(nn, INS_THREADS, 8)), dim3(INS_THREADS), 0, stream,
                           (const uint64_t *)d_keys, nn, s->d_low, s->d_high, s->d_masks, s->hp,
                           d_counts);
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
    CK(hipMemset(s->d_viol, 0, sizeof(uint32_t) * (size_t)((s->p.n_low + 31) / 32)));
    int rc = reset_state(s, next);
    if (rc)
        return rc;
    CK(hipDeviceSynchronize());
    s->sticky_err = 0;
    s->keys_inserted = 0;
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
