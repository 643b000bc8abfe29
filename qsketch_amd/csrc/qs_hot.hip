// qs_hot.hip -- hot-key filter of a plain insert of 32-bit keys.
//
// Under skew most of a stream belongs to a few thousand keys, and once such a
// key's bucket is linked every insert of it is the same commutative update:
// g[w][j] += 1 for the three counters j of its mask (v3:111-129).  Only the
// inserts into unlinked buckets depend on their order (the promotion point,
// v3:160-190).  So a long insert call runs as:
//   1. the candidates: the keys occurring at least HOT_MIN_RUN times among the
//      first HOT_SAMPLE keys (a sort of that sample, then its runs), ordered
//      by count (ties by key) and capped at HOT_MAX; on a second stream, at
//      the same time as
//   2. the first HOT_PREFIX keys through the ordinary pipeline (qs_insert.hip),
//      which links the buckets of the hot keys;
//   3. the hot set: the candidates whose bucket is linked now, in a table of
//      HOT_BUCKETS buckets of 4 slots where each key may sit in one of two
//      buckets (the emptier one when it is placed), so that a lookup is two
//      16-byte reads and no loop; a key that finds both buckets full is left
//      out of the set (it then takes the ordered path, which is always exact);
//   4. the rest of the call in one pass: a hot key is counted in LDS (per
//      persistent block, summed over blocks at the end and added to its high
//      row once), every other key is written, in stream order, to a cold
//      array (packed per chunk, then moved together after a scan);
//   5. the cold array through the ordinary pipeline.
// The result is the sequential one: the hot keys' inserts commute with
// everything after their bucket is linked, and the cold keys keep their
// stream order (so the allocation ids, given in stream order of the
// promoting inserts, are unchanged).  A call whose sample shows no skew
// (candidates covering less than HOT_MIN_COVER of it) runs the rest of the
// call after the prefix through the ordinary pipeline.
//
// qs_hot_split_keys serves a sink whose inserts all commute (the SF sketch,
// qs_sf.hip): no prefix and no link condition; every candidate enters the
// table, and the caller gets the cold keys and the per-slot counts.
#include "qs_internal.h"

#include <cstring>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess)                                                                   \
            return QS_ERR_HIP;                                                                   \
    } while (0)

#define HOT_SAMPLE (1u << 20)              // candidates come from the first HOT_SAMPLE keys
#define HOT_MIN_RUN 8u                     // a candidate occurs at least this often there
#define HOT_CAND_MAX (HOT_SAMPLE / HOT_MIN_RUN)
#define HOT_MAX 8192u                      // hot keys at most
#define HOT_BUCKETS 4096u                  // buckets of 4 slots (load at most 1/2)
#define HOT_SLOTS (HOT_BUCKETS * 4u)
#define HOT_EMPTY 0xffffffffu              // empty slot (that key is never hot)
#define HOT_THREADS 1024
#define HOT_IPT 16u                        // rows of 64 keys per chunk
#define HOT_CHUNK (64u * HOT_IPT)          // keys per chunk (one wavefront's unit of work)
#define HOT_PREFIX (1ull << 24)            // keys inserted the ordinary way first
#define HOT_MIN_KEYS (1ull << 26)          // shorter calls are not filtered
#define HOT_MIN_COVER_NUM 3u               // filter when the candidates cover at least
#define HOT_MIN_COVER_DEN 10u              //   3/10 of the sample

// device control words
#define HC_NCAND 0   // candidates found
#define HC_COVER 1   // sample keys the first HOT_MAX candidates cover
#define HC_NHOT 2    // keys in the hot table
#define HC_WORDS 4

struct qs_hot {
    uint32_t *sk[2], *sv[2];   // sample sort (keys, unused values)
    uint32_t *ck[2], *cv[2];   // candidate sorts
    uint32_t *hist;            // radix histograms of both sorts
    uint32_t *cflag;           // candidate flags of the sorted sample, scanned in place (+ total)
    uint32_t *cbsum;           // block sums of that scan (+ total)
    uint32_t *ctl;             // HC_* words
    uint32_t *h_ctl;           // pinned copy
    uint32_t *tkey;            // hot table: key per slot
    uint32_t *tinfo;           // its link id
    uint32_t *tmask;           // its mask
    uint32_t *tfill;           // entries claimed per bucket while placing
    uint32_t *part;            // per persistent block and slot: hot inserts counted
    uint32_t nblk;             // persistent blocks (rows of part)
    uint32_t *ccold;           // cold keys per chunk, scanned in place (+ total)
    uint32_t *stage;           // per chunk of HOT_CHUNK slots: its cold keys, packed at the front
    uint32_t *bsum;            // block sums of the chunk-count scan (+ total)
    uint64_t chunk_cap;
    uint32_t *cold;            // the cold keys of the call, in stream order
    uint64_t cold_cap;
    uint32_t seq;              // sequence number of the last publication to h_ctl
    const uint32_t *cand;      // the candidates, most frequent first (set by hot_cand_launch)
    hipStream_t side;          // the candidate phase of a plain insert runs here (default priority)
    hipEvent_t ev_fork;        // the caller's stream up to the call
    hipEvent_t ev_cand;        // the candidate phase done
    bool prev_flat;            // the previous call's sample showed no skew
};

// The host reads a few device words without a copy or a synchronisation
// call: one wavefront writes them into coherent pinned host memory and then,
// after a system-scope fence, a new sequence number into the last word; the
// host polls that word.  (The filter's two waits sit between launches; going
// through the runtime's copy and completion path there left the GPU idle for
// milliseconds in some calls.)
#define HOT_PUB_WORDS 8u // h_ctl[0 .. 7) data, h_ctl[7] sequence

__global__ void k_hot_publish(const uint32_t *__restrict__ src, uint32_t nwords, uint32_t *dst, uint32_t seq)
{
    const uint32_t i = threadIdx.x;
    if (i < nwords)
        __hip_atomic_store(dst + i, src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    __syncthreads();
    if (i == 0)
        __hip_atomic_store(dst + HOT_PUB_WORDS - 1, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// publish nwords (< HOT_PUB_WORDS) device words to h->h_ctl and wait for them
static int hot_read(qs_hot *h, const uint32_t *d_src, uint32_t nwords, hipStream_t st)
{
    const uint32_t seq = ++h->seq;
    hipLaunchKernelGGL(k_hot_publish, dim3(1), dim3(64), 0, st, d_src, nwords, h->h_ctl, seq);
    CK(hipGetLastError());
    for (uint64_t it = 1; __atomic_load_n(h->h_ctl + HOT_PUB_WORDS - 1, __ATOMIC_ACQUIRE) != seq; ++it) {
        if ((it & 0xfffffu) == 0) { // now and then: has the stream failed, or finished without it?
            const hipError_t e = hipStreamQuery(st);
            if (e != hipErrorNotReady &&
                __atomic_load_n(h->h_ctl + HOT_PUB_WORDS - 1, __ATOMIC_ACQUIRE) != seq)
                return QS_ERR_HIP;
        }
    }
    return QS_OK;
}

// the two buckets a key may sit in (distinct)
__device__ static inline void hot_buckets(uint32_t key, uint32_t &b1, uint32_t &b2)
{
    b1 = (key * 0x9E3779B1u) >> 20; // HOT_BUCKETS = 2^12
    b2 = (key * 0x85EBCA77u) >> 20;
    b2 = b2 == b1 ? b1 ^ 1u : b2;
}

__global__ void k_hot_sample(const uint32_t *__restrict__ keys, uint32_t n, uint32_t *__restrict__ sk,
                             uint32_t *__restrict__ sv, uint32_t *__restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        sk[i] = keys[i];
        sv[i] = 0;
    }
    if (i < HC_WORDS)
        ctl[i] = 0;
}

// The first index of the run that ends at i in the sorted sample a (the run
// is at least HOT_MIN_RUN long): gallop back from the run end, then bisect;
// the loads stay near i.
__device__ static inline uint32_t run_start(const uint32_t *__restrict__ a, uint32_t i)
{
    const uint32_t k = a[i];
    uint32_t hi = i + 1u - HOT_MIN_RUN, lo = hi, step = HOT_MIN_RUN;
    while (lo > 0) { // a[lo] < k or lo = 0 ends the gallop
        lo = lo > step ? lo - step : 0u;
        if (a[lo] < k)
            break;
        hi = lo;
        step <<= 1;
    }
    if (a[lo] == k) // lo = 0 holds k: the run starts at 0
        hi = lo;
    while (lo + 1u < hi) { // a[lo] < k <= a[hi]
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < k)
            lo = mid;
        else
            hi = mid;
    }
    return hi;
}

// whether i ends a run of the sorted sample that is at least HOT_MIN_RUN long
__device__ static inline bool cand_end(const uint32_t *__restrict__ a, uint32_t n, uint32_t i)
{
    const uint32_t k = a[i];
    return !(i + 1 < n && a[i + 1] == k) && i + 1 >= HOT_MIN_RUN && a[i + 1 - HOT_MIN_RUN] == k;
}

// flag[i] = 1 when i ends a candidate run (scanned next into offsets)
__global__ void k_hot_cand_flag(const uint32_t *__restrict__ a, uint32_t n, uint32_t *__restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flag[i] = cand_end(a, n, i) ? 1u : 0u;
}

// Each candidate run's key and length at its scanned offset (off: exclusive
// offsets, total at off[n]), so the candidates come out in key order, with no
// atomic on a shared counter.
__global__ void k_hot_cand(const uint32_t *__restrict__ a, uint32_t n, const uint32_t *__restrict__ off,
                           uint32_t *__restrict__ ck, uint32_t *__restrict__ cv, uint32_t *__restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        ctl[HC_NCAND] = off[n];
    if (i >= n || off[i + 1] == off[i])
        return;
    ck[off[i]] = a[i];
    cv[off[i]] = i - run_start(a, i) + 1;
}

// candidates in key order -> (HOT_SAMPLE - count, key), so that a stable
// sort of the first orders them by count, ties by key
__global__ void k_hot_rekey(const uint32_t *__restrict__ ck, const uint32_t *__restrict__ cv,
                            uint32_t *__restrict__ k2, uint32_t *__restrict__ v2, const uint32_t *__restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ctl[HC_NCAND]) {
        k2[i] = HOT_SAMPLE - cv[i];
        v2[i] = ck[i];
    }
}

// sample keys the first HOT_MAX candidates cover (one block)
__global__ __launch_bounds__(1024) void k_hot_cover(const uint32_t *__restrict__ k2, uint32_t *__restrict__ ctl)
{
    __shared__ uint32_t s;
    if (threadIdx.x == 0)
        s = 0;
    __syncthreads();
    const uint32_t m = min(ctl[HC_NCAND], HOT_MAX);
    uint32_t sum = 0;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x)
        sum += HOT_SAMPLE - k2[i];
    atomicAdd(&s, sum);
    __syncthreads();
    if (threadIdx.x == 0)
        ctl[HC_COVER] = s;
}

__global__ void k_hot_clear(uint32_t *__restrict__ tkey, uint32_t *__restrict__ tfill, uint32_t *__restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < HOT_SLOTS)
        tkey[i] = HOT_EMPTY;
    if (i < HOT_BUCKETS)
        tfill[i] = 0;
    if (i == 0)
        ctl[HC_NHOT] = 0;
}

// the candidates whose bucket is linked now enter the hot table
__global__ void k_hot_build(const uint32_t *__restrict__ cand, qs_sdesc d0, const uint32_t *__restrict__ low,
                            const uint32_t *__restrict__ masks, uint32_t *__restrict__ tkey,
                            uint32_t *__restrict__ tinfo, uint32_t *__restrict__ tmask, uint32_t *tfill,
                            uint32_t *ctl)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= min(ctl[HC_NCAND], HOT_MAX))
        return;
    const uint32_t k = cand[c];
    if (k == HOT_EMPTY)
        return;
    const uint32_t h = k * d0.seed;
    const uint32_t b = qs_fastmod(h, d0.fd_low);
    const uint32_t w = low[(size_t)b * 8u + 7u];
    if (w <= QS_LINK_START)
        return;
    const uint32_t M = masks[qs_fastmod(h, d0.fd_masks)];
    uint32_t b1, b2;
    hot_buckets(k, b1, b2);
    // the emptier bucket first (a racy look: only the placement depends on it)
    if (__hip_atomic_load(tfill + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
        __hip_atomic_load(tfill + b1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        const uint32_t t = b1;
        b1 = b2;
        b2 = t;
    }
    uint32_t pos = atomicAdd(tfill + b1, 1u), hb = b1;
    if (pos >= 4u) {
        pos = atomicAdd(tfill + b2, 1u);
        hb = b2;
    }
    if (pos >= 4u)
        return; // both full: the key stays on the ordered path
    const uint32_t slot = hb * 4u + pos;
    tkey[slot] = k;
    tinfo[slot] = w;
    tmask[slot] = M;
    atomicAdd(ctl + HC_NHOT, 1u);
}

// every candidate enters the hot table (a sink whose inserts all commute)
__global__ void k_hot_build_any(const uint32_t *__restrict__ cand, uint32_t *__restrict__ tkey, uint32_t *tfill,
                                uint32_t *ctl)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= min(ctl[HC_NCAND], HOT_MAX))
        return;
    const uint32_t k = cand[c];
    if (k == HOT_EMPTY)
        return;
    uint32_t b1, b2;
    hot_buckets(k, b1, b2);
    if (__hip_atomic_load(tfill + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
        __hip_atomic_load(tfill + b1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        const uint32_t t = b1;
        b1 = b2;
        b2 = t;
    }
    uint32_t pos = atomicAdd(tfill + b1, 1u), hb = b1;
    if (pos >= 4u) {
        pos = atomicAdd(tfill + b2, 1u);
        hb = b2;
    }
    if (pos >= 4u)
        return; // both full: the key stays cold
    tkey[hb * 4u + pos] = k;
    atomicAdd(ctl + HC_NHOT, 1u);
}

// slot of key k in the LDS copy of the hot table, or -1: both buckets read
// at once, no loop (the empty marker is never looked up: it is never hot)
__device__ static inline int hot_find(const uint4 *skey4, uint32_t k)
{
    uint32_t b1, b2;
    hot_buckets(k, b1, b2);
    const uint4 p = skey4[b1], q = skey4[b2];
    int s = -1;
    s = p.x == k ? (int)(b1 * 4u + 0u) : s;
    s = p.y == k ? (int)(b1 * 4u + 1u) : s;
    s = p.z == k ? (int)(b1 * 4u + 2u) : s;
    s = p.w == k ? (int)(b1 * 4u + 3u) : s;
    s = q.x == k ? (int)(b2 * 4u + 0u) : s;
    s = q.y == k ? (int)(b2 * 4u + 1u) : s;
    s = q.z == k ? (int)(b2 * 4u + 2u) : s;
    s = q.w == k ? (int)(b2 * 4u + 3u) : s;
    return k == HOT_EMPTY ? -1 : s;
}

// Pass 1 over the keys after the prefix.  A wavefront owns chunks of
// HOT_CHUNK consecutive keys (HOT_IPT rows of 64), loads the next chunk while
// it works on this one, counts hot keys per slot in LDS and writes the
// chunk's cold keys, in stream order, to the front of the chunk's own
// HOT_CHUNK slots of `stage`, and their number to ccold.  Persistent blocks
// (one per CU: the two LDS arrays take 128 KB) take chunks in turn and write
// their row of per-slot counts at the end.
__global__ __launch_bounds__(HOT_THREADS) void k_hot_count(const uint32_t *__restrict__ keys, uint64_t n,
                                                           uint32_t nchunks, const uint32_t *__restrict__ tkey,
                                                           uint32_t *__restrict__ part,
                                                           uint32_t *__restrict__ stage,
                                                           uint32_t *__restrict__ ccold)
{
    __shared__ uint4 skey4[HOT_BUCKETS];
    __shared__ uint32_t scnt[HOT_SLOTS];
    for (uint32_t e = threadIdx.x; e < HOT_BUCKETS; e += HOT_THREADS)
        skey4[e] = reinterpret_cast<const uint4 *>(tkey)[e];
    for (uint32_t e = threadIdx.x; e < HOT_SLOTS; e += HOT_THREADS)
        scnt[e] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t nw = gridDim.x * (HOT_THREADS / 64);
    uint32_t c = blockIdx.x * (HOT_THREADS / 64) + (threadIdx.x >> 6);
    uint32_t k[HOT_IPT], kn[HOT_IPT];
#pragma unroll
    for (uint32_t r = 0; r < HOT_IPT; ++r) {
        const uint64_t i = (uint64_t)c * HOT_CHUNK + r * 64u + lane;
        k[r] = c < nchunks && i < n ? __builtin_nontemporal_load(keys + i) : HOT_EMPTY;
    }
    for (; c < nchunks; c += nw) {
        const uint32_t cn = c + nw; // the next chunk, loaded while this one is probed
#pragma unroll
        for (uint32_t r = 0; r < HOT_IPT; ++r) {
            const uint64_t i = (uint64_t)cn * HOT_CHUNK + r * 64u + lane;
            kn[r] = cn < nchunks && i < n ? __builtin_nontemporal_load(keys + i) : HOT_EMPTY;
        }
        uint32_t cold = 0;
        uint32_t *dst = stage + (size_t)c * HOT_CHUNK;
#pragma unroll
        for (uint32_t r = 0; r < HOT_IPT; ++r) {
            const uint64_t i = (uint64_t)c * HOT_CHUNK + r * 64u + lane;
            const int s = hot_find(skey4, k[r]); // the key past the end is HOT_EMPTY: -1
            if (s >= 0)
                atomicAdd(&scnt[s], 1u);
            const bool is_cold = (i < n) & (s < 0);
            const uint64_t m = __ballot(is_cold); // every lane, outside any branch
            if (is_cold)
                dst[cold + (uint32_t)__popcll(m & below)] = k[r];
            cold += (uint32_t)__popcll(m);
        }
        if (lane == 0)
            ccold[c] = cold;
#pragma unroll
        for (uint32_t r = 0; r < HOT_IPT; ++r)
            k[r] = kn[r];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < HOT_SLOTS; e += HOT_THREADS)
        part[(size_t)blockIdx.x * HOT_SLOTS + e] = scnt[e];
}

// Pass 2: every chunk moves its packed cold keys from `stage` to its scanned
// offset (coff[c] .. coff[c + 1]), so the cold keys end up contiguous in
// stream order.  Only the cold keys are read again, not the whole stream.
__global__ __launch_bounds__(256) void k_hot_move(const uint32_t *__restrict__ stage, uint32_t nchunks,
                                                  const uint32_t *__restrict__ coff, uint32_t *__restrict__ cold)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t c = blockIdx.x * 4u + (threadIdx.x >> 6); c < nchunks; c += nw) {
        const uint32_t off = coff[c], cnt = coff[c + 1] - off; // cnt <= HOT_CHUNK
        const uint32_t *src = stage + (size_t)c * HOT_CHUNK;
        uint32_t v[HOT_IPT];
#pragma unroll
        for (uint32_t r = 0; r < HOT_IPT; ++r)
            v[r] = r * 64u + lane < cnt ? __builtin_nontemporal_load(src + r * 64u + lane) : 0u;
#pragma unroll
        for (uint32_t r = 0; r < HOT_IPT; ++r)
            if (r * 64u + lane < cnt)
                cold[(size_t)off + r * 64u + lane] = v[r];
    }
}

// Exclusive scan of the per-chunk cold counts, in place, with the total after
// them: block sums over SCAN_SPAN entries, one block scans those, and every
// block scans its span again from its offset.
#define SCAN_SPAN 8192u // entries per block: 1024 threads x 8

__device__ static inline uint32_t block_excl_scan(uint32_t x, uint32_t *s_w, uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, o, 64);
        inc += lane >= (uint32_t)o ? y : 0u;
    }
    if (lane == 63)
        s_w[wv] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
        before += w < wv ? s_w[w] : 0u;
        total += s_w[w];
    }
    __syncthreads();
    return before + inc - x;
}

__global__ __launch_bounds__(1024) void k_scan_reduce(const uint32_t *__restrict__ a, uint32_t n,
                                                      uint32_t *__restrict__ bsum)
{
    __shared__ uint32_t s_w[16];
    const uint32_t i0 = blockIdx.x * SCAN_SPAN + threadIdx.x * 8u;
    uint32_t x = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k)
        x += i0 + k < n ? a[i0 + k] : 0u;
    uint32_t total;
    block_excl_scan(x, s_w, total);
    if (threadIdx.x == 0)
        bsum[blockIdx.x] = total;
}

// one block: exclusive scan of the nb block sums (nb <= 8192), the total to bsum[nb]
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t *__restrict__ bsum, uint32_t nb)
{
    __shared__ uint32_t s_w[16];
    const uint32_t i0 = threadIdx.x * 8u;
    uint32_t v[8], x = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) {
        v[k] = i0 + k < nb ? bsum[i0 + k] : 0u;
        x += v[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(x, s_w, total);
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) {
        if (i0 + k < nb)
            bsum[i0 + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0)
        bsum[nb] = total;
}

__global__ __launch_bounds__(1024) void k_scan_apply(uint32_t *__restrict__ a, uint32_t n,
                                                     const uint32_t *__restrict__ bsum, uint32_t nb)
{
    __shared__ uint32_t s_w[16];
    const uint32_t i0 = blockIdx.x * SCAN_SPAN + threadIdx.x * 8u;
    uint32_t v[8], x = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) {
        v[k] = i0 + k < n ? a[i0 + k] : 0u;
        x += v[k];
    }
    uint32_t total;
    uint32_t run = bsum[blockIdx.x] + block_excl_scan(x, s_w, total);
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) {
        if (i0 + k < n)
            a[i0 + k] = run;
        run += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        a[n] = bsum[nb];
}

// the hot counts summed over the persistent blocks and added to the high
// rows: a block takes 64 slots, its four wavefronts a quarter of the rows each
#define FLUSH_WAVES 16u // wavefronts per block, each summing every 16th row of part
__global__ __launch_bounds__(64 * FLUSH_WAVES) void k_hot_flush(const uint32_t *__restrict__ part, uint32_t nblk,
                                                                const uint32_t *__restrict__ tkey,
                                                                const uint32_t *__restrict__ tinfo,
                                                                const uint32_t *__restrict__ tmask,
                                                                uint32_t *__restrict__ high)
{
    __shared__ uint32_t s_sum[FLUSH_WAVES][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t s = blockIdx.x * 64u + lane;
    uint32_t c = 0;
    for (uint32_t b = wv; b < nblk; b += FLUSH_WAVES)
        c += part[(size_t)b * HOT_SLOTS + s];
    s_sum[wv][lane] = c;
    __syncthreads();
    if (wv != 0)
        return;
    c = 0;
#pragma unroll
    for (uint32_t w = 0; w < FLUSH_WAVES; ++w)
        c += s_sum[w][lane];
    if (!c || tkey[s] == HOT_EMPTY)
        return;
    uint32_t *g = high + (size_t)(tinfo[s] - QS_LINK_START) * 32u;
    for (uint32_t M = tmask[s]; M; M &= M - 1u)
        atomicAdd(g + __builtin_ctz(M), c);
}

// ===========================================================================
// host
// ===========================================================================
void qs_hot_state_free(void *state)
{
    qs_hot *h = (qs_hot *)state;
    if (!h)
        return;
    for (int k = 0; k < 2; ++k) {
        (void)hipFree(h->sk[k]);
        (void)hipFree(h->sv[k]);
        (void)hipFree(h->ck[k]);
        (void)hipFree(h->cv[k]);
    }
    (void)hipFree(h->hist);
    (void)hipFree(h->cflag);
    (void)hipFree(h->cbsum);
    (void)hipFree(h->ctl);
    (void)hipFree(h->tkey);
    (void)hipFree(h->tinfo);
    (void)hipFree(h->tmask);
    (void)hipFree(h->tfill);
    (void)hipFree(h->part);
    (void)hipFree(h->ccold);
    (void)hipFree(h->stage);
    (void)hipFree(h->bsum);
    (void)hipFree(h->cold);
    if (h->h_ctl)
        (void)hipHostFree(h->h_ctl);
    if (h->ev_fork)
        (void)hipEventDestroy(h->ev_fork);
    if (h->ev_cand)
        (void)hipEventDestroy(h->ev_cand);
    if (h->side)
        (void)hipStreamDestroy(h->side);
    free(h);
}

void qs_hot_free(qs_sketch *s)
{
    qs_hot_state_free(s->hot);
    s->hot = nullptr;
}

// the filter's scratch in *state (created on first use), sized for `rest`
// keys after the part inserted the ordinary way
static int hot_alloc(void **state, uint64_t rest)
{
    qs_hot *h = (qs_hot *)*state;
    if (!h) {
        h = (qs_hot *)calloc(1, sizeof(qs_hot));
        if (!h)
            return QS_ERR_NOMEM;
        *state = h;
        int dev = 0, cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 256;
        h->nblk = (uint32_t)cu;
        bool ok = true;
        for (int k = 0; k < 2; ++k)
            ok = ok && hipMalloc(&h->sk[k], 4ull * HOT_SAMPLE) == hipSuccess &&
                 hipMalloc(&h->sv[k], 4ull * HOT_SAMPLE) == hipSuccess &&
                 hipMalloc(&h->ck[k], 4ull * HOT_CAND_MAX) == hipSuccess &&
                 hipMalloc(&h->cv[k], 4ull * HOT_CAND_MAX) == hipSuccess;
        ok = ok && hipMalloc(&h->hist, 4ull * qs_sort_hist_words(HOT_SAMPLE)) == hipSuccess &&
             hipMalloc(&h->cflag, 4ull * (HOT_SAMPLE + 1)) == hipSuccess &&
             hipMalloc(&h->cbsum, 4ull * (HOT_SAMPLE / SCAN_SPAN + 1)) == hipSuccess &&
             hipMalloc(&h->ctl, 4ull * HC_WORDS) == hipSuccess &&
             hipMalloc(&h->tkey, 4ull * HOT_SLOTS) == hipSuccess &&
             hipMalloc(&h->tinfo, 4ull * HOT_SLOTS) == hipSuccess &&
             hipMalloc(&h->tmask, 4ull * HOT_SLOTS) == hipSuccess &&
             hipMalloc(&h->tfill, 4ull * HOT_BUCKETS) == hipSuccess &&
             hipMalloc(&h->part, 4ull * HOT_SLOTS * h->nblk) == hipSuccess &&
             hipHostMalloc(&h->h_ctl, 4ull * HOT_PUB_WORDS, hipHostMallocCoherent) == hipSuccess &&
             hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&h->ev_cand, hipEventDisableTiming) == hipSuccess;
        if (ok) {
            memset(h->h_ctl, 0, 4ull * HOT_PUB_WORDS);
            h->seq = 0;
        }
        if (!ok)
            return QS_ERR_NOMEM;
    }
    const uint64_t chunks = (rest + HOT_CHUNK - 1) / HOT_CHUNK;
    if (h->chunk_cap < chunks) {
        (void)hipFree(h->ccold);
        (void)hipFree(h->stage);
        (void)hipFree(h->bsum);
        h->ccold = nullptr;
        h->stage = nullptr;
        h->bsum = nullptr;
        h->chunk_cap = 0;
        if (hipMalloc(&h->ccold, 4ull * (chunks + 1)) != hipSuccess ||
            hipMalloc(&h->stage, 4ull * HOT_CHUNK * chunks) != hipSuccess ||
            hipMalloc(&h->bsum, 4ull * ((chunks + SCAN_SPAN - 1) / SCAN_SPAN + 1)) != hipSuccess)
            return QS_ERR_NOMEM;
        h->chunk_cap = chunks;
    }
    if (h->cold_cap < rest) {
        (void)hipFree(h->cold);
        h->cold = nullptr;
        h->cold_cap = 0;
        if (hipMalloc(&h->cold, 4ull * rest) != hipSuccess)
            return QS_ERR_NOMEM;
        h->cold_cap = rest;
    }
    return QS_OK;
}

// Candidates from the first HOT_SAMPLE keys (step 1), queued on `stream`:
// h->cand = the candidate keys, most frequent first (at most HOT_MAX are
// used).  hot_cand_result then waits for their cover.
static int hot_cand_launch(qs_hot *h, const uint32_t *keys, hipStream_t stream)
{
    hipLaunchKernelGGL(k_hot_sample, dim3(HOT_SAMPLE / 256), dim3(256), 0, stream, keys, HOT_SAMPLE, h->sk[0],
                       h->sv[0], h->ctl);
    const int r0 = qs_sort_pairs(h->sk, h->sv, h->hist, HOT_SAMPLE, 32, stream);
    if (r0 < 0)
        return QS_ERR_HIP;
    // candidates in key order: flags, a scan, each at its offset
    const uint32_t snb = HOT_SAMPLE / SCAN_SPAN;
    hipLaunchKernelGGL(k_hot_cand_flag, dim3(HOT_SAMPLE / 256), dim3(256), 0, stream, h->sk[r0], HOT_SAMPLE,
                       h->cflag);
    hipLaunchKernelGGL(k_scan_reduce, dim3(snb), dim3(1024), 0, stream, h->cflag, HOT_SAMPLE, h->cbsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, h->cbsum, snb);
    hipLaunchKernelGGL(k_scan_apply, dim3(snb), dim3(1024), 0, stream, h->cflag, HOT_SAMPLE, h->cbsum, snb);
    hipLaunchKernelGGL(k_hot_cand, dim3(HOT_SAMPLE / 256), dim3(256), 0, stream, h->sk[r0], HOT_SAMPLE, h->cflag,
                       h->ck[0], h->cv[0], h->ctl);
    uint32_t *k2[2] = {h->ck[1], h->ck[0]}, *v2[2] = {h->cv[1], h->cv[0]};
    hipLaunchKernelGGL(k_hot_rekey, dim3(HOT_CAND_MAX / 256), dim3(256), 0, stream, h->ck[0], h->cv[0], k2[0],
                       v2[0], h->ctl);
    const int r2 = qs_sort_pairs_dev(k2, v2, h->hist, h->ctl + HC_NCAND, HOT_CAND_MAX, 21, stream);
    if (r2 < 0)
        return QS_ERR_HIP;
    hipLaunchKernelGGL(k_hot_cover, dim3(1), dim3(1024), 0, stream, k2[r2], h->ctl);
    CK(hipGetLastError());
    h->cand = v2[r2];
    return QS_OK;
}

// *skewed = whether the candidates cover at least HOT_MIN_COVER of the
// sample (read after the candidate kernels queued on `stream`)
static int hot_cand_result(qs_hot *h, hipStream_t stream, bool *skewed)
{
    *skewed = false;
    const int rc = hot_read(h, h->ctl, HC_WORDS, stream);
    if (rc)
        return rc;
    const uint64_t cover = h->h_ctl[HC_COVER];
    *skewed = h->h_ctl[HC_NCAND] != 0 && cover * HOT_MIN_COVER_DEN >= (uint64_t)HOT_SAMPLE * HOT_MIN_COVER_NUM;
    return QS_OK;
}

// Blocking insert of n 32-bit keys into a plain sketch through the hot-key
// filter.  *used = 0 (and nothing inserted) when the call is too short or its
// sample shows too little skew: the caller then runs the ordinary insert.
int qs_insert_hot(qs_sketch *s, const void *d_keys, uint64_t n, hipStream_t stream, int *used)
{
    *used = 0;
    if (s->nsk != 1 || s->key_bytes != 4 || n < HOT_MIN_KEYS || n - HOT_PREFIX > 0xffffffffull)
        return QS_OK;
    const uint32_t *keys = (const uint32_t *)d_keys;
    const uint64_t rest = n - HOT_PREFIX;
    int rc = hot_alloc(&s->hot, rest);
    if (rc)
        return rc;
    qs_hot *h = (qs_hot *)s->hot;
    const qs_sdesc d0 = s->h_desc[0];
    qs_evpair e1, e2;
    const bool timed = s->timing != 0;

    // The insert scratch is sized before the prefix is queued, for the whole
    // call (at least every key after the prefix, the most the cold part can
    // hold): the cold count varies a little between calls (the hot table's
    // placement races), and a call without skew after one with skew runs as
    // one chunk; growing the scratch between launches, or in the next call,
    // would free and allocate.
    auto size_scratch = [&]() -> int {
        const uint64_t offs_n[2] = {0, n};
        const uint64_t cap = qs_chunk_keys(s, offs_n);
        return qs_ensure_scratch(s, n < cap ? n : cap);
    };
    const uint64_t offs_p[2] = {0, HOT_PREFIX};
    bool skewed = false;
    if (!h->prev_flat) {
        // 2. the prefix goes through the ordinary pipeline on the caller's
        // stream while 1. the candidates run on the side stream, from the
        // point the caller's stream had reached at this call (the two share
        // no memory but the keys, which both only read).  The prefix is
        // queued first: its launches are what keeps the GPU busy while the
        // host queues the candidates' short kernels.  The prefix is inserted
        // whatever the sample shows: a prefix and then the rest is the
        // sequential result, as for any call split into chunks.
        rc = size_scratch();
        if (rc)
            return rc;
        CK(hipEventRecord(h->ev_fork, stream));
        *used = 1;
        rc = qs_insert_set(s, keys, offs_p, nullptr, stream, false, false);
        if (rc)
            return rc;
        CK(hipStreamWaitEvent(h->side, h->ev_fork, 0));
        rc = hot_cand_launch(h, keys, h->side);
        if (rc)
            return rc;
        CK(hipEventRecord(h->ev_cand, h->side));
        rc = hot_cand_result(h, h->side, &skewed); // the GPU runs the prefix meanwhile
        if (rc)
            return rc;
        h->prev_flat = !skewed;
        if (!skewed) { // not skewed enough to pay for the filter: the rest the ordinary way
            CK(hipStreamWaitEvent(stream, h->ev_cand, 0));
            const uint64_t offs_r[2] = {0, rest};
            return qs_insert_set(s, keys + HOT_PREFIX, offs_r, nullptr, stream, true, false);
        }
    } else {
        // The previous call showed no skew: 1. the candidates first, on the
        // caller's stream, so that a call without skew again runs the
        // ordinary pipeline as one chunk (split at the prefix, every bucket
        // the prefix touches would be read and written once more: 0.7 ms
        // of a 2^28-key Zipf 0.8 call).  Either order gives the same state.
        // The candidates are timed with the filter here: the call waits for
        // them.
        qs_evpair e0;
        if (timed)
            qs_tim_begin(s, e0, QS_TK_HOT, 0, stream);
        rc = hot_cand_launch(h, keys, stream);
        if (timed)
            qs_tim_end(s, e0, stream);
        if (rc)
            return rc;
        rc = hot_cand_result(h, stream, &skewed);
        if (rc)
            return rc;
        h->prev_flat = !skewed;
        if (!skewed)
            return QS_OK; // *used = 0: the caller runs the ordinary insert
        rc = size_scratch();
        if (rc)
            return rc;
        *used = 1;
        rc = qs_insert_set(s, keys, offs_p, nullptr, stream, false, false); // 2. the prefix
        if (rc)
            return rc;
        CK(hipEventRecord(h->ev_cand, stream)); // (the wait below is then already met)
    }
    const uint32_t *cand = h->cand;

    // 3. the hot set, 4. the filter
    const uint32_t nchunks = (uint32_t)((rest + HOT_CHUNK - 1) / HOT_CHUNK);
    if (timed)
        qs_tim_begin(s, e1, QS_TK_HOT, 0, stream); // a wait for the candidates counts here
    CK(hipStreamWaitEvent(stream, h->ev_cand, 0));
    hipLaunchKernelGGL(k_hot_clear, dim3(HOT_SLOTS / 256), dim3(256), 0, stream, h->tkey, h->tfill, h->ctl);
    hipLaunchKernelGGL(k_hot_build, dim3(HOT_MAX / 256), dim3(256), 0, stream, cand, d0, s->d_low, s->d_masks,
                       h->tkey, h->tinfo, h->tmask, h->tfill, h->ctl);
    hipLaunchKernelGGL(k_hot_count, dim3(h->nblk), dim3(HOT_THREADS), 0, stream, keys + HOT_PREFIX, rest,
                       nchunks, h->tkey, h->part, h->stage, h->ccold);
    const uint32_t nb = (nchunks + SCAN_SPAN - 1) / SCAN_SPAN; // <= 8192 for any u32 chunk count of a call
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(1024), 0, stream, h->ccold, nchunks, h->bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, h->bsum, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(1024), 0, stream, h->ccold, nchunks, h->bsum, nb);
    hipLaunchKernelGGL(k_hot_move, dim3(h->nblk * 8u), dim3(256), 0, stream, h->stage, nchunks, h->ccold, h->cold);
    hipLaunchKernelGGL(k_hot_flush, dim3(HOT_SLOTS / 64), dim3(64 * FLUSH_WAVES), 0, stream, h->part, h->nblk, h->tkey,
                       h->tinfo, h->tmask, s->d_high);
    CK(hipGetLastError());
    if (timed)
        qs_tim_end(s, e1, stream);
    rc = hot_read(h, h->ccold + nchunks, 1, stream);
    if (rc)
        return rc;
    const uint64_t ncold = h->h_ctl[0];
    const uint64_t nhot = rest - ncold;
    s->hot_keys += nhot;
    s->keys_inserted += nhot;
    if (timed) { // the filtered keys count as inserted keys of the timed path
        qs_tim_begin(s, e2, QS_TIMING_KERNELS, nhot, stream);
        qs_tim_end(s, e2, stream);
    }

    // 5. the cold keys the ordinary way (blocking: reports a pool error)
    if (ncold == 0) {
        CK(hipMemcpyAsync(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        if (s->h_st->err) {
            s->sticky_err = qs_err_code(s->h_st->err);
            return s->sticky_err;
        }
        return QS_OK;
    }
    const uint64_t offs_c[2] = {0, ncold};
    return qs_insert_set(s, h->cold, offs_c, nullptr, stream, true, false);
}

// For a sink whose inserts all commute (the SF sketch): the n keys split into
// hot keys, counted per table slot and persistent block, and cold keys in
// stream order.  out->split = 0 (nothing done) when n < HOT_MIN_KEYS or the
// sample shows too little skew.
int qs_hot_split_keys(void **state, const uint32_t *keys, uint64_t n, hipStream_t stream, qs_hot_split *out)
{
    memset(out, 0, sizeof(*out));
    if (n < HOT_MIN_KEYS || n > 0xffffffffull)
        return QS_OK;
    int rc = hot_alloc(state, n);
    if (rc)
        return rc;
    qs_hot *h = (qs_hot *)*state;
    bool skewed = false;
    rc = hot_cand_launch(h, keys, stream);
    if (rc)
        return rc;
    rc = hot_cand_result(h, stream, &skewed);
    if (rc || !skewed)
        return rc;
    const uint32_t *cand = h->cand;
    const uint32_t nchunks = (uint32_t)((n + HOT_CHUNK - 1) / HOT_CHUNK);
    hipLaunchKernelGGL(k_hot_clear, dim3(HOT_SLOTS / 256), dim3(256), 0, stream, h->tkey, h->tfill, h->ctl);
    hipLaunchKernelGGL(k_hot_build_any, dim3(HOT_MAX / 256), dim3(256), 0, stream, cand, h->tkey, h->tfill, h->ctl);
    hipLaunchKernelGGL(k_hot_count, dim3(h->nblk), dim3(HOT_THREADS), 0, stream, keys, n, nchunks, h->tkey, h->part,
                       h->stage, h->ccold);
    const uint32_t nb = (nchunks + SCAN_SPAN - 1) / SCAN_SPAN;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(1024), 0, stream, h->ccold, nchunks, h->bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, h->bsum, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(1024), 0, stream, h->ccold, nchunks, h->bsum, nb);
    hipLaunchKernelGGL(k_hot_move, dim3(h->nblk * 8u), dim3(256), 0, stream, h->stage, nchunks, h->ccold, h->cold);
    CK(hipGetLastError());
    rc = hot_read(h, h->ccold + nchunks, 1, stream);
    if (rc)
        return rc;
    out->split = 1;
    out->cold = h->cold;
    out->ncold = h->h_ctl[0];
    out->part = h->part;
    out->nblk = h->nblk;
    out->tkey = h->tkey;
    out->nslots = HOT_SLOTS;
    return QS_OK;
}
