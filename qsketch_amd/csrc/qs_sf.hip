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
// Here both tables live in HBM and every insert is a commuting add, so the
// keys can be applied in any order.  An insert call:
//
//   k_sfs_up / k_sfs_down  stable LSD radix passes (8-bit digits at most) sort
//          the keys by bucket, so each bucket's keys are one run
//   k_sf_rows  persistent wavefronts take ranges of the sorted keys from a
//          counter and walk them 64 at a time.  A bucket inside one range
//          has one owner: the lane heading its run applies each run of equal
//          keys with a plain read-modify-write, adding the run's length to
//          the key's three splits and writing each of the three counters as
//          the largest of its splits.  Only a range's first and last bucket
//          can be shared with another range; those take atomic adds and are
//          listed for
//   k_sf_fix  which writes the listed buckets' counters once every add has
//          landed
//
// Splits only grow, so a counter no add touched already holds its largest
// split: after the call every counter is what update() leaves.  Inserts
// shorter than SF_SORT_MIN skip the sort and go to k_sf_runs (runs of equal
// adjacent keys, atomic adds that return the old value, an atomic max).
// Long skewed calls of 4-byte keys first pass the hot-key filter of
// qs_hot.hip (qs_hot_split_keys): the frequent keys are counted in LDS and
// added once each by k_sf_hot_flush, and only the cold keys are sorted.
// k_sf_query reads the key's bucket in 16-byte pieces, m / 4 lanes per key,
// and takes the minimum of the mask's counters.
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
#define SF_WAVES (SF_THREADS / 64)
#define SF_RUN_WAVES 8192u      // wavefronts of k_sf_runs (32 per CU)
#define SF_ROW_RANGE 4096u     // keys per work item of k_sf_rows
#define SF_SORT_MIN (1u << 16)  // shorter inserts are applied unsorted
#define SF_CHUNK4 (1u << 28)    // keys per radix chunk: 4-byte keys
#define SF_CHUNK8 (1u << 27)    //                       8-byte keys

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
    // radix scratch, sized by the longest chunk so far
    void *d_buf[2];
    uint32_t *d_hist;   // digit-major tile counts, then 256 digit totals
    uint32_t *d_err;    // [0] set by a scatter position outside the chunk, [1] work counter
    uint32_t *d_shared; // two buckets per range of k_sf_rows (SF_NO_BUCKET: none)
    uint64_t scratch_keys;
    void *hot;          // the hot-key filter's scratch (qs_hot.hip)
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
    __device__ static inline uint32_t bucket(uint64_t k, const sf_par &p)
    {
        return qs_fastmod((uint32_t)k * p.s32[0], p.fd_n);
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
    __device__ static inline uint32_t bucket(uint64_t key, const sf_par &p)
    {
        return (uint32_t)((key * p.s64[0]) % p.n);
    }
};

// ===========================================================================
// sort: stable LSD radix passes over the four bytes of the sort bits
// ===========================================================================
// A tile is SF_THREADS * (64 / KB) keys (16 KiB); a wavefront's lanes hold
// consecutive keys in every round.
template <int KB> struct sfs_tile {
    static constexpr int IPL = 64 / KB;
    static constexpr uint32_t TILE = SF_THREADS * IPL;
    static constexpr uint32_t WAVE_ITEMS = 64u * IPL;
};

template <int KB>
__device__ static inline uint32_t sfs_digit(KeyT<KB> key, const sf_par &p, int shift, uint32_t dmask)
{
    return (sf_idx<KB>::bucket((uint64_t)key, p) >> shift) & dmask;
}

// per-tile digit counts, digit-major (hist[d * ntiles + tile]); one LDS add
// per run of equal digits on adjacent lanes (sorted and hot keys fill whole
// wavefronts with one digit)
template <int KB>
__global__ __launch_bounds__(SF_THREADS) void k_sfs_up(const KeyT<KB> *__restrict__ in, uint32_t n, sf_par p,
                                                       int shift, uint32_t dmask, uint32_t ntiles,
                                                       uint32_t *__restrict__ hist)
{
    using C = sfs_tile<KB>;
    __shared__ uint32_t s_h[256];
    const uint32_t t = threadIdx.x, lane = t & 63u;
    s_h[t] = 0;
    __syncthreads();
    const uint32_t tb = blockIdx.x * C::TILE;
    KeyT<KB> raw[C::IPL];
#pragma unroll
    for (int r = 0; r < C::IPL; ++r) { // every load first, from clamped positions
        const uint32_t i = tb + r * SF_THREADS + t;
        raw[r] = __builtin_nontemporal_load(in + (i < n ? i : n - 1u));
    }
#pragma unroll
    for (int r = 0; r < C::IPL; ++r) {
        const uint32_t i = tb + r * SF_THREADS + t;
        const bool valid = i < n; // the valid lanes of a wavefront are a prefix
        const uint32_t d = valid ? sfs_digit<KB>(raw[r], p, shift, dmask) : 0u;
        const uint32_t dp = (uint32_t)__shfl_up((int)d, 1, 64);
        const bool head = valid && (lane == 0 || d != dp);
        // both ballots are taken by the whole wavefront, outside any branch
        const uint64_t heads = __ballot(head);
        const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
        const uint64_t later = heads & ~((2ull << lane) - 1ull);
        const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : nvalid;
        if (head)
            atomicAdd(&s_h[d], end - lane);
    }
    __syncthreads();
    if (t <= dmask)
        hist[(size_t)t * ntiles + blockIdx.x] = s_h[t];
}

// exclusive scan over the block of one value per thread
__device__ static inline uint32_t sf_block_excl_scan(uint32_t v, uint32_t *s_red)
{
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if (lane >= (uint32_t)o)
            x += y;
    }
    if (lane == 63)
        s_red[wv] = x;
    __syncthreads();
    uint32_t off = 0;
#pragma unroll
    for (int k = 0; k < SF_WAVES; ++k)
        off += (k < (int)wv) ? s_red[k] : 0u;
    __syncthreads();
    return off + x - v;
}

// Stable scatter of one tile.  Each wavefront ranks its keys round by round:
// one ballot per digit bit gives every lane the lanes sharing its digit, and
// the group's lowest lane advances the wavefront's running count of the
// digit.  The tile is staged digit-sorted in LDS and written as contiguous
// digit runs.  hist holds the exclusive tile prefixes of every digit and
// totals the digit totals (qs_scan_digits).
template <int KB>
__global__ __launch_bounds__(SF_THREADS) void k_sfs_down(const KeyT<KB> *__restrict__ in, uint32_t n, sf_par p,
                                                         int shift, uint32_t dmask, uint32_t ntiles,
                                                         const uint32_t *__restrict__ hist,
                                                         const uint32_t *__restrict__ totals,
                                                         KeyT<KB> *__restrict__ out, uint32_t *__restrict__ err)
{
    using C = sfs_tile<KB>;
    __shared__ KeyT<KB> s_item[C::TILE];
    __shared__ uint32_t s_wh[SF_WAVES][256];
    __shared__ uint32_t s_ls[256];
    __shared__ uint32_t s_goff[256];
    __shared__ uint32_t s_red[SF_WAVES];
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint32_t tile = blockIdx.x;
    const uint32_t tb = tile * C::TILE;
#pragma unroll
    for (int w = 0; w < SF_WAVES; ++w)
        s_wh[w][t] = 0;
    const bool dv = t <= dmask;
    const uint32_t hcur = dv ? hist[(size_t)t * ntiles + tile] : 0u;
    const uint32_t dbase = sf_block_excl_scan(dv ? totals[t] : 0u, s_red);

    const uint64_t lt = (1ull << lane) - 1ull;
    KeyT<KB> key[C::IPL];
    uint32_t dr[C::IPL]; // digit | wave-local position << 8 | valid << 31
#pragma unroll
    for (int r = 0; r < C::IPL; ++r) {
        const uint32_t i = tb + wv * C::WAVE_ITEMS + r * 64u + lane;
        key[r] = __builtin_nontemporal_load(in + (i < n ? i : n - 1u));
    }
#pragma unroll
    for (int r = 0; r < C::IPL; ++r) {
        const uint32_t i = tb + wv * C::WAVE_ITEMS + r * 64u + lane;
        const bool valid = i < n;
        const uint32_t d = sfs_digit<KB>(key[r], p, shift, dmask);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            if ((dmask >> bit) & 1u) { // uniform: the digit width is a launch argument
                const bool on = (d >> bit) & 1u;
                const uint64_t bb = __ballot(on);
                m &= on ? bb : ~bb;
            }
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        const int leader = valid ? (int)__builtin_ctzll(m) : 0;
        uint32_t run = 0;
        if (valid && rank == 0)
            run = atomicAdd(&s_wh[wv][d], (uint32_t)__popcll(m));
        run = (uint32_t)__shfl((int)run, leader, 64);
        dr[r] = d | ((run + rank) << 8) | (valid ? 0x80000000u : 0u);
    }
    __syncthreads();
    {
        // per-wavefront counts -> each wavefront's start within the digit
        uint32_t cnt = 0;
#pragma unroll
        for (int w = 0; w < SF_WAVES; ++w) {
            const uint32_t c = s_wh[w][t];
            s_wh[w][t] = cnt;
            cnt += c;
        }
        const uint32_t lstart = sf_block_excl_scan(cnt, s_red);
        s_ls[t] = lstart;
        s_goff[t] = dbase + hcur - lstart;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < C::IPL; ++r)
        if (dr[r] & 0x80000000u) {
            const uint32_t d = dr[r] & 0xffu;
            s_item[s_ls[d] + s_wh[wv][d] + ((dr[r] >> 8) & 0x7fffffu)] = key[r];
        }
    __syncthreads();
    const uint32_t tn = (n - tb) < C::TILE ? (n - tb) : C::TILE;
    for (uint32_t q = t; q < tn; q += SF_THREADS) {
        const KeyT<KB> k = s_item[q];
        const uint32_t dst = s_goff[sfs_digit<KB>(k, p, shift, dmask)] + q;
        if (dst < n) // the tile counts and the ranking agree; checked, never trusted
            out[dst] = k;
        else
            atomicOr(err, 1u);
    }
}

// ===========================================================================
// runs and flush
// ===========================================================================
// c more inserts of key: add c to one split of each of its three counters and
// raise each counter to its split's new value
template <int KB>
__device__ static inline void sf_add(uint32_t *__restrict__ split, uint32_t *__restrict__ search, const sf_par &p,
                                     const uint32_t *index, uint64_t key, uint32_t c)
{
    uint32_t b, mu, t[3];
    sf_idx<KB>::get(key, p, b, mu, t);
    const uint32_t ix = index[mu];
    uint32_t *row = split + (size_t)b * p.m * p.times;
    uint32_t *srow = search + (size_t)b * p.m;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const uint32_t cj = (ix >> (8 * j)) & 0xffu;
        const uint32_t v = atomicAdd(row + cj * p.times + t[j], c) + c;
        if (v > srow[cj]) // a plain read is never above the counter's current value
            atomicMax(srow + cj, v);
    }
}

// Wavefront w takes keys [w * per_wave, (w + 1) * per_wave) and walks them
// in steps of 64 (SF_RUN_PF steps loaded ahead).  In a step, a lane heads a
// run when its key differs from its left neighbour's (lane 0 compares with
// the carried run); a run that ends inside the step is flushed by its head,
// the step's last run is carried on (wave-uniform key and length), and the
// carried run is flushed when a step starts with another key and at the end.
#define SF_RUN_PF 8
template <int KB>
__global__ __launch_bounds__(SF_THREADS) void k_sf_runs(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                        uint64_t per_wave, sf_par p,
                                                        const uint32_t *__restrict__ index,
                                                        uint32_t *__restrict__ split, uint32_t *__restrict__ search)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = (uint64_t)blockIdx.x * SF_WAVES + (threadIdx.x >> 6);
    const uint64_t r0 = wave * per_wave;
    if (r0 >= n)
        return;
    const uint64_t r1 = r0 + per_wave < n ? r0 + per_wave : n;
    KeyT<KB> ck = 0;  // carried run: key and length (0: none)
    uint32_t cc = 0;
    for (uint64_t g0 = r0; g0 < r1; g0 += 64u * SF_RUN_PF) {
        KeyT<KB> kk[SF_RUN_PF];
#pragma unroll
        for (int s = 0; s < SF_RUN_PF; ++s) { // every load first, from clamped positions
            const uint64_t i = g0 + (uint64_t)s * 64u + lane;
            kk[s] = __builtin_nontemporal_load(keys + (i < r1 ? i : r1 - 1u));
        }
#pragma unroll
        for (int s = 0; s < SF_RUN_PF; ++s) {
            const uint64_t s0 = g0 + (uint64_t)s * 64u;
            if (s0 >= r1) // uniform
                break;
            const uint32_t nv = (uint32_t)(r1 - s0 < 64u ? r1 - s0 : 64u); // valid lanes: a prefix
            const KeyT<KB> k = kk[s];
            const KeyT<KB> left = __shfl_up(k, 1, 64);
            const bool valid = lane < nv;
            const bool head = valid && (lane == 0 ? !(cc && k == ck) : k != left);
            const uint64_t H = __ballot(head);
            if (head) {
                const uint64_t later = H & ~((2ull << lane) - 1ull);
                if (later) // the run ends inside this step
                    sf_add<KB>(split, search, p, index, (uint64_t)k, (uint32_t)__builtin_ctzll(later) - lane);
            }
            if (H == 0) { // the whole step continues the carried run
                cc += nv;
            } else {
                const uint32_t h0 = (uint32_t)__builtin_ctzll(H);
                if (cc && lane == 0) // the carried run ends before the first head
                    sf_add<KB>(split, search, p, index, (uint64_t)ck, cc + h0);
                const uint32_t hl = 63u - (uint32_t)__builtin_clzll(H); // the step's last run
                ck = __shfl(k, (int)hl, 64);
                cc = nv - hl;
            }
        }
    }
    if (cc && lane == 0)
        sf_add<KB>(split, search, p, index, (uint64_t)ck, cc);
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

// c more inserts of key into a bucket this work-item owns alone: plain
// read-modify-write.  For each of the three counters every split is read, c
// is added to the key's split, and the counter is written as the largest
// split.  Later runs of the bucket that touch the counter write it again, so
// it ends as the largest of its final splits.  T = times when 1 <= times <= 4
// (a counter's splits are then one vector load, aligned to its size), else 0.
template <int T> struct sf_vec;
template <> struct sf_vec<1> { using V = uint32_t; };
template <> struct sf_vec<2> { using V = uint2; };
template <> struct sf_vec<3> { using V = uint3; };
template <> struct sf_vec<4> { using V = uint4; };

template <int T> __device__ static inline uint32_t sf_word(const typename sf_vec<T>::V &v, int u)
{
    if constexpr (T == 1)
        return v;
    else if constexpr (T == 2)
        return u == 0 ? v.x : v.y;
    else if constexpr (T == 3)
        return u == 0 ? v.x : (u == 1 ? v.y : v.z);
    else
        return u == 0 ? v.x : (u == 1 ? v.y : (u == 2 ? v.z : v.w));
}

template <int KB, int T>
__device__ static inline void sf_add_owned(uint32_t *__restrict__ split, uint32_t *__restrict__ search,
                                           const sf_par &p, const uint32_t *index, uint64_t key, uint32_t c)
{
    uint32_t b, mu, t[3];
    sf_idx<KB>::get(key, p, b, mu, t);
    const uint32_t ix = index[mu];
    uint32_t *row = split + (size_t)b * p.m * p.times;
    uint32_t *srow = search + (size_t)b * p.m;
    if constexpr (T > 0) {
        using V = typename sf_vec<T>::V;
        V x[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) // every load in flight together
            x[j] = *reinterpret_cast<const V *>(row + ((ix >> (8 * j)) & 0xffu) * T);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint32_t cj = (ix >> (8 * j)) & 0xffu;
            uint32_t mine = 0, v = 0;
#pragma unroll
            for (int u = 0; u < T; ++u) {
                const uint32_t y = sf_word<T>(x[j], u) + ((uint32_t)u == t[j] ? c : 0u);
                mine = (uint32_t)u == t[j] ? y : mine;
                v = max(v, y);
            }
            row[cj * T + t[j]] = mine;
            srow[cj] = v;
        }
    } else {
        const uint32_t last = p.times - 1u;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint32_t cj = (ix >> (8 * j)) & 0xffu;
            uint32_t *sp = row + cj * p.times;
            const uint32_t mine = sp[t[j]] + c;
            uint32_t v = mine;
            for (uint32_t t0 = 0; t0 < p.times; t0 += 8) {
                uint32_t x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) // clamped positions: every load in flight together
                    x[u] = sp[t0 + u < last ? t0 + u : last];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (t0 + u != t[j])
                        v = max(v, x[u]);
            }
            sp[t[j]] = mine;
            srow[cj] = v;
        }
    }
}

// c more inserts of key into a bucket shared with other wavefronts: atomic
// adds without return; the counters are raised afterwards (k_sf_fix)
template <int KB>
__device__ static inline void sf_add_shared(uint32_t *__restrict__ split, const sf_par &p, const uint32_t *index,
                                            uint64_t key, uint32_t c)
{
    uint32_t b, mu, t[3];
    sf_idx<KB>::get(key, p, b, mu, t);
    const uint32_t ix = index[mu];
    uint32_t *row = split + (size_t)b * p.m * p.times;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        atomicAdd(row + ((ix >> (8 * j)) & 0xffu) * p.times + t[j], c);
}

// every bucket listed (SF_NO_BUCKET: none): each counter becomes the largest
// of its splits (a bucket listed twice is written twice with equal values)
#define SF_NO_BUCKET 0xffffffffu
__global__ __launch_bounds__(256) void k_sf_fix(const uint32_t *__restrict__ list, uint32_t count, sf_par p,
                                                const uint32_t *__restrict__ split, uint32_t *__restrict__ search)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count || list[i] == SF_NO_BUCKET)
        return;
    const uint32_t b = list[i];
    const uint32_t *row = split + (size_t)b * p.m * p.times;
    for (uint32_t c = 0; c < p.m; ++c) {
        uint32_t v = 0;
        for (uint32_t t = 0; t < p.times; ++t)
            v = max(v, row[c * p.times + t]);
        search[(size_t)b * p.m + c] = v;
    }
}

// Keys sorted by bucket, cut into ranges of `range` keys that persistent
// wavefronts take from the counter *work (zero at launch) and walk in steps
// of 64.  Each bucket's keys are one run of the sorted keys, so a bucket that
// lies inside one range belongs to that range's wavefront alone; only the
// range's first and last bucket can be shared with the neighbouring ranges:
// those are added with atomics (sf_add_shared) and listed in shared[2 * range
// + 0/1] for k_sf_fix.
//
// In a step, a lane heads a bucket run when its bucket differs from its left
// neighbour's (lane 0 compares with the carried bucket) and a key run when
// its key differs too; each key run's key and length go to LDS.  A bucket
// run's head lane applies the run's key runs one after the other (adjacent
// runs of one key merge into a pending run, added when the key changes or
// the bucket ends).  The step's last bucket is carried: the same lane (the carrier) applies its
// continuation in the following steps, so one work-item does every plain
// read-modify-write of an owned bucket, in program order, and a hot key's
// run costs one add per wavefront range.
template <int KB, int T>
__global__ __launch_bounds__(SF_THREADS) void k_sf_rows(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                        uint64_t range, sf_par p,
                                                        const uint32_t *__restrict__ index_g,
                                                        uint32_t *__restrict__ split, uint32_t *__restrict__ search,
                                                        uint32_t *__restrict__ work, uint32_t *__restrict__ shared)
{
    __shared__ KeyT<KB> s_k[SF_WAVES][64];
    __shared__ uint32_t s_l[SF_WAVES][64];
    __shared__ uint32_t s_index[QS_MAX_MASKS];
    for (uint32_t e = threadIdx.x; e < p.n_masks; e += SF_THREADS)
        s_index[e] = index_g[e];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint64_t nranges = (n + range - 1) / range;
    for (;;) {
        uint32_t rr = 0;
        if (lane == 0)
            rr = atomicAdd(work, 1u);
        rr = (uint32_t)__shfl((int)rr, 0, 64);
        if (rr >= nranges) // every wavefront leaves once the ranges are taken
            break;
        const uint64_t r0 = (uint64_t)rr * range;
        const uint64_t r1 = r0 + range < n ? r0 + range : n;
        const uint32_t bfirst = sf_idx<KB>::bucket((uint64_t)keys[r0], p);
        const uint32_t blast = sf_idx<KB>::bucket((uint64_t)keys[r1 - 1], p);
        const bool sh_first = r0 > 0 && sf_idx<KB>::bucket((uint64_t)keys[r0 - 1], p) == bfirst;
        const bool sh_last = r1 < n && sf_idx<KB>::bucket((uint64_t)keys[r1], p) == blast;
        if (lane == 0) {
            shared[2u * rr] = sh_first ? bfirst : SF_NO_BUCKET;
            shared[2u * rr + 1u] = sh_last ? blast : SF_NO_BUCKET;
        }
        uint32_t cb = 0, co = 0; // carried bucket and its carrier lane
        bool carry = false;      // wave-uniform
        // this lane's pending key run (merged across steps until the key changes)
        KeyT<KB> pk = 0;
        uint32_t pc = 0;
        bool psh = false;

        auto shared_bucket = [&](uint32_t b) { return (b == bfirst && sh_first) || (b == blast && sh_last); };
        auto flush = [&]() {
            if (psh)
                sf_add_shared<KB>(split, p, s_index, (uint64_t)pk, pc);
            else
                sf_add_owned<KB, T>(split, search, p, s_index, (uint64_t)pk, pc);
            pc = 0;
        };
        // apply the key runs at positions [j, end) of the step to bucket b
        auto apply = [&](uint32_t j, uint32_t end, uint32_t b) {
            const bool sh = shared_bucket(b);
            while (j < end) {
                const KeyT<KB> k = s_k[wv][j];
                const uint32_t len = s_l[wv][j];
                if (pc && k == pk) {
                    pc += len;
                } else {
                    if (pc)
                        flush();
                    pk = k;
                    pc = len;
                    psh = sh;
                }
                j += len;
            }
        };
        // the bucket's last key run
        auto finish = [&]() {
            if (pc)
                flush();
        };

        for (uint64_t g0 = r0; g0 < r1; g0 += 64u * SF_RUN_PF) {
            KeyT<KB> kk[SF_RUN_PF];
#pragma unroll
            for (int s = 0; s < SF_RUN_PF; ++s) { // every load first, from clamped positions
                const uint64_t i = g0 + (uint64_t)s * 64u + lane;
                kk[s] = __builtin_nontemporal_load(keys + (i < r1 ? i : r1 - 1u));
            }
#pragma unroll
            for (int s = 0; s < SF_RUN_PF; ++s) {
                const uint64_t s0 = g0 + (uint64_t)s * 64u;
                if (s0 >= r1) // uniform
                    break;
                const uint32_t nv = (uint32_t)(r1 - s0 < 64u ? r1 - s0 : 64u); // valid lanes: a prefix
                const KeyT<KB> k = kk[s];
                const bool valid = lane < nv;
                const uint32_t b = sf_idx<KB>::bucket((uint64_t)k, p);
                const uint32_t bl = (uint32_t)__shfl_up((int)b, 1, 64);
                const KeyT<KB> kl = __shfl_up(k, 1, 64);
                const bool bhead = valid && (lane == 0 ? !(carry && b == cb) : b != bl);
                const bool khead = valid && (lane == 0 || bhead || k != kl);
                const uint64_t BH = __ballot(bhead);
                const uint64_t KH = __ballot(khead);
                {
                    const uint64_t later = KH & ~((2ull << lane) - 1ull);
                    s_k[wv][lane] = k;
                    s_l[wv][lane] = (later ? (uint32_t)__builtin_ctzll(later) : nv) - lane;
                }
                // the bucket at the step's end, and whether the next step
                // starts with it (then it is carried, else it ends here)
                const uint32_t hl = BH ? 63u - (uint32_t)__builtin_clzll(BH) : co;
                const uint32_t bend = BH ? (uint32_t)__shfl((int)b, (int)hl, 64) : cb;
                bool cont = false;
                if (s0 + 64u < r1) {
                    const KeyT<KB> nk = s + 1 < SF_RUN_PF ? __shfl(kk[s + 1 < SF_RUN_PF ? s + 1 : s], 0, 64)
                                                          : keys[s0 + 64u];
                    cont = sf_idx<KB>::bucket((uint64_t)nk, p) == bend;
                }
                __builtin_amdgcn_wave_barrier();
                const uint32_t first = BH ? (uint32_t)__builtin_ctzll(BH) : nv;
                if (carry && lane == co) { // the carried bucket's continuation: positions [0, first)
                    apply(0u, first, cb);
                    if (BH || !cont)
                        finish();
                }
                if (bhead) {
                    const uint64_t later = BH & ~((2ull << lane) - 1ull);
                    const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : nv;
                    apply(lane, end, b);
                    if (later || !cont)
                        finish();
                }
                co = hl;
                cb = bend;
                carry = cont;
                __builtin_amdgcn_wave_barrier(); // the next step rewrites s_k and s_l
            }
        }
        if (carry && lane == co)
            finish();
    }
}

// The hot keys of a filtered call (qs_hot_split_keys): each slot's count
// summed over the persistent blocks and added with sf_add (after the cold
// keys' kernels, so no plain write of k_sf_rows runs beside these atomics).
// A block takes 64 slots, its four wavefronts a quarter of the rows each.
#define SF_FLUSH_WAVES 16u // wavefronts per block, each summing every 16th row of part
__global__ __launch_bounds__(64 * SF_FLUSH_WAVES) void k_sf_hot_flush(const uint32_t *__restrict__ part,
                                                                      uint32_t nblk,
                                                                      const uint32_t *__restrict__ tkey,
                                                                      uint32_t nslots, sf_par p,
                                                                      const uint32_t *__restrict__ index,
                                                                      uint32_t *__restrict__ split,
                                                                      uint32_t *__restrict__ search)
{
    __shared__ uint32_t s_sum[SF_FLUSH_WAVES][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t sl = blockIdx.x * 64u + lane;
    uint32_t c = 0;
    for (uint32_t b = wv; b < nblk; b += SF_FLUSH_WAVES)
        c += part[(size_t)b * nslots + sl];
    s_sum[wv][lane] = c;
    __syncthreads();
    if (wv != 0)
        return;
    c = 0;
#pragma unroll
    for (uint32_t w = 0; w < SF_FLUSH_WAVES; ++w)
        c += s_sum[w][lane];
    if (c) // a slot without a key counts nothing
        sf_add<4>(split, search, p, index, tkey[sl], c);
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
    (void)hipFree(s->d_buf[0]);
    (void)hipFree(s->d_buf[1]);
    (void)hipFree(s->d_hist);
    (void)hipFree(s->d_shared);
    (void)hipFree(s->d_err);
    qs_hot_state_free(s->hot);
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
    if (n > 0xffffffffull)
        return QS_ERR_PARAM; // bucket indexes are 32-bit (0xffffffff marks "no bucket")
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

// radix scratch for chunks of up to `keys` keys (kept across calls)
static int sf_ensure_scratch(qs_sf *s, uint64_t keys)
{
    if (keys <= s->scratch_keys)
        return QS_OK;
    CK(hipDeviceSynchronize());
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(s->d_buf[b]);
        s->d_buf[b] = nullptr;
    }
    (void)hipFree(s->d_hist);
    (void)hipFree(s->d_shared);
    s->d_hist = nullptr;
    s->d_shared = nullptr;
    s->scratch_keys = 0;
    const uint64_t tile = s->key_bytes == 4 ? sfs_tile<4>::TILE : sfs_tile<8>::TILE;
    const uint64_t ntiles = (keys + tile - 1) / tile;
    if (hipMalloc(&s->d_buf[0], keys * s->key_bytes) != hipSuccess ||
        hipMalloc(&s->d_buf[1], keys * s->key_bytes) != hipSuccess ||
        hipMalloc(&s->d_hist, 4ull * 256 * (ntiles + 1)) != hipSuccess ||
        hipMalloc(&s->d_shared, 8ull * ((keys + SF_ROW_RANGE - 1) / SF_ROW_RANGE)) != hipSuccess)
        return QS_ERR_NOMEM;
    if (!s->d_err) {
        if (hipMalloc(&s->d_err, 8) != hipSuccess)
            return QS_ERR_NOMEM;
        CK(hipMemset(s->d_err, 0, 8));
    }
    s->scratch_keys = keys;
    return QS_OK;
}

// k_sf_runs (owned = false: any order, all atomics) or k_sf_rows (owned =
// true: keys sorted by bucket)
template <int KB>
static void sf_launch_flush(qs_sf *s, const sf_par &p, const KeyT<KB> *keys, uint64_t n, bool owned)
{
    const uint64_t steps = (n + 63) / 64;
    const uint64_t per_wave = (steps + SF_RUN_WAVES - 1) / SF_RUN_WAVES * 64; // whole steps
    const uint64_t waves = (n + per_wave - 1) / per_wave;
    const dim3 grid((uint32_t)((waves + SF_WAVES - 1) / SF_WAVES)), block(SF_THREADS);
    if (owned) {
        const uint64_t ranges = (n + SF_ROW_RANGE - 1) / SF_ROW_RANGE;
        const uint64_t w = ranges < SF_RUN_WAVES ? ranges : SF_RUN_WAVES;
        (void)hipMemsetAsync(s->d_err + 1, 0, 4, 0);
        const dim3 g((uint32_t)((w + SF_WAVES - 1) / SF_WAVES));
#define SF_ROWS(T)                                                                                               \
    hipLaunchKernelGGL((k_sf_rows<KB, T>), g, block, 0, 0, keys, n, (uint64_t)SF_ROW_RANGE, p, s->d_index,       \
                       s->d_split, s->d_search, s->d_err + 1, s->d_shared)
        switch (s->times) {
        case 1: SF_ROWS(1); break;
        case 2: SF_ROWS(2); break;
        case 3: SF_ROWS(3); break;
        case 4: SF_ROWS(4); break;
        default: SF_ROWS(0); break;
        }
#undef SF_ROWS
        const uint32_t listed = (uint32_t)(2 * ranges);
        hipLaunchKernelGGL(k_sf_fix, dim3((listed + 255) / 256), dim3(256), 0, 0, s->d_shared, listed, p, s->d_split,
                           s->d_search);
    } else
        hipLaunchKernelGGL(k_sf_runs<KB>, grid, block, 0, 0, keys, n, per_wave, p, s->d_index, s->d_split,
                           s->d_search);
}

// one chunk: sort by bucket (8-bit digits at most), then flush
template <int KB> static int sf_insert_chunk(qs_sf *s, const sf_par &p, const KeyT<KB> *keys, uint32_t cn);

// One chunk of 4-byte keys through the hot-key filter: the frequent keys are
// counted in LDS and added once each (k_sf_hot_flush), only the cold keys
// are sorted.  *done = 0 when the chunk is short or not skewed.
static int sf_insert_hot(qs_sf *s, const sf_par &p, const uint32_t *keys, uint32_t cn, int *done)
{
    *done = 0;
    qs_hot_split hs;
    int rc = qs_hot_split_keys(&s->hot, keys, cn, 0, &hs);
    if (rc || !hs.split)
        return rc;
    *done = 1;
    if (hs.ncold > cn)
        return QS_ERR_INTERNAL;
    if (hs.ncold) {
        rc = sf_insert_chunk<4>(s, p, hs.cold, (uint32_t)hs.ncold);
        if (rc)
            return rc;
    }
    hipLaunchKernelGGL(k_sf_hot_flush, dim3(hs.nslots / 64u), dim3(64 * SF_FLUSH_WAVES), 0, 0, hs.part, hs.nblk,
                       hs.tkey, hs.nslots, p, s->d_index, s->d_split, s->d_search);
    CK(hipGetLastError());
    return QS_OK;
}

template <int KB> static int sf_insert_chunk(qs_sf *s, const sf_par &p, const KeyT<KB> *keys, uint32_t cn)
{
    const int bits = s->n > 1 ? 64 - __builtin_clzll(s->n - 1) : 0; // bits of the largest bucket
    if (cn < SF_SORT_MIN || bits == 0) { // unsorted (one bucket: nothing to sort)
        sf_launch_flush<KB>(s, p, keys, cn, false);
        CK(hipGetLastError());
        return QS_OK;
    }
    const int passes = (bits + 7) / 8;
    const int dbits = (bits + passes - 1) / passes;
    const uint32_t dmask = (1u << dbits) - 1u;
    const uint32_t ntiles = (cn + sfs_tile<KB>::TILE - 1) / sfs_tile<KB>::TILE;
    uint32_t *totals = s->d_hist + (size_t)256 * ntiles;
    const KeyT<KB> *in = keys;
    for (int ps = 0; ps < passes; ++ps) {
        KeyT<KB> *out = static_cast<KeyT<KB> *>(s->d_buf[ps & 1]);
        hipLaunchKernelGGL(k_sfs_up<KB>, dim3(ntiles), dim3(SF_THREADS), 0, 0, in, cn, p, dbits * ps, dmask,
                           ntiles, s->d_hist);
        qs_scan_digits(s->d_hist, ntiles, dmask + 1u, totals, 0);
        hipLaunchKernelGGL(k_sfs_down<KB>, dim3(ntiles), dim3(SF_THREADS), 0, 0, in, cn, p, dbits * ps, dmask,
                           ntiles, s->d_hist, totals, out, s->d_err);
        in = out;
    }
    sf_launch_flush<KB>(s, p, in, cn, true);
    CK(hipGetLastError());
    return QS_OK;
}

extern "C" int qs_sf_insert(qs_sf *s, const void *d_keys, size_t n)
{
    if (!s || (!d_keys && n))
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK; // the counters already hold the largest of their splits
    const sf_par p = sf_par_of(s);
    const uint64_t chunk = s->key_bytes == 4 ? SF_CHUNK4 : SF_CHUNK8;
    if (n >= SF_SORT_MIN) {
        const int rc = sf_ensure_scratch(s, n < chunk ? n : chunk);
        if (rc != QS_OK)
            return rc;
    }
    for (uint64_t off = 0; off < n; off += chunk) {
        const uint32_t cn = (uint32_t)(n - off < chunk ? n - off : chunk);
        int rc = QS_OK, done = 0;
        if (s->key_bytes == 4) {
            rc = sf_insert_hot(s, p, static_cast<const uint32_t *>(d_keys) + off, cn, &done);
            if (rc == QS_OK && !done)
                rc = sf_insert_chunk<4>(s, p, static_cast<const uint32_t *>(d_keys) + off, cn);
        } else {
            rc = sf_insert_chunk<8>(s, p, static_cast<const uint64_t *>(d_keys) + off, cn);
        }
        if (rc != QS_OK)
            return rc;
    }
    CK(hipDeviceSynchronize());
    if (n >= SF_SORT_MIN) {
        uint32_t err = 0;
        CK(hipMemcpy(&err, s->d_err, 4, hipMemcpyDeviceToHost));
        if (err)
            return QS_ERR_INTERNAL;
    }
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
    if (s->d_err)
        CK(hipMemset(s->d_err, 0, 4));
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
