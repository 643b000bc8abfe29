// qs_insert.hip -- deterministic MSWS insert for gfx950: sort by bucket, then
// every bucket segment has exactly one owner.
//
// Replaces insert_kernel / calculation_s2_insert / atomic_malloc
// (include/sketch/device_sketch/device_sketch_v3.h:68-290).  The reference
// lets warps race on a bucket (a stale read decides the upgrade, v3:160-161,
// 263-264); here the result equals the sequential stream-order semantics:
//
//   c_b[j] += 1 for j in M   while max_{j in M} c_b[j] <= 128 and b unlinked
//   the first insert that sees a selected byte > 128 links b to the next high
//   id (ids in stream order) and adds to g[q(j)]; later inserts add to g[j].
//
// Why sort: on MI355X a random access costs one DRAM activation whatever its
// size, and atomics on one 128-byte line serialise at ~11 ns.  An atomic per
// key is therefore bounded far below the HBM roofline, and a Zipf-hot key turns
// its three counters into a serial queue.  A stable LSD radix sort of the
// chunk by bucket (the bucket is recomputed from the key in every pass, the
// chunk-local stream index rides along) turns the table update into an
// address-ordered sweep with one plain read-modify-write per touched bucket.
//
// Per chunk:
//   sort      4-8 radix passes of (key, index) by bucket digit (8 bits/pass)
//   k_seg     one wavefront per window of 64 sorted items: a segmented scan
//             sums the packed byte increments of each bucket run; the run's
//             last lane owns it.  Linked bucket: add to the high row.  Unlinked
//             and no selected byte would exceed 129: add to the low bytes (no
//             allocation can happen, so order does not matter).  Otherwise the
//             segment goes to the in-order replay.  Runs longer than the window
//             plus one extension window are "long": linked ones go to the tail
//             kernel, unlinked ones to the replay first.
//   k_ordered one wavefront per queued segment replays it in stream order until
//             its allocation (at most ceil(28*129/3) inserts can precede it).
//   k_assign  allocation ids in stream order of the triggering inserts.
//   k_tail    long segments after their allocation point: per-counter ballot
//             popcounts over 16K-item windows, one atomic per counter per
//             window and bucket.
#include "qs_internal.h"

#include <cstring>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess)                                                                   \
            return QS_ERR_HIP;                                                                   \
    } while (0)

#define BS_THREADS 256
#define BS_ITEMS 16
#define BS_TILE (BS_THREADS * BS_ITEMS)
#define BS_DBITS 8
#define BS_BINS (1u << BS_DBITS)
#define SEG_THREADS 256
#define TAIL_WINDOW 16384u
#define NO_BUCKET 0xffffffffu
#define END_BUCKET 0xfffffffeu

static int g_cu = 0;

static uint32_t grid_cap(uint64_t want, uint32_t per_cu)
{
    if (g_cu == 0) {
        int dev = 0;
        hipDeviceProp_t pr;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess)
            g_cu = pr.multiProcessorCount;
        if (g_cu <= 0)
            g_cu = 256;
    }
    uint64_t cap = (uint64_t)g_cu * per_cu;
    if (want > cap)
        want = cap;
    return want ? (uint32_t)want : 1u;
}

// ===========================================================================
// radix sort of the chunk by bucket
// ===========================================================================
template <int KB>
__global__ __launch_bounds__(BS_THREADS) void k_bsort_up(const KeyT<KB> *__restrict__ kin, uint32_t n,
                                                         qs_hashpar hp, int shift, uint32_t ntiles,
                                                         uint32_t *__restrict__ hist)
{
    __shared__ uint32_t s_h[BS_BINS];
    for (uint32_t d = threadIdx.x; d < BS_BINS; d += BS_THREADS)
        s_h[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * BS_TILE;
#pragma unroll 4
    for (int r = 0; r < BS_ITEMS; ++r) {
        uint32_t i = base + r * BS_THREADS + threadIdx.x;
        if (i < n)
            atomicAdd(&s_h[(qs_bucket(kin[i], hp) >> shift) & (BS_BINS - 1u)], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < BS_BINS; d += BS_THREADS)
        hist[(size_t)d * ntiles + blockIdx.x] = s_h[d];
}

// Stable scatter of one tile: rank every item within the tile by digit
// (per-wave digit match with BS_DBITS ballots, per-digit prefix over the
// waves and the rounds), stage the digit-sorted tile in LDS, then write each
// digit run contiguously to its global position.
template <int KB>
__global__ __launch_bounds__(BS_THREADS) void k_bsort_down(
    const KeyT<KB> *__restrict__ kin, const uint32_t *__restrict__ iin, uint32_t n, qs_hashpar hp,
    int shift, uint32_t ntiles, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ totals, KeyT<KB> *__restrict__ kout, uint32_t *__restrict__ iout)
{
    __shared__ KeyT<KB> s_key[BS_TILE];
    __shared__ uint32_t s_idx[BS_TILE];
    __shared__ uint32_t s_off[BS_BINS];   // global position of local slot 0 of digit d
    __shared__ uint32_t s_run[BS_BINS];   // next local slot of digit d
    __shared__ uint32_t s_wc[4][BS_BINS]; // per-wave digit counts of a round
    __shared__ uint32_t s_wo[4][BS_BINS]; // per-wave digit offsets of a round
    __shared__ uint32_t s_red[2][4];
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * BS_TILE;

    // digit d = t (BS_BINS == BS_THREADS): global base and local start
    {
        uint32_t tot = totals[t];
        uint32_t hcur = hist[(size_t)t * ntiles + tile];
        uint32_t hnxt = (tile + 1 < ntiles) ? hist[(size_t)t * ntiles + tile + 1] : tot;
        uint32_t cnt = hnxt - hcur;
        uint32_t xa = tot, xb = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t ya = __shfl_up(xa, o, 64), yb = __shfl_up(xb, o, 64);
            if (lane >= (uint32_t)o) {
                xa += ya;
                xb += yb;
            }
        }
        if (lane == 63) {
            s_red[0][wv] = xa;
            s_red[1][wv] = xb;
        }
        __syncthreads();
        uint32_t oa = 0, ob = 0;
        for (uint32_t k = 0; k < wv; ++k) {
            oa += s_red[0][k];
            ob += s_red[1][k];
        }
        uint32_t dexcl = oa + xa - tot;  // digits before t, all tiles
        uint32_t lstart = ob + xb - cnt; // digits before t, this tile
        s_off[t] = dexcl + hcur - lstart;
        s_run[t] = lstart;
#pragma unroll
        for (int w = 0; w < 4; ++w)
            s_wc[w][t] = 0;
    }
    __syncthreads();

    const uint64_t lt = (1ull << lane) - 1ull;
    for (int r = 0; r < BS_ITEMS; ++r) {
        uint32_t i = base + r * BS_THREADS + t;
        bool valid = i < n;
        KeyT<KB> key = valid ? kin[i] : (KeyT<KB>)0;
        uint32_t idx = valid ? (iin ? iin[i] : i) : 0u;
        uint32_t d = (qs_bucket(key, hp) >> shift) & (BS_BINS - 1u);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < BS_DBITS; ++bit) {
            uint64_t bb = __ballot((d >> bit) & 1u);
            m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0)
            s_wc[wv][d] = (uint32_t)__popcll(m);
        __syncthreads();
        {
            uint32_t run = s_run[t];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                uint32_t c = s_wc[w][t];
                s_wo[w][t] = run;
                s_wc[w][t] = 0;
                run += c;
            }
            s_run[t] = run;
        }
        __syncthreads();
        if (valid) {
            uint32_t lpos = s_wo[wv][d] + rank;
            s_key[lpos] = key;
            s_idx[lpos] = idx;
        }
    }
    __syncthreads();
    const uint32_t tn = (n - base) < BS_TILE ? (n - base) : BS_TILE;
    for (uint32_t q = t; q < tn; q += BS_THREADS) {
        KeyT<KB> key = s_key[q];
        uint32_t d = (qs_bucket(key, hp) >> shift) & (BS_BINS - 1u);
        uint32_t g = s_off[d] + q;
        kout[g] = key;
        iout[g] = s_idx[q];
    }
}

// ===========================================================================
// per-segment application
// ===========================================================================

// packed per-byte increments of one 3-bit mask into 7 words
__device__ static inline void mask_to_words(uint32_t M, uint32_t a[7])
{
#pragma unroll
    for (int k = 0; k < 7; ++k)
        a[k] = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        if (M) {
            uint32_t j = (uint32_t)__builtin_ctz(M);
            M &= M - 1u;
            uint32_t inc = 1u << ((j & 3u) * 8u);
#pragma unroll
            for (int k = 0; k < 7; ++k)
                a[k] += ((j >> 2) == (uint32_t)k) ? inc : 0u;
        }
    }
}

__device__ static inline uint32_t sel7(const uint32_t a[7], uint32_t k)
{
    uint32_t r = a[0];
#pragma unroll
    for (int q = 1; q < 7; ++q)
        r = (k == (uint32_t)q) ? a[q] : r;
    return r;
}

// Owner of a complete segment (packed per-byte counts a[], each <= 128):
// returns true when the segment must go to the in-order replay.
__device__ static inline bool apply_segment(uint32_t *__restrict__ low, uint32_t *__restrict__ high,
                                            uint32_t b, const uint32_t a[7])
{
    uint4 *bp = reinterpret_cast<uint4 *>(low + (size_t)b * 8u);
    uint4 x = bp[0], y = bp[1];
    uint32_t W[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
    uint32_t w = W[7];
    if (w > QS_LINK_START) {
        uint32_t *row = high + (size_t)(w - QS_LINK_START) * 32u;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            if (a[k]) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    uint32_t c = (a[k] >> (8 * e)) & 0xffu;
                    if (c)
                        row[4 * k + e] += c;
                }
            }
        }
        return false;
    }
    bool risky = false;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint32_t c = (a[k] >> (8 * e)) & 0xffu;
            uint32_t v = (W[k] >> (8 * e)) & 0xffu;
            risky |= (c != 0u) && (v + c > QS_LOW_LIMIT + 1u);
        }
    }
    if (risky)
        return true;
    x.x += a[0];
    x.y += a[1];
    x.z += a[2];
    x.w += a[3];
    y.x += a[4];
    y.y += a[5];
    y.z += a[6];
    bp[0] = x;
    bp[1] = y;
    return false;
}

// Append `flag` lanes' values to a global list, one atomic per wavefront.
__device__ static inline void wave_append(bool flag, uint32_t value, uint32_t *list, uint32_t *count)
{
    uint64_t m = __ballot(flag);
    if (!m)
        return;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t leader = (uint32_t)__builtin_ctzll(m);
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader, 64);
    if (flag)
        list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = value;
}

template <int KB>
__global__ __launch_bounds__(SEG_THREADS) void k_seg(
    const KeyT<KB> *__restrict__ K, uint32_t n, uint32_t *__restrict__ low,
    uint32_t *__restrict__ high, const uint32_t *__restrict__ masks, qs_hashpar hp,
    uint32_t *__restrict__ longbm, uint32_t *__restrict__ ord, uint32_t *__restrict__ lng,
    qs_dev_state *__restrict__ st)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t i = threadIdx.x; i < hp.n_masks; i += SEG_THREADS)
        s_masks[i] = masks[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t le = (lane == 63u) ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t gw = blockIdx.x * (SEG_THREADS / 64u) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (SEG_THREADS / 64u);
    for (uint64_t w0 = (uint64_t)gw * 64u; w0 < n; w0 += (uint64_t)nw * 64u) {
        const uint32_t p = (uint32_t)w0 + lane;
        const bool valid = p < n;
        KeyT<KB> key = valid ? K[p] : (KeyT<KB>)0;
        uint32_t b = valid ? qs_bucket(key, hp) : NO_BUCKET;
        uint32_t M = valid ? s_masks[qs_maskidx(key, hp)] : 0u;
        uint32_t bl = __shfl(b, (int)(lane == 0 ? 0 : lane - 1u), 64);
        uint32_t br = __shfl(b, (int)(lane == 63 ? 63 : lane + 1u), 64);
        if (lane == 0)
            bl = (w0 > 0) ? qs_bucket(K[w0 - 1], hp) : NO_BUCKET;
        if (lane == 63)
            br = (p + 1u < n) ? qs_bucket(K[p + 1u], hp) : END_BUCKET;
        const bool head = valid && b != bl;
        const bool last = valid && b != br;
        uint32_t a[7];
        mask_to_words(M, a);
        // segmented inclusive scan of the packed increments over runs
        bool f = head || lane == 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            bool fu = __shfl_up((int)f, o, 64) != 0;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                uint32_t y = __shfl_up(a[k], o, 64);
                if (lane >= (uint32_t)o && !f)
                    a[k] += y;
            }
            if (lane >= (uint32_t)o)
                f = f || fu;
        }
        const uint64_t starts = __ballot(head || lane == 0);
        const uint64_t heads = __ballot(head);
        const uint32_t s = 63u - (uint32_t)__builtin_clzll(starts & le); // first lane of my run
        const bool owned = valid && (s > 0u || (heads & 1ull));
        // complete runs owned by this window
        bool risky = false;
        if (owned && last)
            risky = apply_segment(low, high, b, a);
        wave_append(risky, (uint32_t)w0 + s, ord, &st->n_ord);
        // the last run continues past the window: look one window further
        const bool ext = __shfl((int)(owned && !last), 63, 64) != 0;
        if (ext) {
            const uint32_t b63 = __shfl(b, 63, 64);
            const uint32_t s63 = __shfl(s, 63, 64);
            uint32_t e[7];
#pragma unroll
            for (int k = 0; k < 7; ++k)
                e[k] = __shfl(a[k], 63, 64);
            const uint32_t q = (uint32_t)w0 + 64u + lane;
            bool in = false;
            uint32_t Mq = 0;
            if (q < n) {
                KeyT<KB> kq = K[q];
                in = qs_bucket(kq, hp) == b63;
                Mq = in ? s_masks[qs_maskidx(kq, hp)] : 0u;
            }
            const uint64_t im = __ballot(in);
            const uint32_t head_pos = (uint32_t)w0 + s63;
            // the run is a prefix of the next window; its end is inside iff not full
            if (im != ~0ull) {
                uint32_t c[7];
                mask_to_words(Mq, c);
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    uint32_t v = c[k];
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1)
                        v += __shfl_xor(v, o, 64);
                    e[k] += v;
                }
                bool r2 = false;
                if (lane == 0)
                    r2 = apply_segment(low, high, b63, e);
                wave_append(r2, head_pos, ord, &st->n_ord);
            } else {
                // long segment: tail kernel (after the in-order replay if unlinked)
                bool unl = false;
                if (lane == 0) {
                    atomicOr(longbm + (b63 >> 5), 1u << (b63 & 31u));
                    unl = low[(size_t)b63 * 8u + 7u] <= QS_LINK_START;
                }
                wave_append(lane == 0, head_pos, lng, &st->n_long);
                wave_append(unl, head_pos, ord, &st->n_ord);
            }
        }
    }
}

__device__ static inline uint32_t kth_set_bit(uint64_t m, uint32_t k)
{
    // position of the k-th (0-based) set bit of m; caller guarantees popc(m) > k
    uint32_t pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        uint64_t lowm = (1ull << w) - 1ull;
        uint32_t c = (uint32_t)__popcll(m & lowm);
        if (k >= c) {
            k -= c;
            m >>= w;
            pos += (uint32_t)w;
        }
    }
    return pos;
}

// In-order replay of one queued segment per wavefront, up to its allocation.
// Lane j (< 28) holds byte counter j; a batch of 64 inserts is transposed into
// 28 ballots B_j, and the first insert that would see a selected byte > 128 is
// the (129 - c_j)-th set bit of B_j, minimised over j.
template <int KB>
__global__ __launch_bounds__(256) void k_ordered(const KeyT<KB> *__restrict__ K,
                                                 const uint32_t *__restrict__ I, uint32_t n,
                                                 const uint32_t *__restrict__ ord,
                                                 uint32_t *__restrict__ low,
                                                 const uint32_t *__restrict__ masks, qs_hashpar hp,
                                                 uint32_t *__restrict__ longbm,
                                                 uint32_t *__restrict__ lng,
                                                 qs_trig *__restrict__ trig, uint32_t trig_cap,
                                                 qs_dev_state *__restrict__ st, int paper)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (blockDim.x / 64u);
    const uint32_t n_ord = st->n_ord;
    for (uint32_t h = gw; h < n_ord; h += nw) {
        const uint32_t p0 = ord[h];
        const uint32_t b = qs_bucket(K[p0], hp);
        uint32_t *bk = low + (size_t)b * 8u;
        uint32_t word = lane < 7u ? bk[lane] : 0u;
        uint32_t c = __shfl(word, (int)(lane >> 2), 64);
        c = lane < 28u ? ((c >> ((lane & 3u) * 8u)) & 0xffu) : 0u;
        uint32_t hc = 0;
        bool fired = false, more = false;
        uint32_t t_idx = 0, tail_start = 0;
        for (uint32_t p = p0;; p += 64u) {
            const uint32_t q = p + lane;
            bool valid = false;
            uint32_t M = 0, idx = 0;
            if (q < n) {
                KeyT<KB> kq = K[q];
                valid = qs_bucket(kq, hp) == b;
                if (valid) {
                    M = masks[qs_maskidx(kq, hp)];
                    idx = I[q];
                }
            }
            const uint64_t vm = __ballot(valid);
            if (vm == 0)
                break;
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
                if (vm != ~0ull)
                    break;
                continue;
            }
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
            more = (vm == ~0ull);
            break;
        }
        uint32_t w4 = (uint32_t)__shfl(c, (int)(4u * lane) & 63, 64) |
                      ((uint32_t)__shfl(c, (int)(4u * lane + 1u) & 63, 64) << 8) |
                      ((uint32_t)__shfl(c, (int)(4u * lane + 2u) & 63, 64) << 16) |
                      ((uint32_t)__shfl(c, (int)(4u * lane + 3u) & 63, 64) << 24);
        if (lane < 7u)
            bk[lane] = w4;
        if (fired) {
            uint32_t slot = 0;
            if (lane == 0) {
                slot = atomicAdd(&st->n_trig, 1u);
                // the segment continues after the batch of its allocation: make
                // sure the tail kernel sees it
                if (more && tail_start < n) {
                    uint32_t old = atomicOr(longbm + (b >> 5), 1u << (b & 31u));
                    if (!(old & (1u << (b & 31u))))
                        lng[atomicAdd(&st->n_long, 1u)] = p0;
                }
            }
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

// Ids in stream order of the triggering inserts (the sequential bump allocator
// of v3:68-85, next starting at 1025).
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
        const uint32_t id = base + r;
        const uint32_t row = id - QS_LINK_START;
        if (lane < 32u)
            high[(size_t)row * 32u + lane] = tr->hc[lane];
        if (lane == 0) {
            low[(size_t)tr->bucket * 8u + 7u] = id;
            tail_of[row] = tr->tail_start;
        }
    }
}

// Long segments: every insert of a bucket linked before this chunk, and the
// inserts after the allocation batch of a bucket linked in this chunk, are
// commutative high adds.  A wavefront walks a window of TAIL_WINDOW sorted
// positions and keeps per-counter sums (lane j = counter j) while the bucket
// stays the same, flushing them with one atomic per counter.
template <int KB>
__global__ __launch_bounds__(256) void k_tail(const KeyT<KB> *__restrict__ K, uint32_t n,
                                              const uint32_t *__restrict__ low,
                                              uint32_t *__restrict__ high,
                                              const uint32_t *__restrict__ masks, qs_hashpar hp,
                                              const uint32_t *__restrict__ longbm,
                                              const uint32_t *__restrict__ tail_of,
                                              const qs_dev_state *__restrict__ st)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * 4u;
    const uint32_t first_new = st->next;
    for (uint64_t w0 = (uint64_t)gw * TAIL_WINDOW; w0 < n; w0 += (uint64_t)nw * TAIL_WINDOW) {
        const uint64_t wend = (w0 + TAIL_WINDOW < n) ? w0 + TAIL_WINDOW : n;
        uint32_t cur_w = 0, acc = 0;
        for (uint64_t s0 = w0; s0 < wend; s0 += 64u) {
            const uint32_t p = (uint32_t)s0 + lane;
            bool act = false;
            uint32_t b = NO_BUCKET, w = 0, M = 0;
            if (p < wend) {
                KeyT<KB> kp = K[p];
                b = qs_bucket(kp, hp);
                if ((longbm[b >> 5] >> (b & 31u)) & 1u) {
                    w = low[(size_t)b * 8u + 7u];
                    act = w > QS_LINK_START && (w < first_new || p >= tail_of[w - QS_LINK_START]);
                    if (act)
                        M = masks[qs_maskidx(kp, hp)];
                }
            }
            uint64_t am = __ballot(act);
            while (am) {
                const uint32_t lead = (uint32_t)__builtin_ctzll(am);
                const uint32_t wl = __shfl(w, (int)lead, 64);
                const bool in = act && w == wl;
                const uint64_t grp = __ballot(in);
                if (wl != cur_w) {
                    if (cur_w && lane < 28u && acc)
                        atomicAdd(high + (size_t)(cur_w - QS_LINK_START) * 32u + lane, acc);
                    cur_w = wl;
                    acc = 0;
                }
#pragma unroll
                for (int j = 0; j < 28; ++j) {
                    uint64_t bal = __ballot(in && ((M >> j) & 1u));
                    if (lane == (uint32_t)j)
                        acc += (uint32_t)__popcll(bal);
                }
                am &= ~grp;
            }
        }
        if (cur_w && lane < 28u && acc)
            atomicAdd(high + (size_t)(cur_w - QS_LINK_START) * 32u + lane, acc);
    }
}

__global__ void k_clear_long(const uint32_t *__restrict__ lng, const uint32_t *__restrict__ K32,
                             const uint64_t *__restrict__ K64, qs_hashpar hp,
                             uint32_t *__restrict__ longbm, const qs_dev_state *__restrict__ st)
{
    const uint32_t n_long = st->n_long;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_long; i += gridDim.x * blockDim.x) {
        uint32_t p = lng[i];
        uint32_t b = K32 ? qs_bucket(K32[p], hp) : qs_bucket(K64[p], hp);
        atomicAnd(longbm + (b >> 5), ~(1u << (b & 31u)));
    }
}

__global__ void k_finalize(qs_dev_state *st, uint32_t n_trig)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (st->err == 0)
            st->next += n_trig;
        st->n_ord = 0;
        st->n_long = 0;
        st->n_trig = 0;
    }
}

// ===========================================================================
// host side
// ===========================================================================
uint64_t qs_scratch_hist_words(uint64_t chunk)
{
    uint64_t tiles = (chunk + BS_TILE - 1) / BS_TILE;
    if (tiles == 0)
        tiles = 1;
    uint64_t a = BS_BINS * tiles + BS_BINS;
    uint64_t b = qs_sort_hist_words((uint32_t)(chunk < 0xffffffffull ? chunk : 0xffffffffull));
    return a > b ? a : b;
}

void qs_free_scratch(qs_sketch *s)
{
    for (int k = 0; k < 2; ++k) {
        hipFree(s->d_k[k]);
        hipFree(s->d_i[k]);
        hipFree(s->d_tk[k]);
        hipFree(s->d_tv[k]);
        s->d_k[k] = nullptr;
        s->d_i[k] = s->d_tk[k] = s->d_tv[k] = nullptr;
    }
    hipFree(s->d_hist);
    hipFree(s->d_ord);
    hipFree(s->d_long);
    hipFree(s->d_trig);
    s->d_hist = s->d_ord = s->d_long = nullptr;
    s->d_trig = nullptr;
    s->chunk = 0;
}

int qs_ensure_scratch(qs_sketch *s, uint64_t chunk)
{
    if (s->chunk >= chunk)
        return QS_OK;
    qs_free_scratch(s);
    const size_t kb = s->p.key_bytes;
    uint64_t tcap = s->p.n_high < chunk ? s->p.n_high : chunk;
    s->trig_cap = (uint32_t)tcap;
    for (int k = 0; k < 2; ++k) {
        if (hipMalloc(&s->d_k[k], kb * chunk) != hipSuccess ||
            hipMalloc(&s->d_i[k], sizeof(uint32_t) * chunk) != hipSuccess ||
            hipMalloc(&s->d_tk[k], sizeof(uint32_t) * tcap) != hipSuccess ||
            hipMalloc(&s->d_tv[k], sizeof(uint32_t) * tcap) != hipSuccess)
            return QS_ERR_NOMEM;
    }
    if (hipMalloc(&s->d_hist, sizeof(uint32_t) * qs_scratch_hist_words(chunk)) != hipSuccess ||
        hipMalloc(&s->d_ord, sizeof(uint32_t) * chunk) != hipSuccess ||
        hipMalloc(&s->d_long, sizeof(uint32_t) * (chunk / 64 + 64)) != hipSuccess ||
        hipMalloc(&s->d_trig, sizeof(qs_trig) * tcap) != hipSuccess)
        return QS_ERR_NOMEM;
    s->chunk = chunk;
    return QS_OK;
}

template <int KB>
static int sort_chunk(qs_sketch *s, const KeyT<KB> *keys, uint32_t cn, hipStream_t stream, int &res)
{
    const uint32_t ntiles = (cn + BS_TILE - 1) / BS_TILE;
    uint32_t *totals = s->d_hist + (size_t)BS_BINS * ntiles;
    int passes = (s->bucket_bits + BS_DBITS - 1) / BS_DBITS;
    if (passes < 1)
        passes = 1; // n_low == 1: one pass still yields the (identity) index array
    KeyT<KB> *kb[2] = {(KeyT<KB> *)s->d_k[0], (KeyT<KB> *)s->d_k[1]};
    const KeyT<KB> *kin = keys;
    const uint32_t *iin = nullptr;
    int out = 0;
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = BS_DBITS * ps;
        qs_evpair e0, e1, e2;
        qs_tim_begin(s, e0, QS_TK_SORT_UP, cn, stream);
        hipLaunchKernelGGL(k_bsort_up<KB>, dim3(ntiles), dim3(BS_THREADS), 0, stream, kin, cn, s->hp,
                           shift, ntiles, s->d_hist);
        qs_tim_end(s, e0, stream);
        qs_tim_begin(s, e1, QS_TK_SORT_SCAN, cn, stream);
        qs_scan_digits(s->d_hist, ntiles, BS_BINS, totals, stream);
        qs_tim_end(s, e1, stream);
        qs_tim_begin(s, e2, QS_TK_SORT_DOWN, cn, stream);
        hipLaunchKernelGGL(k_bsort_down<KB>, dim3(ntiles), dim3(BS_THREADS), 0, stream, kin, iin, cn,
                           s->hp, shift, ntiles, s->d_hist, totals, kb[out], s->d_i[out]);
        qs_tim_end(s, e2, stream);
        kin = kb[out];
        iin = s->d_i[out];
        out ^= 1;
    }
    CK(hipGetLastError());
    res = out ^ 1;
    return QS_OK;
}

static int read_state(qs_sketch *s, hipStream_t stream)
{
    CK(hipMemcpyAsync(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    return QS_OK;
}

template <int KB>
static int insert_chunk(qs_sketch *s, const KeyT<KB> *keys, uint32_t cn, hipStream_t stream)
{
    int res = 0, rc;
    rc = sort_chunk<KB>(s, keys, cn, stream, res);
    if (rc)
        return rc;
    const KeyT<KB> *K = (const KeyT<KB> *)s->d_k[res];
    const uint32_t *I = s->d_i[res];
    qs_evpair es;
    qs_tim_begin(s, es, QS_TK_SEG, cn, stream);
    hipLaunchKernelGGL(k_seg<KB>, dim3(grid_cap(((uint64_t)cn + 255) / 256, 8)), dim3(SEG_THREADS), 0,
                       stream, K, cn, s->d_low, s->d_high, s->d_masks, s->hp, s->d_longbm, s->d_ord,
                       s->d_long, s->d_st);
    qs_tim_end(s, es, stream);
    CK(hipGetLastError());
    rc = read_state(s, stream);
    if (rc)
        return rc;
    uint32_t n_ord = s->h_st->n_ord;
    uint32_t n_trig = 0;
    if (n_ord) {
        s->chunks_fix++;
        qs_evpair eo;
        qs_tim_begin(s, eo, QS_TK_ORDERED, cn, stream);
        hipLaunchKernelGGL(k_ordered<KB>, dim3(grid_cap(((uint64_t)n_ord + 3) / 4, 8)), dim3(256), 0,
                           stream, K, I, cn, s->d_ord, s->d_low, s->d_masks, s->hp, s->d_longbm,
                           s->d_long, s->d_trig, s->trig_cap, s->d_st,
                           s->p.semantics == QS_PAPER ? 1 : 0);
        qs_tim_end(s, eo, stream);
        CK(hipGetLastError());
        rc = read_state(s, stream);
        if (rc)
            return rc;
        n_trig = s->h_st->n_trig;
        if (n_trig > s->trig_cap)
            n_trig = s->trig_cap;
        if (n_trig) {
            qs_evpair ea;
            qs_tim_begin(s, ea, QS_TK_ASSIGN, cn, stream);
            hipLaunchKernelGGL(k_trig_keys, dim3((n_trig + 255) / 256), dim3(256), 0, stream, s->d_trig,
                               n_trig, s->d_tk[0], s->d_tv[0]);
            int tbits = 32 - __builtin_clz(cn > 1 ? cn - 1 : 1);
            int tc = qs_sort_pairs(s->d_tk, s->d_tv, s->d_hist, n_trig, tbits, stream);
            if (tc < 0)
                return QS_ERR_HIP;
            hipLaunchKernelGGL(k_assign, dim3(grid_cap(((uint64_t)n_trig + 3) / 4, 8)), dim3(256), 0,
                               stream, s->d_trig, s->d_tv[tc], n_trig, s->d_low, s->d_high,
                               s->d_tail_of, (uint64_t)s->p.n_high, s->d_st);
            qs_tim_end(s, ea, stream);
        }
    }
    if (s->h_st->n_long) {
        qs_evpair et;
        qs_tim_begin(s, et, QS_TK_TAIL, cn, stream);
        hipLaunchKernelGGL(k_tail<KB>, dim3(grid_cap(((uint64_t)cn + TAIL_WINDOW * 4 - 1) / (TAIL_WINDOW * 4), 8)),
                           dim3(256), 0, stream, K, cn, s->d_low, s->d_high, s->d_masks, s->hp,
                           s->d_longbm, s->d_tail_of, s->d_st);
        hipLaunchKernelGGL(k_clear_long, dim3(grid_cap((s->h_st->n_long + 255) / 256, 4)), dim3(256), 0,
                           stream, s->d_long, KB == 4 ? (const uint32_t *)K : nullptr,
                           KB == 8 ? (const uint64_t *)K : nullptr, s->hp, s->d_longbm, s->d_st);
        qs_tim_end(s, et, stream);
    }
    s->fix_keys += n_ord;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, stream, s->d_st, n_trig);
    CK(hipGetLastError());
    rc = read_state(s, stream);
    if (rc)
        return rc;
    if (s->h_st->err) {
        s->sticky_err = QS_ERR_POOL;
        return QS_ERR_POOL;
    }
    return QS_OK;
}

template <int KB>
static int insert_impl(qs_sketch *s, const void *d_keys, size_t n, hipStream_t stream)
{
    const KeyT<KB> *keys = (const KeyT<KB> *)d_keys;
    uint64_t cap = s->p.chunk_keys ? s->p.chunk_keys : (KB == 4 ? (1ull << 28) : (1ull << 27));
    if (cap > 0x7fffffffull)
        cap = 0x7fffffffull;
    size_t done = 0;
    while (done < n) {
        uint64_t cn64 = n - done < cap ? n - done : cap;
        int rc = qs_ensure_scratch(s, cn64);
        if (rc)
            return rc;
        const uint32_t cn = (uint32_t)cn64;
        qs_evpair ev;
        qs_tim_begin(s, ev, QS_TIMING_KERNELS, cn, stream);
        qs_tim_end(s, ev, stream);
        rc = insert_chunk<KB>(s, keys + done, cn, stream);
        if (rc)
            return rc;
        s->chunks++;
        s->keys_inserted += cn;
        done += cn;
    }
    return QS_OK;
}

int qs_insert_chunks(qs_sketch *s, const void *d_keys, size_t n, hipStream_t stream)
{
    return s->p.key_bytes == 4 ? insert_impl<4>(s, d_keys, n, stream)
                               : insert_impl<8>(s, d_keys, n, stream);
}
