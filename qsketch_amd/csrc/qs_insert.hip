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
// chunk by bucket turns the table update into an address-ordered sweep with one
// plain read-modify-write per touched bucket.
//
// Each insert is hashed once, in the first radix pass, into a 64-bit item
// (bucket | mask index | chunk-local stream index); later passes and kernels
// only extract bit fields.  Digit ranks inside a wavefront come from an LDS
// match (every lane ORs its lane bit into the slot of its digit and reads the
// group mask back), which is stable and costs three LDS operations per item.
//
// Per chunk:
//   sort      per pass: k_bs_up (tile digit counts), a scan over tiles,
//             k_bs_down (rank, stage digit-sorted in LDS, write digit runs)
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
//   k_tail    long segments after their allocation point: per-lane LDS counter
//             rows over 16K-item windows, one atomic per counter per window and
//             bucket.
#include "qs_internal.h"

#include <cstring>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess)                                                                   \
            return QS_ERR_HIP;                                                                   \
    } while (0)

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
#define OS_THREADS 256
#define OS_WAVES 4
#define OS_IPL 16                         // items per lane
#define OS_WAVE_ITEMS (64 * OS_IPL)       // 1024 items per wave
#define OS_TILE (OS_THREADS * OS_IPL)     // 4096 items per tile
#define OS_BINS 256                       // LDS capacity for digits up to 8 bits

template <int KB, bool FIRST>
__device__ static inline uint64_t load_item(const void *__restrict__ in, uint32_t i,
                                            const qs_hashpar &hp, const qs_pack &pk)
{
    if (FIRST) {
        const KeyT<KB> key = reinterpret_cast<const KeyT<KB> *>(in)[i];
        return (uint64_t)qs_bucket(key, hp) | ((uint64_t)qs_maskidx(key, hp) << pk.bb) |
               ((uint64_t)i << pk.ishift);
    }
    return reinterpret_cast<const uint64_t *>(in)[i];
}

// Per-tile digit counts, digit-major: hist[d * ntiles + tile].
template <int KB, bool FIRST>
__global__ __launch_bounds__(OS_THREADS) void k_bs_up(const void *__restrict__ in, uint32_t n,
                                                      qs_hashpar hp, qs_pack pk, int shift,
                                                      uint32_t dmask, uint32_t nbins,
                                                      uint32_t ntiles, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t s_h[OS_BINS];
    const uint32_t t = threadIdx.x;
    s_h[t] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * OS_TILE;
#pragma unroll 4
    for (int r = 0; r < OS_IPL; ++r) {
        const uint32_t i = base + r * OS_THREADS + t;
        if (i < n) {
            uint32_t b;
            if (FIRST)
                b = qs_bucket(reinterpret_cast<const KeyT<KB> *>(in)[i], hp);
            else
                b = it_b(reinterpret_cast<const uint64_t *>(in)[i], pk);
            atomicAdd(&s_h[(b >> shift) & dmask], 1u);
        }
    }
    __syncthreads();
    if (t < nbins)
        hist[(size_t)t * ntiles + blockIdx.x] = s_h[t];
}

// exclusive block scan of one value per thread (256 threads)
__device__ static inline uint32_t block_excl_scan256(uint32_t v, uint32_t *s_red)
{
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o)
            x += y;
    }
    if (lane == 63)
        s_red[wv] = x;
    __syncthreads();
    uint32_t off = 0;
#pragma unroll
    for (int k = 0; k < OS_WAVES; ++k)
        off += (k < (int)wv) ? s_red[k] : 0u;
    __syncthreads();
    return off + x - v;
}

// Stable scatter of one tile.  Wave w ranks its 1024 consecutive items in 16
// rounds: the LDS match gives the lanes sharing a digit, the lowest of them
// advances the wave's running count.  The tile is then staged digit-sorted in
// LDS and written out as contiguous digit runs.
template <int KB, bool FIRST>
__global__ __launch_bounds__(OS_THREADS) void k_bs_down(
    const void *__restrict__ in, uint32_t n, qs_hashpar hp, qs_pack pk, int shift, uint32_t dmask,
    uint32_t nbins, uint32_t ntiles, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ totals, uint64_t *__restrict__ out)
{
    __shared__ uint64_t s_item[OS_TILE];
    __shared__ unsigned long long s_match[OS_WAVES][OS_BINS];
    __shared__ uint32_t s_wh[OS_WAVES][OS_BINS];
    __shared__ uint32_t s_ls[OS_BINS];
    __shared__ uint32_t s_goff[OS_BINS];
    __shared__ uint32_t s_red[OS_WAVES];
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * OS_TILE;
#pragma unroll
    for (int w = 0; w < OS_WAVES; ++w) {
        s_wh[w][t] = 0;
        s_match[w][t] = 0;
    }
    const bool dv = t < nbins;
    const uint32_t tot = dv ? totals[t] : 0u;
    const uint32_t hcur = dv ? hist[(size_t)t * ntiles + tile] : 0u;
    const uint32_t hnxt = dv ? ((tile + 1 < ntiles) ? hist[(size_t)t * ntiles + tile + 1] : tot) : 0u;
    const uint32_t dbase = block_excl_scan256(tot, s_red);
    const uint32_t lstart = block_excl_scan256(hnxt - hcur, s_red);
    s_ls[t] = lstart;
    s_goff[t] = dbase + hcur - lstart;
    __syncthreads();

    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t item[OS_IPL];
    uint32_t dr[OS_IPL]; // digit | wave-local position << 8 | valid << 31
#pragma unroll
    for (int r = 0; r < OS_IPL; ++r) {
        const uint32_t i = base + wv * OS_WAVE_ITEMS + r * 64u + lane;
        const bool valid = i < n;
        item[r] = valid ? load_item<KB, FIRST>(in, i, hp, pk) : 0ull;
        const uint32_t d = (it_b(item[r], pk) >> shift) & dmask;
        if (valid)
            atomicOr(&s_match[wv][d], (unsigned long long)(1ull << lane));
        const uint64_t m = s_match[wv][d];
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        const uint32_t run = s_wh[wv][d];
        if (valid && rank == 0) {
            s_wh[wv][d] = run + (uint32_t)__popcll(m);
            s_match[wv][d] = 0;
        }
        dr[r] = d | ((run + rank) << 8) | (valid ? 0x80000000u : 0u);
    }
    __syncthreads();
    {
        uint32_t c0 = s_wh[0][t], c1 = s_wh[1][t], c2 = s_wh[2][t];
        s_wh[0][t] = 0;
        s_wh[1][t] = c0;
        s_wh[2][t] = c0 + c1;
        s_wh[3][t] = c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < OS_IPL; ++r) {
        if (dr[r] & 0x80000000u) {
            const uint32_t d = dr[r] & 0xffu;
            s_item[s_ls[d] + s_wh[wv][d] + ((dr[r] >> 8) & 0x7fffffu)] = item[r];
        }
    }
    __syncthreads();
    const uint32_t tn = (n - base) < OS_TILE ? (n - base) : OS_TILE;
    for (uint32_t q = t; q < tn; q += OS_THREADS) {
        const uint64_t it = s_item[q];
        out[s_goff[(it_b(it, pk) >> shift) & dmask] + q] = it;
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

__global__ __launch_bounds__(SEG_THREADS) void k_seg(
    const uint64_t *__restrict__ IT, uint32_t n, uint32_t *__restrict__ low,
    uint32_t *__restrict__ high, const uint32_t *__restrict__ masks, qs_pack pk, uint32_t n_masks,
    uint32_t *__restrict__ longbm, uint32_t *__restrict__ ord, uint32_t *__restrict__ lng,
    qs_dev_state *__restrict__ st)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t i = threadIdx.x; i < n_masks; i += SEG_THREADS)
        s_masks[i] = masks[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t le = (lane == 63u) ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t gw = blockIdx.x * (SEG_THREADS / 64u) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (SEG_THREADS / 64u);
    for (uint64_t w0 = (uint64_t)gw * 64u; w0 < n; w0 += (uint64_t)nw * 64u) {
        const uint32_t p = (uint32_t)w0 + lane;
        const bool valid = p < n;
        const uint64_t it = valid ? IT[p] : 0ull;
        const uint32_t b = valid ? it_b(it, pk) : NO_BUCKET;
        uint32_t bl = __shfl(b, (int)(lane == 0 ? 0 : lane - 1u), 64);
        if (lane == 0)
            bl = (w0 > 0) ? it_b(IT[w0 - 1], pk) : NO_BUCKET;
        // a window inside a segment that started earlier has nothing to own
        if (__ballot(valid && b != bl) == 0)
            continue;
        const uint32_t M = valid ? s_masks[it_mu(it, pk)] : 0u;
        uint32_t br = __shfl(b, (int)(lane == 63 ? 63 : lane + 1u), 64);
        if (lane == 63)
            br = (p + 1u < n) ? it_b(IT[p + 1u], pk) : END_BUCKET;
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
                const uint64_t iq = IT[q];
                in = it_b(iq, pk) == b63;
                Mq = in ? s_masks[it_mu(iq, pk)] : 0u;
            }
            const uint64_t im = __ballot(in);
            const uint32_t head_pos = (uint32_t)w0 + s63;
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
__global__ __launch_bounds__(256) void k_ordered(const uint64_t *__restrict__ IT, uint32_t n,
                                                 const uint32_t *__restrict__ ord,
                                                 uint32_t *__restrict__ low,
                                                 const uint32_t *__restrict__ masks, qs_pack pk,
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
        const uint32_t b = it_b(IT[p0], pk);
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
                const uint64_t iq = IT[q];
                valid = it_b(iq, pk) == b;
                if (valid) {
                    M = masks[it_mu(iq, pk)];
                    idx = it_idx(iq, pk);
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
// positions; each lane counts its inserts' counters in its own LDS row while
// the link stays the same, and the rows are summed and flushed with one atomic
// per counter when it changes.
#define TAIL_ROW 29 // 28 counters, odd stride against bank conflicts

__device__ static inline void tail_flush(uint32_t (*rows)[TAIL_ROW], uint32_t lane,
                                         uint32_t *__restrict__ high, uint32_t cur_w)
{
    if (lane < 28u) {
        uint32_t sum = 0;
        for (uint32_t l = 0; l < 64u; ++l)
            sum += rows[l][lane];
        if (sum)
            atomicAdd(high + (size_t)(cur_w - QS_LINK_START) * 32u + lane, sum);
    }
#pragma unroll
    for (int j = 0; j < 28; ++j)
        rows[lane][j] = 0;
}

__device__ static inline void tail_add(uint32_t (*rows)[TAIL_ROW], uint32_t lane, uint32_t M)
{
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        if (M) {
            rows[lane][__builtin_ctz(M)] += 1u;
            M &= M - 1u;
        }
    }
}

__global__ __launch_bounds__(256) void k_tail(const uint64_t *__restrict__ IT, uint32_t n,
                                              const uint32_t *__restrict__ low,
                                              uint32_t *__restrict__ high,
                                              const uint32_t *__restrict__ masks, qs_pack pk,
                                              uint32_t n_masks, const uint32_t *__restrict__ longbm,
                                              const uint32_t *__restrict__ tail_of,
                                              const qs_dev_state *__restrict__ st)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    __shared__ uint32_t s_rows[4][64][TAIL_ROW];
    for (uint32_t i = threadIdx.x; i < n_masks; i += blockDim.x)
        s_masks[i] = masks[i];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t(*rows)[TAIL_ROW] = s_rows[wv];
#pragma unroll
    for (int j = 0; j < 28; ++j)
        rows[lane][j] = 0;
    __syncthreads();
    const uint32_t gw = blockIdx.x * 4u + wv;
    const uint32_t nw = gridDim.x * 4u;
    const uint32_t first_new = st->next;
    for (uint64_t w0 = (uint64_t)gw * TAIL_WINDOW; w0 < n; w0 += (uint64_t)nw * TAIL_WINDOW) {
        const uint64_t wend = (w0 + TAIL_WINDOW < n) ? w0 + TAIL_WINDOW : n;
        uint32_t cur_w = 0;
        // lookup cache for the bucket of the previous single-bucket step
        uint32_t c_b = NO_BUCKET, c_w = 0, c_tail = 0;
        bool c_long = false;
        for (uint64_t s0 = w0; s0 < wend; s0 += 64u) {
            const uint32_t p = (uint32_t)s0 + lane;
            const bool valid = p < wend;
            const uint64_t it = valid ? IT[p] : 0ull;
            const uint32_t b = valid ? it_b(it, pk) : NO_BUCKET;
            const uint32_t b_first = __shfl(b, 0, 64);
            const uint32_t last_lane = (wend - s0 < 64u) ? (uint32_t)(wend - s0 - 1u) : 63u;
            const uint32_t b_last = __shfl(b, (int)last_lane, 64);
            if (b_first == b_last) {
                // the whole step is one bucket (the items are sorted by bucket)
                if (b_first != c_b) {
                    c_b = b_first;
                    c_long = (longbm[c_b >> 5] >> (c_b & 31u)) & 1u;
                    c_w = c_long ? low[(size_t)c_b * 8u + 7u] : 0u;
                    c_tail = (c_w > QS_LINK_START && c_w >= first_new) ? tail_of[c_w - QS_LINK_START] : 0u;
                }
                if (!c_long || c_w <= QS_LINK_START)
                    continue;
                const bool act = valid && p >= c_tail;
                if (!__ballot(act))
                    continue;
                if (c_w != cur_w) {
                    if (cur_w)
                        tail_flush(rows, lane, high, cur_w);
                    cur_w = c_w;
                }
                if (act)
                    tail_add(rows, lane, s_masks[it_mu(it, pk)]);
                continue;
            }
            bool act = false;
            uint32_t w = 0;
            if (valid && ((longbm[b >> 5] >> (b & 31u)) & 1u)) {
                w = low[(size_t)b * 8u + 7u];
                act = w > QS_LINK_START && (w < first_new || p >= tail_of[w - QS_LINK_START]);
            }
            uint64_t am = __ballot(act);
            while (am) {
                const uint32_t lead = (uint32_t)__builtin_ctzll(am);
                const uint32_t wl = __shfl(w, (int)lead, 64);
                const bool in = act && w == wl;
                if (wl != cur_w) {
                    if (cur_w)
                        tail_flush(rows, lane, high, cur_w);
                    cur_w = wl;
                }
                if (in)
                    tail_add(rows, lane, s_masks[it_mu(it, pk)]);
                am &= ~__ballot(in);
            }
        }
        if (cur_w)
            tail_flush(rows, lane, high, cur_w);
    }
}

__global__ void k_clear_long(const uint32_t *__restrict__ lng, const uint64_t *__restrict__ IT,
                             qs_pack pk, uint32_t *__restrict__ longbm,
                             const qs_dev_state *__restrict__ st)
{
    const uint32_t n_long = st->n_long;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_long; i += gridDim.x * blockDim.x) {
        const uint32_t b = it_b(IT[lng[i]], pk);
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
    uint64_t tiles = (chunk + OS_TILE - 1) / OS_TILE;
    if (tiles == 0)
        tiles = 1;
    uint64_t a = OS_BINS * tiles + OS_BINS;
    uint64_t b = qs_sort_hist_words((uint32_t)(chunk < 0xffffffffull ? chunk : 0xffffffffull));
    return a > b ? a : b;
}

void qs_free_scratch(qs_sketch *s)
{
    for (int k = 0; k < 2; ++k) {
        hipFree(s->d_it[k]);
        hipFree(s->d_tk[k]);
        hipFree(s->d_tv[k]);
        s->d_it[k] = nullptr;
        s->d_tk[k] = s->d_tv[k] = nullptr;
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
    uint64_t tcap = s->p.n_high < chunk ? s->p.n_high : chunk;
    s->trig_cap = (uint32_t)tcap;
    for (int k = 0; k < 2; ++k) {
        if (hipMalloc(&s->d_it[k], sizeof(uint64_t) * chunk) != hipSuccess ||
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
    const uint32_t ntiles = (cn + OS_TILE - 1) / OS_TILE;
    const int bits = (int)s->pk.bb;
    const int passes = (bits + 7) / 8;
    const int dbits = (bits + passes - 1) / passes; // <= 8
    const uint32_t nbins = 1u << dbits, dmask = nbins - 1u;
    uint32_t *totals = s->d_hist + (size_t)OS_BINS * ntiles;
    const void *in = keys;
    int out = 0;
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = dbits * ps;
        qs_evpair e0, e1, e2;
        qs_tim_begin(s, e0, QS_TK_SORT_UP, cn, stream);
        if (ps == 0)
            hipLaunchKernelGGL((k_bs_up<KB, true>), dim3(ntiles), dim3(OS_THREADS), 0, stream, in, cn,
                               s->hp, s->pk, shift, dmask, nbins, ntiles, s->d_hist);
        else
            hipLaunchKernelGGL((k_bs_up<KB, false>), dim3(ntiles), dim3(OS_THREADS), 0, stream, in, cn,
                               s->hp, s->pk, shift, dmask, nbins, ntiles, s->d_hist);
        qs_tim_end(s, e0, stream);
        qs_tim_begin(s, e1, QS_TK_SORT_SCAN, cn, stream);
        qs_scan_digits(s->d_hist, ntiles, nbins, totals, stream);
        qs_tim_end(s, e1, stream);
        qs_tim_begin(s, e2, QS_TK_SORT_DOWN, cn, stream);
        if (ps == 0)
            hipLaunchKernelGGL((k_bs_down<KB, true>), dim3(ntiles), dim3(OS_THREADS), 0, stream, in, cn,
                               s->hp, s->pk, shift, dmask, nbins, ntiles, s->d_hist, totals,
                               s->d_it[out]);
        else
            hipLaunchKernelGGL((k_bs_down<KB, false>), dim3(ntiles), dim3(OS_THREADS), 0, stream, in, cn,
                               s->hp, s->pk, shift, dmask, nbins, ntiles, s->d_hist, totals,
                               s->d_it[out]);
        qs_tim_end(s, e2, stream);
        in = s->d_it[out];
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
    int res = 0;
    int rc = sort_chunk<KB>(s, keys, cn, stream, res);
    if (rc)
        return rc;
    const uint64_t *IT = s->d_it[res];
    qs_evpair es;
    qs_tim_begin(s, es, QS_TK_SEG, cn, stream);
    hipLaunchKernelGGL(k_seg, dim3(grid_cap(((uint64_t)cn + 255) / 256, 8)), dim3(SEG_THREADS), 0, stream,
                       IT, cn, s->d_low, s->d_high, s->d_masks, s->pk, s->p.n_masks, s->d_longbm,
                       s->d_ord, s->d_long, s->d_st);
    qs_tim_end(s, es, stream);
    CK(hipGetLastError());
    rc = read_state(s, stream);
    if (rc)
        return rc;
    const uint32_t n_ord = s->h_st->n_ord;
    uint32_t n_trig = 0;
    if (n_ord) {
        s->chunks_fix++;
        qs_evpair eo;
        qs_tim_begin(s, eo, QS_TK_ORDERED, cn, stream);
        hipLaunchKernelGGL(k_ordered, dim3(grid_cap(((uint64_t)n_ord + 3) / 4, 8)), dim3(256), 0, stream,
                           IT, cn, s->d_ord, s->d_low, s->d_masks, s->pk, s->d_longbm, s->d_long,
                           s->d_trig, s->trig_cap, s->d_st, s->p.semantics == QS_PAPER ? 1 : 0);
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
        hipLaunchKernelGGL(k_tail, dim3(grid_cap(((uint64_t)cn + TAIL_WINDOW * 4 - 1) / (TAIL_WINDOW * 4), 8)),
                           dim3(256), 0, stream, IT, cn, s->d_low, s->d_high, s->d_masks, s->pk,
                           s->p.n_masks, s->d_longbm, s->d_tail_of, s->d_st);
        hipLaunchKernelGGL(k_clear_long, dim3(grid_cap((s->h_st->n_long + 255) / 256, 4)), dim3(256), 0,
                           stream, s->d_long, IT, s->pk, s->d_longbm, s->d_st);
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
    // chunk limit: the stream index field of a packed item, and 2^30
    uint64_t cap = s->p.chunk_keys ? s->p.chunk_keys : (1ull << 28);
    const uint32_t ibits = 64u - s->pk.ishift;
    if (ibits < 63u && cap > (1ull << ibits))
        cap = 1ull << ibits;
    if (cap > (1ull << 30))
        cap = 1ull << 30;
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
