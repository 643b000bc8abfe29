// qs_insert.hip -- deterministic MSWS insert for gfx950: sort by bucket, then
// every bucket segment has exactly one owner.  One pipeline serves a plain
// sketch and a batch of independent sketches (the multi-sketch form).
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
// chunk by (sketch, bucket) turns the table update into an address-ordered
// sweep with one plain read-modify-write per touched bucket.
//
// Each insert is hashed once, in the first radix pass, into a 64-bit item
// (bucket | sketch | mask index | stream index); later passes and kernels
// only extract bit fields.  Digit ranks inside a wavefront come from a ballot
// match (one ballot per digit bit gives each lane the mask of lanes sharing
// its digit), which is stable; the group's lowest lane advances the digit's
// running count with one LDS atomic.
//
// Per chunk (a fixed launch sequence with no host synchronisation, so it can
// be captured in a hipGraph; all counts stay on the device):
//   sort      per pass: tile digit counts (k_bs_up from the keys in the first
//             pass, k_bs_up_dig from the one-byte digits the previous
//             scatter wrote in later ones), a scan over tiles, k_bs_down
//             (rank, stage digit-sorted in LDS, write digit runs and the
//             next pass's digit bytes)
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
//   k_assign  allocation ids per sketch, in stream order of the triggering inserts.
//   k_tail    long segments after their allocation point: per-lane LDS counter
//             rows over windows of sorted items, one atomic per counter per
//             window and bucket.  Only the runs at the two ends of a 64-item
//             step can be long, so a step looks up at most two rows.
#include "qs_internal.h"

#include <cstring>
#include <type_traits>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess)                                                                   \
            return QS_ERR_HIP;                                                                   \
    } while (0)

#define SEG_THREADS 256
#define TAIL_WINDOW 4096u
#define NO_CELL 0xffffffffffffffffull
#define END_CELL 0xfffffffffffffffeull

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

// sketch owning batch position pos: offs[sid] <= pos < offs[sid + 1]
__device__ static inline uint32_t find_sid(const uint64_t *__restrict__ offs, uint32_t nsk, uint64_t pos)
{
    uint32_t lo = 0, hi = nsk; // invariant: offs[lo] <= pos < offs[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (offs[mid] <= pos)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ===========================================================================
// radix sort of the chunk by cell = (sketch, bucket)
// ===========================================================================
#ifndef QS_OS_THREADS
#define QS_OS_THREADS 256
#endif
#define OS_THREADS QS_OS_THREADS
#define OS_WAVES (OS_THREADS / 64)
// items per lane: 8 measured best on MI355X (2^28-key Zipf insert: 16 items
// per lane left the scatter at 4 waves/SIMD, 4 grew the scan of the tile
// histograms more than the scatter gained)
#ifndef QS_OS_IPL
#define QS_OS_IPL 8
#endif
#define OS_IPL QS_OS_IPL
#define OS_WAVE_ITEMS (64 * OS_IPL)       // items per wave
#define OS_TILE (OS_THREADS * OS_IPL)     // items per tile
#define OS_BINS 256                       // LDS capacity for digits up to 8 bits

// First pass: hash key number i of the chunk (batch position base + i).  A
// single sketch (MULTI false) hashes with d0, passed by value; a batch finds
// the owning sketch by binary search over the batch offsets.
// The sketches a tile of a batch spans: their offsets offs[sid0 .. sid0 + cnt]
// copied to LDS by the block (cnt = 0: too many, search the global array).
#define SIDW 64u
struct sid_window {
    const uint64_t *o;
    uint32_t sid0, cnt;
};

__device__ static inline void load_sid_window(const uint64_t *__restrict__ offs, uint32_t nsk, uint64_t lo,
                                              uint64_t hi, uint64_t *s_off, uint32_t *s_w)
{
    if (threadIdx.x == 0) {
        const uint32_t a = find_sid(offs, nsk, lo), b = find_sid(offs, nsk, hi);
        s_w[0] = a;
        s_w[1] = (b - a + 1u) <= SIDW ? (b - a + 1u) : 0u;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; s_w[1] && j <= s_w[1]; j += blockDim.x)
        s_off[j] = offs[s_w[0] + j];
    __syncthreads();
}

// bucket and mask fields of a key's item: h % n_low and h % n_masks, or in the
// quotient form h % n_low and h / n_low (32-bit keys: h = key * seed mod 2^32)
template <int KB>
__device__ static inline uint64_t bucket_mask_fields(KeyT<KB> key, const qs_sdesc &d, const qs_pack &pk)
{
    if (KB == 4 && pk.quot) {
        const uint32_t h = (uint32_t)key * d.seed;
        const uint32_t q = qs_fastquot(h, d.fd_low);
        return (uint64_t)(h - q * d.n_low) | ((uint64_t)q << pk.mshift);
    }
    return (uint64_t)qs_bucket_d(key, d) | ((uint64_t)qs_maskidx_d(key, d) << pk.mshift);
}

template <int KB, bool MULTI>
__device__ static inline uint64_t make_item(const KeyT<KB> key, uint32_t i, uint64_t base,
                                            const qs_sdesc &d0, const qs_sdesc *__restrict__ D,
                                            const uint64_t *__restrict__ offs, uint32_t nsk,
                                            const qs_pack &pk, const sid_window &win)
{
    if (!MULTI)
        return bucket_mask_fields<KB>(key, d0, pk) | ((uint64_t)i << pk.ishift);
    const uint64_t pos = base + i;
    uint32_t sid;
    uint64_t sstart;
    if (win.cnt) { // win.o[lo] <= pos < win.o[hi]
        uint32_t lo = 0, hi = win.cnt;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (win.o[mid] <= pos)
                lo = mid;
            else
                hi = mid;
        }
        sid = win.sid0 + lo;
        sstart = win.o[lo];
    } else {
        sid = find_sid(offs, nsk, pos);
        sstart = offs[sid];
    }
    const qs_sdesc &d = D[sid];
    const uint64_t start = sstart > base ? sstart : base; // this sketch's part of the chunk
    return bucket_mask_fields<KB>(key, d, pk) | ((uint64_t)sid << pk.bb) | ((pos - start) << pk.ishift);
}

// what a radix pass reads per position: the key (first pass) or the item,
// streamed once (non-temporal); positions past n read the last one
template <int KB, bool FIRST> using RawT = typename std::conditional<FIRST, KeyT<KB>, uint64_t>::type;

template <int KB, bool FIRST>
__device__ static inline RawT<KB, FIRST> load_raw(const void *in, uint32_t i, uint32_t n)
{
    const uint32_t ic = i < n ? i : n - 1u;
    return __builtin_nontemporal_load(reinterpret_cast<const RawT<KB, FIRST> *>(in) + ic);
}

// Tile counts of the first pass, from the keys: each item is built on the way
// (the scatter builds it again rather than reading it back)
template <int KB, bool MULTI>
__global__ __launch_bounds__(OS_THREADS) void k_bs_up(const void *__restrict__ in, uint32_t n, uint64_t base,
                                                      qs_sdesc d0, const qs_sdesc *__restrict__ D,
                                                      const uint64_t *__restrict__ offs, uint32_t nsk,
                                                      qs_pack pk, int shift, uint32_t dmask,
                                                      uint32_t nbins, uint32_t ntiles,
                                                      uint16_t *__restrict__ hist)
{
    __shared__ uint32_t s_h[OS_BINS];
    __shared__ uint64_t s_off[MULTI ? SIDW + 1 : 1];
    __shared__ uint32_t s_w[2];
    const uint32_t t = threadIdx.x;
    if (t < OS_BINS)
        s_h[t] = 0;
    const uint32_t tb = blockIdx.x * OS_TILE;
    sid_window win{s_off, 0u, 0u};
    if (MULTI) {
        const uint32_t te = (n - tb) < OS_TILE ? n : tb + OS_TILE;
        load_sid_window(offs, nsk, base + tb, base + te - 1u, s_off, s_w);
        win.sid0 = s_w[0];
        win.cnt = s_w[1];
    }
    __syncthreads();
    const uint32_t lane = t & 63u;
    // every load first, from clamped positions: a load inside a per-item branch
    // is issued and waited for one at a time
    KeyT<KB> raw[OS_IPL];
#pragma unroll
    for (int r = 0; r < OS_IPL; ++r)
        raw[r] = load_raw<KB, true>(in, tb + r * OS_THREADS + t, n);
#pragma unroll
    for (int r = 0; r < OS_IPL; ++r) {
        const uint32_t i = tb + r * OS_THREADS + t; // a wavefront's lanes hold consecutive items
        const bool valid = i < n;
        uint32_t d = 0;
        if (valid) {
            const uint64_t c = it_cell(make_item<KB, MULTI>(raw[r], i, base, d0, D, offs, nsk, pk, win), pk);
            d = (uint32_t)(c >> shift) & dmask;
        }
        // one LDS add per run of equal digits on adjacent lanes (a Zipf-hot
        // cell fills whole wavefronts, whose adds to one bin would serialise)
        // (both ballots are taken by the whole wavefront, outside any branch)
        const uint32_t dp = (uint32_t)__shfl_up((int)d, 1, 64);
        const bool head = valid && (lane == 0 || d != dp);
        const uint64_t heads = __ballot(head);
        const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid)); // valid lanes are a prefix
        const uint64_t later = heads & ~((2ull << lane) - 1ull);
        const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : nvalid;
        if (head)
            atomicAdd(&s_h[d], end - lane);
    }
    __syncthreads();
    if (t < nbins)
        hist[(size_t)blockIdx.x * nbins + t] = (uint16_t)s_h[t]; // tile-major: one contiguous row per tile
}

// Tile counts of a later pass, from the digit bytes the previous scatter wrote
// (one byte per item read instead of the 8-byte item).  A block counts
// UP_TILES tiles, each into its own LDS row; in each, thread t counts the
// items t * OS_IPL .. t * OS_IPL + OS_IPL - 1, with one LDS add per run of
// equal digits (after the first pass the items of one cell are adjacent);
// runs of threads that hold a single digit are added once per run of lanes.
// The digit buffer holds whole tiles and loads of tiles past the last read the
// last one, so every load is in bounds; items past n are not counted.
#define UP_TILES 4u
static_assert(OS_IPL % 8 == 0, "digit bytes are read 8 at a time");
__global__ __launch_bounds__(OS_THREADS) void k_bs_up_dig(const uint8_t *__restrict__ dig, uint32_t n,
                                                          uint32_t ntiles, uint32_t nbins,
                                                          uint16_t *__restrict__ hist)
{
    __shared__ uint32_t s_h[UP_TILES][OS_BINS];
    const uint32_t t = threadIdx.x;
    const uint32_t tile0 = blockIdx.x * UP_TILES;
    uint64_t w[UP_TILES][OS_IPL / 8];
#pragma unroll
    for (uint32_t k = 0; k < UP_TILES; ++k) {
        const uint32_t tl = tile0 + k < ntiles ? tile0 + k : ntiles - 1u;
        const uint8_t *src = dig + (size_t)tl * OS_TILE + t * OS_IPL;
#pragma unroll
        for (int j = 0; j < OS_IPL / 8; ++j)
            w[k][j] = __builtin_nontemporal_load(reinterpret_cast<const uint64_t *>(src) + j);
    }
    if (t < OS_BINS)
#pragma unroll
        for (uint32_t k = 0; k < UP_TILES; ++k)
            s_h[k][t] = 0;
    __syncthreads();
    const uint32_t lane = t & 63u;
#pragma unroll
    for (uint32_t k = 0; k < UP_TILES; ++k) {
        const uint64_t i0 = (uint64_t)(tile0 + k) * OS_TILE + t * OS_IPL;
        const uint32_t nv = i0 < n ? (n - i0 < OS_IPL ? (uint32_t)(n - i0) : OS_IPL) : 0u; // counted items
        // A thread whose items all carry one digit is counted together with
        // its neighbours: one add per run of such lanes (a Zipf-hot cell fills
        // whole tiles, and one add per lane to one bin would serialise).  The
        // ballots and the shuffle are taken by the whole wavefront.
        const uint32_t d0 = (uint32_t)w[k][0] & 0xffu;
        bool whole = nv == OS_IPL;
#pragma unroll
        for (int j = 0; j < OS_IPL / 8; ++j)
            whole = whole && w[k][j] == 0x0101010101010101ull * d0;
        const uint64_t wm = __ballot(whole);
        const uint32_t dprev = (uint32_t)__shfl_up((int)d0, 1, 64);
        const bool head = whole && !(lane > 0u && ((wm >> (lane - 1u)) & 1ull) && dprev == d0);
        const uint64_t heads = __ballot(head);
        const uint64_t later = (heads | ~wm) & ~((2ull << lane) - 1ull); // lanes that end my run
        const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : 64u;
        if (head)
            atomicAdd(&s_h[k][d0], (end - lane) * OS_IPL);
        if (!whole) {
            uint32_t cur = 0, run = 0;
#pragma unroll
            for (int j = 0; j < OS_IPL; ++j) {
                const uint32_t d = (uint32_t)(w[k][j >> 3] >> (8 * (j & 7))) & 0xffu;
                if ((uint32_t)j < nv) {
                    if (run && d != cur) {
                        atomicAdd(&s_h[k][cur], run);
                        run = 0;
                    }
                    cur = d;
                    ++run;
                }
            }
            if (run)
                atomicAdd(&s_h[k][cur], run);
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < UP_TILES; ++k)
        if (tile0 + k < ntiles && t < nbins)
            hist[(size_t)(tile0 + k) * nbins + t] = (uint16_t)s_h[k][t];
}

// Two-level exclusive scan over tiles of the tile-major digit counts.
// Level 1: each block takes SCAN_TPB consecutive tiles, thread d = digit d
// (reads and writes are contiguous rows), and leaves group-local prefixes plus
// the group's sums.  Level 2: one block per digit scans that digit's group
// sums and writes its total.  The scatter adds the two levels.
#define SCAN_TPB 32u
// The tile counts and their group-local prefixes are 16-bit: a count is at
// most OS_TILE and a prefix at most (SCAN_TPB - 1) * OS_TILE.
static_assert(SCAN_TPB * OS_TILE <= 65536u, "16-bit tile counts");

__global__ __launch_bounds__(256) void k_hist_scan_local(uint16_t *__restrict__ hist, uint32_t ntiles,
                                                         uint32_t nbins, uint32_t *__restrict__ gsum)
{
    const uint32_t d = threadIdx.x;
    if (d >= nbins)
        return;
    const uint32_t t0 = blockIdx.x * SCAN_TPB;
    const uint32_t t1 = min(t0 + SCAN_TPB, ntiles);
    uint32_t run = 0;
#pragma unroll 8
    for (uint32_t t = t0; t < t1; ++t) {
        uint16_t *h = hist + (size_t)t * nbins + d;
        const uint32_t v = *h;
        *h = (uint16_t)run;
        run += v;
    }
    gsum[(size_t)blockIdx.x * nbins + d] = run;
}

// one block per digit: a block-wide scan of the group sums (8 per thread)
#define GSCAN_PT 8u
__global__ __launch_bounds__(1024) void k_hist_scan_groups(uint32_t *__restrict__ gsum, uint32_t ngrp,
                                                           uint32_t nbins, uint32_t *__restrict__ totals)
{
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const uint32_t d = blockIdx.x, t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    if (t == 0)
        s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < ngrp; base += 1024u * GSCAN_PT) {
        const uint32_t g0 = base + t * GSCAN_PT;
        uint32_t v[GSCAN_PT];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < GSCAN_PT; ++k) {
            v[k] = (g0 + k < ngrp) ? gsum[(size_t)(g0 + k) * nbins + d] : 0u;
            sum += v[k];
        }
        uint32_t x = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o)
                x += y;
        }
        if (lane == 63)
            s_w[wv] = x;
        __syncthreads();
        if (t < 16) {
            const uint32_t a0 = s_w[t];
            uint32_t sc = a0;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const uint32_t y = __shfl_up(sc, o, 16);
                if ((t & 15u) >= (uint32_t)o)
                    sc += y;
            }
            s_w[t] = sc - a0;
        }
        __syncthreads();
        const uint32_t carry = s_carry;
        uint32_t run = carry + s_w[wv] + x - sum;
#pragma unroll
        for (uint32_t k = 0; k < GSCAN_PT; ++k) {
            if (g0 + k < ngrp)
                gsum[(size_t)(g0 + k) * nbins + d] = run;
            run += v[k];
        }
        __syncthreads();
        if (t == 1023)
            s_carry = carry + s_w[wv] + x;
        __syncthreads();
    }
    if (t == 0)
        totals[d] = s_carry;
}

// exclusive block scan of one value per thread (OS_THREADS threads)
__device__ static inline uint32_t block_excl_scan(uint32_t v, uint32_t *s_red)
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

// Stable scatter of one tile.  Wave w ranks its OS_WAVE_ITEMS consecutive
// items in OS_IPL rounds: the ballot match gives the lanes sharing a digit, the lowest of them
// advances the wave's running count.  The tile is then staged digit-sorted in
// LDS and written out as contiguous digit runs.
template <int KB, bool FIRST, bool MULTI>
__global__ __launch_bounds__(OS_THREADS) void k_bs_down(
    const void *__restrict__ in, uint32_t n, uint64_t base, qs_sdesc d0, const qs_sdesc *__restrict__ D,
    const uint64_t *__restrict__ offs, uint32_t nsk, qs_pack pk, int shift, uint32_t dmask,
    uint32_t nbins, uint32_t ntiles, const uint16_t *__restrict__ hist, const uint32_t *__restrict__ gsum,
    const uint32_t *__restrict__ totals, uint64_t *__restrict__ out, uint8_t *__restrict__ dig, int nshift,
    qs_dev_state *__restrict__ st)
{
    __shared__ uint64_t s_item[OS_TILE];
    __shared__ uint32_t s_wh[OS_WAVES][OS_BINS];
    __shared__ uint32_t s_ls[OS_BINS];
    __shared__ uint32_t s_goff[OS_BINS];
    __shared__ uint32_t s_red[OS_WAVES];
    __shared__ uint64_t s_off[MULTI ? SIDW + 1 : 1];
    __shared__ uint32_t s_w[2];
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    // Blocks go to the 8 XCDs in turn; give each XCD a contiguous range of
    // tiles, so neighbouring tiles' digit runs (adjacent in the output) meet
    // in one L2 (a bijection of [0, ntiles) whatever the placement)
    const uint32_t xq = ntiles >> 3, xr = ntiles & 7u, xb = blockIdx.x & 7u;
    const uint32_t tile = xb * xq + min(xb, xr) + (blockIdx.x >> 3);
    const uint32_t tb = tile * OS_TILE;
    sid_window win{s_off, 0u, 0u};
    if (FIRST && MULTI) {
        const uint32_t te = (n - tb) < OS_TILE ? n : tb + OS_TILE;
        load_sid_window(offs, nsk, base + tb, base + te - 1u, s_off, s_w);
        win.sid0 = s_w[0];
        win.cnt = s_w[1];
    }
    if (t < OS_BINS)
#pragma unroll
        for (int w = 0; w < OS_WAVES; ++w)
            s_wh[w][t] = 0;
    const bool dv = t < nbins;
    const uint32_t tot = dv ? totals[t] : 0u;
    // items of digit t before this tile (tile-major counts, two-level scan)
    const uint32_t hcur = dv ? hist[(size_t)tile * nbins + t] + gsum[(size_t)(tile / SCAN_TPB) * nbins + t] : 0u;
    const uint32_t dbase = block_excl_scan(tot, s_red);

    const uint64_t lt = (1ull << lane) - 1ull;
    const int dbits = __builtin_popcount(dmask);
    uint64_t item[OS_IPL];
    uint32_t dr[OS_IPL]; // digit | wave-local position << 8 | valid << 31
    {
        // every load first (see k_bs_up)
        RawT<KB, FIRST> raw[OS_IPL];
#pragma unroll
        for (int r = 0; r < OS_IPL; ++r)
            raw[r] = load_raw<KB, FIRST>(in, tb + wv * OS_WAVE_ITEMS + r * 64u + lane, n);
#pragma unroll
        for (int r = 0; r < OS_IPL; ++r) {
            const uint32_t i = tb + wv * OS_WAVE_ITEMS + r * 64u + lane;
            if (FIRST)
                item[r] = i < n ? make_item<KB, MULTI>((KeyT<KB>)raw[r], i, base, d0, D, offs, nsk, pk, win) : 0ull;
            else
                item[r] = i < n ? (uint64_t)raw[r] : 0ull;
        }
    }
#pragma unroll
    for (int r = 0; r < OS_IPL; ++r) {
        const uint32_t i = tb + wv * OS_WAVE_ITEMS + r * 64u + lane;
        const bool valid = i < n;
        const uint32_t d = (uint32_t)(it_cell(item[r], pk) >> shift) & dmask;
        // lanes sharing my digit: one ballot per digit bit (the bits above the
        // digit are 0 in every lane and leave m as it is)
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            if (bit < dbits) { // uniform: the digit width is a launch argument
                const bool on = (d >> bit) & 1u;
                const uint64_t bb = __ballot(on);
                m &= on ? bb : ~bb;
            }
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        const int leader = valid ? (int)__builtin_ctzll(m) : 0;
        // the group's lowest lane advances the wave's running count of the
        // digit (LDS atomics of one wave complete in issue order)
        uint32_t run = 0;
        if (valid && rank == 0)
            run = atomicAdd(&s_wh[wv][d], (uint32_t)__popcll(m));
        run = (uint32_t)__shfl((int)run, leader, 64);
        dr[r] = d | ((run + rank) << 8) | (valid ? 0x80000000u : 0u);
    }
    __syncthreads();
    {
        // this tile's count of digit t comes from the ranking itself; the
        // per-wave counts become per-wave starts within the digit
        uint32_t cnt = 0;
        if (t < OS_BINS)
#pragma unroll
            for (int w = 0; w < OS_WAVES; ++w) {
                const uint32_t c = s_wh[w][t];
                s_wh[w][t] = cnt;
                cnt += c;
            }
        const uint32_t lstart = block_excl_scan(cnt, s_red);
        if (t < OS_BINS) {
            s_ls[t] = lstart;
            s_goff[t] = dbase + hcur - lstart;
        }
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
    const uint32_t tn = (n - tb) < OS_TILE ? (n - tb) : OS_TILE;
    for (uint32_t q = t; q < tn; q += OS_THREADS) {
        const uint64_t it = s_item[q];
        const uint64_t c = it_cell(it, pk);
        const uint32_t dst = s_goff[(uint32_t)(c >> shift) & dmask] + q;
        if (dst < n) { // the tile counts and the ranking agree; checked, never trusted
            out[dst] = it;
            if (nshift >= 0) // the next pass's digit, for its tile counts
                dig[dst] = (uint8_t)((uint32_t)(c >> nshift) & dmask);
        } else
            atomicOr(&st->err, QS_EF_SORT);
    }
}

// ===========================================================================
// per-segment application
// ===========================================================================

// table pointers of one sketch
struct seg_tables {
    uint32_t *low;       // this sketch's low table
    uint32_t *high;      // this sketch's high table
    const uint32_t *masks;
};

// a single sketch's tables start at offset 0 of every pool
template <bool MULTI>
__device__ static inline seg_tables tables_of(uint32_t sid, const qs_sdesc *__restrict__ D,
                                              uint32_t *low_pool, uint32_t *high_pool,
                                              const uint32_t *mask_pool)
{
    seg_tables t;
    t.low = low_pool;
    t.high = high_pool;
    t.masks = mask_pool;
    if (MULTI) {
        const qs_sdesc &d = D[sid];
        t.low += d.low_off;
        t.high += d.high_off;
        t.masks += d.masks_off;
    }
    return t;
}

template <bool MULTI>
__device__ static inline uint32_t mask_of(uint32_t sid, uint32_t mu, const qs_sdesc *__restrict__ D,
                                          const uint32_t *__restrict__ mask_pool, const uint32_t *s_masks,
                                          uint32_t unm)
{
    if (!MULTI)
        return s_masks[mu];
    // equal-sized mask tables: the offset needs no descriptor load
    return mask_pool[(unm ? sid * unm : D[sid].masks_off) + mu];
}

// packed per-byte increments of one 3-bit mask into 7 words
__device__ static inline void mask_to_words(uint32_t M, uint32_t a[7])
{
    // word k holds counters 4k..4k+3: multiplying nibble n by 1 + 2^7 + 2^14 +
    // 2^21 puts its bit i at bit 8i (the four shifted copies do not overlap)
#pragma unroll
    for (int k = 0; k < 7; ++k)
        a[k] = (((M >> (4 * k)) & 0xfu) * 0x00204081u) & 0x01010101u;
}

// Owner of a complete segment (packed per-byte counts a[], each <= 128),
// whose bucket words x, y are already loaded: returns true when the segment
// must go to the in-order replay.
__device__ static inline bool apply_loaded(uint32_t *__restrict__ low, uint32_t *__restrict__ high, uint32_t b,
                                           const uint32_t a[7], uint4 x, uint4 y)
{
    uint4 *bp = reinterpret_cast<uint4 *>(low + (size_t)b * 8u);
    uint32_t W[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
    uint32_t w = W[7];
    if (w > QS_LINK_START) {
        // the owner alone writes this high row: every touched 16-byte group
        // is read first (the loads are in flight together), then written
        uint4 *row4 = reinterpret_cast<uint4 *>(high + (size_t)(w - QS_LINK_START) * 32u);
        uint4 q[7];
#pragma unroll
        for (int k = 0; k < 7; ++k)
            if (a[k])
                q[k] = row4[k];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            if (a[k]) {
                q[k].x += a[k] & 0xffu;
                q[k].y += (a[k] >> 8) & 0xffu;
                q[k].z += (a[k] >> 16) & 0xffu;
                q[k].w += a[k] >> 24;
                row4[k] = q[k];
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

__device__ static inline bool apply_segment(uint32_t *__restrict__ low, uint32_t *__restrict__ high,
                                            uint32_t b, const uint32_t a[7])
{
    const uint4 *bp = reinterpret_cast<const uint4 *>(low + (size_t)b * 8u);
    return apply_loaded(low, high, b, a, bp[0], bp[1]);
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

__device__ static inline uint64_t shfl64(uint64_t v, int src)
{
    uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

template <bool MULTI>
__device__ static inline void set_long(uint32_t *longbm, const qs_sdesc *__restrict__ D, uint32_t sid,
                                       uint32_t b)
{
    const uint64_t bit = (MULTI ? D[sid].bm_off : 0ull) + b;
    atomicOr(longbm + (bit >> 5), 1u << (bit & 31u));
}

template <bool MULTI>
__device__ static inline bool test_long(const uint32_t *longbm, const qs_sdesc *__restrict__ D,
                                        uint32_t sid, uint32_t b)
{
    const uint64_t bit = (MULTI ? D[sid].bm_off : 0ull) + b;
    return (longbm[bit >> 5] >> (bit & 31u)) & 1u;
}

template <bool MULTI>
__global__ __launch_bounds__(SEG_THREADS) void k_seg(
    const uint64_t *__restrict__ IT, uint32_t n, const qs_sdesc *__restrict__ D, uint32_t n_masks0,
    uint32_t *__restrict__ low_pool, uint32_t *__restrict__ high_pool,
    const uint32_t *__restrict__ mask_pool, qs_pack pk, uint32_t *__restrict__ longbm,
    uint32_t *__restrict__ ord, uint32_t *__restrict__ lng, qs_dev_state *__restrict__ st)
{
    // a single sketch keeps its masks in LDS; a batch reads each sketch's masks
    // from the (cache-resident) pool
    __shared__ uint32_t s_masks[MULTI ? 1 : QS_MAX_MASKS];
    if (!MULTI)
        for (uint32_t i = threadIdx.x; i < n_masks0; i += SEG_THREADS)
            s_masks[i] = mask_pool[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t le = (lane == 63u) ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t gw = blockIdx.x * (SEG_THREADS / 64u) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (SEG_THREADS / 64u);
    for (uint64_t w0 = (uint64_t)gw * 64u; w0 < n; w0 += (uint64_t)nw * 64u) {
        const uint32_t p = (uint32_t)w0 + lane;
        const bool valid = p < n;
        const uint64_t it = valid ? IT[p] : 0ull;
        const uint64_t c = valid ? it_cell(it, pk) : NO_CELL;
        uint64_t cl = shfl64(c, lane == 0 ? 0 : (int)lane - 1);
        if (lane == 0)
            cl = (w0 > 0) ? it_cell(IT[w0 - 1], pk) : NO_CELL;
        // a window inside a segment that started earlier has nothing to own
        if (__ballot(valid && c != cl) == 0)
            continue;
        const uint32_t sid = it_sid(it, pk);
        const uint32_t b = it_b(it, pk);
        uint32_t M = 0;
        if (valid)
            M = mask_of<MULTI>(sid, item_mu<MULTI>(it, pk, D), D, mask_pool, s_masks, pk.unm);
        uint64_t cr = shfl64(c, lane == 63 ? 63 : (int)lane + 1);
        if (lane == 63)
            cr = (p + 1u < n) ? it_cell(IT[p + 1u], pk) : END_CELL;
        const bool head = valid && c != cl;
        const bool last = valid && c != cr;
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
        if (owned && last) {
            seg_tables tb = tables_of<MULTI>(sid, D, low_pool, high_pool, mask_pool);
            risky = apply_segment(tb.low, tb.high, b, a);
        }
        wave_append(risky, (uint32_t)w0 + s, ord, &st->n_ord);
        // the last run continues past the window: look one window further
        const bool ext = __shfl((int)(owned && !last), 63, 64) != 0;
        if (ext) {
            const uint64_t c63 = shfl64(c, 63);
            const uint32_t s63 = __shfl(s, 63, 64);
            const uint32_t sid63 = (uint32_t)(c63 >> pk.bb);
            const uint32_t b63 = (uint32_t)(c63 & pk.bmask);
            uint32_t e[7];
#pragma unroll
            for (int k = 0; k < 7; ++k)
                e[k] = __shfl(a[k], 63, 64);
            const uint32_t q = (uint32_t)w0 + 64u + lane;
            bool in = false;
            uint32_t Mq = 0;
            if (q < n) {
                const uint64_t iq = IT[q];
                in = it_cell(iq, pk) == c63;
                if (in)
                    Mq = mask_of<MULTI>(sid63, item_mu<MULTI>(iq, pk, D), D, mask_pool, s_masks, pk.unm);
            }
            const uint64_t im = __ballot(in);
            const uint32_t head_pos = (uint32_t)w0 + s63;
            if (im != ~0ull) {
                uint32_t cc[7];
                mask_to_words(Mq, cc);
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    uint32_t v = cc[k];
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1)
                        v += __shfl_xor(v, o, 64);
                    e[k] += v;
                }
                bool r2 = false;
                if (lane == 0) {
                    seg_tables tb = tables_of<MULTI>(sid63, D, low_pool, high_pool, mask_pool);
                    r2 = apply_segment(tb.low, tb.high, b63, e);
                }
                wave_append(r2, head_pos, ord, &st->n_ord);
            } else {
                // long segment: tail kernel (after the in-order replay if unlinked)
                bool unl = false;
                if (lane == 0) {
                    set_long<MULTI>(longbm, D, sid63, b63);
                    unl = tables_of<MULTI>(sid63, D, low_pool, high_pool, mask_pool).low[(size_t)b63 * 8u + 7u] <=
                          QS_LINK_START;
                }
                wave_append(lane == 0, head_pos, lng, &st->n_long);
                wave_append(unl, head_pos, ord, &st->n_ord);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_segw<W, IPL>: the same per-segment application with a tunable shape.  A
// group of W lanes (a sub-wavefront) owns windows of T = W * IPL sorted items;
// each lane reduces IPL consecutive items on its own (runs that start and end
// inside the lane are applied by it directly), then a W-wide segmented scan
// joins the runs that cross lanes.  A run is owned by the window holding its
// head.  A run that crosses the window end is finished from one extension
// window of T items; one that fills the extension, or that reaches 256 items
// (the packed byte sums hold at most 255), is long.
// ---------------------------------------------------------------------------
template <int W>
__device__ static inline uint64_t shfl64w(uint64_t v, int src)
{
    uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, W);
    uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, W);
    return ((uint64_t)hi << 32) | lo;
}

__device__ static inline void add_words(uint32_t a[7], const uint32_t b[7])
{
#pragma unroll
    for (int k = 0; k < 7; ++k)
        a[k] += b[k];
}

// one owner finishes a run: apply it, or queue it for the replay, or mark it
// long; returns FR_* bits (a value, not references, so that the flags stay
// in registers)
#define FR_RISKY 1u    // to the in-order replay
#define FR_LONG 2u     // to the tail kernel
#define FR_UNLINKED 4u // long and unlinked: to the replay first
template <bool MULTI>
__device__ static inline uint32_t finish_run(bool act, uint32_t sid, uint32_t b, uint64_t len, const uint32_t a[7],
                                             const qs_sdesc *__restrict__ D, uint32_t *low_pool,
                                             uint32_t *high_pool, const uint32_t *mask_pool, uint32_t *longbm)
{
    if (!act)
        return 0u;
    seg_tables tb = tables_of<MULTI>(sid, D, low_pool, high_pool, mask_pool);
    if (len >= 256u) {
        set_long<MULTI>(longbm, D, sid, b);
        return FR_LONG | (tb.low[(size_t)b * 8u + 7u] <= QS_LINK_START ? FR_UNLINKED : 0u);
    }
    return apply_segment(tb.low, tb.high, b, a) ? FR_RISKY : 0u;
}

// The same finish in two phases, so that the bucket loads of all the runs a
// lane finishes in one window are in flight together (the runs are distinct
// cells, so no load depends on another run's store): load_run reads the
// bucket's two 16-byte words, apply_run then does what finish_run does.
template <bool MULTI>
__device__ static inline void load_run(bool act, uint64_t cell, const qs_pack &pk, const qs_sdesc *__restrict__ D,
                                       uint32_t *low_pool, uint32_t *high_pool, const uint32_t *mask_pool,
                                       uint4 &x, uint4 &y)
{
    if (!act)
        return;
    const seg_tables tb = tables_of<MULTI>((uint32_t)(cell >> pk.bb), D, low_pool, high_pool, mask_pool);
    const uint4 *bp = reinterpret_cast<const uint4 *>(tb.low + (size_t)(uint32_t)(cell & pk.bmask) * 8u);
    x = bp[0];
    y = bp[1];
}

template <bool MULTI>
__device__ static inline uint32_t apply_run(bool act, uint64_t cell, uint64_t len, const uint32_t a[7], uint4 x,
                                            uint4 y, const qs_pack &pk, const qs_sdesc *__restrict__ D,
                                            uint32_t *low_pool, uint32_t *high_pool, const uint32_t *mask_pool,
                                            uint32_t *longbm)
{
    if (!act)
        return 0u;
    const uint32_t sid = (uint32_t)(cell >> pk.bb), b = (uint32_t)(cell & pk.bmask);
    if (len >= 256u) {
        set_long<MULTI>(longbm, D, sid, b);
        return FR_LONG | (y.w <= QS_LINK_START ? FR_UNLINKED : 0u);
    }
    const seg_tables tb = tables_of<MULTI>(sid, D, low_pool, high_pool, mask_pool);
    return apply_loaded(tb.low, tb.high, b, a, x, y) ? FR_RISKY : 0u;
}

// (A single sketch's shapes of up to 4 items per lane are held to 168
// registers, three wavefronts per SIMD, which they fit without spilling.)
template <bool MULTI, int W, int IPL>
__global__ __launch_bounds__(SEG_THREADS) __attribute__((amdgpu_waves_per_eu(!MULTI && IPL <= 4 ? 3 : 1))) void k_segw(
    const uint64_t *__restrict__ IT, uint32_t n, const qs_sdesc *__restrict__ D, uint32_t n_masks0,
    uint32_t *__restrict__ low_pool, uint32_t *__restrict__ high_pool,
    const uint32_t *__restrict__ mask_pool, qs_pack pk, uint32_t *__restrict__ longbm,
    uint32_t *__restrict__ ord, uint32_t *__restrict__ lng, qs_dev_state *__restrict__ st)
{
    constexpr uint32_t T = (uint32_t)W * IPL;
    __shared__ uint32_t s_masks[MULTI ? 1 : QS_MAX_MASKS];
    if (!MULTI)
        for (uint32_t i = threadIdx.x; i < n_masks0; i += SEG_THREADS)
            s_masks[i] = mask_pool[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gl = lane & (uint32_t)(W - 1);
    const uint64_t gmask = (W == 64) ? ~0ull : (((1ull << W) - 1ull) << (lane - gl));
    const uint64_t gid = ((uint64_t)blockIdx.x * SEG_THREADS + threadIdx.x) / W;
    const uint64_t ngroups = (uint64_t)gridDim.x * SEG_THREADS / W;
    // The next window's items (and, in the first and last lane of the group,
    // the items either side of it) are loaded while this window is processed.
    // The loads are unconditional, from clamped positions, so that the wait
    // for this window's items does not also wait for them.
    // Each lane also loads one of the W items after the window: a run that
    // continues past the window end and stops within them is finished
    // without another load.
    uint64_t nit[IPL], nedge, nxit;
    const uint64_t last = (uint64_t)n - 1u;
    auto edge_pos = [&](uint64_t w) -> uint64_t {
        const uint64_t e = gl == 0 ? (w > 0 ? w - 1u : 0u) : w + T; // only lanes 0 and W - 1 use it
        return e < last ? e : last;
    };
    auto ext_pos = [&](uint64_t w) -> uint64_t {
        const uint64_t e = w + T + gl;
        return e < last ? e : last;
    };
    {
        const uint64_t w0 = gid * T, p0 = w0 + (uint64_t)gl * IPL;
#pragma unroll
        for (int k = 0; k < IPL; ++k)
            nit[k] = IT[p0 + k < last ? p0 + k : last];
        nedge = IT[edge_pos(w0)];
        nxit = IT[ext_pos(w0)];
    }
    for (uint64_t w0 = gid * T; w0 < n; w0 += ngroups * T) {
        const uint64_t p0 = w0 + (uint64_t)gl * IPL;
        uint64_t c[IPL];
        uint32_t M[IPL];
        uint64_t itk[IPL];
        const uint64_t edge = nedge, xit = nxit;
#pragma unroll
        for (int k = 0; k < IPL; ++k)
            itk[k] = nit[k];
        {
            const uint64_t w1 = w0 + ngroups * T, p1 = w1 + (uint64_t)gl * IPL;
#pragma unroll
            for (int k = 0; k < IPL; ++k)
                nit[k] = IT[p1 + k < last ? p1 + k : last];
            nedge = IT[edge_pos(w1)];
            nxit = IT[ext_pos(w1)];
        }
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            const uint64_t p = p0 + k;
            const bool v = p < n;
            const uint64_t it = itk[k];
            c[k] = v ? it_cell(it, pk) : NO_CELL;
            M[k] = v ? mask_of<MULTI>(it_sid(it, pk), item_mu<MULTI>(it, pk, D), D, mask_pool, s_masks, pk.unm) : 0u;
        }
        // cell before my first item and after my last one
        uint64_t cprev = shfl64w<W>(c[IPL - 1], (int)gl - 1);
        if (gl == 0)
            cprev = (w0 > 0) ? it_cell(edge, pk) : NO_CELL;
        uint64_t cnext = shfl64w<W>(c[0], (int)gl + 1);
        if (gl == (uint32_t)(W - 1))
            cnext = (w0 + T < n) ? it_cell(edge, pk) : END_CELL;
        const bool head0 = c[0] != NO_CELL && c[0] != cprev;
        bool anyhead = head0;
#pragma unroll
        for (int k = 1; k < IPL; ++k)
            anyhead |= c[k] != NO_CELL && c[k] != c[k - 1];
        if ((__ballot(anyhead) & gmask) == 0)
            continue; // the window lies inside a run owned by an earlier window

        // lane-local runs: first run [0, f1), interior runs, last run [rs, IPL).
        // An interior run [rs, k) with rs >= 1 ends at item k - 1 (1 <= k - 1
        // <= IPL - 2); it is recorded in slot k - 2 and applied further down
        // with the lane's other runs of this window.
        constexpr int NI = IPL > 2 ? IPL - 2 : 1;
        uint32_t cur[7], af[7], tmp[7], ia[NI][7];
        uint64_t icell[NI];
        uint32_t ipos[NI];
        bool iact[NI];
#pragma unroll
        for (int j = 0; j < NI; ++j)
            iact[j] = false;
        mask_to_words(M[0], cur);
#pragma unroll
        for (int k = 0; k < 7; ++k)
            af[k] = 0;
        uint32_t rs = 0, f1 = IPL;
#pragma unroll
        for (int k = 1; k < IPL; ++k) {
            const bool bnd = c[k] != c[k - 1];
            if (k >= 2) { // (rs is 0 at k = 1: no interior run ends there)
                const int j = k >= 2 ? k - 2 : 0;
                iact[j] = bnd && rs > 0 && c[k - 1] != NO_CELL;
                icell[j] = c[k - 1];
                ipos[j] = (uint32_t)(p0 + rs);
#pragma unroll
                for (int e = 0; e < 7; ++e)
                    ia[j][e] = cur[e];
            }
            if (bnd && rs == 0) {
#pragma unroll
                for (int j = 0; j < 7; ++j)
                    af[j] = cur[j];
                f1 = (uint32_t)k;
            }
            mask_to_words(M[k], tmp);
            if (bnd) {
#pragma unroll
                for (int j = 0; j < 7; ++j)
                    cur[j] = tmp[j];
                rs = (uint32_t)k;
            } else {
                add_words(cur, tmp);
            }
        }
        const bool last_head = rs > 0 || head0; // my last run starts in this lane
        // W-wide segmented inclusive scan of the last-run sums, and an inclusive
        // max-scan of the run heads (position + 1; 0 = before the window)
        uint32_t x[7];
#pragma unroll
        for (int k = 0; k < 7; ++k)
            x[k] = cur[k];
        uint64_t hp = last_head ? (p0 + rs + 1u) : 0ull;
        bool f = last_head || gl == 0;
#pragma unroll
        for (int o = 1; o < W; o <<= 1) {
            const bool fu = __shfl_up((int)f, o, W) != 0;
            const uint64_t hu = shfl64w<W>(hp, (int)gl - o);
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const uint32_t y = __shfl_up(x[k], o, W);
                if (gl >= (uint32_t)o && !f)
                    x[k] += y;
            }
            if (gl >= (uint32_t)o) {
                hp = hu > hp ? hu : hp;
                f = f || fu;
            }
        }
        // (a) my first run ends inside this lane
        uint32_t carry[7];
#pragma unroll
        for (int k = 0; k < 7; ++k)
            carry[k] = __shfl_up(x[k], 1, W);
        uint64_t hprev = shfl64w<W>(hp, (int)gl - 1);
        if (gl == 0)
            hprev = 0;
        const uint64_t head = head0 ? p0 + 1u : hprev;
        const bool act_a = f1 < (uint32_t)IPL && c[0] != NO_CELL && head != 0;
        if (!head0) {
#pragma unroll
            for (int k = 0; k < 7; ++k)
                af[k] += carry[k];
        }
        // (b) my last run ends with my last item
        const uint64_t cl = c[IPL - 1];
        const bool act_b = cl != NO_CELL && cl != cnext && hp != 0;
        // (c) the group's last run continues past the window: lane 0 of the
        // group finishes it below
        const bool ext = __shfl((int)(cl != NO_CELL && cl == cnext && hp != 0), W - 1, W) != 0;
        const uint64_t ce = shfl64w<W>(cl, W - 1);
        // the buckets of (a), the interior runs, (b) and (c), loaded together
        uint4 xa, ya, xb, yb, xe, ye, xi[NI], yi[NI];
        load_run<MULTI>(ext && gl == 0, ce, pk, D, low_pool, high_pool, mask_pool, xe, ye);
        load_run<MULTI>(act_a, c[0], pk, D, low_pool, high_pool, mask_pool, xa, ya);
#pragma unroll
        for (int j = 0; j < NI; ++j)
            load_run<MULTI>(iact[j], icell[j], pk, D, low_pool, high_pool, mask_pool, xi[j], yi[j]);
        load_run<MULTI>(act_b, cl, pk, D, low_pool, high_pool, mask_pool, xb, yb);
        const uint32_t fr_a = apply_run<MULTI>(act_a, c[0], act_a ? p0 + f1 - (head - 1u) : 0u, af, xa, ya, pk, D,
                                               low_pool, high_pool, mask_pool, longbm);
        uint32_t fr_i[NI];
#pragma unroll
        for (int j = 0; j < NI; ++j) // an interior run is shorter than IPL items: never long
            fr_i[j] = apply_run<MULTI>(iact[j], icell[j], 1u, ia[j], xi[j], yi[j], pk, D, low_pool, high_pool,
                                       mask_pool, longbm);
        const uint32_t fr_b = apply_run<MULTI>(act_b, cl, act_b ? p0 + IPL - (hp - 1u) : 0u, x, xb, yb, pk, D,
                                               low_pool, high_pool, mask_pool, longbm);
        // (c): the run's items in the extension window [w0 + T, w0 + 2T) are
        // a prefix of it (the items are sorted by cell).  When fewer than W
        // of the W items loaded after the window belong to the run, they are
        // all of it; otherwise the extension window is read.
        uint32_t fr_e = 0u, hpos_e = 0u;
        if (ext) {
            const uint64_t he = shfl64w<W>(hp, W - 1);
            uint32_t e[7];
#pragma unroll
            for (int k = 0; k < 7; ++k)
                e[k] = __shfl(x[k], W - 1, W);
            const uint32_t sid = (uint32_t)(ce >> pk.bb), b = (uint32_t)(ce & pk.bmask);
            uint32_t cnt = 0, sum[7];
#pragma unroll
            for (int k = 0; k < 7; ++k)
                sum[k] = 0;
            const bool in1 = w0 + T + gl < n && it_cell(xit, pk) == ce;
            if ((__ballot(in1) & gmask) != gmask) {
                if (in1) {
                    mask_to_words(mask_of<MULTI>(sid, item_mu<MULTI>(xit, pk, D), D, mask_pool, s_masks, pk.unm), sum);
                    cnt = 1;
                }
            } else {
#pragma unroll
                for (int k = 0; k < IPL; ++k) {
                    const uint64_t q = w0 + T + (uint64_t)gl * IPL + k;
                    if (q < n) {
                        const uint64_t iq = IT[q];
                        if (it_cell(iq, pk) == ce) {
                            mask_to_words(mask_of<MULTI>(sid, item_mu<MULTI>(iq, pk, D), D, mask_pool, s_masks, pk.unm),
                                          tmp);
                            add_words(sum, tmp);
                            ++cnt;
                        }
                    }
                }
            }
#pragma unroll
            for (int o = W / 2; o > 0; o >>= 1) {
                cnt += (uint32_t)__shfl_xor((int)cnt, o, W);
#pragma unroll
                for (int k = 0; k < 7; ++k)
                    sum[k] += (uint32_t)__shfl_xor((int)sum[k], o, W);
            }
            add_words(e, sum);
            const uint64_t len = w0 + T + cnt - (he - 1u);
            const bool all_in = cnt == T;
            if (gl == 0) {
                if (all_in) {
                    set_long<MULTI>(longbm, D, sid, b);
                    fr_e = FR_LONG | (ye.w <= QS_LINK_START ? FR_UNLINKED : 0u);
                } else {
                    fr_e = apply_run<MULTI>(true, ce, len, e, xe, ye, pk, D, low_pool, high_pool, mask_pool, longbm);
                }
            }
            hpos_e = (uint32_t)(he - 1u);
        }
        // The window's list entries (segment heads): the replay list takes the
        // risky runs and the long unlinked ones, the long list the long ones.
        // One 64-bit atomic on the adjacent counters (n_ord, n_long) reserves
        // the whole wavefront's entries; neither count can carry into the
        // other (each list holds at most one chunk, under 2^31 entries).
        const uint32_t hpos_a = (uint32_t)(head - 1u), hpos_b = (uint32_t)(hp - 1u);
        const bool oa = (fr_a & (FR_RISKY | FR_UNLINKED)) != 0u, la = (fr_a & FR_LONG) != 0u;
        const bool ob = (fr_b & (FR_RISKY | FR_UNLINKED)) != 0u, lb = (fr_b & FR_LONG) != 0u;
        const bool oe = (fr_e & (FR_RISKY | FR_UNLINKED)) != 0u, le = (fr_e & FR_LONG) != 0u;
        uint32_t n_o = (uint32_t)oa + (uint32_t)ob + (uint32_t)oe;
#pragma unroll
        for (int j = 0; j < NI; ++j)
            n_o += (fr_i[j] & FR_RISKY) != 0u ? 1u : 0u;
        const uint32_t n_l = (uint32_t)la + (uint32_t)lb + (uint32_t)le;
        static_assert(IPL + 1 < 16, "replay entries per lane fit 4 bits");
        // Prefix counts from one ballot per count bit: lanes of groups that
        // skipped this window are inactive here and count as zero.
        if (__ballot((n_o | n_l) != 0u)) {
            const uint64_t below = (1ull << lane) - 1ull;
            uint32_t eo = 0, el = 0, to = 0, tl = 0;
#pragma unroll
            for (int bit = 0; bit < 4; ++bit) {
                const uint64_t m = __ballot(((n_o >> bit) & 1u) != 0u);
                eo += (uint32_t)__popcll(m & below) << bit;
                to += (uint32_t)__popcll(m) << bit;
            }
#pragma unroll
            for (int bit = 0; bit < 2; ++bit) {
                const uint64_t m = __ballot(((n_l >> bit) & 1u) != 0u);
                el += (uint32_t)__popcll(m & below) << bit;
                tl += (uint32_t)__popcll(m) << bit;
            }
            const uint32_t leader = (uint32_t)__builtin_ctzll(__ballot(true));
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(reinterpret_cast<unsigned long long *>(&st->n_ord),
                                 ((unsigned long long)tl << 32) | to);
            base = shfl64(base, (int)leader);
            uint32_t po = (uint32_t)base + eo, pl = (uint32_t)(base >> 32) + el;
            if (oa)
                ord[po++] = hpos_a;
#pragma unroll
            for (int j = 0; j < NI; ++j)
                if ((fr_i[j] & FR_RISKY) != 0u)
                    ord[po++] = ipos[j];
            if (ob)
                ord[po++] = hpos_b;
            if (oe)
                ord[po] = hpos_e;
            if (la)
                lng[pl++] = hpos_a;
            if (lb)
                lng[pl++] = hpos_b;
            if (le)
                lng[pl] = hpos_e;
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
// A wavefront stages its allocations and newly long segments in LDS and
// reserves their list slots TBUF at a time: one global atomic per segment on
// a single counter would serialise (tens of thousands of segments replay in a
// chunk that links the hot buckets).
#define TBUF 8u

// copy a wave's ntb staged allocations to list slots base, base + 1, ...
__device__ static inline void write_trig(qs_trig (*s_tr)[TBUF], uint32_t wv, uint32_t lane, uint32_t ntb,
                                         uint32_t base, qs_trig *__restrict__ trig, uint32_t trig_cap,
                                         qs_dev_state *__restrict__ st)
{
    for (uint32_t e = 0; e < ntb; ++e) {
        const uint32_t slot = base + e;
        if (slot < trig_cap) {
            const uint32_t *src = reinterpret_cast<const uint32_t *>(&s_tr[wv][e]);
            uint32_t *dst = reinterpret_cast<uint32_t *>(trig + slot);
            if (lane < sizeof(qs_trig) / 4u)
                dst[lane] = src[lane];
        } else if (lane == 0) {
            atomicOr(&st->err, QS_EF_POOL); // more allocations than the pools hold
        }
    }
}

__device__ static inline void flush_trig(qs_trig (*s_tr)[TBUF], uint32_t wv, uint32_t lane, uint32_t &ntb,
                                         qs_trig *__restrict__ trig, uint32_t trig_cap,
                                         qs_dev_state *__restrict__ st)
{
    if (!ntb)
        return;
    uint32_t base = 0;
    if (lane == 0)
        base = atomicAdd(&st->n_trig, ntb);
    base = __shfl(base, 0, 64);
    write_trig(s_tr, wv, lane, ntb, base, trig, trig_cap, st);
    ntb = 0;
}

__device__ static inline void flush_long(uint32_t (*s_ln)[TBUF], uint32_t wv, uint32_t lane, uint32_t &nlb,
                                         uint32_t *__restrict__ lng, qs_dev_state *__restrict__ st)
{
    if (!nlb)
        return;
    uint32_t base = 0;
    if (lane == 0)
        base = atomicAdd(&st->n_long, nlb);
    base = __shfl(base, 0, 64);
    if (lane < nlb)
        lng[base + lane] = s_ln[wv][lane];
    nlb = 0;
}
__global__ __launch_bounds__(256) void k_ordered(const uint64_t *__restrict__ IT, uint32_t n,
                                                 const uint32_t *__restrict__ ord,
                                                 const qs_sdesc *__restrict__ D,
                                                 uint32_t *__restrict__ low_pool,
                                                 const uint32_t *__restrict__ mask_pool, qs_pack pk,
                                                 uint32_t *__restrict__ longbm,
                                                 uint32_t *__restrict__ lng,
                                                 qs_trig *__restrict__ trig, uint32_t trig_cap,
                                                 uint32_t one_masks, qs_dev_state *__restrict__ st)
{
    __shared__ qs_trig s_tr[4][TBUF];
    __shared__ uint32_t s_ln[4][TBUF];
    __shared__ uint32_t s_masks[QS_MAX_MASKS]; // a single sketch's masks (one_masks of them)
    for (uint32_t i = threadIdx.x; i < one_masks; i += blockDim.x)
        s_masks[i] = mask_pool[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (blockDim.x / 64u);
    const uint32_t n_ord = st->n_ord;
    uint32_t ntb = 0, nlb = 0;
    unsigned long long replayed_all = 0;
    // Loaded ahead, unconditionally from clamped positions so that waits do
    // not cover them: the next segment's first batch of items, the head of
    // the segment after it, and the next batch of this segment's items.
    const uint32_t last = n - 1u, olast = n_ord ? n_ord - 1u : 0u;
    auto head_at = [&](uint32_t h) -> uint32_t { return ord[h < olast ? h : olast]; };
    uint32_t p0n = head_at(gw), p0nn = head_at(gw + nw);
    uint64_t iqn = IT[p0n + lane < last ? p0n + lane : last];
    for (uint32_t h = gw; h < n_ord; h += nw) {
        const uint32_t p0 = p0n;
        uint64_t iq = iqn; // the first batch; lane 0 holds the head
        p0n = p0nn;
        p0nn = head_at(h + 2u * nw);
        iqn = IT[p0n + lane < last ? p0n + lane : last];
        const uint64_t cell = it_cell(shfl64(iq, 0), pk);
        const uint32_t b = (uint32_t)(cell & pk.bmask);
        const uint32_t sid = (uint32_t)(cell >> pk.bb);
        const qs_sdesc &d = D[sid];
        uint32_t *bk = low_pool + d.low_off + (size_t)b * 8u;
        const uint32_t *masks = mask_pool + d.masks_off;
        uint32_t word = lane < 7u ? bk[lane] : 0u;
        uint32_t c = __shfl(word, (int)(lane >> 2), 64);
        c = lane < 28u ? ((c >> ((lane & 3u) * 8u)) & 0xffu) : 0u;
        uint32_t hc = 0;
        bool fired = false, more = false;
        uint32_t t_idx = 0, tail_start = 0, replayed = 0;
        for (uint32_t p = p0;; p += 64u) {
            const uint64_t iqc = iq; // this batch's item
            iq = IT[p + 64u + lane < last ? p + 64u + lane : last];
            const uint32_t q = p + lane;
            const bool valid = q < n && it_cell(iqc, pk) == cell;
            uint32_t M = 0, idx = 0;
            if (valid) {
                const uint32_t mu = pk.quot ? qs_fastmod(it_mu(iqc, pk) * d.n_low + b, d.fd_masks) : it_mu(iqc, pk);
                M = one_masks ? s_masks[mu] : masks[mu];
                idx = it_idx(iqc, pk);
            }
            const uint64_t vm = __ballot(valid);
            if (vm == 0)
                break;
            replayed += (uint32_t)__popcll(vm);
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
            if (d.paper)
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
        replayed_all += replayed;
        if (fired) {
            // stage the allocation; its list slot is reserved later
            if (lane < 32u)
                s_tr[wv][ntb].hc[lane] = hc;
            if (lane == 0) {
                s_tr[wv][ntb].idx = t_idx;
                s_tr[wv][ntb].bucket = b;
                s_tr[wv][ntb].tail_start = tail_start;
                s_tr[wv][ntb].sid = sid;
            }
            ++ntb;
            // the segment continues after the batch of its allocation: make
            // sure the tail kernel sees it
            if (more && tail_start < n) {
                uint32_t newly = 0;
                if (lane == 0) {
                    const uint64_t bit = d.bm_off + b;
                    const uint32_t old = atomicOr(longbm + (bit >> 5), 1u << (bit & 31u));
                    newly = (old & (1u << (bit & 31u))) ? 0u : 1u;
                }
                if (__shfl((int)newly, 0, 64)) {
                    if (lane == 0)
                        s_ln[wv][nlb] = p0;
                    ++nlb;
                }
            }
            if (ntb == TBUF)
                flush_trig(s_tr, wv, lane, ntb, trig, trig_cap, st);
            if (nlb == TBUF)
                flush_long(s_ln, wv, lane, nlb, lng, st);
        }
    }
    // Final flush, one reservation per list per block: same-address atomics
    // serialise (~11 ns each), and one per wavefront and list costs more than
    // the replay itself.
    __shared__ uint32_t s_nt[4], s_nl[4], s_base[2];
    __shared__ unsigned long long s_rep[4];
    if (lane == 0) {
        s_nt[wv] = ntb;
        s_nl[wv] = nlb;
        s_rep[wv] = replayed_all;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tt = s_nt[0] + s_nt[1] + s_nt[2] + s_nt[3];
        const uint32_t tl = s_nl[0] + s_nl[1] + s_nl[2] + s_nl[3];
        const unsigned long long rk = s_rep[0] + s_rep[1] + s_rep[2] + s_rep[3];
        s_base[0] = tt ? atomicAdd(&st->n_trig, tt) : 0u;
        s_base[1] = tl ? atomicAdd(&st->n_long, tl) : 0u;
        if (rk)
            atomicAdd(&st->replay_keys, rk);
    }
    __syncthreads();
    uint32_t bt = s_base[0], bl = s_base[1];
    for (uint32_t w = 0; w < wv; ++w) {
        bt += s_nt[w];
        bl += s_nl[w];
    }
    write_trig(s_tr, wv, lane, ntb, bt, trig, trig_cap, st);
    if (lane < nlb)
        lng[bl + lane] = s_ln[wv][lane];
}

// trigger sort key: sketch << ib | stream index (fits 32 bits by construction)
__global__ void k_trig_keys(const qs_trig *__restrict__ trig, const qs_dev_state *__restrict__ st,
                            uint32_t trig_cap, uint32_t ib, uint32_t *__restrict__ tk,
                            uint32_t *__restrict__ tv)
{
    const uint32_t n = min(st->n_trig, trig_cap);
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        tk[r] = (trig[r].sid << ib) | trig[r].idx;
        tv[r] = r;
    }
}

// first position of sketch sid among the sorted trigger keys
__device__ static inline uint32_t group_start(const uint32_t *__restrict__ sk, uint32_t n, uint32_t key0)
{
    uint32_t lo = 0, hi = n; // first index with sk[i] >= key0
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (sk[mid] < key0)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Ids in stream order of the triggering inserts, per sketch (the sequential
// bump allocator of v3:68-85, next starting at 1025).
__global__ __launch_bounds__(256) void k_assign(const qs_trig *__restrict__ trig,
                                                const uint32_t *__restrict__ skey,
                                                const uint32_t *__restrict__ order, uint32_t trig_cap,
                                                uint32_t ib, const qs_sdesc *__restrict__ D,
                                                uint32_t *__restrict__ low_pool,
                                                uint32_t *__restrict__ high_pool,
                                                uint32_t *__restrict__ tail_pool,
                                                const uint32_t *__restrict__ next,
                                                qs_dev_state *__restrict__ st)
{
    const uint32_t n_trig = min(st->n_trig, trig_cap);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * 4u;
    for (uint32_t r = gw; r < n_trig; r += nw) {
        const qs_trig *tr = trig + order[r];
        const uint32_t sid = tr->sid;
        // sketch 0's keys sort first: its group starts at 0 (every trigger of a
        // single sketch skips the search)
        const uint32_t start = sid == 0 ? 0u : group_start(skey, n_trig, sid << ib);
        const qs_sdesc &d = D[sid];
        const uint32_t id = next[sid] + (r - start);
        const uint32_t row = id - QS_LINK_START;
        if (row >= d.n_high) {
            if (lane == 0)
                atomicOr(&st->err, QS_EF_POOL);
            continue;
        }
        if (lane < 32u)
            high_pool[d.high_off + (size_t)row * 32u + lane] = tr->hc[lane];
        if (lane == 0) {
            low_pool[d.low_off + (size_t)tr->bucket * 8u + 7u] = id;
            tail_pool[d.tail_off + row] = tr->tail_start;
        }
    }
}

// Long segments: every insert of a bucket linked before this chunk, and the
// inserts after the allocation batch of a bucket linked in this chunk, are
// commutative high adds.  A wavefront walks a window of TAIL_WINDOW sorted
// positions in steps of 64 (loading TAIL_PF steps ahead); each lane counts its
// inserts' counters in its own LDS row, two 16-bit counters per word (at most
// one insert per step, so a counter holds at most 64 in a window), with one
// LDS add per selected counter; when the bucket changes the wavefront sums the
// rows (at most 64 x 64 per counter) and adds them with one atomic per counter.
// (Expanding each mask into packed bytes in registers cost more VALU issue:
// 0.99 ms against 0.87 ms for this form at 2^28 Zipf keys.)
#define TAIL_PF 4

#define TROW 15 // 28 counters per lane, two 16-bit halves per word, odd stride
__device__ static inline void tail_flush(uint32_t (*rows)[TROW], uint32_t lane, uint32_t *row)
{
    __builtin_amdgcn_wave_barrier();
    if (lane < 28u) {
        uint32_t v = 0;
#pragma unroll 8
        for (uint32_t l = 0; l < 64u; ++l)
            v += rows[l][lane >> 1];
        v = (lane & 1u) ? (v >> 16) : (v & 0xffffu);
        if (v)
            atomicAdd(row + lane, v);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = 0; j < 14; ++j)
        rows[lane][j] = 0;
    __builtin_amdgcn_wave_barrier();
}

// high row that receives the tail inserts of cell (sid, b), or null; tail =
// first sorted position that goes to it
template <bool MULTI>
__device__ static inline uint32_t *tail_row(uint32_t sid, uint32_t b, const qs_sdesc *__restrict__ D,
                                            const uint32_t *__restrict__ low_pool,
                                            uint32_t *__restrict__ high_pool,
                                            const uint32_t *__restrict__ tail_pool,
                                            const uint32_t *__restrict__ longbm,
                                            const uint32_t *__restrict__ first_new, uint32_t &tail)
{
    tail = 0xffffffffu;
    if (!test_long<MULTI>(longbm, D, sid, b))
        return nullptr;
    const uint64_t low_off = MULTI ? D[sid].low_off : 0ull;
    const uint32_t w = low_pool[low_off + (size_t)b * 8u + 7u];
    if (w <= QS_LINK_START)
        return nullptr;
    const uint64_t tail_off = MULTI ? D[sid].tail_off : 0ull;
    const uint64_t high_off = MULTI ? D[sid].high_off : 0ull;
    tail = (w >= first_new[MULTI ? sid : 0u]) ? tail_pool[tail_off + (w - QS_LINK_START)] : 0u;
    return high_pool + high_off + (size_t)(w - QS_LINK_START) * 32u;
}

// a cell shuffle in the cell's own width
template <typename CT> __device__ static inline CT shfl_cell(CT v, int src);
template <> __device__ inline uint32_t shfl_cell<uint32_t>(uint32_t v, int src)
{
    return (uint32_t)__shfl((int)v, src, 64);
}
template <> __device__ inline uint64_t shfl_cell<uint64_t>(uint64_t v, int src)
{
    return shfl64(v, src);
}

// CT: the cell type, uint32_t when bucket and sketch bits fit in 31 (then the
// sentinels NO_CELL, END_CELL truncated to 32 bits are above every cell)
template <bool MULTI, typename CT>
__global__ __launch_bounds__(256) void k_tail(const uint64_t *__restrict__ IT, uint32_t n,
                                              const qs_sdesc *__restrict__ D, uint32_t n_masks0,
                                              const uint32_t *__restrict__ low_pool,
                                              uint32_t *__restrict__ high_pool,
                                              const uint32_t *__restrict__ mask_pool, qs_pack pk,
                                              const uint32_t *__restrict__ longbm,
                                              const uint32_t *__restrict__ tail_pool,
                                              const uint32_t *__restrict__ first_new,
                                              const uint32_t *__restrict__ lng, bool short_runs,
                                              const qs_dev_state *__restrict__ st)
{
    if (st->n_long == 0)
        return;
    // a single sketch's masks, n_masks0 words of dynamic LDS (sized by the
    // launch: 16 KiB for the largest table would hold the kernel to five
    // wavefronts per SIMD)
    extern __shared__ uint32_t s_masks[];
    if (!MULTI)
        for (uint32_t i = threadIdx.x; i < n_masks0; i += blockDim.x)
            s_masks[i] = mask_pool[i];
    __shared__ uint32_t s_rows[4][64][TROW];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t(*rows)[TROW] = s_rows[wv];
#pragma unroll
    for (int j = 0; j < 14; ++j)
        rows[lane][j] = 0;
    __syncthreads();
    const uint32_t gw = blockIdx.x * 4u + wv;
    const uint32_t nw = gridDim.x * 4u;
    for (uint64_t w0 = (uint64_t)gw * TAIL_WINDOW; w0 < n; w0 += (uint64_t)nw * TAIL_WINDOW) {
        const uint64_t wend = (w0 + TAIL_WINDOW < n) ? w0 + TAIL_WINDOW : n;
        uint32_t *cur = nullptr;
        // lookup cache: the last cell whose row was fetched
        CT c_c = (CT)NO_CELL;
        uint32_t *c_row = nullptr;
        uint32_t c_tail = 0;
        // cell of the item before the step
        CT c_prev = (w0 > 0) ? (CT)it_cell(IT[w0 - 1u], pk) : (CT)NO_CELL;
        // unconditional loads from clamped positions (validity is applied at
        // use), so that waiting for one group does not wait for the next
        const uint64_t last = (uint64_t)n - 1u;
        uint64_t buf[TAIL_PF];
#pragma unroll
        for (int k = 0; k < TAIL_PF; ++k) {
            const uint64_t q = w0 + 64u * k + lane;
            buf[k] = IT[q < last ? q : last];
        }
        for (uint64_t g = w0; g < wend; g += 64u * TAIL_PF) {
            // the next TAIL_PF steps' items are loaded while these are processed
            uint64_t nb[TAIL_PF];
#pragma unroll
            for (int k = 0; k < TAIL_PF; ++k) {
                const uint64_t q = g + 64u * (TAIL_PF + k) + lane;
                nb[k] = IT[q < last ? q : last];
            }
#pragma unroll
            for (int k = 0; k < TAIL_PF; ++k) {
                const uint64_t s0 = g + 64u * k;
                if (s0 >= wend)
                    break;
                const uint32_t p = (uint32_t)s0 + lane;
                const bool valid = p < wend;
                const uint64_t it = buf[k];
                const uint64_t nx = (k + 1 < TAIL_PF) ? buf[k + 1] : nb[0]; // the next step's items
                const CT c = valid ? (CT)it_cell(it, pk) : (CT)NO_CELL;
                const uint32_t last_lane = (wend - s0 < 64u) ? (uint32_t)(wend - s0 - 1u) : 63u;
                const CT c_first = shfl_cell<CT>(c, 0);
                const CT c_last = shfl_cell<CT>(c, (int)last_lane);
                const CT nxc = (CT)it_cell(nx, pk);
                const CT nx0 = shfl_cell<CT>(nxc, 0); // taken by the whole wavefront
                const CT c_after = (s0 + 64u < wend) ? nx0
                                   : (wend < n)      ? (CT)it_cell(IT[wend], pk)
                                                     : (CT)END_CELL;
                // This pass takes the runs that cross a step boundary (hold two
                // positions 64k - 1 and 64k); in every step such a run touches it
                // holds the first or the last item.  Runs inside one step are
                // left to the list pass below.
                const bool cross_f = c_first == c_prev || (c_first == c_last && c_first == c_after);
                const bool cross_l = c_last != c_first && c_last == c_after;
                c_prev = c_last;
                uint32_t *row_f = nullptr, *row_l = nullptr;
                uint32_t tail_f = 0, tail_l = 0;
                if (cross_f) {
                    if (c_first != c_c) {
                        c_c = c_first;
                        c_row = tail_row<MULTI>((uint32_t)(c_c >> pk.bb), (uint32_t)(c_c & pk.bmask), D, low_pool,
                                                high_pool, tail_pool, longbm, first_new, c_tail);
                    }
                    row_f = c_row;
                    tail_f = c_tail;
                }
                if (cross_l) {
                    if (c_last != c_c) {
                        // Unless the segment pass marks runs under 65 items long,
                        // a run of at most 64 items has no tail inserts (the
                        // replay's tails start 64 items after the head): its
                        // length is its items here plus the next step's leading
                        // ones, when the run ends there.
                        bool need = true;
                        if (!short_runs && s0 + 64u < wend) {
                            const bool nv = s0 + 64u + lane < wend;
                            const uint64_t nm = __ballot(nv && nxc == c_last);
                            const uint32_t b = (~nm) ? (uint32_t)__builtin_ctzll(~nm) : 64u;
                            const uint32_t a = (uint32_t)__popcll(__ballot(valid && c == c_last));
                            need = a + b > 64u;
                        }
                        c_c = c_last;
                        c_row = nullptr;
                        c_tail = 0;
                        if (need)
                            c_row = tail_row<MULTI>((uint32_t)(c_c >> pk.bb), (uint32_t)(c_c & pk.bmask), D,
                                                    low_pool, high_pool, tail_pool, longbm, first_new, c_tail);
                    }
                    row_l = c_row;
                    tail_l = c_tail;
                }
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    uint32_t *const row = side ? row_l : row_f;
                    if (!row)
                        continue;
                    const bool act = valid && c == (side ? c_last : c_first) && p >= (side ? tail_l : tail_f);
                    if (!__ballot(act))
                        continue;
                    if (row != cur) {
                        if (cur)
                            tail_flush(rows, lane, cur);
                        cur = row;
                    }
                    if (act) {
                        uint32_t Mi = mask_of<MULTI>((uint32_t)(c >> pk.bb), item_mu<MULTI>(it, pk, D), D, mask_pool,
                                                     s_masks, pk.unm);
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            if (Mi) {
                                const uint32_t j = (uint32_t)__builtin_ctz(Mi);
                                atomicAdd(&rows[lane][j >> 1], 1u << (16u * (j & 1u))); // this lane's own row
                                Mi &= Mi - 1u;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < TAIL_PF; ++k)
                buf[k] = nb[k];
        }
        if (cur)
            tail_flush(rows, lane, cur);
    }
    // List pass: a long run inside one step.  The segment pass marks runs of
    // at least T + 1 items long (T = its window), so these exist only for
    // windows under 64 items.  The replay also marks runs that continue after
    // their allocation batch, but their tail starts 64 items after the head,
    // outside a run that fits in one step.
    if (!short_runs)
        return;
    const uint32_t n_long = st->n_long;
    for (uint32_t e = gw; e < n_long; e += nw) {
        const uint32_t h = lng[e]; // head of the run
        const uint32_t p = h + lane;
        const uint64_t it = p < n ? IT[p] : 0ull;
        const uint64_t c = p < n ? it_cell(it, pk) : NO_CELL;
        const uint64_t ch = shfl64(c, 0);
        const uint64_t out = __ballot(c != ch);
        const uint32_t len = out ? (uint32_t)__builtin_ctzll(out) : 64u;
        bool cross = (h >> 6) != ((h + len - 1u) >> 6);
        if (len == 64u && !cross)
            cross = h + 64u < n && it_cell(IT[h + 64u], pk) == ch;
        if (cross)
            continue; // taken by the window pass
        uint32_t tail = 0;
        uint32_t *row = tail_row<MULTI>((uint32_t)(ch >> pk.bb), (uint32_t)(ch & pk.bmask), D, low_pool, high_pool,
                                        tail_pool, longbm, first_new, tail);
        if (!row)
            continue;
        const bool act = lane < len && p >= tail;
        const uint32_t M = act ? mask_of<MULTI>((uint32_t)(ch >> pk.bb), item_mu<MULTI>(it, pk, D), D, mask_pool, s_masks, pk.unm)
                               : 0u;
        uint32_t cnt = 0;
#pragma unroll
        for (int j = 0; j < 28; ++j) {
            const uint32_t k = (uint32_t)__popcll(__ballot((M >> j) & 1u));
            cnt = lane == (uint32_t)j ? k : cnt;
        }
        if (lane < 28u && cnt)
            atomicAdd(row + lane, cnt);
    }
}

__global__ void k_clear_long(const uint32_t *__restrict__ lng, const uint64_t *__restrict__ IT,
                             const qs_sdesc *__restrict__ D, qs_pack pk, uint32_t *__restrict__ longbm,
                             const qs_dev_state *__restrict__ st)
{
    const uint32_t n_long = st->n_long;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_long; i += gridDim.x * blockDim.x) {
        const uint64_t it = IT[lng[i]];
        const uint64_t bit = D[it_sid(it, pk)].bm_off + it_b(it, pk);
        atomicAnd(longbm + (bit >> 5), ~(1u << (bit & 31u)));
    }
}

// next[sid] += allocations of sid in this chunk: the last trigger of each
// sketch group writes its id + 1.
__global__ void k_next(const qs_trig *__restrict__ trig, const uint32_t *__restrict__ skey,
                       const uint32_t *__restrict__ order, uint32_t trig_cap, uint32_t ib,
                       uint32_t *__restrict__ next, const qs_dev_state *__restrict__ st)
{
    const uint32_t n_trig = min(st->n_trig, trig_cap);
    if (st->err)
        return;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n_trig; r += gridDim.x * blockDim.x) {
        const uint32_t sid = skey[r] >> ib;
        if (r + 1 == n_trig || (skey[r + 1] >> ib) != sid) {
            const uint32_t start = group_start(skey, n_trig, sid << ib);
            next[sid] += r - start + 1;
        }
    }
}

__global__ void k_reset(qs_dev_state *st)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->ord_segments += st->n_ord;
        st->long_segments += st->n_long;
        st->chunks_ord += st->n_ord ? 1u : 0u;
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
    // tile-major counts + group sums + digit totals
    uint64_t a = OS_BINS * tiles + OS_BINS * ((tiles + SCAN_TPB - 1) / SCAN_TPB) + OS_BINS;
    uint64_t b = qs_sort_hist_words((uint32_t)(chunk < 0xffffffffull ? chunk : 0xffffffffull));
    return a > b ? a : b;
}

void qs_free_scratch(qs_sketch *s)
{
    qs_scratch &sc = s->sc;
    for (int k = 0; k < 2; ++k) {
        hipFree(sc.it[k]);
        hipFree(sc.tk[k]);
        hipFree(sc.tv[k]);
    }
    hipFree(sc.hist);
    hipFree(sc.dig);
    hipFree(sc.ord);
    hipFree(sc.lng);
    hipFree(sc.trig);
    memset(&sc, 0, sizeof(sc));
}

uint32_t qs_trig_cap(const qs_sketch *s, uint64_t chunk)
{
    uint64_t pool = 0;
    for (uint32_t i = 0; i < s->nsk; ++i)
        pool += s->ps[i].n_high;
    return (uint32_t)(pool < chunk ? pool : chunk);
}

uint64_t qs_scratch_bytes(uint64_t chunk, uint32_t tcap, qs_scratch *sc, uint8_t *base)
{
    // 256-byte aligned pieces, in this order
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        uint8_t *p = base ? base + off : nullptr;
        off += (bytes + 255) & ~255ull;
        return p;
    };
    qs_scratch t;
    memset(&t, 0, sizeof(t));
    for (int k = 0; k < 2; ++k)
        t.it[k] = (uint64_t *)take(sizeof(uint64_t) * chunk);
    t.hist = (uint32_t *)take(sizeof(uint32_t) * qs_scratch_hist_words(chunk));
    t.dig = take((chunk + OS_TILE - 1) / OS_TILE * OS_TILE);
    t.ord = (uint32_t *)take(sizeof(uint32_t) * chunk);
    t.lng = (uint32_t *)take(sizeof(uint32_t) * (chunk / 8 + 64));
    t.trig = (qs_trig *)take(sizeof(qs_trig) * tcap);
    for (int k = 0; k < 2; ++k) {
        t.tk[k] = (uint32_t *)take(sizeof(uint32_t) * tcap);
        t.tv[k] = (uint32_t *)take(sizeof(uint32_t) * tcap);
    }
    t.chunk = chunk;
    t.trig_cap = tcap;
    if (sc)
        *sc = t;
    return off;
}

int qs_ensure_scratch(qs_sketch *s, uint64_t chunk)
{
    grid_cap(0, 1); // query the CU count now: never while a launch sequence is captured
    if (s->sc.chunk >= chunk)
        return QS_OK;
    qs_free_scratch(s);
    const uint32_t tcap = qs_trig_cap(s, chunk);
    qs_scratch &sc = s->sc;
    sc.trig_cap = tcap;
    for (int k = 0; k < 2; ++k) {
        if (hipMalloc(&sc.it[k], sizeof(uint64_t) * chunk) != hipSuccess ||
            hipMalloc(&sc.tk[k], sizeof(uint32_t) * tcap) != hipSuccess ||
            hipMalloc(&sc.tv[k], sizeof(uint32_t) * tcap) != hipSuccess)
            return QS_ERR_NOMEM;
    }
    if (hipMalloc(&sc.hist, sizeof(uint32_t) * qs_scratch_hist_words(chunk)) != hipSuccess ||
        hipMalloc(&sc.dig, (chunk + OS_TILE - 1) / OS_TILE * OS_TILE) != hipSuccess ||
        hipMalloc(&sc.ord, sizeof(uint32_t) * chunk) != hipSuccess ||
        hipMalloc(&sc.lng, sizeof(uint32_t) * (chunk / 8 + 64)) != hipSuccess ||
        hipMalloc(&sc.trig, sizeof(qs_trig) * tcap) != hipSuccess)
        return QS_ERR_NOMEM;
    sc.chunk = chunk;
    return QS_OK;
}

// launch context of one batch
struct batch_ctx {
    const uint64_t *d_offs; // device batch offsets (batches of several sketches)
    bool timed;             // record timing events (never while capturing)
    const qs_scratch *sc;   // the scratch this batch runs in
};

template <int KB, bool MULTI>
static int sort_chunk(qs_sketch *s, const KeyT<KB> *keys, uint32_t cn, uint64_t base, const batch_ctx &bc,
                      hipStream_t stream, int &res)
{
    const uint32_t ntiles = (cn + OS_TILE - 1) / OS_TILE;
    const int bits = (int)(s->pk.bb + s->pk.sb);
    const int passes = (bits + 7) / 8;
    const int dbits = (bits + passes - 1) / passes; // <= 8
    const uint32_t nbins = 1u << dbits, dmask = nbins - 1u;
    const uint32_t ngrp = (ntiles + SCAN_TPB - 1) / SCAN_TPB;
    const qs_scratch &sc = *bc.sc;
    uint16_t *tcount = reinterpret_cast<uint16_t *>(sc.hist); // 16-bit tile counts (half the region)
    uint32_t *gsum = sc.hist + (size_t)OS_BINS * ntiles;
    uint32_t *totals = gsum + (size_t)OS_BINS * ngrp;
    const qs_sdesc d0 = s->h_desc[0];
    const void *in = keys;
    int out = 0;
    qs_evpair ep; // the previous pass's scatter, ended by the next pass's first record
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = dbits * ps;
        const int nshift = ps + 1 < passes ? shift + dbits : -1; // the scatter writes the next digit
        qs_evpair e0, e1, e2;
        if (bc.timed) {
            if (ps == 0)
                qs_tim_begin(s, e0, QS_TK_SORT_UP, cn, stream);
            else
                qs_tim_next(s, ep, e0, QS_TK_SORT_UP, cn, stream);
        }
        if (ps == 0)
            hipLaunchKernelGGL((k_bs_up<KB, MULTI>), dim3(ntiles), dim3(OS_THREADS), 0, stream, in, cn,
                               base, d0, s->d_desc, bc.d_offs, s->nsk, s->pk, shift, dmask, nbins, ntiles,
                               tcount);
        else
            hipLaunchKernelGGL(k_bs_up_dig, dim3((ntiles + UP_TILES - 1) / UP_TILES), dim3(OS_THREADS), 0, stream,
                               sc.dig, cn, ntiles, nbins, tcount);
        if (bc.timed)
            qs_tim_next(s, e0, e1, QS_TK_SORT_SCAN, cn, stream);
        hipLaunchKernelGGL(k_hist_scan_local, dim3(ngrp), dim3(256), 0, stream, tcount, ntiles, nbins, gsum);
        hipLaunchKernelGGL(k_hist_scan_groups, dim3(nbins), dim3(1024), 0, stream, gsum, ngrp, nbins, totals);
        if (bc.timed)
            qs_tim_next(s, e1, e2, QS_TK_SORT_DOWN, cn, stream);
        if (ps == 0)
            hipLaunchKernelGGL((k_bs_down<KB, true, MULTI>), dim3(ntiles), dim3(OS_THREADS), 0, stream, in, cn,
                               base, d0, s->d_desc, bc.d_offs, s->nsk, s->pk, shift, dmask, nbins, ntiles,
                               tcount, gsum, totals, sc.it[out], sc.dig, nshift, s->d_st);
        else
            hipLaunchKernelGGL((k_bs_down<KB, false, false>), dim3(ntiles), dim3(OS_THREADS), 0, stream, in, cn,
                               base, d0, s->d_desc, bc.d_offs, s->nsk, s->pk, shift, dmask, nbins, ntiles,
                               tcount, gsum, totals, sc.it[out], sc.dig, nshift, s->d_st);
        if (bc.timed) {
            if (ps + 1 == passes)
                qs_tim_end(s, e2, stream);
            else
                ep = e2;
        }
        in = sc.it[out];
        out ^= 1;
    }
    CK(hipGetLastError());
    res = out ^ 1;
    return QS_OK;
}

template <bool MULTI, int W, int IPL>
static void launch_segw(qs_sketch *s, const qs_scratch &sc, const uint64_t *IT, uint32_t cn, hipStream_t stream)
{
    const uint64_t groups = ((uint64_t)cn + W * IPL - 1) / (W * IPL);
    hipLaunchKernelGGL((k_segw<MULTI, W, IPL>), dim3(grid_cap((groups * W + SEG_THREADS - 1) / SEG_THREADS, 8)),
                       dim3(SEG_THREADS), 0, stream, IT, cn, s->d_desc, s->h_desc[0].n_masks, s->d_low, s->d_high,
                       s->d_masks, s->pk, s->d_longbm, sc.ord, sc.lng, s->d_st);
}

template <bool MULTI, int W>
static void launch_segw_ipl(qs_sketch *s, const qs_scratch &sc, const uint64_t *IT, uint32_t cn, hipStream_t stream)
{
    switch (s->seg_ipl) {
    case 1: launch_segw<MULTI, W, 1>(s, sc, IT, cn, stream); break;
    case 2: launch_segw<MULTI, W, 2>(s, sc, IT, cn, stream); break;
    case 4: launch_segw<MULTI, W, 4>(s, sc, IT, cn, stream); break;
    default: launch_segw<MULTI, W, 8>(s, sc, IT, cn, stream); break;
    }
}

bool qs_seg_shape_ok(uint32_t w, uint32_t ipl)
{
    if (w == 0 && ipl == 0)
        return true; // the 64-lane, one-item k_seg
    return (w == 8 || w == 16 || w == 32 || w == 64) && (ipl == 1 || ipl == 2 || ipl == 4 || ipl == 8);
}

template <bool MULTI>
static void launch_seg(qs_sketch *s, const qs_scratch &sc, const uint64_t *IT, uint32_t cn, hipStream_t stream)
{
    switch (s->seg_w) {
    case 8: launch_segw_ipl<MULTI, 8>(s, sc, IT, cn, stream); break;
    case 16: launch_segw_ipl<MULTI, 16>(s, sc, IT, cn, stream); break;
    case 32: launch_segw_ipl<MULTI, 32>(s, sc, IT, cn, stream); break;
    case 64: launch_segw_ipl<MULTI, 64>(s, sc, IT, cn, stream); break;
    default:
        hipLaunchKernelGGL(k_seg<MULTI>, dim3(grid_cap(((uint64_t)cn + 255) / 256, 8)), dim3(SEG_THREADS), 0,
                           stream, IT, cn, s->d_desc, s->h_desc[0].n_masks, s->d_low, s->d_high, s->d_masks, s->pk,
                           s->d_longbm, sc.ord, sc.lng, s->d_st);
        break;
    }
}

// One chunk: a fixed launch sequence with no host synchronisation; every
// count the later kernels need stays in device memory.
template <int KB, bool MULTI>
static int insert_chunk(qs_sketch *s, const KeyT<KB> *keys, uint32_t cn, uint64_t base, const batch_ctx &bc,
                        hipStream_t stream)
{
    int res = 0;
    int rc = sort_chunk<KB, MULTI>(s, keys, cn, base, bc, stream, res);
    if (rc)
        return rc;
    const qs_scratch &sc = *bc.sc;
    const uint64_t *IT = sc.it[res];
    const uint32_t nm0 = s->h_desc[0].n_masks;
    qs_evpair es, eo, ea, et;
    if (bc.timed)
        qs_tim_begin(s, es, QS_TK_SEG, cn, stream);
    launch_seg<MULTI>(s, sc, IT, cn, stream);
    if (bc.timed) {
        qs_tim_next(s, es, eo, QS_TK_ORDERED, cn, stream);
    }
    hipLaunchKernelGGL(k_ordered, dim3(grid_cap(((uint64_t)cn + 255) / 256, 8)), dim3(256), 0, stream, IT, cn,
                       sc.ord, s->d_desc, s->d_low, s->d_masks, s->pk, s->d_longbm, sc.lng, sc.trig,
                       sc.trig_cap, s->nsk == 1 ? s->h_desc[0].n_masks : 0u, s->d_st);
    if (bc.timed) {
        qs_tim_next(s, eo, ea, QS_TK_ASSIGN, cn, stream);
    }
    hipLaunchKernelGGL(k_trig_keys, dim3(grid_cap(((uint64_t)sc.trig_cap + 255) / 256, 4)), dim3(256), 0, stream,
                       sc.trig, s->d_st, sc.trig_cap, s->pk.ib, sc.tk[0], sc.tv[0]);
    // a single sketch's trigger keys are stream indices below cn (one digit
    // pass fewer for a chunk of at most 2^24 keys)
    const int tbits = s->nsk == 1 ? (int)qs_bits_for(cn) : (int)(s->pk.sb + s->pk.ib);
    uint32_t *tk[2] = {sc.tk[0], sc.tk[1]}, *tv[2] = {sc.tv[0], sc.tv[1]};
    const int tc = qs_sort_pairs_dev(tk, tv, sc.hist, &s->d_st->n_trig, sc.trig_cap,
                                     tbits > 32 ? 32 : tbits, stream);
    if (tc < 0)
        return QS_ERR_HIP;
    hipLaunchKernelGGL(k_assign, dim3(grid_cap(((uint64_t)sc.trig_cap + 3) / 4, 8)), dim3(256), 0, stream,
                       sc.trig, sc.tk[tc], sc.tv[tc], sc.trig_cap, s->pk.ib, s->d_desc, s->d_low,
                       s->d_high, s->d_tail_of, s->d_next, s->d_st);
    if (bc.timed) {
        qs_tim_next(s, ea, et, QS_TK_TAIL, cn, stream);
    }
    // k_segw windows under 64 items mark runs long that fit inside one step
    const bool short_runs = s->seg_w != 0 && s->seg_w * s->seg_ipl < 64u;
    const dim3 tgrid(grid_cap(((uint64_t)cn + TAIL_WINDOW * 4 - 1) / (TAIL_WINDOW * 4), 8));
    const size_t tail_lds = MULTI ? 0 : (size_t)nm0 * sizeof(uint32_t); // nm0 <= QS_MAX_MASKS
    if (s->pk.bb + s->pk.sb <= 31u)
        hipLaunchKernelGGL((k_tail<MULTI, uint32_t>), tgrid, dim3(256), tail_lds, stream, IT, cn, s->d_desc, nm0,
                           s->d_low, s->d_high, s->d_masks, s->pk, s->d_longbm, s->d_tail_of, s->d_next, sc.lng,
                           short_runs, s->d_st);
    else
        hipLaunchKernelGGL((k_tail<MULTI, uint64_t>), tgrid, dim3(256), tail_lds, stream, IT, cn, s->d_desc, nm0,
                           s->d_low, s->d_high, s->d_masks, s->pk, s->d_longbm, s->d_tail_of, s->d_next, sc.lng,
                           short_runs, s->d_st);
    hipLaunchKernelGGL(k_clear_long, dim3(grid_cap(((uint64_t)cn / 64 + 255) / 256, 4)), dim3(256), 0, stream,
                       sc.lng, IT, s->d_desc, s->pk, s->d_longbm, s->d_st);
    hipLaunchKernelGGL(k_next, dim3(grid_cap(((uint64_t)sc.trig_cap + 255) / 256, 4)), dim3(256), 0, stream,
                       sc.trig, sc.tk[tc], sc.tv[tc], sc.trig_cap, s->pk.ib, s->d_next, s->d_st);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, s->d_st);
    if (bc.timed)
        qs_tim_end(s, et, stream);
    CK(hipGetLastError());
    return QS_OK;
}

// Keys per chunk.  A chunk holds at most 2^ib keys of one sketch (the stream
// index field) and at most 2^30 keys.  A batch of several sketches whose every
// stream fits the index field is one chunk when it fits 2^30 keys; sketch 0's
// chunk_keys, when set, overrides that (for a set of any size).
uint64_t qs_chunk_keys(const qs_sketch *s, const uint64_t *offs)
{
    const uint64_t n = offs[s->nsk];
    const uint64_t forced = s->ps[0].chunk_keys;
    if (s->nsk > 1 && !forced && n <= (1ull << 30)) {
        uint64_t longest = 0;
        for (uint32_t i = 0; i < s->nsk; ++i)
            longest = offs[i + 1] - offs[i] > longest ? offs[i + 1] - offs[i] : longest;
        if (longest <= s->max_stream)
            return n;
    }
    uint64_t cap = forced ? forced : (1ull << 28);
    if (cap > s->max_stream)
        cap = s->max_stream;
    if (cap > (1ull << 30))
        cap = 1ull << 30;
    return cap;
}

template <int KB, bool MULTI>
static int insert_impl(qs_sketch *s, const void *d_keys, const uint64_t *offs, const uint64_t *d_offs,
                       hipStream_t stream, bool blocking, bool capture, const qs_scratch *fixed)
{
    const KeyT<KB> *keys = (const KeyT<KB> *)d_keys;
    const uint64_t n = offs[s->nsk];
    if (offs[0] != 0)
        return QS_ERR_PARAM;
    for (uint32_t i = 0; i < s->nsk; ++i)
        if (offs[i + 1] < offs[i])
            return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    if (MULTI && !d_offs)
        return QS_ERR_PARAM;
    const uint64_t cap = qs_chunk_keys(s, offs);
    const uint64_t need = n < cap ? n : cap;
    int rc = QS_OK;
    if (fixed) {
        // caller-provided scratch: it must hold the whole batch as one chunk
        // and every allocation the chunk can trigger
        if (n > cap || n > fixed->chunk || qs_trig_cap(s, n) > fixed->trig_cap)
            return QS_ERR_PARAM;
    } else {
        if (s->sc.chunk < need && s->graphs > 0)
            return QS_ERR_STATE; // a captured graph holds the current scratch buffers
        rc = qs_ensure_scratch(s, need);
        if (rc)
            return rc;
    }
    batch_ctx bc;
    bc.d_offs = d_offs;
    bc.timed = s->timing && !capture && !fixed;
    bc.sc = fixed ? fixed : &s->sc;
    for (uint64_t done = 0; done < n;) {
        const uint64_t cn64 = n - done < cap ? n - done : cap;
        const uint32_t cn = (uint32_t)cn64;
        if (bc.timed) {
            qs_evpair ev;
            qs_tim_begin(s, ev, QS_TIMING_KERNELS, cn, stream);
            qs_tim_end(s, ev, stream);
        }
        rc = insert_chunk<KB, MULTI>(s, keys + done, cn, done, bc, stream);
        if (rc)
            return rc;
        if (!capture) {
            s->chunks++;
            s->keys_inserted += cn;
        }
        done += cn;
    }
    if (blocking) {
        CK(hipMemcpyAsync(s->h_st, s->d_st, sizeof(qs_dev_state), hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        if (s->h_st->err) {
            s->sticky_err = qs_err_code(s->h_st->err);
            return s->sticky_err;
        }
    }
    return QS_OK;
}

int qs_insert_set(qs_sketch *s, const void *d_keys, const uint64_t *offs, const uint64_t *d_offs,
                  hipStream_t stream, bool blocking, bool capture, const qs_scratch *fixed)
{
    if (s->nsk == 1)
        return s->key_bytes == 4
                   ? insert_impl<4, false>(s, d_keys, offs, d_offs, stream, blocking, capture, fixed)
                   : insert_impl<8, false>(s, d_keys, offs, d_offs, stream, blocking, capture, fixed);
    return s->key_bytes == 4 ? insert_impl<4, true>(s, d_keys, offs, d_offs, stream, blocking, capture, fixed)
                             : insert_impl<8, true>(s, d_keys, offs, d_offs, stream, blocking, capture, fixed);
}
