// qs_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value) pairs.
//
// Used by the ordered fix-up of the insert path: the keys routed to the fix-up
// arrive in stream order and are sorted by low bucket, so that each bucket's
// inserts form one segment still in stream order (LSD radix sort is stable).
// It is also used to order the allocations of one chunk by stream position.
//
// One pass = 8 key bits = three kernels:
//   k_hist       per-tile digit histogram (LDS atomics), digit-major layout
//   k_scan_digit one workgroup per digit: exclusive scan over the tiles
//   k_scatter    per tile: digit base (scan of the 256 digit totals) + the
//                tile's offset + a stable in-tile rank (per-wave digit match
//                with 8 ballots, then a per-digit prefix over the 4 waves)
#include "qs_internal.h"

#define SORT_THREADS 256
#define SORT_ITEMS 8
#define SORT_TILE (SORT_THREADS * SORT_ITEMS)
#define SORT_WAVES (SORT_THREADS / QS_WAVE)

size_t qs_sort_hist_words(uint32_t max_count)
{
    size_t tiles = (max_count + SORT_TILE - 1) / SORT_TILE;
    if (tiles == 0)
        tiles = 1;
    return 256 * tiles + 256; // + per-digit totals
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_hist(const uint32_t *__restrict__ keys,
                                                            uint32_t count, uint32_t ntiles,
                                                            int shift, uint32_t *__restrict__ hist,
                                                            const uint32_t *__restrict__ d_count)
{
    __shared__ uint32_t s_cnt[256];
    const uint32_t t = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    if (d_count)
        count = *d_count < count ? *d_count : count;
    s_cnt[t] = 0;
    __syncthreads();
    const uint32_t base = tile * SORT_TILE;
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        uint32_t i = base + r * SORT_THREADS + t;
        if (i < count)
            atomicAdd(&s_cnt[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[(size_t)t * ntiles + tile] = s_cnt[t];
}

// Block d scans hist[d][0..ntiles) in place (exclusive) and writes the digit
// total to totals[d].
#define SCAN_PER_THREAD 8u

__global__ __launch_bounds__(1024) void k_sort_scan_digit(uint32_t *__restrict__ hist,
                                                          uint32_t ntiles,
                                                          uint32_t *__restrict__ totals)
{
    // each thread scans SCAN_PER_THREAD consecutive tiles, the block scans the
    // thread sums: 8192 tiles per iteration
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u, wv = t >> 6;
    uint32_t *row = hist + (size_t)blockIdx.x * ntiles;
    if (t == 0)
        s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < ntiles; base += 1024u * SCAN_PER_THREAD) {
        const uint32_t i0 = base + t * SCAN_PER_THREAD;
        uint32_t v[SCAN_PER_THREAD];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
            v[k] = (i0 + k < ntiles) ? row[i0 + k] : 0u;
            sum += v[k];
        }
        // inclusive wave scan of the thread sums
        uint32_t x = sum;
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
            uint32_t a = s_w[t];
            uint32_t sc = a;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                uint32_t y = __shfl_up(sc, o, 16);
                if ((t & 15u) >= (uint32_t)o)
                    sc += y;
            }
            s_w[t] = sc - a; // exclusive wave offsets
        }
        __syncthreads();
        const uint32_t carry = s_carry;
        uint32_t run = carry + s_w[wv] + x - sum; // exclusive prefix of my first tile
#pragma unroll
        for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
            if (i0 + k < ntiles)
                row[i0 + k] = run;
            run += v[k];
        }
        __syncthreads();
        if (t == 1023)
            s_carry = carry + s_w[wv] + x;
        __syncthreads();
    }
    if (t == 0)
        totals[blockIdx.x] = s_carry;
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint32_t count,
    uint32_t ntiles, int shift, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ totals, uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
    const uint32_t *__restrict__ d_count)
{
    __shared__ uint32_t s_base[256];
    if (d_count)
        count = *d_count < count ? *d_count : count;
    __shared__ uint32_t s_cnt[SORT_WAVES][256];
    __shared__ uint32_t s_tmp[SORT_WAVES];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u, wv = t >> 6;
    const uint32_t tile = blockIdx.x;

    // exclusive scan of the 256 digit totals -> digit base
    {
        uint32_t v = totals[t];
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o)
                x += y;
        }
        if (lane == 63)
            s_tmp[wv] = x;
        __syncthreads();
        uint32_t wo = 0;
        for (uint32_t k = 0; k < wv; ++k)
            wo += s_tmp[k];
        s_base[t] = wo + x - v + hist[(size_t)t * ntiles + tile];
#pragma unroll
        for (int k = 0; k < SORT_WAVES; ++k)
            s_cnt[k][t] = 0;
    }
    __syncthreads();

    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t base = tile * SORT_TILE;
    for (int r = 0; r < SORT_ITEMS; ++r) {
        uint32_t i = base + r * SORT_THREADS + t;
        bool valid = i < count;
        uint32_t k = valid ? kin[i] : 0u;
        uint32_t v = valid ? vin[i] : 0u;
        uint32_t d = (k >> shift) & 0xffu;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            uint64_t bb = __ballot((d >> bit) & 1u);
            m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0)
            s_cnt[wv][d] = (uint32_t)__popcll(m);
        __syncthreads();
        // per-digit prefix over waves (thread t owns digit t)
        {
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < SORT_WAVES; ++w) {
                uint32_t c = s_cnt[w][t];
                s_cnt[w][t] = run;
                run += c;
            }
            __syncthreads();
            if (valid) {
                uint32_t pos = s_base[d] + s_cnt[wv][d] + rank;
                kout[pos] = k;
                vout[pos] = v;
            }
            __syncthreads();
            s_base[t] += run;
#pragma unroll
            for (int w = 0; w < SORT_WAVES; ++w)
                s_cnt[w][t] = 0;
        }
        __syncthreads();
    }
}

void qs_scan_digits(uint32_t *hist, uint32_t ntiles, uint32_t nbins, uint32_t *totals,
                    hipStream_t stream)
{
    hipLaunchKernelGGL(k_sort_scan_digit, dim3(nbins), dim3(1024), 0, stream, hist, ntiles, totals);
}

static int sort_pairs_impl(uint32_t *keys[2], uint32_t *vals[2], uint32_t *d_hist, uint32_t count,
                           const uint32_t *d_count, int key_bits, hipStream_t stream)
{
    uint32_t ntiles = (count + SORT_TILE - 1) / SORT_TILE;
    uint32_t *totals = d_hist + (size_t)256 * ntiles;
    int passes = (key_bits + 7) / 8;
    int cur = 0;
    for (int ps = 0; ps < passes; ++ps) {
        int shift = 8 * ps;
        hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(SORT_THREADS), 0, stream, keys[cur],
                           count, ntiles, shift, d_hist, d_count);
        hipLaunchKernelGGL(k_sort_scan_digit, dim3(256), dim3(1024), 0, stream, d_hist, ntiles,
                           totals);
        hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(SORT_THREADS), 0, stream,
                           keys[cur], vals[cur], count, ntiles, shift, d_hist, totals,
                           keys[cur ^ 1], vals[cur ^ 1], d_count);
        cur ^= 1;
    }
    if (hipGetLastError() != hipSuccess)
        return -1;
    return cur;
}

int qs_sort_pairs(uint32_t *keys[2], uint32_t *vals[2], uint32_t *d_hist, uint32_t count,
                  int key_bits, hipStream_t stream)
{
    if (count <= 1)
        return 0;
    return sort_pairs_impl(keys, vals, d_hist, count, nullptr, key_bits, stream);
}

int qs_sort_pairs_dev(uint32_t *keys[2], uint32_t *vals[2], uint32_t *d_hist,
                      const uint32_t *d_count, uint32_t max_count, int key_bits, hipStream_t stream)
{
    if (max_count == 0)
        return 0;
    return sort_pairs_impl(keys, vals, d_hist, max_count, d_count, key_bits, stream);
}
