// qs_calib.hip -- kernels of known HBM traffic, for stating the ceilings the
// sketch kernels run against and for calibrating the FETCH_SIZE / WRITE_SIZE
// counters.  Measurement tools only: nothing in the sketch calls them.
//   k_calib_copy    streaming copy, 16 bytes per lane per step
//   k_calib_fill    streaming fill, 16 bytes per lane per step (plain or
//                   non-temporal stores)
//   k_calib_gather  n random reads of G bytes (G / 16 lanes, 16 bytes each)
//                   from a buffer of G-byte lines; the group's xor goes to a
//                   4-byte result per read, like a query's estimate
//   k_calib_rmw     n read-modify-writes of G bytes (every word + 1), at
//                   random lines or in ascending order with random gaps (the
//                   segment pass's pattern: distinct buckets, sorted)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qsketch_calib.h"

typedef unsigned int cu32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_calib_copy(const cu32x4 *__restrict__ src, cu32x4 *__restrict__ dst,
                                                    uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = __builtin_nontemporal_load(src + i);
}

__device__ static inline uint64_t calib_mix(uint64_t seed, uint64_t i)
{
    uint64_t z = seed * 0xD1B54A32D192ED03ull + (i + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <bool NT>
__global__ __launch_bounds__(256) void k_calib_fill(cu32x4 *__restrict__ dst, uint64_t n16, uint32_t value)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const cu32x4 v = {value, value, value, value};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
        if (NT)
            __builtin_nontemporal_store(v, dst + i);
        else
            dst[i] = v;
    }
}

template <int LANES>
__global__ __launch_bounds__(256) void k_calib_gather(const cu32x4 *__restrict__ buf, uint64_t n_lines, uint64_t n,
                                                      uint64_t seed, uint32_t *__restrict__ out)
{
    const uint32_t part = threadIdx.x % LANES;
    const uint64_t per_block = 256 / LANES;
    const uint64_t stride = (uint64_t)gridDim.x * per_block;
    for (uint64_t i = (uint64_t)blockIdx.x * per_block + threadIdx.x / LANES; i < n; i += stride) {
        const uint64_t line = __umul64hi(calib_mix(seed, i), n_lines); // uniform in [0, n_lines)
        const cu32x4 v = __builtin_nontemporal_load(buf + line * LANES + part);
        uint32_t x = v[0] ^ v[1] ^ v[2] ^ v[3];
#pragma unroll
        for (int d = 1; d < LANES; d <<= 1)
            x ^= (uint32_t)__shfl_xor((int)x, d, 64);
        if (part == 0)
            out[i] = x;
    }
}

template <int LANES>
__global__ __launch_bounds__(256) void k_calib_rmw(cu32x4 *__restrict__ buf, uint64_t n_lines, uint64_t n,
                                                   uint64_t seed, int sorted)
{
    const uint32_t part = threadIdx.x % LANES;
    const uint64_t per_block = 256 / LANES;
    const uint64_t stride = (uint64_t)gridDim.x * per_block;
    const uint64_t gap = n_lines / n; // >= 1: the host checks n <= n_lines
    for (uint64_t i = (uint64_t)blockIdx.x * per_block + threadIdx.x / LANES; i < n; i += stride) {
        const uint64_t h = calib_mix(seed, i);
        // sorted: line i * gap + (h mod gap) <= n * gap - 1 < n_lines
        const uint64_t line = sorted ? i * gap + h % gap : __umul64hi(h, n_lines);
        cu32x4 *p = buf + line * LANES + part;
        cu32x4 v = *p;
        v += 1u;
        *p = v;
    }
}

extern "C" int qs_calib_copy(const void *d_src, void *d_dst, uint64_t bytes, void *stream)
{
    if (!d_src || !d_dst || bytes % 16)
        return QS_ERR_PARAM;
    const uint64_t n16 = bytes / 16;
    uint64_t blocks = (n16 + 255) / 256;
    if (blocks > 256 * 64)
        blocks = 256 * 64;
    hipLaunchKernelGGL(k_calib_copy, dim3((uint32_t)(blocks ? blocks : 1)), dim3(256), 0, (hipStream_t)stream,
                       (const cu32x4 *)d_src, (cu32x4 *)d_dst, n16);
    return hipGetLastError() == hipSuccess ? QS_OK : QS_ERR_HIP;
}

extern "C" int qs_calib_fill(void *d_dst, uint64_t bytes, uint32_t value, int nontemporal, void *stream)
{
    if (!d_dst || bytes % 16)
        return QS_ERR_PARAM;
    const uint64_t n16 = bytes / 16;
    uint64_t blocks = (n16 + 255) / 256;
    if (blocks > 256 * 64)
        blocks = 256 * 64;
    const dim3 grid((uint32_t)(blocks ? blocks : 1)), block(256);
    if (nontemporal)
        hipLaunchKernelGGL(k_calib_fill<true>, grid, block, 0, (hipStream_t)stream, (cu32x4 *)d_dst, n16, value);
    else
        hipLaunchKernelGGL(k_calib_fill<false>, grid, block, 0, (hipStream_t)stream, (cu32x4 *)d_dst, n16, value);
    return hipGetLastError() == hipSuccess ? QS_OK : QS_ERR_HIP;
}

extern "C" int qs_calib_gather(const void *d_buf, uint64_t buf_bytes, uint32_t access_bytes, uint64_t n,
                               uint64_t seed, uint32_t *d_out, void *stream)
{
    if (!d_buf || !d_out || (access_bytes != 16 && access_bytes != 32 && access_bytes != 64 && access_bytes != 128))
        return QS_ERR_PARAM;
    const uint64_t n_lines = buf_bytes / access_bytes;
    if (n_lines == 0)
        return QS_ERR_PARAM;
    const uint32_t lanes = access_bytes / 16;
    uint64_t blocks = (n * lanes + 255) / 256;
    if (blocks > 256 * 8)
        blocks = 256 * 8;
    const dim3 grid((uint32_t)(blocks ? blocks : 1)), block(256);
    const hipStream_t st = (hipStream_t)stream;
    const cu32x4 *b = (const cu32x4 *)d_buf;
    switch (lanes) {
    case 1: hipLaunchKernelGGL(k_calib_gather<1>, grid, block, 0, st, b, n_lines, n, seed, d_out); break;
    case 2: hipLaunchKernelGGL(k_calib_gather<2>, grid, block, 0, st, b, n_lines, n, seed, d_out); break;
    case 4: hipLaunchKernelGGL(k_calib_gather<4>, grid, block, 0, st, b, n_lines, n, seed, d_out); break;
    default: hipLaunchKernelGGL(k_calib_gather<8>, grid, block, 0, st, b, n_lines, n, seed, d_out); break;
    }
    return hipGetLastError() == hipSuccess ? QS_OK : QS_ERR_HIP;
}

extern "C" int qs_calib_rmw(void *d_buf, uint64_t buf_bytes, uint32_t access_bytes, uint64_t n, uint64_t seed,
                            int sorted, void *stream)
{
    if (!d_buf || (access_bytes != 16 && access_bytes != 32 && access_bytes != 64 && access_bytes != 128))
        return QS_ERR_PARAM;
    const uint64_t n_lines = buf_bytes / access_bytes;
    if (n > n_lines)
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    const uint32_t lanes = access_bytes / 16;
    uint64_t blocks = (n * lanes + 255) / 256;
    if (blocks > 256 * 8)
        blocks = 256 * 8;
    const dim3 grid((uint32_t)blocks), block(256);
    const hipStream_t st = (hipStream_t)stream;
    cu32x4 *b = (cu32x4 *)d_buf;
    switch (lanes) {
    case 1: hipLaunchKernelGGL(k_calib_rmw<1>, grid, block, 0, st, b, n_lines, n, seed, sorted); break;
    case 2: hipLaunchKernelGGL(k_calib_rmw<2>, grid, block, 0, st, b, n_lines, n, seed, sorted); break;
    case 4: hipLaunchKernelGGL(k_calib_rmw<4>, grid, block, 0, st, b, n_lines, n, seed, sorted); break;
    default: hipLaunchKernelGGL(k_calib_rmw<8>, grid, block, 0, st, b, n_lines, n, seed, sorted); break;
    }
    return hipGetLastError() == hipSuccess ? QS_OK : QS_ERR_HIP;
}
