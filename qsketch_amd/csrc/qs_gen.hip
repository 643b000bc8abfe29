// qs_gen.hip -- counter-based synthetic key streams (uniform, Zipf, noisy
// 64-bit Zipf).  Sampling is integer-only, so host and device agree bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/qsketch.h"
#include "../../include/qsketch_gen.h"

#define NOISE_HALF 12 // noise support [-12, 12]

struct qs_zipf {
    uint64_t universe;
    double alpha;
    uint32_t *h_thr;
    uint32_t *h_alias;
    uint32_t *d_thr;
    uint32_t *d_alias;
};

__host__ __device__ static inline uint64_t gen_mix(uint64_t seed, uint64_t i)
{
    uint64_t z = seed * 0xD1B54A32D192ED03ull + (i + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ static inline uint32_t gen_mulhi(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * b) >> 32);
}

__host__ __device__ static inline uint32_t zipf_rank(const uint32_t *thr, const uint32_t *alias,
                                                     uint32_t universe_m1, uint64_t x)
{
    // column uniform in [0, U): hi32(x) * U >> 32 (U = universe_m1 + 1 <= 2^32)
    uint64_t U = (uint64_t)universe_m1 + 1ull;
    uint32_t col = (uint32_t)(((uint64_t)(uint32_t)(x >> 32) * U) >> 32);
    uint32_t coin = (uint32_t)x;
    uint32_t t = thr[col];
    return (t == 0xffffffffu || coin < t) ? col : alias[col];
}

__host__ __device__ static inline uint32_t feistel32(uint64_t seed, uint32_t x)
{
    uint32_t L = x >> 16, R = x & 0xffffu;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        uint32_t k = (uint32_t)(gen_mix(seed, 0x100u + r) & 0xffffffffu);
        uint32_t f = (R * 0x9E3779B1u) ^ k;
        f ^= f >> 15;
        f *= 0x85EBCA6Bu;
        f ^= f >> 13;
        uint32_t nl = R;
        R = (L ^ f) & 0xffffu;
        L = nl;
    }
    return (L << 16) | R;
}

__host__ __device__ static inline uint64_t feistel48(uint64_t seed, uint64_t x)
{
    uint64_t L = (x >> 24) & 0xffffffull, R = x & 0xffffffull;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        uint64_t k = gen_mix(seed, 0x200u + r);
        uint64_t f = (R * 0x9E3779B97F4A7C15ull) ^ k;
        f ^= f >> 29;
        f *= 0xBF58476D1CE4E5B9ull;
        f ^= f >> 32;
        uint64_t nl = R;
        R = (L ^ f) & 0xffffffull;
        L = nl;
    }
    return (L << 24) | R;
}

// round(log2(x)) for x >= 1, integer-exact: k = floor(log2 x); +1 when x^2 >= 2^(2k+1)
__host__ __device__ static inline uint32_t round_log2(uint64_t x)
{
    uint32_t k = 63u - (uint32_t)__builtin_clzll(x);
    if (k < 31u && x * x >= (1ull << (2u * k + 1u)))
        ++k;
    return k;
}

__host__ __device__ static inline uint64_t noisy_key(uint64_t seed, uint32_t rank, const uint32_t *ncdf)
{
    uint32_t u = (uint32_t)(gen_mix(seed ^ 0x5bd1e995ull, rank) >> 32);
    int nz = -NOISE_HALF;
    for (int k = 0; k < 2 * NOISE_HALF; ++k)
        nz += (u >= ncdf[k]) ? 1 : 0;
    int v = (int)round_log2((uint64_t)rank + 1ull) + nz;
    v = v < 0 ? 0 : (v > 31 ? 31 : v);
    return ((uint64_t)v << 48) | feistel48(seed, rank);
}

extern "C" uint32_t qs_rank_to_key32(uint64_t seed, uint32_t rank) { return feistel32(seed, rank); }

// ---------------------------------------------------------------------------
// alias table (Vose), host double precision
// ---------------------------------------------------------------------------
extern "C" int qs_zipf_create(double alpha, uint64_t universe, int with_device, qs_zipf **out)
{
    if (!out || universe < 1 || universe > (1ull << 32) || !(alpha >= 0.0))
        return QS_ERR_PARAM;
    *out = nullptr;
    qs_zipf *z = (qs_zipf *)calloc(1, sizeof(qs_zipf));
    if (!z)
        return QS_ERR_NOMEM;
    z->universe = universe;
    z->alpha = alpha;
    size_t U = (size_t)universe;
    z->h_thr = (uint32_t *)malloc(sizeof(uint32_t) * U);
    z->h_alias = (uint32_t *)malloc(sizeof(uint32_t) * U);
    std::vector<double> p(U);
    if (!z->h_thr || !z->h_alias) {
        qs_zipf_destroy(z);
        return QS_ERR_NOMEM;
    }
    double sum = 0.0;
    for (size_t r = 0; r < U; ++r) {
        p[r] = std::pow((double)(r + 1), -alpha);
        sum += p[r];
    }
    std::vector<uint32_t> small, large;
    small.reserve(U);
    large.reserve(U);
    for (size_t r = 0; r < U; ++r) {
        p[r] = p[r] * (double)U / sum;
        if (p[r] < 1.0)
            small.push_back((uint32_t)r);
        else
            large.push_back((uint32_t)r);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t s = small.back();
        small.pop_back();
        uint32_t l = large.back();
        double ps = p[s];
        double t = ps * 4294967296.0;
        z->h_thr[s] = t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
        z->h_alias[s] = l;
        p[l] = (p[l] + ps) - 1.0;
        if (p[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (uint32_t r : large) {
        z->h_thr[r] = 0xffffffffu;
        z->h_alias[r] = r;
    }
    for (uint32_t r : small) {
        z->h_thr[r] = 0xffffffffu;
        z->h_alias[r] = r;
    }
    if (with_device) {
        if (hipMalloc(&z->d_thr, sizeof(uint32_t) * U) != hipSuccess ||
            hipMalloc(&z->d_alias, sizeof(uint32_t) * U) != hipSuccess ||
            hipMemcpy(z->d_thr, z->h_thr, sizeof(uint32_t) * U, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(z->d_alias, z->h_alias, sizeof(uint32_t) * U, hipMemcpyHostToDevice) !=
                hipSuccess) {
            qs_zipf_destroy(z);
            return QS_ERR_HIP;
        }
    }
    *out = z;
    return QS_OK;
}

extern "C" void qs_zipf_destroy(qs_zipf *z)
{
    if (!z)
        return;
    free(z->h_thr);
    free(z->h_alias);
    if (z->d_thr)
        hipFree(z->d_thr);
    if (z->d_alias)
        hipFree(z->d_alias);
    free(z);
}

// cumulative thresholds of round(N(0, sigma)) over [-NOISE_HALF, NOISE_HALF]
static void noise_cdf(double sigma, uint32_t *c)
{
    for (int k = 0; k < 2 * NOISE_HALF; ++k) {
        double edge = (double)(k - NOISE_HALF) + 0.5; // boundary between k-H and k-H+1
        double phi = sigma > 0 ? 0.5 * std::erfc(-edge / (sigma * std::sqrt(2.0))) : (edge > 0 ? 1.0 : 0.0);
        double t = phi * 4294967296.0;
        c[k] = t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_gen_uniform(uint64_t seed, uint64_t offset, uint64_t n, uint32_t *out)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint32_t)(gen_mix(seed, offset + i) >> 32);
}

__global__ void k_gen_zipf(const uint32_t *thr, const uint32_t *alias, uint32_t um1, uint64_t seed,
                           uint64_t offset, uint64_t n, uint32_t *out)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = feistel32(seed, zipf_rank(thr, alias, um1, gen_mix(seed, offset + i)));
}

struct ncdf_t {
    uint32_t c[2 * NOISE_HALF];
};

__global__ void k_gen_zipf_noisy(const uint32_t *thr, const uint32_t *alias, uint32_t um1,
                                 uint64_t seed, uint64_t offset, uint64_t n, ncdf_t nc, uint64_t *out)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t r = zipf_rank(thr, alias, um1, gen_mix(seed, offset + i));
        out[i] = noisy_key(seed, r, nc.c);
    }
}

static unsigned gen_blocks(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 4096)
        b = 4096;
    return b ? (unsigned)b : 1u;
}

extern "C" int qs_gen_uniform_u32(uint64_t seed, uint64_t offset, uint64_t n, uint32_t *d_out, void *stream)
{
    if (!d_out && n)
        return QS_ERR_PARAM;
    if (!n)
        return QS_OK;
    hipLaunchKernelGGL(k_gen_uniform, dim3(gen_blocks(n)), dim3(256), 0, (hipStream_t)stream, seed, offset, n,
                       d_out);
    return hipGetLastError() == hipSuccess ? QS_OK : QS_ERR_HIP;
}

extern "C" int qs_gen_uniform_u32_host(uint64_t seed, uint64_t offset, uint64_t n, uint32_t *out)
{
    if (!out && n)
        return QS_ERR_PARAM;
    for (uint64_t i = 0; i < n; ++i)
        out[i] = (uint32_t)(gen_mix(seed, offset + i) >> 32);
    return QS_OK;
}

extern "C" int qs_gen_zipf_u32(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n, uint32_t *d_out,
                               void *stream)
{
    if (!z || !z->d_thr || (!d_out && n))
        return QS_ERR_PARAM;
    if (!n)
        return QS_OK;
    hipLaunchKernelGGL(k_gen_zipf, dim3(gen_blocks(n)), dim3(256), 0, (hipStream_t)stream, z->d_thr,
                       z->d_alias, (uint32_t)(z->universe - 1), seed, offset, n, d_out);
    return hipGetLastError() == hipSuccess ? QS_OK : QS_ERR_HIP;
}

extern "C" int qs_gen_zipf_u32_host(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                                    uint32_t *out)
{
    if (!z || (!out && n))
        return QS_ERR_PARAM;
    uint32_t um1 = (uint32_t)(z->universe - 1);
    for (uint64_t i = 0; i < n; ++i)
        out[i] = feistel32(seed, zipf_rank(z->h_thr, z->h_alias, um1, gen_mix(seed, offset + i)));
    return QS_OK;
}

extern "C" int qs_gen_zipf_rank_host(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                                     uint32_t *out)
{
    if (!z || (!out && n))
        return QS_ERR_PARAM;
    uint32_t um1 = (uint32_t)(z->universe - 1);
    for (uint64_t i = 0; i < n; ++i)
        out[i] = zipf_rank(z->h_thr, z->h_alias, um1, gen_mix(seed, offset + i));
    return QS_OK;
}

extern "C" int qs_gen_zipf_u64_noisy(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                                     double sigma, uint64_t *d_out, void *stream)
{
    if (!z || !z->d_thr || (!d_out && n))
        return QS_ERR_PARAM;
    if (!n)
        return QS_OK;
    ncdf_t nc;
    noise_cdf(sigma, nc.c);
    hipLaunchKernelGGL(k_gen_zipf_noisy, dim3(gen_blocks(n)), dim3(256), 0, (hipStream_t)stream, z->d_thr,
                       z->d_alias, (uint32_t)(z->universe - 1), seed, offset, n, nc, d_out);
    return hipGetLastError() == hipSuccess ? QS_OK : QS_ERR_HIP;
}

extern "C" int qs_gen_zipf_u64_noisy_host(const qs_zipf *z, uint64_t seed, uint64_t offset, uint64_t n,
                                          double sigma, uint64_t *out)
{
    if (!z || (!out && n))
        return QS_ERR_PARAM;
    ncdf_t nc;
    noise_cdf(sigma, nc.c);
    uint32_t um1 = (uint32_t)(z->universe - 1);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t r = zipf_rank(z->h_thr, z->h_alias, um1, gen_mix(seed, offset + i));
        out[i] = noisy_key(seed, r, nc.c);
    }
    return QS_OK;
}
