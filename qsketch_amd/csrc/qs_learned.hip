// qs_learned.hip -- learned heavy-hitter routing in front of the MSWS sketch.
//
// No counterpart in the reference (it has no learned component).  The network
// is a 64 -> 64 -> 64 -> 1 MLP with ReLU on the 64 bits of a key, BF16 weights
// and activations, FP32 accumulation:
//   h1 = bf16(relu(W1 x + b1)),  h2 = bf16(relu(W2 h1 + b2)),  y = w3 . h2 + b3
// Training kernels (VALU, one key per lane):
//   k_mlp_grad  forward + backward of a minibatch, gradients summed per block
//               in LDS-staged small GEMMs and added to an FP32 gradient buffer
//   k_adam      Adam on the FP32 master weights; writes the BF16 copies and
//               clears the gradient buffer
// Serving kernels:
//   k_mlp_mfma  the same forward on the matrix cores: a wavefront scores 32
//               keys with v_mfma_f32_32x32x8_bf16, computing H^T = W X^T so that
//               the accumulator layout of one layer is the B-operand layout of
//               the next (no transpose between layers)
//   k_route     k_mlp_mfma + routing: keys with score >= threshold are counted
//               in a per-block LDS table and flushed to the exact GPU table with
//               one atomic per distinct key and block; the others are flagged
//               for a stable compaction into the sketch's input
//   k_reflag    after a routing pass the table could not hold: heavy keys
//               without a slot go to the sketch instead
//   k_table_query / k_learned_query  the query: the table count of a key
//               held in the table, else the sketch estimate (the second
//               kernel scores keys first and probes only heavy-scored ones)
#include "qs_internal.h"

#include "../../include/qsketch_learned.h"

#include <cmath>
#include <cstring>
#include <vector>

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess)                                                                   \
            return QS_ERR_HIP;                                                                   \
    } while (0)

#define OFF_W1 0
#define OFF_B1 4096
#define OFF_W2 4160
#define OFF_B2 8256
#define OFF_W3 8320
#define OFF_B3 8384
#define NPAR QS_MLP_PARAMS

#define T_EMPTY 0xffffffffffffffffull
#define T_NONE 0xffffffffu
#define T_PROBES 4096u

#define ROUTE_THREADS 256
#define ROUTE_TILE 8192u // keys per routing block: 4 waves x 64 steps x 32 keys
#define LT_SLOTS 4096u   // per-block LDS aggregation table
#define LT_PROBES 32u

struct qs_learned {
    qs_sketch *sk;
    uint32_t key_bytes;
    float *d_p;        // FP32 master parameters
    uint16_t *d_pb;    // BF16 copies
    float *d_m, *d_v;  // Adam moments
    float *d_g;        // gradient accumulator
    float *d_loss;
    uint64_t step;
    float thr;
    uint32_t tlog2;
    uint64_t tcap;
    uint64_t *d_tkeys;
    uint32_t *d_tcnt;
    uint32_t *d_err;   // [0] set by k_route when the table turned a key away, [1] slots in use
    uint64_t inserted; // keys inserted since create / clear
    uint32_t full;     // the table has turned a key away since create / clear
    uint32_t used;     // table slots in use after the last insert
    // routing scratch
    uint64_t rcap;     // keys
    uint8_t *d_flag;   // 1 = goes to the sketch
    uint32_t *d_tiles; // light keys per routing tile (+ scan totals)
    void *d_light;
    uint32_t *h_pin;   // pinned: light total, error flag, slots in use
};

// ---------------------------------------------------------------------------
// BF16 (round to nearest even) and key bits
// ---------------------------------------------------------------------------
__device__ static inline uint16_t f2bf(float f)
{
    // v_cvt_pk_bf16_f32: round to nearest even, as f2bf_host's bit arithmetic
    // for every finite value
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ static inline float bf2f(uint16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}
static inline uint16_t f2bf_host(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// the key at i, or at n - 1 past the end (n > 0): loads that prefetch the next
// tile's keys stay unconditional, so the wait for this tile's keys does not also
// wait for them
template <int KB> __device__ static inline uint64_t key64(const KeyT<KB> *keys, uint64_t i);
template <int KB>
__device__ static inline uint64_t key_at(const KeyT<KB> *keys, uint64_t i, uint64_t n)
{
    return key64<KB>(keys, i < n ? i : n - 1u);
}

template <int KB>
__device__ static inline uint64_t key64(const KeyT<KB> *keys, uint64_t i)
{
    return (uint64_t)keys[i];
}

// ---------------------------------------------------------------------------
// training: forward + backward of 64 keys per block (one per lane)
// ---------------------------------------------------------------------------
#define GB 64 // keys per gradient block

template <int KB>
__global__ __launch_bounds__(GB) void k_mlp_grad(const KeyT<KB> *__restrict__ keys, const float *__restrict__ tgt,
                                                 uint32_t n, const uint16_t *__restrict__ pb,
                                                 const float *__restrict__ p, float inv_n, float *__restrict__ g,
                                                 float *__restrict__ loss)
{
    __shared__ uint16_t sW1[64 * 64], sW2[64 * 64];
    __shared__ float sb1[64], sb2[64], sw3[64];
    __shared__ uint16_t sh1[GB][66], sh2[GB][66];
    __shared__ float sd[GB][65]; // dh2, later dh1
    __shared__ unsigned long long sx[GB];
    __shared__ float sdy[GB];
    const uint32_t t = threadIdx.x;
    for (uint32_t e = t; e < 4096u; e += GB) {
        sW1[e] = pb[OFF_W1 + e];
        sW2[e] = pb[OFF_W2 + e];
    }
    sb1[t] = p[OFF_B1 + t];
    sb2[t] = p[OFF_B2 + t];
    sw3[t] = bf2f(pb[OFF_W3 + t]);
    const float b3 = p[OFF_B3];
    __syncthreads();

    const uint32_t i = blockIdx.x * GB + t;
    const bool valid = i < n;
    const uint64_t x = valid ? key64<KB>(keys, i) : 0ull;
    // forward; this lane's activations live in its LDS rows, the ReLU masks
    // in registers
    uint64_t m1 = 0, m2 = 0;
    for (int j = 0; j < 64; ++j) {
        float acc = sb1[j];
        for (int k = 0; k < 64; ++k)
            acc += ((x >> k) & 1ull) ? bf2f(sW1[j * 64 + k]) : 0.f;
        m1 |= (uint64_t)(acc > 0.f) << j;
        sh1[t][j] = f2bf(fmaxf(acc, 0.f));
    }
    float y = b3;
    for (int j = 0; j < 64; ++j) {
        float acc = sb2[j];
        for (int k = 0; k < 64; ++k)
            acc = fmaf(bf2f(sW2[j * 64 + k]), bf2f(sh1[t][k]), acc);
        m2 |= (uint64_t)(acc > 0.f) << j;
        const uint16_t h2 = f2bf(fmaxf(acc, 0.f));
        sh2[t][j] = h2;
        y = fmaf(sw3[j], bf2f(h2), y);
    }
    const float diff = valid ? (y - tgt[i]) : 0.f;
    const float dy = diff * inv_n;
    // backward to the hidden layers
    for (int j = 0; j < 64; ++j)
        sd[t][j] = ((m2 >> j) & 1ull) ? dy * sw3[j] : 0.f;
    float dh1[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) {
        float acc = 0.f;
        for (int j = 0; j < 64; ++j)
            acc = fmaf(bf2f(sW2[j * 64 + k]), sd[t][j], acc);
        dh1[k] = ((m1 >> k) & 1ull) ? acc : 0.f;
    }
    sx[t] = x;
    sdy[t] = dy;
    // loss: sum over the block, one atomic
    float l = 0.5f * diff * diff * inv_n;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        l += __shfl_xor(l, o, 64);
    if (t == 0)
        atomicAdd(loss, l);
    __syncthreads();
    // W2 (row j, column k), b2, w3, b3
    for (uint32_t e = t; e < 4225u; e += GB) {
        float acc = 0.f;
        if (e < 4096u) {
            const uint32_t j = e >> 6, k = e & 63u;
            for (uint32_t s = 0; s < GB; ++s)
                acc = fmaf(sd[s][j], bf2f(sh1[s][k]), acc);
            atomicAdd(g + OFF_W2 + e, acc);
        } else if (e < 4160u) {
            const uint32_t j = e - 4096u;
            for (uint32_t s = 0; s < GB; ++s)
                acc += sd[s][j];
            atomicAdd(g + OFF_B2 + j, acc);
        } else if (e < 4224u) {
            const uint32_t j = e - 4160u;
            for (uint32_t s = 0; s < GB; ++s)
                acc = fmaf(sdy[s], bf2f(sh2[s][j]), acc);
            atomicAdd(g + OFF_W3 + j, acc);
        } else {
            for (uint32_t s = 0; s < GB; ++s)
                acc += sdy[s];
            atomicAdd(g + OFF_B3, acc);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 64; ++k)
        sd[t][k] = dh1[k];
    __syncthreads();
    // W1 (row j, column = key bit k), b1
    for (uint32_t e = t; e < 4160u; e += GB) {
        float acc = 0.f;
        if (e < 4096u) {
            const uint32_t j = e >> 6, k = e & 63u;
            for (uint32_t s = 0; s < GB; ++s)
                acc += ((sx[s] >> k) & 1ull) ? sd[s][j] : 0.f;
            atomicAdd(g + OFF_W1 + e, acc);
        } else {
            const uint32_t j = e - 4096u;
            for (uint32_t s = 0; s < GB; ++s)
                acc += sd[s][j];
            atomicAdd(g + OFF_B1 + j, acc);
        }
    }
}

// Adam on the master weights; refreshes the BF16 copies and clears the gradients
__global__ void k_adam(float *__restrict__ p, uint16_t *__restrict__ pb, float *__restrict__ m,
                       float *__restrict__ v, float *__restrict__ g, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < NPAR; e += gridDim.x * blockDim.x) {
        const float gr = g[e];
        g[e] = 0.f;
        const float mm = b1 * m[e] + (1.f - b1) * gr;
        const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
        m[e] = mm;
        v[e] = vv;
        const float pp = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
        p[e] = pp;
        pb[e] = f2bf(pp);
    }
}

__global__ void k_to_bf16(const float *__restrict__ p, uint16_t *__restrict__ pb)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < NPAR; e += gridDim.x * blockDim.x)
        pb[e] = f2bf(p[e]);
}

__global__ void k_zero_f32(float *__restrict__ a, uint32_t n)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        a[e] = 0.f;
}

// per-key VALU forward (the training forward)
template <int KB>
__global__ __launch_bounds__(256) void k_mlp_fwd(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                 const uint16_t *__restrict__ pb, const float *__restrict__ p,
                                                 float *__restrict__ out)
{
    __shared__ uint16_t sW1[64 * 64], sW2[64 * 64];
    __shared__ float sb1[64], sb2[64], sw3[64];
    for (uint32_t e = threadIdx.x; e < 4096u; e += 256u) {
        sW1[e] = pb[OFF_W1 + e];
        sW2[e] = pb[OFF_W2 + e];
    }
    if (threadIdx.x < 64) {
        sb1[threadIdx.x] = p[OFF_B1 + threadIdx.x];
        sb2[threadIdx.x] = p[OFF_B2 + threadIdx.x];
        sw3[threadIdx.x] = bf2f(pb[OFF_W3 + threadIdx.x]);
    }
    __syncthreads();
    const float b3 = p[OFF_B3];
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) {
        const uint64_t x = key64<KB>(keys, i);
        float h1[64];
#pragma unroll
        for (int j = 0; j < 64; ++j) {
            float acc = sb1[j];
            for (int k = 0; k < 64; ++k)
                acc += ((x >> k) & 1ull) ? bf2f(sW1[j * 64 + k]) : 0.f;
            h1[j] = bf2f(f2bf(fmaxf(acc, 0.f)));
        }
        float y = b3;
        for (int j = 0; j < 64; ++j) {
            float acc = sb2[j];
#pragma unroll
            for (int k = 0; k < 64; ++k)
                acc = fmaf(bf2f(sW2[j * 64 + k]), h1[k], acc);
            y = fmaf(sw3[j], bf2f(f2bf(fmaxf(acc, 0.f))), y);
        }
        out[i] = y;
    }
}

// ---------------------------------------------------------------------------
// matrix-core forward.  v_mfma_f32_32x32x16_bf16 (8 BF16 per lane, gfx950):
//   A (32 x 16):  lane l holds A[l % 32][8 (l / 32) + i],   i = 0..7
//   B (16 x 32):  lane l holds B[8 (l / 32) + i][l % 32]
//   D (32 x 32):  lane l, register r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// Layer 1 computes H1^T (64 units x 32 keys) = W1 X^T as 2 row tiles x 4 K
// steps; lane l's key is key l % 32, and K step s, half h = l / 32 is key byte
// 2 s + h.  Layer 2 sums over the units in the order the accumulators hold
// them: K step s takes registers 8 (s % 2) .. 8 (s % 2) + 7 of row tile s / 2,
// so slot i of half h is unit 32 (s / 2) + 8 (2 (s % 2) + i / 4) + 4 h + i % 4
// (every unit once), and W2's columns are laid out in that order.  The output
// layer is one more row tile in the same K order: W3 is its row 0 (the other
// rows are zero), so the score of key l is register 0 of lane l < 32.
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// The whole network lives in LDS: the weight operands in MFMA A-operand order
// (layer, row tile, K step, lane -> 8 BF16), so each MFMA reads its operand
// with one 16-byte LDS load, plus the biases and output weights.  Registers
// then hold only accumulators and activations, for several waves per SIMD.
struct mlp_lds {
    bf16x8 a[2][2][4][64];
    bf16x8 a3[4][64]; // output layer: W3 in row 0 of the A operand, other rows 0
    bf16x4 nib[16];   // layer-1 B operand of a key nibble: bit i -> 1.0 or 0.0 in BF16
    float b1[64], b2[64], b3;
};

// BF16(max(lo, 0)), BF16(max(hi, 0)) packed: one conversion of the pair, then
// max(., 0) on the two 16-bit halves (rounding keeps the sign, and negative
// BF16 values and -0 are negative 16-bit integers: the same bits as rounding
// max(x, 0))
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
__device__ static inline uint32_t relu_bf16x2(float lo, float hi)
{
    const bf16x2v b = __builtin_convertvector((f32x2v){lo, hi}, bf16x2v); // v_cvt_pk_bf16_f32
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, b), (s16x2)0));
}

// a layer's activations as the next layer's B operands: K step s takes
// registers 8 (s % 2) .. 8 (s % 2) + 7 of row tile s / 2
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
__device__ static inline void acts_to_b(const float *acc0, const float *acc1, bf16x8 hb[4])
{
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const float *a = s < 2 ? acc0 : acc1;
        u32x4v w;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int r = 8 * (s % 2) + 2 * p;
            w[p] = relu_bf16x2(a[r], a[r + 1]);
        }
        hb[s] = __builtin_bit_cast(bf16x8, w);
    }
}

// input column of K slot i, half h, K step s of layer 0 / 1
__device__ static inline uint32_t kcol(uint32_t layer, uint32_t s, uint32_t h, uint32_t i)
{
    return layer == 0 ? 16u * s + 8u * h + i : 32u * (s / 2u) + 8u * (2u * (s % 2u) + i / 4u) + 4u * h + i % 4u;
}

__device__ static inline void load_mlp(mlp_lds &S, const uint16_t *__restrict__ pb, const float *__restrict__ p)
{
    for (uint32_t e = threadIdx.x; e < 2u * 2u * 4u * 64u; e += blockDim.x) {
        const uint32_t l = e & 63u, s = (e >> 6) & 3u, t = (e >> 8) & 1u, layer = e >> 9;
        const uint32_t h = l >> 5, c = l & 31u;
        const uint16_t *w = pb + (layer ? OFF_W2 : OFF_W1) + (32u * t + c) * 64u;
        bf16x8 v;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            v[i] = (short)w[kcol(layer, s, h, (uint32_t)i)];
        S.a[layer][t][s][l] = v;
    }
    for (uint32_t e = threadIdx.x; e < 4u * 64u; e += blockDim.x) {
        const uint32_t l = e & 63u, s = e >> 6, h = l >> 5, c = l & 31u;
        bf16x8 v;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            v[i] = c == 0 ? (short)pb[OFF_W3 + kcol(1, s, h, (uint32_t)i)] : (short)0;
        S.a3[s][l] = v;
    }
    for (uint32_t e = threadIdx.x; e < 16u; e += blockDim.x) {
        bf16x4 v;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            v[i] = ((e >> i) & 1u) ? (short)0x3f80 : (short)0;
        S.nib[e] = v;
    }
    for (uint32_t e = threadIdx.x; e < 64u; e += blockDim.x) {
        S.b1[e] = p[OFF_B1 + e];
        S.b2[e] = p[OFF_B2 + e];
    }
    if (threadIdx.x == 0)
        S.b3 = p[OFF_B3];
    __syncthreads();
}

// D row held by register r of row tile t in this lane
__device__ static inline uint32_t drow(int t, int r, uint32_t h)
{
    return 32u * t + 8u * (r / 4) + 4u * h + (r % 4);
}

__device__ static inline f32x16 mfma16(const bf16x8 &a, const bf16x8 &b, const f32x16 &c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), c,
                                                   0, 0, 0);
}

// score of key x: lane l < 32 returns the score of its key (lanes 32..63
// return another value of the output tile)
__device__ static inline float mlp_mfma(const mlp_lds &S, uint64_t x)
{
    const uint32_t lane = threadIdx.x & 63u, h = lane >> 5;
    // reload the network from LDS for every key tile instead of letting the
    // compiler hoist it into (occupancy-limiting) registers
    asm volatile("" ::: "memory");
    f32x16 acc0, acc1;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        acc0[r] = S.b1[drow(0, r, h)];
        acc1[r] = S.b1[drow(1, r, h)];
    }
    const uint64_t xs = x >> (8u * h); // this half-wave's bytes: byte 2 s + h at bits 16 s
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const uint32_t byte = (uint32_t)(xs >> (16 * s)) & 0xffu;
        const bf16x4 lo = S.nib[byte & 0xfu], hi = S.nib[byte >> 4];
        const bf16x8 b = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        acc0 = mfma16(S.a[0][0][s][lane], b, acc0);
        acc1 = mfma16(S.a[0][1][s][lane], b, acc1);
    }
    bf16x8 hb[4];
    {
        float a0[16], a1[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            a0[r] = acc0[r];
            a1[r] = acc1[r];
        }
        acts_to_b(a0, a1, hb);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        acc0[r] = S.b2[drow(0, r, h)];
        acc1[r] = S.b2[drow(1, r, h)];
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        acc0 = mfma16(S.a[1][0][s][lane], hb[s], acc0);
        acc1 = mfma16(S.a[1][1][s][lane], hb[s], acc1);
    }
    {
        float a0[16], a1[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            a0[r] = acc0[r];
            a1[r] = acc1[r];
        }
        acts_to_b(a0, a1, hb);
    }
    f32x16 acc3 = {};
#pragma unroll
    for (int s = 0; s < 4; ++s)
        acc3 = mfma16(S.a3[s][lane], hb[s], acc3);
    return acc3[0] + S.b3;
}

template <int KB>
__global__ __launch_bounds__(256) void k_mlp_mfma(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                  const uint16_t *__restrict__ pb, const float *__restrict__ p,
                                                  float *__restrict__ out)
{
    __shared__ mlp_lds S;
    load_mlp(S, pb, p);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t gw = ((uint64_t)blockIdx.x * 256u + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * 256u) >> 6;
    if (n == 0)
        return;
    uint64_t i = gw * 32u + (lane & 31u);
    uint64_t x = key_at<KB>(keys, i, n);
    for (uint64_t base = gw * 32u; base < n; base += nw * 32u) {
        // the next key is loaded while this one is scored
        const uint64_t i2 = i + nw * 32u;
        const uint64_t x2 = key_at<KB>(keys, i2, n);
        const float y = mlp_mfma(S, x);
        if (i < n && lane < 32u)
            out[i] = y;
        i = i2;
        x = x2;
    }
}

// ---------------------------------------------------------------------------
// exact table: open addressing with linear probing, u64 keys (all-ones =
// empty; that key value is never routed to the table), u32 counts
// ---------------------------------------------------------------------------
__device__ static inline uint32_t t_home(uint64_t key, uint32_t tlog2)
{
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64u - tlog2));
}

// slot of key, claiming a free one if it is absent (*used counts claims)
__device__ static inline uint32_t t_insert(uint64_t *tkeys, uint32_t tlog2, uint64_t key, uint32_t *used)
{
    const uint32_t mask = (1u << tlog2) - 1u;
    uint32_t s = t_home(key, tlog2);
    for (uint32_t pr = 0; pr < T_PROBES; ++pr, s = (s + 1u) & mask) {
        uint64_t cur = __hip_atomic_load(tkeys + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == key)
            return s;
        if (cur == T_EMPTY) {
            unsigned long long prev =
                atomicCAS((unsigned long long *)(tkeys + s), (unsigned long long)T_EMPTY, (unsigned long long)key);
            if (prev == T_EMPTY)
                atomicAdd(used, 1u);
            if (prev == T_EMPTY || prev == key)
                return s;
        }
    }
    return T_NONE;
}

__device__ static inline uint32_t t_find(const uint64_t *tkeys, uint32_t tlog2, uint64_t key)
{
    const uint32_t mask = (1u << tlog2) - 1u;
    uint32_t s = t_home(key, tlog2);
    for (uint32_t pr = 0; pr < T_PROBES; ++pr, s = (s + 1u) & mask) {
        const uint64_t cur = tkeys[s];
        if (cur == key)
            return s;
        if (cur == T_EMPTY)
            return T_NONE;
    }
    return T_NONE;
}

// err[0]: set when the table has no slot for key; err[1]: slots in use
__device__ static inline void t_add(uint64_t *tkeys, uint32_t *tcnt, uint32_t tlog2, uint64_t key, uint32_t c,
                                    uint32_t *err)
{
    const uint32_t s = t_insert(tkeys, tlog2, key, err + 1);
    if (s == T_NONE)
        atomicOr(err, 1u);
    else
        atomicAdd(tcnt + s, c);
}

// Route one tile of ROUTE_TILE keys per block: score, count heavy keys in the
// block's LDS table (flushed to the exact table at the end), flag the others.
template <int KB>
__global__ __launch_bounds__(ROUTE_THREADS) void k_route(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                         const uint16_t *__restrict__ pb,
                                                         const float *__restrict__ p, float thr,
                                                         uint64_t *__restrict__ tkeys, uint32_t *__restrict__ tcnt,
                                                         uint32_t tlog2, uint8_t *__restrict__ flag,
                                                         uint32_t *__restrict__ tiles, uint32_t *__restrict__ err)
{
    __shared__ unsigned long long s_key[LT_SLOTS];
    __shared__ uint32_t s_cnt[LT_SLOTS];
    __shared__ uint32_t s_light;
    for (uint32_t e = threadIdx.x; e < LT_SLOTS; e += ROUTE_THREADS) {
        s_key[e] = T_EMPTY;
        s_cnt[e] = 0;
    }
    if (threadIdx.x == 0)
        s_light = 0;
    __shared__ mlp_lds S;
    load_mlp(S, pb, p);
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint64_t tb = (uint64_t)blockIdx.x * ROUTE_TILE;
    uint32_t light = 0;
    uint64_t xn = 0;
    {
        const uint64_t i0 = tb + (uint64_t)wv * 32u + (lane & 31u);
        xn = key_at<KB>(keys, i0, n);
    }
    for (uint32_t st = 0; st < ROUTE_TILE / 128u; ++st) {
        const uint64_t i = tb + (uint64_t)(st * 4u + wv) * 32u + (lane & 31u);
        const bool valid = i < n;
        const uint64_t x = xn;
        // the next step's key is loaded while this one is scored
        const uint64_t in = i + 128u;
        xn = key_at<KB>(keys, in, n);
        const float y = mlp_mfma(S, x);
        if (lane >= 32u || !valid)
            continue;
        const bool heavy = y >= thr && x != T_EMPTY;
        flag[i] = heavy ? 0 : 1;
        if (!heavy) {
            ++light;
            continue;
        }
        uint32_t sl = (uint32_t)((x * 0x9E3779B97F4A7C15ull) >> 52) & (LT_SLOTS - 1u);
        bool done = false;
        for (uint32_t pr = 0; pr < LT_PROBES && !done; ++pr, sl = (sl + 1u) & (LT_SLOTS - 1u)) {
            unsigned long long cur = s_key[sl];
            if (cur == T_EMPTY)
                cur = atomicCAS(&s_key[sl], (unsigned long long)T_EMPTY, (unsigned long long)x);
            if (cur == T_EMPTY || cur == x) {
                atomicAdd(&s_cnt[sl], 1u);
                done = true;
            }
        }
        if (!done)
            t_add(tkeys, tcnt, tlog2, x, 1u, err);
    }
    atomicAdd(&s_light, light);
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < LT_SLOTS; e += ROUTE_THREADS)
        if (s_cnt[e])
            t_add(tkeys, tcnt, tlog2, s_key[e], s_cnt[e], err);
    if (threadIdx.x == 0)
        tiles[blockIdx.x] = s_light;
}

// stable compaction of the flagged keys of each routing tile
template <int KB>
__global__ __launch_bounds__(ROUTE_THREADS) void k_light_scatter(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                                 const uint8_t *__restrict__ flag,
                                                                 const uint32_t *__restrict__ tile_off,
                                                                 KeyT<KB> *__restrict__ out)
{
    __shared__ uint32_t s_w[ROUTE_THREADS / 64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint64_t tb = (uint64_t)blockIdx.x * ROUTE_TILE;
    uint32_t base = tile_off[blockIdx.x];
    for (uint32_t st = 0; st < ROUTE_TILE / ROUTE_THREADS; ++st) {
        const uint64_t i = tb + (uint64_t)st * ROUTE_THREADS + threadIdx.x;
        const bool f = i < n && flag[i];
        const uint64_t m = __ballot(f);
        if (lane == 0)
            s_w[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = 0, tot = 0;
#pragma unroll
        for (uint32_t w = 0; w < ROUTE_THREADS / 64; ++w) {
            off += w < wv ? s_w[w] : 0u;
            tot += s_w[w];
        }
        if (f)
            out[base + off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = keys[i];
        base += tot;
        __syncthreads();
    }
}

// After a routing pass in which the table turned keys away (*err set): every
// key flagged heavy that is not in the table goes to the sketch instead, and
// the tile's light count is redone.  A key is either in the table for all of
// its occurrences or for none: slots only fill, so a key's probe window that
// had no room for it never gains room, and once the key holds a slot every
// later probe finds it.  Blocks exit at once when *err is clear.
template <int KB>
__global__ __launch_bounds__(ROUTE_THREADS) void k_reflag(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                          const uint64_t *__restrict__ tkeys, uint32_t tlog2,
                                                          uint8_t *__restrict__ flag, uint32_t *__restrict__ tiles,
                                                          const uint32_t *__restrict__ err)
{
    if (*err == 0)
        return;
    __shared__ uint32_t s_light;
    if (threadIdx.x == 0)
        s_light = 0;
    __syncthreads();
    const uint64_t tb = (uint64_t)blockIdx.x * ROUTE_TILE;
    uint32_t light = 0;
    for (uint32_t st = 0; st < ROUTE_TILE / ROUTE_THREADS; ++st) {
        const uint64_t i = tb + (uint64_t)st * ROUTE_THREADS + threadIdx.x;
        if (i >= n)
            break;
        uint8_t f = flag[i];
        if (!f && t_find(tkeys, tlog2, key64<KB>(keys, i)) == T_NONE) {
            f = 1;
            flag[i] = 1;
        }
        light += f;
    }
    atomicAdd(&s_light, light);
    __syncthreads();
    if (threadIdx.x == 0)
        tiles[blockIdx.x] = s_light;
}

// Queries of the learned form.  A key is counted entirely in one part: only
// keys the network scores heavy ever enter the table, and a heavy key either
// holds a slot for all its occurrences or for none (k_reflag).  So a key found
// in the table is answered with its table count, and every other key with
// the sketch estimate.  Two kernels give that same answer:
//   k_table_query    looks every key up in the table first (no network); the
//                    probe ends at the first empty slot, so it is used while
//                    the table is at most half full and has never filled
//   k_learned_query  scores 32 keys per wavefront on the matrix cores and
//                    probes the table only for heavy-scored keys
// Both read a key's sketch bucket only when the table does not answer it.
typedef unsigned int lq_u32x4 __attribute__((ext_vector_type(4)));

__device__ static inline uint32_t lq_sel4(const uint4 &v, uint32_t k)
{
    const uint32_t lo = (k & 1u) ? v.y : v.x;
    const uint32_t hi = (k & 1u) ? v.w : v.z;
    return (k & 2u) ? hi : lo;
}

// Sketch estimate of key x by a pair of lanes PAIR apart, one 16-byte half of
// the bucket each (as k_query); need is the same on both lanes of a pair,
// and the exchanges stay outside the branches.
template <int KB, int PAIR>
__device__ static inline uint32_t lq_sketch(bool need, uint64_t x, uint32_t half, const qs_sdesc &d0,
                                            const uint32_t *s_masks, const uint32_t *__restrict__ low,
                                            const uint32_t *__restrict__ high)
{
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    uint32_t Mh = 0;
    if (need) {
        const KeyT<KB> key = (KeyT<KB>)x;
        const uint32_t M = s_masks[qs_maskidx_d(key, d0)];
        const uint32_t b = qs_bucket_d(key, d0);
        const lq_u32x4 vv = __builtin_nontemporal_load(reinterpret_cast<const lq_u32x4 *>(low + (size_t)b * 8u) + half);
        v = make_uint4(vv[0], vv[1], vv[2], vv[3]);
        Mh = half ? (M >> 16) : (M & 0xffffu);
    }
    uint32_t lo = 0xffffffffu;
    for (uint32_t m = Mh; m; m &= m - 1u) {
        const uint32_t jj = (uint32_t)__builtin_ctz(m);
        lo = min(lo, (lq_sel4(v, jj >> 2) >> ((jj & 3u) * 8u)) & 0xffu);
    }
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, PAIR, 64));
    const uint32_t wo = (uint32_t)__shfl_xor((int)v.w, PAIR, 64);
    const uint32_t w = half ? v.w : wo;
    uint32_t hi = 0xffffffffu;
    if (need && w > QS_LINK_START) {
        const uint32_t *g = high + (size_t)(w - QS_LINK_START) * 32u + 16u * half;
        for (uint32_t m = Mh; m; m &= m - 1u)
            hi = min(hi, g[__builtin_ctz(m)]);
    }
    hi = min(hi, (uint32_t)__shfl_xor((int)hi, PAIR, 64));
    return lo + (w > QS_LINK_START ? hi : 0u);
}

template <int KB>
__global__ __launch_bounds__(256) void k_table_query(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                     const uint64_t *__restrict__ tkeys,
                                                     const uint32_t *__restrict__ tcnt, uint32_t tlog2, qs_sdesc d0,
                                                     const uint32_t *__restrict__ low,
                                                     const uint32_t *__restrict__ high,
                                                     const uint32_t *__restrict__ masks, uint32_t *__restrict__ out)
{
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t e = threadIdx.x; e < d0.n_masks; e += blockDim.x)
        s_masks[e] = masks[e];
    __syncthreads();
    const uint32_t half = threadIdx.x & 1u;
    const uint64_t stride = (uint64_t)gridDim.x * 128u;
    for (uint64_t i = (uint64_t)blockIdx.x * 128u + (threadIdx.x >> 1); i < n; i += stride) {
        const uint64_t x = key64<KB>(keys, i);
        // the all-ones key is the empty-slot marker and is never in the table
        const uint32_t s = x != T_EMPTY ? t_find(tkeys, tlog2, x) : T_NONE;
        const bool need = s == T_NONE;
        const uint32_t est = lq_sketch<KB, 1>(need, x, half, d0, s_masks, low, high);
        if (!half)
            out[i] = need ? est : tcnt[s];
    }
}

template <int KB>
__global__ __launch_bounds__(256) void k_learned_query(const KeyT<KB> *__restrict__ keys, uint64_t n,
                                                       const uint16_t *__restrict__ pb, const float *__restrict__ p,
                                                       float thr, const uint64_t *__restrict__ tkeys,
                                                       const uint32_t *__restrict__ tcnt, uint32_t tlog2,
                                                       qs_sdesc d0, const uint32_t *__restrict__ low,
                                                       const uint32_t *__restrict__ high,
                                                       const uint32_t *__restrict__ masks,
                                                       uint32_t *__restrict__ out)
{
    __shared__ mlp_lds S;
    __shared__ uint32_t s_masks[QS_MAX_MASKS];
    for (uint32_t e = threadIdx.x; e < d0.n_masks; e += blockDim.x)
        s_masks[e] = masks[e];
    load_mlp(S, pb, p); // ends with a barrier
    const uint32_t lane = threadIdx.x & 63u, l32 = lane & 31u, half = lane >> 5;
    const uint64_t gw = ((uint64_t)blockIdx.x * 256u + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * 256u) >> 6;
    if (n == 0)
        return;
    uint64_t xn = key_at<KB>(keys, gw * 32u + l32, n);
    for (uint64_t base = gw * 32u; base < n; base += nw * 32u) {
        const uint64_t i = base + l32;
        const bool valid = i < n;
        const uint64_t x = xn;
        xn = key_at<KB>(keys, i + nw * 32u, n); // the next tile's key, loaded while this one is scored
        const float y = mlp_mfma(S, x);         // lanes 0..31: the scores of keys 0..31
        uint32_t s = T_NONE;
        if (lane < 32u && valid && y >= thr && x != T_EMPTY)
            s = t_find(tkeys, tlog2, x);
        s = (uint32_t)__shfl((int)s, (int)l32, 64); // lane l + 32 takes key l's answer
        const bool need = valid && s == T_NONE;
        const uint32_t est = lq_sketch<KB, 32>(need, x, half, d0, s_masks, low, high);
        if (lane < 32u && valid)
            out[i] = need ? est : tcnt[s];
    }
}

__global__ void k_table_clear(uint64_t *__restrict__ tkeys, uint32_t *__restrict__ tcnt, uint64_t cap)
{
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cap; e += (uint64_t)gridDim.x * blockDim.x) {
        tkeys[e] = T_EMPTY;
        tcnt[e] = 0;
    }
}

// ===========================================================================
// host
// ===========================================================================
static uint32_t lgrid(uint64_t want, uint32_t cap)
{
    return (uint32_t)(want < 1 ? 1 : (want > cap ? cap : want));
}

extern "C" void qs_learned_destroy(qs_learned *L)
{
    if (!L)
        return;
    hipDeviceSynchronize();
    qs_destroy(L->sk);
    hipFree(L->d_p);
    hipFree(L->d_pb);
    hipFree(L->d_m);
    hipFree(L->d_v);
    hipFree(L->d_g);
    hipFree(L->d_loss);
    hipFree(L->d_tkeys);
    hipFree(L->d_tcnt);
    hipFree(L->d_err);
    hipFree(L->d_flag);
    hipFree(L->d_tiles);
    hipFree(L->d_light);
    if (L->h_pin)
        hipHostFree(L->h_pin);
    free(L);
}

extern "C" qs_sketch *qs_learned_sketch(qs_learned *L)
{
    return L ? L->sk : nullptr;
}

extern "C" int qs_learned_create(const qs_params *p, uint32_t table_log2, uint64_t init_seed, qs_learned **out)
{
    if (!out || !p || table_log2 < 10 || table_log2 > 28)
        return QS_ERR_PARAM;
    *out = nullptr;
    qs_learned *L = (qs_learned *)calloc(1, sizeof(qs_learned));
    if (!L)
        return QS_ERR_NOMEM;
    int rc = qs_create(p, &L->sk);
    if (rc) {
        free(L);
        return rc;
    }
    L->key_bytes = p->key_bytes;
    L->tlog2 = table_log2;
    L->tcap = 1ull << table_log2;
    L->thr = INFINITY;
    bool ok = hipMalloc(&L->d_p, NPAR * 4) == hipSuccess && hipMalloc(&L->d_pb, NPAR * 2) == hipSuccess &&
              hipMalloc(&L->d_m, NPAR * 4) == hipSuccess && hipMalloc(&L->d_v, NPAR * 4) == hipSuccess &&
              hipMalloc(&L->d_g, NPAR * 4) == hipSuccess && hipMalloc(&L->d_loss, 4) == hipSuccess &&
              hipMalloc(&L->d_tkeys, L->tcap * 8) == hipSuccess && hipMalloc(&L->d_tcnt, L->tcap * 4) == hipSuccess &&
              hipMalloc(&L->d_err, 8) == hipSuccess &&
              hipHostMalloc(&L->h_pin, 16, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        qs_learned_destroy(L);
        return QS_ERR_NOMEM;
    }
    // initial weights: uniform in +-1/sqrt(64) from a SplitMix64 stream, biases 0
    std::vector<float> h(NPAR, 0.f);
    uint64_t st = init_seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto uni = [&]() { return (float)((double)(next() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * 0.125f; };
    for (int e = 0; e < 4096; ++e)
        h[OFF_W1 + e] = uni();
    for (int e = 0; e < 4096; ++e)
        h[OFF_W2 + e] = uni();
    for (int e = 0; e < 64; ++e)
        h[OFF_W3 + e] = uni();
    rc = qs_mlp_set_params(L, h.data());
    if (rc == QS_OK)
        rc = qs_learned_clear(L);
    if (rc) {
        qs_learned_destroy(L);
        return rc;
    }
    *out = L;
    return QS_OK;
}

extern "C" int qs_mlp_get_params(qs_learned *L, float *hp)
{
    if (!L || !hp)
        return QS_ERR_PARAM;
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(hp, L->d_p, NPAR * 4, hipMemcpyDeviceToHost));
    return QS_OK;
}

extern "C" int qs_mlp_set_params(qs_learned *L, const float *hp)
{
    if (!L || !hp)
        return QS_ERR_PARAM;
    if (L->inserted)
        return QS_ERR_STATE; // the routing of the keys already in would change
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(L->d_p, hp, NPAR * 4, hipMemcpyHostToDevice));
    CK(hipMemset(L->d_m, 0, NPAR * 4));
    CK(hipMemset(L->d_v, 0, NPAR * 4));
    CK(hipMemset(L->d_g, 0, NPAR * 4));
    hipLaunchKernelGGL(k_to_bf16, dim3(33), dim3(256), 0, 0, L->d_p, L->d_pb);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    L->step = 0;
    return QS_OK;
}

static int grad_launch(qs_learned *L, const void *d_keys, const float *d_t, uint32_t n)
{
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, 0, L->d_loss, 1u);
    const dim3 grid((n + GB - 1) / GB), block(GB);
    const float inv = 1.0f / (float)n;
    if (L->key_bytes == 4)
        hipLaunchKernelGGL(k_mlp_grad<4>, grid, block, 0, 0, (const uint32_t *)d_keys, d_t, n, L->d_pb, L->d_p, inv,
                           L->d_g, L->d_loss);
    else
        hipLaunchKernelGGL(k_mlp_grad<8>, grid, block, 0, 0, (const uint64_t *)d_keys, d_t, n, L->d_pb, L->d_p, inv,
                           L->d_g, L->d_loss);
    CK(hipGetLastError());
    return QS_OK;
}

extern "C" int qs_mlp_grad(qs_learned *L, const void *d_keys, const float *d_t, uint32_t n, float *loss,
                           float *h_grad)
{
    if (!L || !d_keys || !d_t || n == 0)
        return QS_ERR_PARAM;
    int rc = grad_launch(L, d_keys, d_t, n);
    if (rc)
        return rc;
    CK(hipDeviceSynchronize());
    if (loss)
        CK(hipMemcpy(loss, L->d_loss, 4, hipMemcpyDeviceToHost));
    if (h_grad)
        CK(hipMemcpy(h_grad, L->d_g, NPAR * 4, hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_zero_f32, dim3(33), dim3(256), 0, 0, L->d_g, (uint32_t)NPAR);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_mlp_step(qs_learned *L, const void *d_keys, const float *d_t, uint32_t n, float lr, float beta1,
                           float beta2, float eps, float *loss, float *h_grad)
{
    if (!L || !d_keys || !d_t || n == 0)
        return QS_ERR_PARAM;
    if (L->inserted)
        return QS_ERR_STATE; // the routing of the keys already in would change
    int rc = grad_launch(L, d_keys, d_t, n);
    if (rc)
        return rc;
    if (h_grad) {
        CK(hipDeviceSynchronize());
        CK(hipMemcpy(h_grad, L->d_g, NPAR * 4, hipMemcpyDeviceToHost));
    }
    L->step++;
    const float c1 = 1.f - powf(beta1, (float)L->step), c2 = 1.f - powf(beta2, (float)L->step);
    hipLaunchKernelGGL(k_adam, dim3(33), dim3(256), 0, 0, L->d_p, L->d_pb, L->d_m, L->d_v, L->d_g, lr, beta1, beta2,
                       eps, c1, c2);
    CK(hipGetLastError());
    if (loss) {
        CK(hipMemcpy(loss, L->d_loss, 4, hipMemcpyDeviceToHost));
    }
    return QS_OK;
}

extern "C" int qs_mlp_score(qs_learned *L, const void *d_keys, uint64_t n, float *d_scores, int mfma)
{
    if (!L || (!d_keys && n) || (!d_scores && n))
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    if (mfma) {
        const dim3 grid(lgrid((n + 127) / 128, 4096)), block(256);
        if (L->key_bytes == 4)
            hipLaunchKernelGGL(k_mlp_mfma<4>, grid, block, 0, 0, (const uint32_t *)d_keys, n, L->d_pb, L->d_p,
                               d_scores);
        else
            hipLaunchKernelGGL(k_mlp_mfma<8>, grid, block, 0, 0, (const uint64_t *)d_keys, n, L->d_pb, L->d_p,
                               d_scores);
    } else {
        const dim3 grid(lgrid((n + 255) / 256, 4096)), block(256);
        if (L->key_bytes == 4)
            hipLaunchKernelGGL(k_mlp_fwd<4>, grid, block, 0, 0, (const uint32_t *)d_keys, n, L->d_pb, L->d_p,
                               d_scores);
        else
            hipLaunchKernelGGL(k_mlp_fwd<8>, grid, block, 0, 0, (const uint64_t *)d_keys, n, L->d_pb, L->d_p,
                               d_scores);
    }
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_learned_set_threshold(qs_learned *L, float thr)
{
    if (!L || std::isnan(thr))
        return QS_ERR_PARAM;
    if (L->inserted && thr != L->thr)
        return QS_ERR_STATE; // the routing of the keys already in would change
    L->thr = thr;
    return QS_OK;
}

extern "C" float qs_learned_threshold(const qs_learned *L)
{
    return L ? L->thr : NAN;
}

extern "C" int qs_learned_clear(qs_learned *L)
{
    if (!L)
        return QS_ERR_PARAM;
    int rc = qs_clear(L->sk);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_table_clear, dim3(lgrid((L->tcap + 255) / 256, 2048)), dim3(256), 0, 0, L->d_tkeys,
                       L->d_tcnt, L->tcap);
    CK(hipGetLastError());
    CK(hipMemset(L->d_err, 0, 8));
    CK(hipDeviceSynchronize());
    L->inserted = 0;
    L->full = 0;
    L->used = 0;
    return QS_OK;
}

static int ensure_route(qs_learned *L, uint64_t n)
{
    if (L->rcap >= n)
        return QS_OK;
    hipFree(L->d_flag);
    hipFree(L->d_tiles);
    hipFree(L->d_light);
    L->d_flag = nullptr;
    L->d_tiles = nullptr;
    L->d_light = nullptr;
    L->rcap = 0;
    const uint64_t tiles = (n + ROUTE_TILE - 1) / ROUTE_TILE;
    if (hipMalloc(&L->d_flag, n) != hipSuccess || hipMalloc(&L->d_tiles, (tiles + 1) * 4) != hipSuccess ||
        hipMalloc(&L->d_light, n * L->key_bytes) != hipSuccess)
        return QS_ERR_NOMEM;
    L->rcap = n;
    return QS_OK;
}

template <int KB>
static int learned_insert(qs_learned *L, const KeyT<KB> *keys, uint64_t n, uint64_t *n_heavy)
{
    int rc = ensure_route(L, n);
    if (rc)
        return rc;
    const uint32_t tiles = (uint32_t)((n + ROUTE_TILE - 1) / ROUTE_TILE);
    CK(hipMemsetAsync(L->d_err, 0, 4, 0)); // this call's flag
    hipLaunchKernelGGL(k_route<KB>, dim3(tiles), dim3(ROUTE_THREADS), 0, 0, keys, n, L->d_pb, L->d_p, L->thr,
                       L->d_tkeys, L->d_tcnt, L->tlog2, L->d_flag, L->d_tiles, L->d_err);
    // keys the table turned away go to the sketch (no work unless it did)
    hipLaunchKernelGGL(k_reflag<KB>, dim3(tiles), dim3(ROUTE_THREADS), 0, 0, keys, n, L->d_tkeys, L->tlog2,
                       L->d_flag, L->d_tiles, L->d_err);
    qs_scan_digits(L->d_tiles, tiles, 1, L->d_tiles + tiles, 0);
    hipLaunchKernelGGL(k_light_scatter<KB>, dim3(tiles), dim3(ROUTE_THREADS), 0, 0, keys, n, L->d_flag, L->d_tiles,
                       (KeyT<KB> *)L->d_light);
    CK(hipGetLastError());
    CK(hipMemcpy(L->h_pin, L->d_tiles + tiles, 4, hipMemcpyDeviceToHost));
    CK(hipMemcpy(L->h_pin + 1, L->d_err, 8, hipMemcpyDeviceToHost));
    const uint64_t n_light = L->h_pin[0];
    if (L->h_pin[1])
        L->full = 1;
    L->used = L->h_pin[2];
    L->inserted += n;
    if (n_heavy)
        *n_heavy = n - n_light;
    if (n_light == 0)
        return QS_OK;
    return qs_insert(L->sk, L->d_light, n_light);
}

extern "C" int qs_learned_insert(qs_learned *L, const void *d_keys, uint64_t n, uint64_t *n_heavy)
{
    if (!L || (!d_keys && n))
        return QS_ERR_PARAM;
    if (n == 0) {
        if (n_heavy)
            *n_heavy = 0;
        return QS_OK;
    }
    if (n > 0xffffffffull)
        return QS_ERR_PARAM; // tile offsets are 32-bit
    return L->key_bytes == 4 ? learned_insert<4>(L, (const uint32_t *)d_keys, n, n_heavy)
                             : learned_insert<8>(L, (const uint64_t *)d_keys, n, n_heavy);
}

extern "C" int qs_learned_query(qs_learned *L, const void *d_keys, uint64_t n, uint32_t *d_counts)
{
    if (!L || ((!d_keys || !d_counts) && n))
        return QS_ERR_PARAM;
    if (n == 0)
        return QS_OK;
    if (n > 0xffffffffull)
        return QS_ERR_PARAM; // the sketch's query bound
    const qs_sketch *s = L->sk;
    const dim3 block(256);
    if (!L->full && 2ull * L->used <= L->tcap) {
        // probes of absent keys stay short: look the table up first
        const dim3 grid(lgrid((2 * n + 255) / 256, 2048));
        if (L->key_bytes == 4)
            hipLaunchKernelGGL(k_table_query<4>, grid, block, 0, 0, (const uint32_t *)d_keys, n, L->d_tkeys,
                               L->d_tcnt, L->tlog2, s->h_desc[0], s->d_low, s->d_high, s->d_masks, d_counts);
        else
            hipLaunchKernelGGL(k_table_query<8>, grid, block, 0, 0, (const uint64_t *)d_keys, n, L->d_tkeys,
                               L->d_tcnt, L->tlog2, s->h_desc[0], s->d_low, s->d_high, s->d_masks, d_counts);
    } else {
        const dim3 grid(lgrid((n + 127) / 128, 8192));
        if (L->key_bytes == 4)
            hipLaunchKernelGGL(k_learned_query<4>, grid, block, 0, 0, (const uint32_t *)d_keys, n, L->d_pb, L->d_p,
                               L->thr, L->d_tkeys, L->d_tcnt, L->tlog2, s->h_desc[0], s->d_low, s->d_high,
                               s->d_masks, d_counts);
        else
            hipLaunchKernelGGL(k_learned_query<8>, grid, block, 0, 0, (const uint64_t *)d_keys, n, L->d_pb, L->d_p,
                               L->thr, L->d_tkeys, L->d_tcnt, L->tlog2, s->h_desc[0], s->d_low, s->d_high,
                               s->d_masks, d_counts);
    }
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    return QS_OK;
}

extern "C" int qs_learned_table(qs_learned *L, uint64_t *h_keys, uint32_t *h_counts, uint64_t cap,
                                uint64_t *n_used)
{
    if (!L || !n_used)
        return QS_ERR_PARAM;
    std::vector<uint64_t> k(L->tcap);
    std::vector<uint32_t> c(L->tcap);
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(k.data(), L->d_tkeys, L->tcap * 8, hipMemcpyDeviceToHost));
    CK(hipMemcpy(c.data(), L->d_tcnt, L->tcap * 4, hipMemcpyDeviceToHost));
    uint64_t u = 0;
    for (uint64_t e = 0; e < L->tcap; ++e) {
        if (k[e] == T_EMPTY)
            continue;
        if (u < cap && h_keys && h_counts) {
            h_keys[u] = k[e];
            h_counts[u] = c[e];
        }
        ++u;
    }
    *n_used = u;
    return QS_OK;
}

extern "C" int qs_learned_table_full(const qs_learned *L)
{
    return L ? (int)L->full : QS_ERR_PARAM;
}

extern "C" int qs_learned_memory(const qs_learned *L, uint64_t *table_bytes, uint64_t *mlp_bytes,
                                 uint64_t *sketch_bytes)
{
    if (!L)
        return QS_ERR_PARAM;
    if (table_bytes)
        *table_bytes = L->tcap * 12u;
    if (mlp_bytes)
        *mlp_bytes = (4096u * 2 + 64u) * 2u + (64u * 2 + 1u) * 4u; // BF16 weights + FP32 biases
    if (sketch_bytes)
        *sketch_bytes = L->sk->low_words * 4u + L->sk->high_words * 4u;
    return QS_OK;
}
