// CPU check of the 64-bit insert item packing (qs_make_pack): every field
// round-trips at its extreme values, and the 32-bit allocation-order key
// (sketch << ib | stream index) decodes to the sketch with the GPU's shift
// semantics (a 32-bit shift uses the low 5 bits of its count).
#include <cstdio>
#include <cstdint>
#include <vector>

#include "../../qsketch_amd/csrc/qs_internal.h"

static uint32_t gpu_shl32(uint32_t x, uint32_t s) { return x << (s & 31u); }
static uint32_t gpu_shr32(uint32_t x, uint32_t s) { return x >> (s & 31u); }

int main()
{
    std::vector<uint64_t> lows = {1, 2, 3, 13, 61, 761, 1021, 4093, 12281, 65521, 786431, 1572853,
                                  3145721, 201326557, (1ull << 29), (1ull << 31) - 1};
    std::vector<uint32_t> nsks = {1, 2, 3, 24, 64, 4096, 4097, 65536, 1u << 20};
    std::vector<uint64_t> masks = {1, 2, 61, 509, 1021, 1024, 1025, 4096};
    long bad = 0, valid = 0, refused = 0, quot = 0;
    uint64_t rng = 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
        rng ^= rng << 13;
        rng ^= rng >> 7;
        rng ^= rng << 17;
        return rng;
    };
    for (uint32_t kb : {4u, 8u})
    for (uint64_t L : lows)
        for (uint32_t n : nsks)
            for (uint64_t m : masks) {
                qs_pack pk;
                // a set's smallest table decides the quotient width
                const uint64_t Lmin = n > 1 ? (L > 13 ? 13 : L) : L;
                if (!qs_make_pack(L, Lmin, n, m, kb, &pk)) {
                    ++refused;
                    continue;
                }
                ++valid;
                auto fail = [&](const char *what) {
                    if (bad < 20)
                        printf("violation %s: n_low=%llu nsk=%u n_masks=%llu bb=%u sb=%u bm=%u ib=%u\n", what,
                               (unsigned long long)L, n, (unsigned long long)m, pk.bb, pk.sb, pk.bm, pk.ib);
                    ++bad;
                };
                const uint64_t field_max = pk.quot ? 0xffffffffull / Lmin : m - 1; // largest mask-field value
                if ((1ull << pk.bb) < L || (n > 1 && (1ull << pk.sb) < n) || (1ull << pk.bm) <= field_max)
                    fail("field too narrow");
                if (pk.quot && (kb != 4 || (1ull << pk.bm) >= m))
                    fail("quotient form where the mask index is as narrow");
                if (!pk.quot && kb == 4 && (0xffffffffull / Lmin) < (1ull << (pk.bm - 1)) && pk.bm > 1)
                    fail("mask index where the quotient is narrower");
                if (pk.quot) {
                    // the mask index recovered from (quotient, bucket) equals h % n_masks
                    ++quot;
                    pk.q_low = (uint32_t)L;
                    pk.q_fd = qs_fastdiv_make((uint32_t)m);
                    const qs_fastdiv fl = qs_fastdiv_make((uint32_t)L);
                    qs_sdesc D[2] = {};
                    D[0].n_low = (uint32_t)Lmin;
                    D[0].fd_masks = qs_fastdiv_make((uint32_t)m);
                    D[1].n_low = (uint32_t)L;
                    D[1].fd_masks = qs_fastdiv_make((uint32_t)(m > 7 ? m - 6 : m));
                    for (int r = 0; r < 200; ++r) {
                        const uint32_t h = r < 2 ? (r ? 0xffffffffu : 0u) : (uint32_t)next();
                        const uint32_t q = qs_fastquot(h, fl), b = h - q * (uint32_t)L;
                        if (q != h / (uint32_t)L || b != h % (uint32_t)L)
                            fail("quotient");
                        const uint64_t it = b | ((uint64_t)q << pk.mshift) | (1ull << pk.ishift);
                        if (item_mu<false>(it, pk, nullptr) != h % (uint32_t)m)
                            fail("mask index from quotient");
                        if (n > 1) {
                            for (uint32_t sd = 0; sd < 2; ++sd) {
                                const uint32_t Ls = D[sd].n_low, ms = D[sd].fd_masks.d;
                                const uint32_t qs = h / Ls, bs = h % Ls;
                                const uint64_t its = bs | ((uint64_t)sd << pk.bb) | ((uint64_t)qs << pk.mshift);
                                if (item_mu<true>(its, pk, D) != h % ms)
                                    fail("set mask index from quotient");
                            }
                        }
                    }
                }
                if (pk.bb + pk.sb + pk.bm + pk.ib > 64 || pk.sb + pk.ib > 32 || pk.ib > 30 || pk.ib < 12)
                    fail("widths");
                const uint64_t b = L - 1, sid = n - 1, mu = field_max, idx = (1ull << pk.ib) - 1;
                const uint64_t it = b | (sid << pk.bb) | (mu << pk.mshift) | (idx << pk.ishift);
                if (it_b(it, pk) != b || it_sid(it, pk) != sid || it_mu(it, pk) != mu || it_idx(it, pk) != idx ||
                    it_cell(it, pk) != (b | (sid << pk.bb)))
                    fail("round trip");
                for (uint32_t s : {0u, 1u, (uint32_t)sid / 2, (uint32_t)sid}) {
                    if (s > sid)
                        continue;
                    const uint32_t key = gpu_shl32(s, pk.ib) | (uint32_t)idx;
                    if (gpu_shr32(key, pk.ib) != s)
                        fail("order key decode");
                    if (s > 0 && gpu_shl32(s, pk.ib) <= gpu_shl32(s - 1, pk.ib) + (uint32_t)idx)
                        fail("order key monotone");
                }
            }
    printf("packings: %ld valid (%ld in the quotient form), %ld refused\n", valid, quot, refused);
    printf("pack violations: %ld\n", bad);
    return bad ? 1 : 0;
}
