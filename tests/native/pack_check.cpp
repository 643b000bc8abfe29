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
    long bad = 0, valid = 0, refused = 0;
    for (uint64_t L : lows)
        for (uint32_t n : nsks)
            for (uint64_t m : masks) {
                qs_pack pk;
                if (!qs_make_pack(L, n, m, &pk)) {
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
                if ((1ull << pk.bb) < L || (n > 1 && (1ull << pk.sb) < n) || (1ull << pk.bm) < m)
                    fail("field too narrow");
                if (pk.bb + pk.sb + pk.bm + pk.ib > 64 || pk.sb + pk.ib > 32 || pk.ib > 30 || pk.ib < 12)
                    fail("widths");
                const uint64_t b = L - 1, sid = n - 1, mu = m - 1, idx = (1ull << pk.ib) - 1;
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
    printf("packings: %ld valid, %ld refused\n", valid, refused);
    printf("pack violations: %ld\n", bad);
    return bad ? 1 : 0;
}
