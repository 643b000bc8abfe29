// Exhaustive / randomized check of qs_fastmod (the bucket and mask index
// reduction used by every kernel) against the % operator, on the host.
//   usage: fastmod_check [full]   -- "full" sweeps all 2^32 dividends for a
//   fixed divisor list; otherwise random dividends for many divisors.
#include "../../qsketch_amd/csrc/qs_internal.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char **argv)
{
    bool full = argc > 1 && !strcmp(argv[1], "full");
    const uint32_t ds[] = {1u, 2u, 3u, 7u, 1021u, 3145721u, 100663291u, 201326557u, 402653171u,
                           536870909u, 2147483647u, 2147483648u, 4294967291u, 4294967295u};
    unsigned long long bad = 0;
    for (uint32_t d : ds) {
        qs_fastdiv f = qs_fastdiv_make(d);
        if (full && d >= 1021u && d <= 402653171u) {
            for (uint64_t n = 0; n <= 0xffffffffull; ++n)
                if (qs_fastmod((uint32_t)n, f) != (uint32_t)n % d)
                    ++bad;
        }
        uint64_t x = 88172645463325252ull;
        for (int k = 0; k < 20000000; ++k) {
            x ^= x << 13; x ^= x >> 7; x ^= x << 17;
            uint32_t n = (uint32_t)x;
            if (qs_fastmod(n, f) != n % d) ++bad;
        }
        const uint32_t edges[] = {0u, 1u, d - 1u, d, d + 1u, 0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu};
        for (uint32_t n : edges)
            if (qs_fastmod(n, f) != n % d) ++bad;
    }
    // random divisors
    uint64_t y = 0x9E3779B97F4A7C15ull;
    for (int k = 0; k < 200000; ++k) {
        y ^= y << 13; y ^= y >> 7; y ^= y << 17;
        uint32_t d = (uint32_t)(y >> 32) >> (y & 31);
        if (d == 0) d = 1;
        qs_fastdiv f = qs_fastdiv_make(d);
        for (int r = 0; r < 64; ++r) {
            y ^= y << 13; y ^= y >> 7; y ^= y << 17;
            uint32_t n = (uint32_t)y;
            if (qs_fastmod(n, f) != n % d) ++bad;
        }
    }
    printf("fastmod mismatches: %llu\n", bad);
    return bad ? 1 : 0;
}
