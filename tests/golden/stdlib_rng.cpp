// Known-answer generator for the oracle's <random> restatement: runs the
// reference's call pattern (include/tool/tool.h:189-220, lib/hashmask.cu:59-79)
// on this toolchain's libstdc++ and prints the results as JSON.
//   g++ -O2 -std=c++17 stdlib_rng.cpp -o stdlib_rng && ./stdlib_rng > stdlib_rng.json
#include <climits>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

static void seed_buffer(uint64_t rng_seed, uint64_t stream, char out[128])
{
    uint64_t st = rng_seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 16; ++i) {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        for (int k = 0; k < 8; ++k)
            out[i * 8 + k] = (char)(uint8_t)(z >> (8 * k));
    }
}

static void print_list(const std::vector<unsigned long long> &v)
{
    printf("[");
    for (size_t i = 0; i < v.size(); ++i)
        printf("%s%llu", i ? ", " : "", v[i]);
    printf("]");
}

int main()
{
    printf("{\n");
    // 1. raw minstd_rand0 streams
    printf("\"minstd\": {");
    const unsigned long long seeds[] = {0ull, 1ull, 42ull, 2147483647ull, 2147483648ull, 1099511627781ull};
    for (int s = 0; s < 6; ++s) {
        std::minstd_rand0 e(seeds[s]);
        std::vector<unsigned long long> v;
        for (int i = 0; i < 8; ++i)
            v.push_back(e());
        printf("%s\"%llu\": ", s ? ", " : "", seeds[s]);
        print_list(v);
    }
    printf("},\n");
    // 2. seed_seq::generate
    printf("\"seed_seq\": [");
    {
        std::vector<std::vector<uint32_t>> ins = {{}, {1u}, {1u, 2u, 3u}, {0xffffffffu, 0x80000000u, 7u, 9u, 11u}};
        std::vector<size_t> ns = {1, 4, 7, 10, 40};
        bool first = true;
        for (auto &in : ins)
            for (size_t n : ns) {
                std::seed_seq sq(in.begin(), in.end());
                std::vector<uint32_t> out(n);
                sq.generate(out.begin(), out.end());
                std::vector<unsigned long long> iv(in.begin(), in.end()), ov(out.begin(), out.end());
                printf("%s{\"in\": ", first ? "" : ", ");
                print_list(iv);
                printf(", \"n\": %zu, \"out\": ", n);
                print_list(ov);
                printf("}");
                first = false;
            }
    }
    printf("],\n");
    // 3. engines seeded like CPU_Tools (seed_seq over 128 chars) and the two
    //    distributions the reference uses, plus the full create sequence.
    printf("\"create\": [");
    const unsigned long long rs[] = {1ull, 2ull, 12345ull, 0xdeadbeefcafef00dull};
    for (int k = 0; k < 4; ++k) {
        char buf[128];
        seed_buffer(rs[k], 0, buf);
        std::default_random_engine seng;
        {
            std::seed_seq sq(buf, buf + 128);
            seng.seed(sq);
        }
        std::default_random_engine probe = seng;
        std::vector<unsigned long long> raw;
        for (int i = 0; i < 4; ++i)
            raw.push_back(probe());
        std::uniform_int_distribution<unsigned int> dis_seed(1, UINT_MAX);
        std::vector<unsigned long long> sd;
        for (int i = 0; i < 6; ++i)
            sd.push_back(dis_seed(seng));
        seed_buffer(rs[k], 1, buf);
        std::default_random_engine meng;
        {
            std::seed_seq sq(buf, buf + 128);
            meng.seed(sq);
        }
        std::uniform_int_distribution<unsigned int> dis(0, 27);
        std::vector<unsigned long long> masks;
        for (int i = 0; i < 1021; ++i) {
            unsigned char hm[4] = {0, 0, 0, 0};
            int j = 0;
            while (j < 3) {
                unsigned id = dis(meng);
                unsigned char b = (unsigned char)(1u << (id % 8));
                if ((hm[id / 8] & b) == 0) {
                    hm[id / 8] |= b;
                    j++;
                }
            }
            masks.push_back((unsigned long long)hm[0] | ((unsigned long long)hm[1] << 8) |
                            ((unsigned long long)hm[2] << 16) | ((unsigned long long)hm[3] << 24));
        }
        printf("%s{\"rng_seed\": %llu, \"seed_engine_raw\": ", k ? ", " : "", rs[k]);
        print_list(raw);
        printf(", \"seeds\": ");
        print_list(sd);
        printf(", \"masks\": ");
        print_list(masks);
        printf("}");
    }
    printf("],\n");
    // 4. uniform_int_distribution on a plain-seeded engine, several ranges
    printf("\"uniform\": [");
    {
        const unsigned ranges[][2] = {{0u, 27u}, {1u, UINT_MAX}, {0u, 2147483645u}, {5u, 5u},
                                      {0u, 2147483646u}, {3u, 4000000000u}};
        for (int r = 0; r < 6; ++r) {
            std::minstd_rand0 e(777u + r);
            std::uniform_int_distribution<unsigned int> d(ranges[r][0], ranges[r][1]);
            std::vector<unsigned long long> v;
            for (int i = 0; i < 16; ++i)
                v.push_back(d(e));
            printf("%s{\"seed\": %u, \"a\": %u, \"b\": %u, \"out\": ", r ? ", " : "", 777u + r,
                   ranges[r][0], ranges[r][1]);
            print_list(v);
            printf("}");
        }
    }
    printf("]\n}\n");
    return 0;
}
