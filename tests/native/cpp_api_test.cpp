// Drives the C++ interface (include/qsketch.hpp) the way the reference's
// Test harness drives Sketch<Key_T, Count_T> (include/test/test.h:137-184):
// construct a Sketch_GPU_Sub_Warp_Level sized by number_of_buckets, insert
// device key batches, search, clear -- and checks every estimate against the
// CPU oracle; then the same for Sketch_GPU_Sub_Warp and
// Sketch_GPU_SF_Host_Sub_Warp.  Exit status 0 = all estimates bit-exact.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/qsketch.hpp"
#include "../../include/qsketch_gen.h"
#include "../../oracle/qs_oracle.h"

int main()
{
    using namespace qsketch;
    const size_t N = 1 << 20, batch = 1 << 18;
    const size_t n_low = Sketch_GPU_Sub_Warp_Level<unsigned, unsigned>::number_of_buckets(N, 8, 0.5);
    std::printf("n_low %zu\n", n_low);
    qs_zipf *z = nullptr;
    if (qs_zipf_create(1.1, 1 << 14, 0, &z))
        return 2;
    std::vector<unsigned> keys(N);
    qs_gen_zipf_u32_host(z, 5, 0, N, keys.data());
    qs_zipf_destroy(z);

    unsigned *d_keys = nullptr, *d_counts = nullptr;
    if (hipMalloc(&d_keys, N * 4) != hipSuccess || hipMalloc(&d_counts, N * 4) != hipSuccess)
        return 3;
    hipMemcpy(d_keys, keys.data(), N * 4, hipMemcpyHostToDevice);

    Sketch_GPU_Sub_Warp_Level<unsigned, unsigned, unsigned, unsigned> map(
        n_low, 8, 1 << 14, default_values::WARP_SIZE, default_values::seed_sz,
        default_values::HASH_MASK_TABLE_SZ, /*rng_seed=*/3);
    Sketch<unsigned, unsigned> &sk = map;
    for (size_t off = 0; off < N; off += batch)
        sk.insert(d_keys + off, batch);
    sk.search(d_keys, N, d_counts);
    std::vector<unsigned> est(N);
    hipMemcpy(est.data(), d_counts, N * 4, hipMemcpyDeviceToHost);

    orq_sketch o;
    if (orq_create(&o, n_low, 1 << 14, 1021, 1, 4, ORQ_LITERAL, 3))
        return 4;
    orq_insert(&o, keys.data(), N);
    std::vector<unsigned> ref(N);
    orq_query(&o, keys.data(), N, ref.data());
    size_t bad = 0;
    for (size_t i = 0; i < N; ++i)
        bad += est[i] != ref[i];
    std::printf("estimates differing from the oracle: %zu of %zu; linked buckets %u\n", bad, N,
                o.next - 1025);

    sk.clear();
    sk.search(d_keys, N, d_counts);
    hipMemcpy(est.data(), d_counts, N * 4, hipMemcpyDeviceToHost);
    for (size_t i = 0; i < N; ++i)
        bad += est[i] != 0;
    bool threw = false;
    try {
        Sketch_GPU_Sub_Warp_Level<unsigned, unsigned> tiny(1021, 8, 4, 32);
        tiny.insert(d_keys, N); // 3 usable high buckets cannot hold the hot keys
    } catch (const error &e) {
        threw = e.code == QS_ERR_POOL;
        std::printf("pool exhaustion reported: %s\n", e.what());
    }
    orq_destroy(&o);

    // the single-level and the SF sketch through the same Sketch& interface
    {
        Sketch_GPU_Sub_Warp<unsigned, unsigned> sl(40009, 16, 1, 1021, /*rng_seed=*/5);
        Sketch_GPU_SF_Host_Sub_Warp<unsigned, unsigned> sf(40009, 8, 3, 1, 1021, /*rng_seed=*/5);
        Sketch<unsigned, unsigned> *both[2] = {&sl, &sf};
        orq_sl osl;
        orq_sf osf;
        if (orq_sl_create(&osl, 40009, 16, 1021, 1, 4, 5) || orq_sf_create(&osf, 40009, 8, 3, 1021, 1, 4, 5))
            return 5;
        orq_sl_insert(&osl, keys.data(), N);
        orq_sf_insert(&osf, keys.data(), N);
        for (int v = 0; v < 2; ++v) {
            both[v]->insert(d_keys, N);
            both[v]->search(d_keys, N, d_counts);
            hipMemcpy(est.data(), d_counts, N * 4, hipMemcpyDeviceToHost);
            if (v == 0)
                orq_sl_query(&osl, keys.data(), N, ref.data());
            else
                orq_sf_query(&osf, keys.data(), N, ref.data());
            size_t diff = 0;
            for (size_t i = 0; i < N; ++i)
                diff += est[i] != ref[i];
            std::printf("%s: estimates differing from the oracle: %zu of %zu\n",
                        v == 0 ? "Sketch_GPU_Sub_Warp" : "Sketch_GPU_SF_Host_Sub_Warp", diff, N);
            bad += diff;
        }
        orq_sl_destroy(&osl);
        orq_sf_destroy(&osf);
    }
    hipFree(d_keys);
    hipFree(d_counts);
    std::printf("%s\n", bad == 0 && threw ? "OK" : "FAIL");
    return bad == 0 && threw ? 0 : 1;
}
