// qsketch.hpp -- C++ host interface over the C ABI (include/qsketch.h),
// mirroring the reference's plugin interface so that code written against it
// (e.g. its Test harness, include/test/test.h:4-15) keeps compiling:
//
//   qsketch::Sketch<Key_T, Count_T>            include/sketch/sketch.h:5-28
//   qsketch::Sketch_GPU_Sub_Warp_Level<...>    include/sketch/device_sketch/device_sketch_v3.h:524-634
//   qsketch::Sketch_GPU_Sub_Warp<...>          include/sketch/device_sketch/device_sketch_v2.h:289-410
//   qsketch::Sketch_GPU_SF_Host_Sub_Warp<...>  include/sketch/device_sketch/device_sketch_v42.h:207-480
//   qsketch::default_values                    include/default.h, lib/default.cu:29-63
//   Sketch_GPU::number_of_buckets              include/sketch/device_sketch/device_sketch.h:371-376
//
// Same argument meaning as the reference: keys and counts are device
// pointers, insert/search are blocking and return 0.  Unlike the reference
// (which ignores errors in release builds, include/qsketch.h:104-108), a
// failing call throws qsketch::error carrying the qs_* code.  Header-only;
// link with libqsketch_amd.so.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "qsketch.h"
#include "qsketch_sf.h"
#include "qsketch_single.h"

namespace qsketch {

struct error : std::runtime_error {
    int code;
    explicit error(int c, const char *what)
        : std::runtime_error(std::string(what) + ": " + qs_strerror(c)), code(c) {}
};

inline int check(int rc, const char *what)
{
    if (rc != QS_OK)
        throw error(rc, what);
    return rc;
}

namespace default_values {
// lib/default.cu:39-54
constexpr std::size_t PRIME_NUMBER = 33554393;
constexpr std::size_t NEXT_LEVEL_ID_START = 1024;
constexpr unsigned WARP_SIZE = 32;
constexpr std::size_t seed_sz = 1;
constexpr std::size_t HASH_MASK_ONES = 3;
constexpr std::size_t HASH_MASK_TABLE_SZ = 1021;
} // namespace default_values

inline std::size_t find_greatest_prime(std::size_t upper_bound)
{
    return (std::size_t)qs_find_greatest_prime(upper_bound);
}

// include/sketch/sketch.h:5-28
template <typename Key_T, typename Count_T>
struct Sketch {
    virtual int insert(Key_T *, std::size_t) = 0;
    virtual int search(Key_T *, std::size_t, Count_T *) = 0;
    virtual unsigned char *pre_cal(Key_T *, std::size_t, void *) { return nullptr; }
    virtual void clear() = 0;
    virtual ~Sketch() {}
};

// The two-level MSWS.  Hashed_T / Seed_T / Hash_Function are accepted for
// source compatibility with the reference's template signature; the hash is
// always hash_mul_add (device_hash_function.h:4-18) with Hashed_T = the key
// width.  Count_T must be a 32-bit unsigned type.
template <typename Key_T, typename Count_T, typename Hashed_T = Key_T, typename Seed_T = unsigned,
          typename Hash_Function = void>
struct Sketch_GPU_Sub_Warp_Level : public Sketch<Key_T, Count_T> {
    static_assert(sizeof(Key_T) == 4 || sizeof(Key_T) == 8, "keys are 32- or 64-bit");
    static_assert(sizeof(Count_T) == 4 && std::is_unsigned<Count_T>::value, "counts are uint32");

    qs_sketch *h = nullptr;
    qs_params params{};

    // v3:544-547
    Sketch_GPU_Sub_Warp_Level(std::size_t _n_low, std::size_t _m_low, std::size_t _n_high,
                              std::size_t _m_high, std::size_t _ss = default_values::seed_sz,
                              std::size_t _n_hash_mask_table = default_values::HASH_MASK_TABLE_SZ,
                              std::uint64_t rng_seed = 1, int semantics = QS_LITERAL)
    {
        check(qs_default_params(&params, 1u << 20, 0.5), "qs_default_params");
        params.n_low = _n_low;
        params.m_low = (std::uint32_t)_m_low;
        params.n_high = _n_high;
        params.m_high = (std::uint32_t)_m_high;
        params.seed_sz = (std::uint32_t)_ss;
        params.n_masks = (std::uint32_t)_n_hash_mask_table;
        params.key_bytes = sizeof(Key_T);
        params.rng_seed = rng_seed;
        params.semantics = semantics;
        check(qs_create(&params, &h), "qs_create");
    }
    explicit Sketch_GPU_Sub_Warp_Level(const qs_params &p) : params(p)
    {
        check(qs_create(&params, &h), "qs_create");
    }
    Sketch_GPU_Sub_Warp_Level(const Sketch_GPU_Sub_Warp_Level &) = delete;
    Sketch_GPU_Sub_Warp_Level &operator=(const Sketch_GPU_Sub_Warp_Level &) = delete;
    ~Sketch_GPU_Sub_Warp_Level() override { qs_destroy(h); }

    // device_sketch.h:371-376 (m = 8 words, HASH_MASK_ONES = 3)
    static std::size_t number_of_buckets(std::size_t insert_keys_sz, std::size_t m, double factor)
    {
        double row = ((insert_keys_sz / factor) * default_values::HASH_MASK_ONES) / (double)m;
        return find_greatest_prime((std::size_t)row);
    }

    int insert(Key_T *keys, std::size_t keys_sz) override
    {
        return check(qs_insert(h, keys, keys_sz), "qs_insert");
    }
    int search(Key_T *keys, std::size_t keys_sz, Count_T *count) override
    {
        return check(qs_query(h, keys, keys_sz, reinterpret_cast<std::uint32_t *>(count)), "qs_query");
    }
    void clear() override { check(qs_clear(h), "qs_clear"); }

    // device_sketch.h:403-407: MiB of the low + high tables
    double memory_usage_mib() const
    {
        return ((double)params.n_low * 32.0 + (double)params.n_high * 128.0) / (1 << 20);
    }
};

// The single-level sub-warp sketch: n buckets of m counters, count-min.
template <typename Key_T, typename Count_T, typename Hashed_T = Key_T, typename Seed_T = unsigned,
          typename Hash_Function = void>
struct Sketch_GPU_Sub_Warp : public Sketch<Key_T, Count_T> {
    static_assert(sizeof(Key_T) == 4 || sizeof(Key_T) == 8, "keys are 32- or 64-bit");
    static_assert(sizeof(Count_T) == 4 && std::is_unsigned<Count_T>::value, "counts are uint32");

    qs_sl *h = nullptr;

    // v2:307-309
    Sketch_GPU_Sub_Warp(std::size_t _n, std::size_t _m, std::size_t _ss = default_values::seed_sz,
                        std::size_t _n_hash_mask_table = default_values::HASH_MASK_TABLE_SZ,
                        std::uint64_t rng_seed = 1)
    {
        check(qs_sl_create(_n, (std::uint32_t)_m, (std::uint32_t)_ss, (std::uint32_t)_n_hash_mask_table,
                           sizeof(Key_T), rng_seed, &h),
              "qs_sl_create");
    }
    Sketch_GPU_Sub_Warp(const Sketch_GPU_Sub_Warp &) = delete;
    Sketch_GPU_Sub_Warp &operator=(const Sketch_GPU_Sub_Warp &) = delete;
    ~Sketch_GPU_Sub_Warp() override { qs_sl_destroy(h); }

    int insert(Key_T *keys, std::size_t keys_sz) override
    {
        return check(qs_sl_insert(h, keys, keys_sz), "qs_sl_insert");
    }
    int search(Key_T *keys, std::size_t keys_sz, Count_T *count) override
    {
        return check(qs_sl_query(h, keys, keys_sz, reinterpret_cast<std::uint32_t *>(count)), "qs_sl_query");
    }
    void clear() override { check(qs_sl_clear(h), "qs_sl_clear"); }
};

// The SF sketch: n buckets of m counters, each split `times` ways; insert and
// the update of the counters run on the GPU.
template <typename Key_T, typename Count_T, typename Hashed_T = Key_T, typename Seed_T = unsigned,
          typename Hash_Function = void>
struct Sketch_GPU_SF_Host_Sub_Warp : public Sketch<Key_T, Count_T> {
    static_assert(sizeof(Key_T) == 4 || sizeof(Key_T) == 8, "keys are 32- or 64-bit");
    static_assert(sizeof(Count_T) == 4 && std::is_unsigned<Count_T>::value, "counts are uint32");

    qs_sf *h = nullptr;

    // v42:242-243
    Sketch_GPU_SF_Host_Sub_Warp(std::size_t _n, std::size_t _m, unsigned _times,
                                std::size_t _ss = default_values::seed_sz,
                                std::size_t _n_hash_mask_table = default_values::HASH_MASK_TABLE_SZ,
                                std::uint64_t rng_seed = 1)
    {
        check(qs_sf_create(_n, (std::uint32_t)_m, _times, (std::uint32_t)_ss, (std::uint32_t)_n_hash_mask_table,
                           sizeof(Key_T), rng_seed, &h),
              "qs_sf_create");
    }
    Sketch_GPU_SF_Host_Sub_Warp(const Sketch_GPU_SF_Host_Sub_Warp &) = delete;
    Sketch_GPU_SF_Host_Sub_Warp &operator=(const Sketch_GPU_SF_Host_Sub_Warp &) = delete;
    ~Sketch_GPU_SF_Host_Sub_Warp() override { qs_sf_destroy(h); }

    int insert(Key_T *keys, std::size_t keys_sz) override
    {
        return check(qs_sf_insert(h, keys, keys_sz), "qs_sf_insert");
    }
    int search(Key_T *keys, std::size_t keys_sz, Count_T *count) override
    {
        return check(qs_sf_query(h, keys, keys_sz, reinterpret_cast<std::uint32_t *>(count)), "qs_sf_query");
    }
    void clear() override { check(qs_sf_clear(h), "qs_sf_clear"); }
};

} // namespace qsketch
