// SPDX-License-Identifier: Apache-2.0
// Seeded synthetic permutations for tests and benchmarks.
//
// The reference seeds std::shuffle with std::mt19937_64 (tests/test_benes.cpp:
// 60-68, tools/benchmark.cpp:20-28), whose shuffle is implementation-defined;
// these generators fix every draw instead: splitmix64 outputs, reduced to a
// range with Lemire's multiply-shift rejection method.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/benes_hip.h"
#include "keyed.hpp"

namespace {

struct SplitMix64 {
    uint64_t state;
    uint64_t next() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Uniform in [0, range), range >= 1 (Lemire 2019).
    uint64_t below(uint64_t range) {
        unsigned __int128 m = static_cast<unsigned __int128>(next()) * range;
        uint64_t low = static_cast<uint64_t>(m);
        if (low < range) {
            const uint64_t threshold = (0 - range) % range;
            while (low < threshold) {
                m = static_cast<unsigned __int128>(next()) * range;
                low = static_cast<uint64_t>(m);
            }
        }
        return static_cast<uint64_t>(m >> 64);
    }
};

template <typename T>
void fill(uint32_t kind, uint64_t n, uint64_t seed, uint32_t rate_ppm, T* out) {
    SplitMix64 rng{seed};
    if (kind == BENES_HIP_PERM_KEYED) {
        for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<T>(benes_hip::keyed_perm(i, n, seed));
        return;
    }
    if (kind == BENES_HIP_PERM_REVERSE) {
        for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<T>(n - 1 - i);
        return;
    }
    for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<T>(i);
    if (kind == BENES_HIP_PERM_FISHER_YATES) {
        for (uint64_t i = n; i-- > 1;) {
            const uint64_t j = rng.below(i + 1);
            const T tmp = out[i];
            out[i] = out[j];
            out[j] = tmp;
        }
    } else if (kind == BENES_HIP_PERM_SINGLE_CYCLE) {
        for (uint64_t i = n; i-- > 1;) {
            const uint64_t j = rng.below(i);
            const T tmp = out[i];
            out[i] = out[j];
            out[j] = tmp;
        }
    } else if (kind == BENES_HIP_PERM_NEAR_IDENTITY && n >= 2) {
        const uint64_t swaps = static_cast<uint64_t>(
            (static_cast<unsigned __int128>(n) * rate_ppm) / 1000000u);
        for (uint64_t k = 0; k < swaps; ++k) {
            const uint64_t i = rng.below(n);
            uint64_t j = rng.below(n - 1);
            if (j >= i) ++j;
            const T tmp = out[i];
            out[i] = out[j];
            out[j] = tmp;
        }
    }
}

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform random permutation drawn on every host thread (Rao 1961,
// Sandelius 1962): each index goes to one of 2^k buckets by its own draw,
// the buckets are laid out one after another (indices in increasing order
// within a bucket), and each bucket is then shuffled by Fisher-Yates with a
// stream of its own.  Every arrangement of 0..n-1 is equally likely, as with
// one Fisher-Yates pass over the whole array, and the result depends only on
// (n, seed), not on the number of threads.
template <typename T>
void fill_uniform_parallel(uint64_t n, uint64_t seed, T* out) {
    uint32_t k = 0;
    while (k < 20 && (n >> (k + 18)) != 0) ++k;  // about 2^18 entries per bucket
    const uint64_t buckets = uint64_t{1} << k;
    const uint64_t key = mix64(seed ^ 0x5EEDB0C7u);
    auto bucket_of = [&](uint64_t i) {
        return k == 0 ? uint64_t{0} : mix64(key + i * 0x9E3779B97F4A7C15ull) >> (64 - k);
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const uint64_t threads = std::min<uint64_t>(std::min<uint64_t>(hw, 64), std::max<uint64_t>(1, n >> 16));
    auto run = [&](auto&& body) {
        std::vector<std::thread> pool;
        for (uint64_t t = 1; t < threads; ++t) pool.emplace_back(body, t);
        body(0);
        for (auto& th : pool) th.join();
    };
    // Per-thread bucket counts over its contiguous index range, then cursors.
    std::vector<uint64_t> cursor(threads * buckets, 0);
    auto range = [&](uint64_t t, uint64_t& lo, uint64_t& hi) {
        lo = n / threads * t + std::min<uint64_t>(t, n % threads);
        hi = lo + n / threads + (t < n % threads ? 1 : 0);
    };
    run([&](uint64_t t) {
        uint64_t lo, hi;
        range(t, lo, hi);
        uint64_t* c = cursor.data() + t * buckets;
        for (uint64_t i = lo; i < hi; ++i) ++c[bucket_of(i)];
    });
    std::vector<uint64_t> start(buckets + 1, 0);
    uint64_t at = 0;
    for (uint64_t b = 0; b < buckets; ++b) {
        start[b] = at;
        for (uint64_t t = 0; t < threads; ++t) {
            const uint64_t c = cursor[t * buckets + b];
            cursor[t * buckets + b] = at;
            at += c;
        }
    }
    start[buckets] = n;
    run([&](uint64_t t) {
        uint64_t lo, hi;
        range(t, lo, hi);
        uint64_t* c = cursor.data() + t * buckets;
        for (uint64_t i = lo; i < hi; ++i) out[c[bucket_of(i)]++] = static_cast<T>(i);
    });
    run([&](uint64_t t) {
        for (uint64_t b = t; b < buckets; b += threads) {
            SplitMix64 rng{mix64(key ^ (b * 0xD1B54A32D192ED03ull))};
            T* const p = out + start[b];
            const uint64_t len = start[b + 1] - start[b];
            for (uint64_t i = len; i-- > 1;) {
                const uint64_t j = rng.below(i + 1);
                const T tmp = p[i];
                p[i] = p[j];
                p[j] = tmp;
            }
        }
    });
}

}  // namespace

extern "C" int benes_hip_generate(uint32_t kind, uint64_t n, uint64_t seed, uint32_t rate_ppm,
                                  void* out, uint32_t width_bytes) {
    if (kind > BENES_HIP_PERM_UNIFORM_PARALLEL) return BENES_HIP_INVALID_ARGUMENT;
    if (width_bytes == 4) {
        if (n > 0x100000000ull) return BENES_HIP_INVALID_ARGUMENT;
        if (kind == BENES_HIP_PERM_UNIFORM_PARALLEL) {
            fill_uniform_parallel(n, seed, static_cast<uint32_t*>(out));
        } else {
            fill(kind, n, seed, rate_ppm, static_cast<uint32_t*>(out));
        }
    } else if (width_bytes == 8) {
        if (kind == BENES_HIP_PERM_UNIFORM_PARALLEL) {
            fill_uniform_parallel(n, seed, static_cast<uint64_t*>(out));
        } else {
            fill(kind, n, seed, rate_ppm, static_cast<uint64_t*>(out));
        }
    } else {
        return BENES_HIP_INVALID_ARGUMENT;
    }
    return BENES_HIP_OK;
}
