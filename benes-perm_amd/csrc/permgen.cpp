// SPDX-License-Identifier: Apache-2.0
// Seeded synthetic permutations for tests and benchmarks.
//
// The reference seeds std::shuffle with std::mt19937_64 (tests/test_benes.cpp:
// 60-68, tools/benchmark.cpp:20-28), whose shuffle is implementation-defined;
// these generators fix every draw instead: splitmix64 outputs, reduced to a
// range with Lemire's multiply-shift rejection method.
#include <cstdint>
#include <cstring>

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

}  // namespace

extern "C" int benes_hip_generate(uint32_t kind, uint64_t n, uint64_t seed, uint32_t rate_ppm,
                                  void* out, uint32_t width_bytes) {
    if (kind > BENES_HIP_PERM_KEYED) return BENES_HIP_INVALID_ARGUMENT;
    if (width_bytes == 4) {
        if (n > 0x100000000ull) return BENES_HIP_INVALID_ARGUMENT;
        fill(kind, n, seed, rate_ppm, static_cast<uint32_t*>(out));
    } else if (width_bytes == 8) {
        fill(kind, n, seed, rate_ppm, static_cast<uint64_t*>(out));
    } else {
        return BENES_HIP_INVALID_ARGUMENT;
    }
    return BENES_HIP_OK;
}
