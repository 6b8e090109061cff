// SPDX-License-Identifier: Apache-2.0
// A keyed pseudo-random bijection of [0, n), computable entry by entry on the
// host or the device: a four-round Feistel network over the smallest even
// number of bits 2h with 2^(2h) >= n (round functions are splitmix64
// finalisers of the right half and a per-round key), walked along its own
// cycles until the image falls inside [0, n).  The walk ends because x itself
// lies on its cycle.  It stands in for a seeded Fisher-Yates draw where n is
// too large for a sequential shuffle (4.3e9 entries).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define BH_KEYED_HD __host__ __device__ __forceinline__
#else
#define BH_KEYED_HD inline
#endif

namespace benes_hip {

BH_KEYED_HD uint64_t keyed_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

BH_KEYED_HD uint64_t keyed_perm(uint64_t x, uint64_t n, uint64_t seed) {
    uint32_t bits = 0;
    while (bits < 64 && (n - 1) >> bits) ++bits;  // bit width of n - 1
    const uint32_t h = bits < 2 ? 1 : (bits + 1) / 2;
    const uint64_t mask = (uint64_t{1} << h) - 1;
    do {
        uint64_t l = x >> h, r = x & mask;
        for (uint64_t round = 0; round < 4; ++round) {
            const uint64_t f = keyed_mix(r ^ keyed_mix(seed + round)) & mask;
            const uint64_t t = l ^ f;
            l = r;
            r = t;
        }
        x = (l << h) | r;
    } while (x >= n);
    return x;
}

}  // namespace benes_hip
