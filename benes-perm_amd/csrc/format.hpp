// SPDX-License-Identifier: Apache-2.0
// Format core of the compact version-1 Benes blob, shared by host and device.
//
// The reference walks a recursive SectionAddress object down the tree
// (src/internal.hpp:475-718) and memoises metrics in a FormatSizeCache
// (src/internal.hpp:225-313).  Here the same quantities are closed forms of
// (depth, path):
//
//  * every node at depth d has floor(N/2^d) entries, plus one when the
//    bit-reversed path is below N mod 2^d (node_count / node_large);
//  * metrics of a count only ever need the two counts present at each depth,
//    so the host builds per-depth tables bottom-up (Plan) and the device reads
//    them; no 96/128-byte per-node descriptors are stored.
//
// The tables hold, per depth and per "large" flag: switch bits, rank bits,
// the residual (frontier-region) bits a left sibling leaves behind, and for
// cluster-group roots the packet sizes.  node_address() then reproduces the
// reference's SwitchBitAddress (internal.hpp:456-473) for any node in O(depth).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define BH_HD __host__ __device__ __forceinline__
#else
#define BH_HD inline
#endif

namespace benes_hip {

constexpr int kMaxDepth = 64;

// Per-side description of one depth of the tree.
struct SideDepth {
    // Residual bits of a left sibling at this depth when the walk turns right
    // into its brother: sec(c) - packet_switch_bits(c, levels left in the
    // parent's cluster group), indexed by the sibling's large flag.
    uint64_t residual[2];
    // Only meaningful when this depth starts a cluster group:
    uint64_t group_switch_bits[2];  // packet_switch_bits(root, group depth)
    uint64_t full_packet_bits[2];   // bits in any packet but the last
    uint8_t group_root;             // 1 when a cluster group starts here
    uint8_t group_depth;            // levels of the group containing this depth
    uint8_t relative_level;         // depth - group root depth
    uint8_t root_depth;             // depth of the group root
};

struct Plan {
    uint64_t n;                 // entries
    uint32_t leaf_log2;         // middle_group_log2
    uint32_t leaf;              // 1 << leaf_log2
    uint32_t depth;             // routing depth D: nodes at depth D are leaves
    uint32_t placement;         // 0 separate, 1 embedded input, 2 embedded output
    uint64_t input_offset, output_offset, middle_offset, end_offset;  // bytes
    uint64_t switch_bits_total, middle_bits_total;
    uint64_t middle_bits[kMaxDepth + 1][2];   // rank bits of a subtree at depth d
    SideDepth side[2][kMaxDepth + 1];
};

// Address of the switch bits of one node on one side (internal.hpp:456-473).
struct SideAddr {
    uint64_t packet_base;
    uint64_t full_packet_bits;
    uint64_t full_prefix;
    uint64_t tail_prefix;
    uint64_t tail_packet;
    uint32_t span_log2;
    uint32_t pad;
};

BH_HD uint64_t bit_of(const SideAddr& a, uint64_t pair) {
    const uint64_t packet = pair >> a.span_log2;
    return a.packet_base + packet * a.full_packet_bits +
           (packet == a.tail_packet ? a.tail_prefix : a.full_prefix) +
           (pair & ((uint64_t{1} << a.span_log2) - 1));
}

BH_HD uint64_t min_u64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// Reverse the low `bits` bits of v.
BH_HD uint64_t reverse_bits(uint64_t v, uint32_t bits) {
    if (bits == 0) return 0;
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse64(v) >> (64 - bits);
#else
    uint64_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
#endif
}

// Among the 2^level nodes of one level below a root of count C, the m = C mod
// 2^level nodes with bit-reversed path < m hold one extra entry.  This counts
// those "large" nodes among the first k paths: split on the first branch (the
// top bit of the reversed path); even paths come first in reversed order.
BH_HD uint64_t large_before(uint64_t k, uint64_t m, uint32_t level) {
    uint64_t included = 0;
    for (uint32_t bits = level; bits != 0; --bits) {
        const uint64_t half = uint64_t{1} << (bits - 1);
        const uint64_t even = (k + 1) >> 1;  // paths with first branch left
        if (m <= half) {
            k = even;
        } else {
            included += even;
            k >>= 1;
            m -= half;
        }
    }
    return included + ((k != 0 && m != 0) ? 1 : 0);
}

BH_HD bool node_large(uint64_t n, uint32_t d, uint64_t path) {
    if (d == 0) return false;
    const uint64_t m = n & ((uint64_t{1} << d) - 1);
    return reverse_bits(path, d) < m;
}

BH_HD uint64_t node_count(uint64_t n, uint32_t d, uint64_t path) {
    return (n >> d) + (node_large(n, d, path) ? 1 : 0);
}

// Offset of node `path` at depth d in the level array: nodes tile [0, n) in
// path order.
BH_HD uint64_t node_offset(uint64_t n, uint32_t d, uint64_t path) {
    if (d == 0) return 0;
    const uint64_t m = n & ((uint64_t{1} << d) - 1);
    return path * (n >> d) + large_before(path, m, d);
}

// Switch bits a node of count c contributes to packet P at relative level l of
// a group of depth g (the reference's `contribution`, internal.hpp:617-629).
BH_HD uint64_t packet_share(uint64_t c, uint64_t packet, uint32_t span_log2,
                            uint32_t leaf) {
    if (c <= leaf) return 0;
    const uint64_t pairs = c >> 1;
    const uint64_t first = packet << span_log2;
    if (first >= pairs) return 0;
    return min_u64(uint64_t{1} << span_log2, pairs - first);
}

// Bits of packet P at relative level l contributed by the first k nodes of
// that level in a group whose root holds C entries.
BH_HD uint64_t level_prefix_bits(uint64_t C, uint32_t g, uint32_t l,
                                 uint64_t packet, uint64_t k, uint32_t leaf) {
    const uint64_t small = C >> l;
    const uint64_t m = C & ((uint64_t{1} << l) - 1);
    const uint32_t span_log2 = g - l - 1;
    // All m large nodes lie among all 2^l paths.
    const uint64_t inc = k == (uint64_t{1} << l) ? m : large_before(k, m, l);
    return inc * packet_share(small + 1, packet, span_log2, leaf) +
           (k - inc) * packet_share(small, packet, span_log2, leaf);
}

// Bits of packet P that precede node `path` at relative level r.
BH_HD uint64_t packet_prefix(uint64_t C, uint32_t g, uint32_t r, uint64_t path,
                             uint64_t packet, uint32_t leaf) {
    uint64_t bits = 0;
    for (uint32_t l = 0; l < r; ++l) {
        bits += level_prefix_bits(C, g, l, packet, uint64_t{1} << l, leaf);
    }
    return bits + level_prefix_bits(C, g, r, packet, path, leaf);
}

// Result of walking from the root to node (d, path) on one side.
struct Walk {
    uint64_t packet_base;   // start of the enclosing group's section
    uint64_t preceding;     // frontier-region bits before this node in its group
    uint64_t root_count;    // entries of the enclosing group root
    uint64_t group_switch_bits;
    uint64_t full_packet_bits;
    uint32_t root_depth;
    uint32_t group_depth;
    uint32_t relative_level;
    uint32_t pad;
};

// State of a walk standing at the root.
BH_HD Walk walk_root(const Plan& p, int side) {
    Walk w{};
    w.root_count = p.n;
    const SideDepth* sd = p.side[side];
    if (p.n > p.leaf) {
        w.group_switch_bits = sd[0].group_switch_bits[0];
        w.full_packet_bits = sd[0].full_packet_bits[0];
        w.group_depth = sd[0].group_depth;
    }
    return w;
}

// Continues a walk standing at depth `from` on the way to node (d, path),
// accumulating what the reference's SectionAddress::child() accumulates
// (internal.hpp:553-601): the residual bits of every left sibling passed on
// the way, and each new group's base.
BH_HD Walk walk_from(const Plan& p, int side, Walk w, uint32_t from, uint32_t d, uint64_t path) {
    const SideDepth* sd = p.side[side];
    for (uint32_t i = from + 1; i <= d; ++i) {
        const uint64_t here = path >> (d - i);
        if (here & 1) {
            const uint64_t sibling = here - 1;
            w.preceding += sd[i].residual[node_large(p.n, i, sibling) ? 1 : 0];
        }
        if (sd[i].group_root) {
            // The walk leaves the current group: the child starts a new section.
            w.packet_base += w.group_switch_bits + w.preceding;
            w.preceding = 0;
            w.root_depth = i;
            const int large = node_large(p.n, i, here) ? 1 : 0;
            w.root_count = (p.n >> i) + large;
            if (w.root_count > p.leaf) {
                w.group_switch_bits = sd[i].group_switch_bits[large];
                w.full_packet_bits = sd[i].full_packet_bits[large];
                w.group_depth = sd[i].group_depth;
            } else {
                w.group_switch_bits = 0;
                w.full_packet_bits = 0;
                w.group_depth = 0;
            }
        }
    }
    w.relative_level = d - w.root_depth;
    return w;
}

BH_HD Walk walk_to(const Plan& p, int side, uint32_t d, uint64_t path) {
    return walk_from(p, side, walk_root(p, side), 0, d, path);
}

// SwitchBitAddress of internal node (d, path) on one side, from the walk
// that reached it.
BH_HD SideAddr address_at(const Plan& p, const Walk& w, uint32_t d, uint64_t path) {
    SideAddr a{};
    const uint32_t g = w.group_depth, r = w.relative_level;
    const uint64_t local = path & ((uint64_t{1} << r) - 1);
    const uint64_t count = node_count(p.n, d, path);
    a.packet_base = w.packet_base;
    a.full_packet_bits = w.full_packet_bits;
    a.span_log2 = g - r - 1;
    const uint64_t pairs = count >> 1;
    a.tail_packet = pairs == 0 ? 0 : (pairs - 1) >> a.span_log2;
    a.full_prefix = packet_prefix(w.root_count, g, r, local, 0, p.leaf);
    a.tail_prefix = packet_prefix(w.root_count, g, r, local, a.tail_packet, p.leaf);
    return a;
}

// SwitchBitAddress of internal node (d, path) on one side.
BH_HD SideAddr node_address(const Plan& p, int side, uint32_t d, uint64_t path) {
    return address_at(p, walk_to(p, side, d, path), d, path);
}

// Bit offset of the factoradic rank of leaf (d, path) within its section:
// the middle section for separate placement (left-to-right concatenation,
// benes.cpp:317-319), or the embedding side's section (internal.hpp:506-517).
BH_HD uint64_t leaf_rank_bit(const Plan& p, uint32_t d, uint64_t path) {
    if (p.placement == 0) {
        uint64_t bit = 0;
        for (uint32_t i = 1; i <= d; ++i) {
            const uint64_t here = path >> (d - i);
            if (here & 1) bit += p.middle_bits[i][node_large(p.n, i, here - 1) ? 1 : 0];
        }
        return bit;
    }
    const Walk w = walk_to(p, p.placement == 1 ? 0 : 1, d, path);
    if (w.root_count <= p.leaf) return w.packet_base;
    return w.packet_base + w.group_switch_bits + w.preceding;
}

// Embedded rank offset from a walk that reached the leaf.
BH_HD uint64_t rank_bit_at(const Walk& w, uint32_t leaf) {
    if (w.root_count <= leaf) return w.packet_base;
    return w.packet_base + w.group_switch_bits + w.preceding;
}

BH_HD uint64_t leaf_section_offset(const Plan& p) {
    return p.placement == 0 ? p.middle_offset
         : p.placement == 1 ? p.input_offset : p.output_offset;
}

// Width in bits of the rank of a c-entry leaf: bit width of c! - 1.
BH_HD uint32_t rank_bits(uint32_t c) {
    // c = 0..32
    constexpr uint8_t table[33] = {0, 0, 1, 3, 5, 7, 10, 13, 16, 19, 22, 26, 29,
                                   33, 37, 41, 45, 49, 53, 57, 62, 66, 70, 75,
                                   80, 84, 89, 94, 98, 103, 108, 113, 118};
    return table[c];
}

}  // namespace benes_hip
