// SPDX-License-Identifier: Apache-2.0
// gfx950 kernels of the compact-v1 compression path.
//
// Level array layout (HBM): at every depth d the nodes tile [0, n) in path
// order and each node holds its node-local permutation (values < its count);
// a node's two children occupy exactly its range (left = first ceil(c/2)).
// Levels ping-pong between two u32 arrays, so routing is a pure streaming
// pass; only the union-find touches memory at random.
//
// Every launch serves a batch: item i's entries live at [base_i, base_i + n_i)
// of the shared arrays and its blob at word blob_i of the shared blob.  A
// launch carries a table of segments (one per item taking part) giving the
// item's first block; a block finds its segment by binary search.  A plain
// compress is a batch of one.
//
// Before the first level:
//   k_scatter_inverse, k_check_inverse
//                    validation and the inverse permutation, no atomics
// Per level, for every depth above the fusion depth:
//   k_describe       one thread per node: offsets and both sides' addresses
// Nodes larger than one workgroup's LDS (HBM levels):
//   k_hook_first, k_union
//                    one thread per output pair: parity union-find over the
//                    node's input slots in HBM (plain-store first hooks, then
//                    one CAS per remaining hook; long walks deferred)
//   k_union_plain, k_union_verify, k_union_rejoin
//                    instead of k_union where a node keeps to one XCD and its
//                    words fit in 8 MB: the hooks are plain stores, each XCD
//                    uniting about 2 MB of nodes at a time in its L2; the
//                    check then collects every relation a concurrent hook
//                    overwrote, and the re-join joins them again
//   k_hook_bin, k_hook_scatter
//                    the first hooks of very large nodes instead: staged by
//                    word bin, then written bin by bin, each bin by one XCD
//   k_flatten        pointer-jumping flatten and the deferred edges, only
//                    when k_union deferred one
//   k_route_inputs   one thread per input pair: switch values, routing of P,
//                    values ballot-packed into node-major bit rows
//   k_route_outputs  one thread per output pair: output switch values from
//                    the input rows, routing of the inverse, root words for
//                    the children's first hooks
// Nodes of at most 2^16 entries (LDS levels):
//   k_level_lds      one workgroup per node: the same three steps in LDS
// Both kinds of level, per side:
//   k_pack           threads per (cluster group, packet): gather a packet's
//                    run of switch bits from the rows into whole 64-bit words
//   k_pack_narrow    the packets giving each node at most 4 bits, by ballots
//                    over whole row words
// Below the fusion depth:
//   k_subtree   one workgroup per subtree: every remaining level in LDS, then
//               the factoradic ranks of its leaves.
#pragma once

#include <hip/hip_runtime.h>

#include "format.hpp"
#include "uf.hpp"

namespace benes_hip {

constexpr uint32_t kBlock = 256;          // threads per block of level kernels
// Subtree shapes (entries log2, threads): k_subtree is instantiated for each;
// the host picks one per call (smaller subtrees use less LDS, so more
// workgroups share a CU, at the price of one more level outside LDS).
constexpr uint32_t kSubLog2Max = 14;       // largest subtree handled in LDS

struct NodeDesc {
    uint64_t off;  // absolute offset of the node's entries in the shared arrays
    uint64_t cnt;
    uint64_t row;  // first word of the node's bit rows
    SideAddr a[2];
};

// One permutation of a batch.
struct ItemCtx {
    const void* src;    // its entries (device), width bytes each
    const Plan* plan;   // its format tables (device)
    uint64_t n;
    uint64_t base;      // offset of its entries in the shared arrays
    uint64_t blob;      // offset of its blob in the shared blob, in words
    uint32_t width;     // 4 or 8
    uint32_t df;        // fusion depth: first depth handled by k_subtree
    uint32_t header_bytes;
    uint8_t header[152];  // 20 fixed bytes + both cluster lists (at most 64 each)
};

// One item's share of a launch.
struct Seg {
    uint64_t block0;     // first block of the item in the launch
    uint64_t blocks;     // blocks of the item (a multiple of 8 when XCD-mapped)
    uint64_t nodes;      // nodes of the item at this depth
    uint64_t chunks;     // blocks per node, or packets per group (k_pack)
    uint64_t desc0;      // index of the item's first NodeDesc at this depth
    uint64_t row0;       // first row word of the item at this depth
    uint64_t row_words;  // row words per node
    uint32_t item;
    uint32_t depth;      // the depth a k_describe segment describes
    uint32_t side;       // the side a k_pack / k_pack_narrow segment packs
    // Lane groups of a level launch (lanes > 0): this entry is one node of
    // its item (nodes == 1, its descriptor at desc0), lane `lane` of `lanes`
    // consecutive entries that share block0; lane l runs on XCD l % 8, each
    // XCD taking its lanes lane_g at a time (level_block).
    uint32_t lane;
    uint32_t lanes;
    uint32_t lane_g;
    // Level launches: the item's children at this depth are coloured in LDS
    // (at most 2^16 entries each), so they are stored as 16-bit entries
    // (narrow_at).
    uint32_t narrow_out;
};

// Index of the segment of block b: the last one whose first block is <= b.
__device__ __forceinline__ uint32_t find_seg_index(const Seg* segs, uint32_t nsegs, uint64_t b) {
    uint32_t lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ const Seg& find_seg_at(const Seg* segs, uint32_t nsegs, uint64_t b) {
    return segs[find_seg_index(segs, nsegs, b)];
}

__device__ __forceinline__ const Seg& find_seg(const Seg* segs, uint32_t nsegs) {
    return find_seg_at(segs, nsegs, blockIdx.x);
}

// Entry i of an item's level-0 permutation.
__device__ __forceinline__ uint64_t load_src(const ItemCtx& it, uint64_t i) {
    return it.width == 4 ? static_cast<const uint32_t*>(it.src)[i]
                         : static_cast<const uint64_t*>(it.src)[i];
}

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64)

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }

// Inclusive prefix sum over the lanes of a wave.
__device__ __forceinline__ uint64_t wave_inclusive_sum(uint64_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint64_t o = __shfl_up(v, off);
        if (lane >= off) v += o;
    }
    return v;
}

// ORs the low `len` (<= 64) bits of `v` into blob bits [dst, dst + len).
__device__ __forceinline__ void or_word(uint64_t* blob, uint64_t dst, uint64_t v) {
    if (v == 0) return;
    auto* w = reinterpret_cast<unsigned long long*>(blob) + (dst >> 6);
    const uint32_t sh = dst & 63;
    atomicOr(w, static_cast<unsigned long long>(v << sh));
    if (sh && (v >> (64 - sh))) atomicOr(w + 1, static_cast<unsigned long long>(v >> (64 - sh)));
}

// Bits [first, first + len) (len <= 64) of the bit row starting at word `row`.
__device__ __forceinline__ uint64_t row_range(const uint64_t* rows, uint64_t row, uint64_t first,
                                              uint32_t len) {
    const uint64_t w = row + (first >> 6);
    const uint32_t sh = first & 63;
    uint64_t v = rows[w] >> sh;
    if (sh && sh + len > 64) v |= rows[w + 1] << (64 - sh);
    return len >= 64 ? v : v & ((uint64_t{1} << len) - 1);
}

__device__ __forceinline__ uint32_t row_bit(const uint64_t* rows, uint64_t row, uint64_t pair) {
    return static_cast<uint32_t>((rows[row + (pair >> 6)] >> (pair & 63)) & 1);
}

// ---------------------------------------------------------------------------
// Input validation and the inverse (benes.cpp:152-173).
//
// inv starts as all-ones; every in-range entry stores its index at its value.
// Per-entry kernels stride over their item with the blocks of its segment, so
// a launch stays within 2^32 work-items whatever the batch holds.
// A value out of range raises the item's error flag; a duplicate leaves some
// other slot unwritten, which k_check_inverse reports.  No atomics: on this
// device every global atomic costs several random stores.

template <typename VI>
__global__ __launch_bounds__(kBlock) void k_scatter_inverse(const ItemCtx* __restrict__ items,
                                                            const Seg* __restrict__ segs,
                                                            uint32_t nsegs, VI* __restrict__ inv,
                                                            int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    const ItemCtx it = items[sg.item];
    for (uint64_t i = (blockIdx.x - sg.block0) * kBlock + threadIdx.x; i < it.n;
         i += sg.blocks * kBlock) {
        const uint64_t v = load_src(it, i);
        if (v >= it.n) {
            errs[sg.item] = 1;
            return;
        }
        inv[it.base + v] = static_cast<VI>(i);
    }
}

// A slot of the inverse still all-ones was reached by no entry: some value
// occurred twice.
template <typename VI>
__global__ __launch_bounds__(kBlock) void k_check_inverse(const ItemCtx* __restrict__ items,
                                                          const Seg* __restrict__ segs,
                                                          uint32_t nsegs, const VI* __restrict__ inv,
                                                          int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    const ItemCtx it = items[sg.item];
    for (uint64_t i = (blockIdx.x - sg.block0) * kBlock + threadIdx.x; i < it.n;
         i += sg.blocks * kBlock) {
        if (inv[it.base + i] == static_cast<VI>(~VI{0})) errs[sg.item] = 1;
    }
}

// Fills `count` 32-bit words with `value`.  Scratch that must start at a
// known value (the blob, the inverse, error flags and counters) is cleared
// with these kernels rather than memset nodes, so a call queues only kernels
// and copies on its stream.
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint64_t count,
                                                      uint32_t value) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        p[i] = value;
    }
}

// Writes each item's header bytes (internal.hpp:431-454) at the start of its
// zeroed blob, before any switch bit is OR'ed in.
__global__ void k_headers(const ItemCtx* __restrict__ items, uint32_t count,
                          uint8_t* __restrict__ blob_all) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const ItemCtx& it = items[i];
    uint8_t* dst = blob_all + it.blob * 8;
    for (uint32_t b = 0; b < it.header_bytes; ++b) dst[b] = it.header[b];
}

// ---------------------------------------------------------------------------
// Global levels
//
// Each node keeps its permutation P and its inverse I (both node-local).
// The colouring is a parity union-find over the node's input slots only: the
// two entries I[2q], I[2q+1] reaching output pair q must take different
// colours, which links their input slots with label 1 ^ (I[2q] & 1) ^
// (I[2q+1] & 1).  Output switch values then follow from the input ones:
// t_q = colour(I[2q]).  Children of both arrays are streaming passes:
// left P[p] = P[2p + s_p] >> 1 and left I[q] = I[2q + t_q] >> 1.

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// HBM-level segments start on a multiple of 8 blocks.  When an item has at
// least 8 nodes at a depth, blocks are renumbered so that every block of node
// j lands on XCD j % 8, the node's union-find words, arrays and rows then
// living in that one XCD's L2; with 2 or 4 nodes each node gets 4 or 2 XCDs
// of its own.  Blocks of one XCD then either take its nodes one after
// another (`spread` false: the blocks in flight share one node, whose words
// stay in L2), or visit its nodes in turn (`spread` true: the blocks in flight
// cover up to kSpreadNodes of them, and no more than kSpreadChunks chunks'
// worth of nodes, about 16 MB of words).  The union pass spreads: its
// compare-and-swaps go past L2 to memory, and crowding them on one small
// node's words serialises them there; nodes of 16 MB or more are taken one at
// a time, since more of them in flight only widen the span of random accesses
// (at 2^32 + 12345 entries, depths 4-9: 1000 ms with 64 nodes per XCD, 857
// with one).
constexpr uint32_t kXcds = 8;
constexpr uint32_t kSpreadNodes = 64;
constexpr uint32_t kSpreadChunks = 1u << 14;

__device__ __forceinline__ bool block_to_node_at(const Seg& sg, uint64_t block, uint64_t& node,
                                                 uint64_t& chunk, bool spread) {
    // Every quantity here is below 2^32 (a launch has fewer than 2^32
    // work-items), so the divisions are 32-bit ones, far shorter than 64-bit.
    const uint32_t b = static_cast<uint32_t>(block - sg.block0);
    const uint32_t nodes = static_cast<uint32_t>(sg.nodes);
    const uint32_t chunks = static_cast<uint32_t>(sg.chunks);
    uint32_t n, c;
    if (nodes % kXcds != 0) {
        // 1, 2 or 4 nodes (the segment starts on a multiple of 8): block b
        // serves node b mod nodes, so each node keeps to its own XCDs.
        n = b % nodes;
        c = b / nodes;
    } else {
        const uint32_t xcd = b % kXcds, local = b / kXcds, per_xcd = nodes / kXcds;
        if (spread) {
            // At most kSpreadNodes of the XCD's nodes at a time, so that very
            // large items (thousands of nodes per XCD) keep the live words
            // within reach of the address translation caches.
            // Group sizes are powers of two, as per_xcd is, so that the
            // groups tile the XCD's nodes exactly.
            uint32_t cap = kSpreadNodes;
            while (cap > 1 && cap * chunks > kSpreadChunks) cap >>= 1;
            const uint32_t g = per_xcd < cap ? per_xcd : cap;
            const uint32_t grp = local / (g * chunks), in_grp = local - grp * (g * chunks);
            n = (grp * g + in_grp % g) * kXcds + xcd;
            c = in_grp / g;
        } else {
            n = (local / chunks) * kXcds + xcd;
            c = local % chunks;
        }
    }
    node = n;
    chunk = c;
    return n < nodes && c < chunks;
}

__device__ __forceinline__ bool block_to_node(const Seg& sg, uint64_t& node, uint64_t& chunk,
                                              bool spread) {
    return block_to_node_at(sg, blockIdx.x, node, chunk, spread);
}

// Segments of 8 nodes or more: each XCD visits its nodes `group` at a time,
// their blocks interleaved.
__device__ __forceinline__ bool block_to_node_grouped_at(const Seg& sg, uint64_t block,
                                                         uint64_t& node, uint64_t& chunk,
                                                         uint32_t group) {
    const uint32_t b = static_cast<uint32_t>(block - sg.block0);
    const uint32_t nodes = static_cast<uint32_t>(sg.nodes);
    const uint32_t chunks = static_cast<uint32_t>(sg.chunks);
    const uint32_t xcd = b % kXcds, local = b / kXcds, per_xcd = nodes / kXcds;
    const uint32_t g = per_xcd < group ? per_xcd : group;
    const uint32_t grp = local / (g * chunks), in_grp = local - grp * (g * chunks);
    const uint32_t n = (grp * g + in_grp % g) * kXcds + xcd, c = in_grp / g;
    node = n;
    chunk = c;
    return n < nodes && c < chunks;
}

// Block b of a level launch (k_hook_first to k_route_outputs): its segment
// and (node, chunk).  A lane-grouped segment is one node of an item; the
// lanes of a group share its blocks, lane l on XCD l % 8, each XCD taking
// its lanes lane_g at a time with their blocks interleaved, so that the
// small nodes of many items each keep to one XCD.  A block past the lanes of
// its group, or past its lane's chunks, has no work.  Other segments map by
// block_to_node_grouped_at (a nonzero `group`) or block_to_node_at.
__device__ __forceinline__ bool level_block(const Seg* segs, uint32_t nsegs, uint64_t b,
                                            uint32_t group, bool spread, const Seg*& seg,
                                            uint64_t& node, uint64_t& chunk) {
    const uint32_t i = find_seg_index(segs, nsegs, b);
    const Seg& s = segs[i];
    if (s.lanes) {
        const uint32_t rel = static_cast<uint32_t>(b - s.block0);
        const uint32_t x = rel % kXcds, j = rel / kXcds;
        const uint32_t l = (j % s.lane_g) * kXcds + x;
        if (l >= s.lanes) return false;
        seg = segs + (i - s.lane + l);
        node = 0;
        chunk = j / s.lane_g;
        return chunk < seg->chunks;
    }
    seg = &s;
    if (group) return block_to_node_grouped_at(s, b, node, chunk, group);
    return block_to_node_at(s, b, node, chunk, spread);
}

// Node descriptors of every depth above the fusion depth, in one launch:
// each segment is one item at one depth (`depth`), its descriptors at desc0.
__global__ void k_describe(const ItemCtx* __restrict__ items, const Seg* __restrict__ segs,
                           uint32_t nsegs, NodeDesc* __restrict__ desc,
                           const int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    const uint32_t d = sg.depth;
    const ItemCtx it = items[sg.item];
    const Plan& P = *it.plan;
    const uint64_t j = (blockIdx.x - sg.block0) * blockDim.x + threadIdx.x;
    if (j >= sg.nodes) return;
    NodeDesc nd;
    nd.off = it.base + node_offset(P.n, d, j);
    nd.cnt = node_count(P.n, d, j);
    nd.row = sg.row0 + j * sg.row_words;
    nd.a[0] = node_address(P, 0, d, j);
    nd.a[1] = node_address(P, 1, d, j);
    desc[sg.desc0 + j] = nd;
}

// Input slot of entry x of a node of `count` entries: the odd tail is slot 0.
__device__ __forceinline__ uint64_t in_slot(uint64_t x, uint64_t count) {
    if (count & 1) return x == count - 1 ? 0 : (x >> 1) + 1;
    return x >> 1;
}

// The colouring's union-find runs in two passes over the output pairs.
// Every input slot has exactly one even entry (2p for pair p, the even tail
// entry count - 1 for slot 0 of an odd node), and each entry reaches one
// output pair, whose edge joins the entry's slot to the other entry's slot.
//   k_hook_first   the words start as root words (set by the previous
//                  depth's k_route_outputs, or a fill at the root); each slot
//                  whose even entry's edge leads to a smaller slot stores
//                  that relation.  One plain store for about half the slots,
//                  no atomics.
//   k_union        the edges reached by their larger slot's odd entry (the
//                  others were settled by the first pass, which hooked the
//                  larger slot through that very edge).  A larger slot the
//                  first pass left a root is hooked straight under the smaller
//                  one with a compare-and-swap; otherwise (slots larger than
//                  both neighbours, about a third of the edges of a random
//                  permutation) a Rem walk joins the two trees.
// Every stored word is a true parity relation to a smaller slot, so the
// forest invariants of the union-find hold across the passes; an edge whose
// larger slot already holds the edge's own relation is skipped.
struct PairEdge {
    uint64_t lo, hi;    // the two input slots, lo < hi (equal: a 2-cycle)
    uint32_t label;     // value(lo) ^ value(hi)
    bool hi_odd;        // the larger slot reaches this pair through its odd entry
};

// The level arrays are streamed once per kernel and read again only by the
// next one; their loads and stores are non-temporal so that they do not push
// the union-find words out of L2.
template <typename VI>
__device__ __forceinline__ PairEdge pair_edge(const VI* inv, uint64_t off, uint64_t cnt,
                                              uint64_t q) {
    const uint64_t a = static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q));
    const uint64_t b = static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q + 1));
    const uint64_t sa = in_slot(a, cnt), sb = in_slot(b, cnt);
    PairEdge e;
    e.lo = sa < sb ? sa : sb;
    e.hi = sa < sb ? sb : sa;
    e.label = static_cast<uint32_t>(1 ^ ((a ^ b) & 1));
    e.hi_odd = ((sa < sb ? b : a) & 1) != 0;
    return e;
}

// Only slots s with (s >> part_shift) == part are written (a node whose
// words would just overflow an XCD's L2 is hooked in two passes, each
// writing half of them), so every slot is written by at most one pass.
// With kCount (depths whose nodes spread over several XCDs), every eighth
// block counts its slots hooked under the slot just below them (`adjacent`),
// up to just past `chain_threshold` (a wave adds its count only while the
// total is at most the threshold, so the one counter sees few atomics): many
// of them mean chains of neighbouring slots (rotations, reversals), which
// k_route_inputs then walks with loads and stores coherent across XCDs.
template <typename VI, typename W, bool kCount>
__global__ __launch_bounds__(kBlock) void k_hook_first(const Seg* __restrict__ segs,
                                                       uint32_t nsegs,
                                                       const VI* __restrict__ inv,
                                                       W* __restrict__ uf,
                                                       const NodeDesc* __restrict__ desc,
                                                       unsigned long long* __restrict__ adjacent,
                                                       unsigned long long chain_threshold,
                                                       uint32_t part_shift, uint32_t part,
                                                       const int* __restrict__ errs) {
    const Seg* sgp;
    uint64_t node, chunk;
    if (!level_block(segs, nsegs, blockIdx.x, 0u, false, sgp, node, chunk)) return;
    const Seg& sg = *sgp;
    if (errs[sg.item]) return;
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt, h = cnt >> 1;
    const uint64_t q = chunk * kBlock + threadIdx.x;
    W* u = uf + off;
    bool next_to = false;
    if (q < h) {
        const uint64_t a = static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q));
        const uint64_t b = static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q + 1));
        const uint64_t sa = in_slot(a, cnt), sb = in_slot(b, cnt);
        const W label = static_cast<W>(1 ^ ((a ^ b) & 1));
        if (!(a & 1) && sb < sa && (sa >> part_shift) == part) {
            u[sa] = (static_cast<W>(sb) << 1) | label;
        }
        if (!(b & 1) && sa < sb && (sb >> part_shift) == part) {
            u[sb] = (static_cast<W>(sa) << 1) | label;
        }
        if (kCount && (blockIdx.x & 7) == 0) {
            next_to = (!(a & 1) && sb + 1 == sa) || (!(b & 1) && sa + 1 == sb);
        }
    }
    if (!kCount || (blockIdx.x & 7) != 0) return;
    const uint64_t ballot = __ballot(next_to);
    if (ballot && lane_id() == 0 &&
        __hip_atomic_load(adjacent, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= chain_threshold) {
        atomicAdd(adjacent, static_cast<unsigned long long>(__popcll(ballot)));
    }
}

// First hooks of very large nodes, in two passes.  Stored straight into a
// node of many megabytes of words within a union-find array of gigabytes,
// each random first-hook store misses the address translation caches and
// dirties a line that is written back before its neighbours arrive.
// Instead, k_hook_bin stages each store as (word index within its bin, word)
// in the bin of 2^bin_log2 words it falls in: a block sorts its stores by bin
// in LDS, reserves a run in each bin with one atomic, and writes the runs out
// contiguously.  k_hook_scatter then writes the bins, each by the blocks of
// one XCD, so that a bin's words stay within one XCD's L2 and translation
// caches while they are written.  Only the slots that get a parent are
// written (the others start as root words), each once, so bin b
// receives at most 2^bin_log2 stores and its run fits in positions
// [b << bin_log2, (b + 1) << bin_log2) of the staging arrays, which are the
// level arrays this depth writes only after its union-find is done.
// `bin_fill[b]` counts bin b's stores (zero before the depth starts).
constexpr uint32_t kHookBinMinLog2 = 18;  // 1 MB of 32-bit words
constexpr uint32_t kHookBinThreads = 1024;
constexpr uint32_t kHookBinPairs = 4;  // output pairs per thread
constexpr uint32_t kHookBinStores = 2 * kHookBinThreads * kHookBinPairs;
// Bins a node's slot words may touch, a multiple of 64 (one wave scans
// them): a node of s slots touches at most ceil(s / 2^bin_log2) + 1.
constexpr uint32_t kHookBinSpan = 320;
constexpr uint32_t kHookScatterPerThread = 8;
constexpr uint32_t kHookScatterEntries = kBlock * kHookScatterPerThread;  // per block

template <typename VI, bool kCount>
__global__ __launch_bounds__(kHookBinThreads) void k_hook_bin(
    const Seg* __restrict__ segs, uint32_t nsegs, const VI* __restrict__ inv,
    const NodeDesc* __restrict__ desc, uint32_t bin_log2, uint32_t* __restrict__ bin_word,
    uint32_t* __restrict__ bin_value, uint32_t* __restrict__ bin_fill,
    unsigned long long* __restrict__ adjacent, unsigned long long chain_threshold,
    const int* __restrict__ errs) {
    __shared__ uint32_t fill[kHookBinSpan];   // this block's stores per bin
    __shared__ uint32_t start[kHookBinSpan];  // their run's first staging position
    __shared__ uint32_t base[kHookBinSpan];   // their run's first position in the bin
    __shared__ uint32_t stage_rel[kHookBinStores];
    __shared__ uint32_t stage_val[kHookBinStores];
    __shared__ uint32_t total;
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    uint64_t node, chunk;
    if (!block_to_node(sg, node, chunk, false)) return;
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt, h = cnt >> 1;
    // Staged word indices are relative to the node's first bin.
    const uint64_t bin0 = off >> bin_log2;
    const uint64_t rel0 = off - (bin0 << bin_log2);
    for (uint32_t i = threadIdx.x; i < kHookBinSpan; i += kHookBinThreads) fill[i] = 0;
    __syncthreads();
    // Store j of this thread: slot 2k (the pair's first entry) or 2k + 1.
    uint32_t rel[2 * kHookBinPairs], val[2 * kHookBinPairs], rank[2 * kHookBinPairs];
    bool has[2 * kHookBinPairs];
    uint32_t adjacent_hooks = 0;
#pragma unroll
    for (uint32_t k = 0; k < kHookBinPairs; ++k) {
        const uint64_t q = (chunk * kHookBinPairs + k) * kHookBinThreads + threadIdx.x;
        has[2 * k] = has[2 * k + 1] = false;
        bool next_to = false;
        if (q < h) {
            const uint64_t a = static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q));
            const uint64_t b =
                static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q + 1));
            const uint64_t sa = in_slot(a, cnt), sb = in_slot(b, cnt);
            const uint32_t label = static_cast<uint32_t>(1 ^ ((a ^ b) & 1));
            // Only the words that get a parent (the others start as root
            // words, as for k_hook_first).
            if (!(a & 1) && sb < sa) {
                has[2 * k] = true;
                rel[2 * k] = static_cast<uint32_t>(rel0 + sa);
                val[2 * k] = (static_cast<uint32_t>(sb) << 1) | label;
            }
            if (!(b & 1) && sa < sb) {
                has[2 * k + 1] = true;
                rel[2 * k + 1] = static_cast<uint32_t>(rel0 + sb);
                val[2 * k + 1] = (static_cast<uint32_t>(sa) << 1) | label;
            }
            next_to = (!(a & 1) && sb + 1 == sa) || (!(b & 1) && sa + 1 == sb);
        }
        if (kCount) adjacent_hooks += static_cast<uint32_t>(__popcll(__ballot(next_to)));
    }
#pragma unroll
    for (uint32_t j = 0; j < 2 * kHookBinPairs; ++j) {
        if (has[j]) rank[j] = atomicAdd(&fill[rel[j] >> bin_log2], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // Staging runs in bin order; one reservation per bin this block fills.
        uint32_t carry = 0;
        for (uint32_t c = 0; c < kHookBinSpan; c += 64) {
            const uint32_t i = c + lane_id();
            const uint32_t f = fill[i];
            const uint32_t inc = static_cast<uint32_t>(wave_inclusive_sum(f));
            start[i] = carry + inc - f;
            if (f) base[i] = atomicAdd(bin_fill + bin0 + i, f);
            carry += __shfl(inc, 63);
        }
        if (lane_id() == 0) total = carry;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < 2 * kHookBinPairs; ++j) {
        if (has[j]) {
            const uint32_t at = start[rel[j] >> bin_log2] + rank[j];
            stage_rel[at] = rel[j];
            stage_val[at] = val[j];
        }
    }
    __syncthreads();
    const uint32_t mask = (uint32_t{1} << bin_log2) - 1;
    for (uint32_t i = threadIdx.x; i < total; i += kHookBinThreads) {
        const uint32_t r = stage_rel[i], lb = r >> bin_log2;
        const uint64_t at = ((bin0 + lb) << bin_log2) + base[lb] + (i - start[lb]);
        __builtin_nontemporal_store(r & mask, bin_word + at);
        __builtin_nontemporal_store(stage_val[i], bin_value + at);
    }
    if (!kCount || (blockIdx.x & 7) != 0) return;
    if (adjacent_hooks && lane_id() == 0 &&
        __hip_atomic_load(adjacent, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= chain_threshold) {
        atomicAdd(adjacent, static_cast<unsigned long long>(adjacent_hooks));
    }
}

// Bin b's stores, by the blocks of one XCD (block i runs on XCD i % 8).
// Round r takes bins 8r .. 8r + 7, XCD x bin 8r + (x + r) % 8: the XCDs take
// their bins one after another, and rotating the assignment keeps them
// evenly loaded where only some bins of each node hold slots.
__global__ __launch_bounds__(kBlock) void k_hook_scatter(uint32_t* __restrict__ uf,
                                                         const uint32_t* __restrict__ bin_word,
                                                         const uint32_t* __restrict__ bin_value,
                                                         const uint32_t* __restrict__ bin_fill,
                                                         uint64_t bins, uint32_t bin_log2) {
    // Blocks per bin: 2^bin_log2 / kHookScatterEntries.
    const uint32_t chunks = (uint32_t{1} << bin_log2) / kHookScatterEntries;
    const uint32_t xcd = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
    const uint32_t round = local / chunks;
    const uint64_t bin = uint64_t{round} * kXcds + (xcd + round) % kXcds;
    const uint64_t chunk = local % chunks;
    if (bin >= bins) return;
    const uint32_t n = bin_fill[bin];
    const uint64_t first = bin << bin_log2;
#pragma unroll
    for (uint32_t k = 0; k < kHookScatterPerThread; ++k) {
        const uint64_t e = (chunk * kHookScatterPerThread + k) * kBlock + threadIdx.x;
        if (e < n) {
            uf[first + __builtin_nontemporal_load(bin_word + first + e)] =
                __builtin_nontemporal_load(bin_value + first + e);
        }
    }
}

// The level-0 inverse of large items, binned the same way.  Stored straight
// (k_scatter_inverse), each random 4- or 8-byte store of inv[P[i]] = i
// dirties a whole line that is written back long before its neighbours
// arrive.  k_inverse_bin instead stages every store, packed as (value <<
// bin_log2 | position within its bin of 2^bin_log2 entries; values below
// 2^33 and bins of at most 2^31 entries fit), in the bin it falls in: a
// block sorts its entries' stores by bin in LDS, reserves a run per bin with
// one atomic and writes the runs out contiguously.  k_inverse_scatter then
// stores the bins, each by the blocks of one XCD, so that a bin's inverse
// words are written while they sit in that XCD's L2.  A permutation puts
// exactly one store at each position, so bin b's run fits in positions
// [b << bin_log2, (b + 1) << bin_log2) of the staging array (the union-find
// words, unused until depth 0's first hooks); a store that would overflow its
// bin can only come from a repeated value, but bins are shared by
// neighbouring items of a batch, so it cannot tell whose: it raises
// `overflow` instead, and k_inverse_redo then redoes the whole inverse
// directly (returning at once when `overflow` is clear), so that
// k_check_inverse blames exactly the items with repeated values.
constexpr uint32_t kInvBinThreads = 1024;
constexpr uint32_t kInvBinPerThread = 8;
constexpr uint32_t kInvBinEntries = kInvBinThreads * kInvBinPerThread;
constexpr uint32_t kInvBinSpan = 512;  // bins an item may touch, a multiple of 64

template <typename VI>
__global__ __launch_bounds__(kInvBinThreads) void k_inverse_bin(
    const ItemCtx* __restrict__ items, const Seg* __restrict__ segs, uint32_t nsegs,
    uint32_t bin_log2, uint64_t total_n, uint64_t* __restrict__ staged,
    uint32_t* __restrict__ bin_fill, int* __restrict__ overflow, int* __restrict__ errs) {
    __shared__ uint32_t fill[kInvBinSpan];
    __shared__ uint32_t start[kInvBinSpan];
    __shared__ uint32_t base[kInvBinSpan];
    __shared__ uint64_t stage[kInvBinEntries];
    __shared__ uint32_t total;
    const Seg& sg = find_seg(segs, nsegs);
    const ItemCtx it = items[sg.item];
    const uint64_t bin0 = it.base >> bin_log2;
    for (uint32_t i = threadIdx.x; i < kInvBinSpan; i += kInvBinThreads) fill[i] = 0;
    __syncthreads();
    const uint64_t i0 = (blockIdx.x - sg.block0) * uint64_t{kInvBinEntries};
    uint64_t word[kInvBinPerThread];
    uint32_t rank[kInvBinPerThread], lbin[kInvBinPerThread];
    bool has[kInvBinPerThread];
    bool bad = false;
#pragma unroll
    for (uint32_t k = 0; k < kInvBinPerThread; ++k) {
        const uint64_t i = i0 + k * kInvBinThreads + threadIdx.x;
        has[k] = false;
        if (i >= it.n) continue;
        const uint64_t v = load_src(it, i);
        if (v >= it.n) {
            bad = true;
            continue;
        }
        const uint64_t pos = it.base + v;
        has[k] = true;
        lbin[k] = static_cast<uint32_t>((pos >> bin_log2) - bin0);
        word[k] = (static_cast<uint64_t>(i) << bin_log2) | (pos & ((uint64_t{1} << bin_log2) - 1));
        rank[k] = atomicAdd(&fill[lbin[k]], 1u);
    }
    if (bad) errs[sg.item] = 1;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t carry = 0;
        for (uint32_t c = 0; c < kInvBinSpan; c += 64) {
            const uint32_t b = c + lane_id();
            const uint32_t f = fill[b];
            const uint32_t inc = static_cast<uint32_t>(wave_inclusive_sum(f));
            start[b] = carry + inc - f;
            if (f) base[b] = atomicAdd(bin_fill + bin0 + b, f);
            carry += __shfl(inc, 63);
        }
        if (lane_id() == 0) total = carry;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kInvBinPerThread; ++k) {
        if (has[k]) stage[start[lbin[k]] + rank[k]] = word[k];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < total; e += kInvBinThreads) {
        const uint64_t w = stage[e];
        // The bin of staged entry e: the last bin whose run starts at or
        // before e (never an empty one: its successor would start there too,
        // and the last bin's run ends at `total`, past e).
        uint32_t lo = 0, hi = kInvBinSpan - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (start[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        const uint64_t bin = bin0 + lo;
        const uint64_t at = base[lo] + (e - start[lo]);
        const uint64_t first = bin << bin_log2;
        const uint64_t cap = min(uint64_t{1} << bin_log2, total_n - first);
        if (at >= cap) {
            *overflow = 1;  // more stores than positions: a repeated value somewhere
            continue;
        }
        __builtin_nontemporal_store(w, staged + first + at);
    }
}

// Bin b's inverse stores, by the blocks of one XCD (assignment as in
// k_hook_scatter).
template <typename VI>
__global__ __launch_bounds__(kBlock) void k_inverse_scatter(VI* __restrict__ inv,
                                                            const uint64_t* __restrict__ staged,
                                                            const uint32_t* __restrict__ bin_fill,
                                                            uint64_t bins, uint32_t bin_log2,
                                                            uint64_t total_n) {
    const uint32_t chunks = (uint32_t{1} << bin_log2) / kHookScatterEntries;
    const uint32_t xcd = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
    const uint32_t round = local / chunks;
    const uint64_t bin = uint64_t{round} * kXcds + (xcd + round) % kXcds;
    const uint64_t chunk = local % chunks;
    if (bin >= bins) return;
    const uint64_t first = bin << bin_log2;
    const uint64_t mask = (uint64_t{1} << bin_log2) - 1;
    // Runs past the bin's positions were dropped (and flagged) when staged.
    const uint64_t n = min(static_cast<uint64_t>(bin_fill[bin]),
                           min(uint64_t{1} << bin_log2, total_n - first));
#pragma unroll
    for (uint32_t k = 0; k < kHookScatterPerThread; ++k) {
        const uint64_t e = (chunk * kHookScatterPerThread + k) * kBlock + threadIdx.x;
        if (e < n) {
            const uint64_t w = __builtin_nontemporal_load(staged + first + e);
            inv[first + (w & mask)] = static_cast<VI>(w >> bin_log2);
        }
    }
}

// The direct inverse after an overflow of the binned one (nothing when
// *overflow is clear): every item's inverse reset to all-ones, then stored
// directly, as k_scatter_inverse does.  A small grid strides over the
// per-entry launch's virtual blocks, so a clear flag costs one short launch.
template <typename VI>
__global__ __launch_bounds__(kBlock) void k_inverse_redo(const ItemCtx* __restrict__ items,
                                                         const Seg* __restrict__ segs,
                                                         uint32_t nsegs, uint64_t vblocks,
                                                         VI* __restrict__ inv, int pass,
                                                         const int* __restrict__ overflow,
                                                         int* __restrict__ errs) {
    if (*overflow == 0) return;
    for (uint64_t b = blockIdx.x; b < vblocks; b += gridDim.x) {
        const Seg& sg = find_seg_at(segs, nsegs, b);
        const ItemCtx it = items[sg.item];
        for (uint64_t i = (b - sg.block0) * kBlock + threadIdx.x; i < it.n; i += sg.blocks * kBlock) {
            if (pass == 0) {
                inv[it.base + i] = static_cast<VI>(~VI{0});
                continue;
            }
            const uint64_t v = load_src(it, i);
            if (v >= it.n) {
                errs[sg.item] = 1;
                break;
            }
            inv[it.base + v] = static_cast<VI>(i);
        }
    }
}

// Rem walks longer than this many steps are handed to k_flatten.  A walk
// follows one dependent load per step; on random
// permutations the longest walk of a level is a few dozen steps, but the
// first-hook pass chains every slot under its neighbour for some structured
// ones (a reverse of odd length, a rotation by one of even length), and the
// edge closing such a cycle would walk half the node one load at a time.
constexpr uint32_t kWalkCap = 1024;
// Deferred edges per level (the list's capacity is a launch argument, at
// most this); beyond it an edge walks its whole path itself.
constexpr uint64_t kDeferCap = uint64_t{1} << 18;
struct DeferredEdge {
    uint64_t off;       // the node's first union-find word
    uint64_t lo_label;  // smaller slot, label in bit 63
    uint64_t hi;        // larger slot
};

template <typename VI, typename W>
__global__ __launch_bounds__(kBlock) void k_union(const Seg* __restrict__ segs, uint32_t nsegs,
                                                  const VI* __restrict__ inv, W* __restrict__ uf,
                                                  const NodeDesc* __restrict__ desc,
                                                  DeferredEdge* __restrict__ deferred,
                                                  unsigned long long* __restrict__ deferred_count,
                                                  uint64_t defer_cap,
                                                  const int* __restrict__ errs) {
    const Seg* sgp;
    uint64_t node, chunk;
    if (!level_block(segs, nsegs, blockIdx.x, 0u, true, sgp, node, chunk)) return;
    const Seg& sg = *sgp;
    if (errs[sg.item]) return;
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt;
    const uint64_t q = chunk * kBlock + threadIdx.x;
    if (q >= (cnt >> 1)) return;
    const PairEdge e = pair_edge(inv, off, cnt, q);
    if (e.lo == e.hi || !e.hi_odd) return;
    W* u = uf + off;
    const W lo = static_cast<W>(e.lo), hi = static_cast<W>(e.hi);
    W w = DeviceForest::load(u + hi);
    if (w == UfWord<W>::root) {
        w = DeviceForest::hook(u + hi, (lo << 1) | e.label);
        if (w == UfWord<W>::root) return;
    }
    if (w == ((lo << 1) | e.label)) return;
    if (uf_union_rem<DeviceForest>(u, w >> 1, static_cast<uint32_t>(w & 1), lo, 0u, e.label,
                                   kWalkCap)) {
        return;
    }
    // A long path: flattened for the edge by k_flatten, unless the list is full.
    const unsigned long long i = atomicAdd(deferred_count, 1ull);
    if (i < defer_cap) {
        deferred[i] = DeferredEdge{nd.off, e.lo | (static_cast<uint64_t>(e.label) << 63), e.hi};
        return;
    }
    uf_union_rem<DeviceForest>(u, w >> 1, static_cast<uint32_t>(w & 1), lo, 0u, e.label,
                               ~uint64_t{0});
}

// The union pass without compare-and-swaps, for depths whose nodes each keep
// to one XCD and whose words fit its L2.  The blocks of an XCD take its nodes
// one after another, so the walks' loads and the hooks' stores stay in L2,
// where a compare-and-swap would go past it to memory (global atomics on this
// device all execute there, at about a third of the rate of random loads).
// Hooks are plain stores: two walks that find the same root at once may both
// store to it, and only the later relation stays.  Each thread therefore
// records the hook it made, (hooked slot << 32) | word, at
// rec[(off + 2q) >> 1] (distinct for every pair of every node), or
// kNoHook; k_union_verify then finds the relations that were overwritten
// and k_union_rejoin joins them again.  Every stored word is a true relation to a smaller slot, so
// the walks of this pass stay correct whatever mix of old and new words they
// read; a walk that relied on an overwritten relation is covered once that
// relation is joined again, since the relations it followed are facts.
constexpr uint64_t kNoHook = ~uint64_t{0};

template <typename VI>
__global__ __launch_bounds__(kBlock) void k_union_plain(const Seg* __restrict__ segs,
                                                        uint32_t nsegs,
                                                        const VI* __restrict__ inv,
                                                        uint32_t* __restrict__ uf,
                                                        const NodeDesc* __restrict__ desc,
                                                        uint64_t* __restrict__ rec,
                                                        DeferredEdge* __restrict__ deferred,
                                                        unsigned long long* __restrict__ deferred_count,
                                                        uint64_t defer_cap, uint32_t group,
                                                        const int* __restrict__ errs) {
    using W = uint32_t;
    const Seg* sgp;
    uint64_t node, chunk;
    if (!level_block(segs, nsegs, blockIdx.x, group, false, sgp, node, chunk)) return;
    const Seg& sg = *sgp;
    if (errs[sg.item]) return;
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt;
    const uint64_t q = chunk * kBlock + threadIdx.x;
    if (q >= (cnt >> 1)) return;
    const PairEdge e = pair_edge(inv, off, cnt, q);
    uint64_t r = kNoHook;
    if (e.lo != e.hi && e.hi_odd) {
        W* u = uf + off;
        const W lo = static_cast<W>(e.lo), hi = static_cast<W>(e.hi);
        const W rel = (lo << 1) | e.label;
        const W w = DeviceForest::load(u + hi);
        if (w == UfWord<W>::root) {
            DeviceForest::store(u + hi, rel);
            r = (static_cast<uint64_t>(hi) << 32) | rel;
        } else if (w != rel) {
            W hooked = 0, word = 0;
            int got = uf_union_rem_plain<DeviceForest>(u, w >> 1, static_cast<uint32_t>(w & 1), lo,
                                                       0u, e.label, kWalkCap, hooked, word);
            if (got < 0) {
                // A long path: joined after k_flatten, unless the list is
                // full; then walked to the end here (each step lowers the
                // larger of the two vertices, so fewer than 2^32 steps).
                const unsigned long long i = atomicAdd(deferred_count, 1ull);
                if (i < defer_cap) {
                    deferred[i] = DeferredEdge{off, e.lo | (static_cast<uint64_t>(e.label) << 63), e.hi};
                } else {
                    got = uf_union_rem_plain<DeviceForest>(u, w >> 1, static_cast<uint32_t>(w & 1),
                                                           lo, 0u, e.label, 0xFFFFFFFFu, hooked,
                                                           word);
                }
            }
            if (got > 0) r = (static_cast<uint64_t>(hooked) << 32) | word;
        }
    }
    __builtin_nontemporal_store(r, rec + ((off + 2 * q) >> 1));
}

// After k_union_plain: each recorded hook is checked against its word.  A
// relation that a later hook overwrote must be joined again; the check only
// collects them, and k_union_rejoin joins them, so that the check's waves do
// not wait on the walks (about one relation in 70 is lost, nearly one per
// wave).  The blocks take the nodes in the order k_union_plain did, so each
// node's words are read while in L2.  One wave per block of kBlock pairs:
// each lane issues the loads of its four records, then of the four words
// they name, before comparing any, so that the wave waits for memory twice
// instead of eight times.  The lost records of a block are then written,
// packed, over the block's own records (all of which the wave has read by
// then), and their count to lost_counts[block] (0 for a block with no work).
constexpr uint32_t kVerifyThreads = 64;
constexpr uint32_t kRejoinCounters = 1024;
constexpr uint32_t kVerifyPairs = kBlock / kVerifyThreads;

template <typename VI>
__global__ __launch_bounds__(kVerifyThreads) void k_union_verify(
    const Seg* __restrict__ segs, uint32_t nsegs, const uint32_t* __restrict__ uf,
    const NodeDesc* __restrict__ desc, uint64_t* __restrict__ rec,
    uint32_t* __restrict__ lost_counts, uint32_t group, const int* __restrict__ errs) {
    using W = uint32_t;
    const Seg* sgp;
    uint64_t node, chunk;
    if (!level_block(segs, nsegs, blockIdx.x, group, false, sgp, node, chunk) ||
        errs[sgp->item]) {
        if (threadIdx.x == 0) lost_counts[blockIdx.x] = 0;
        return;
    }
    const NodeDesc& nd = desc[sgp->desc0 + node];
    const uint64_t off = nd.off, h = nd.cnt >> 1;
    const W* const u = uf + off;
    // Record of pair q: rec[(off + 2q) >> 1]; the block's pairs are
    // consecutive, from rec[base].
    const uint64_t base = (off + 2 * chunk * kBlock) >> 1;
    uint64_t r[kVerifyPairs];
#pragma unroll
    for (uint32_t k = 0; k < kVerifyPairs; ++k) {
        const uint64_t q = chunk * kBlock + k * kVerifyThreads + threadIdx.x;
        r[k] = q < h ? __builtin_nontemporal_load(rec + base + k * kVerifyThreads + threadIdx.x)
                     : kNoHook;
    }
    W w[kVerifyPairs];
#pragma unroll
    for (uint32_t k = 0; k < kVerifyPairs; ++k) {
        w[k] = r[k] != kNoHook ? DeviceForest::load(u + static_cast<W>(r[k] >> 32)) : 0u;
    }
    uint32_t lost = 0;
#pragma unroll
    for (uint32_t k = 0; k < kVerifyPairs; ++k) {
        lost += r[k] != kNoHook && w[k] != static_cast<W>(r[k]);
    }
    const uint32_t incl = static_cast<uint32_t>(wave_inclusive_sum(lost));
    uint64_t at = base + incl - lost;
#pragma unroll
    for (uint32_t k = 0; k < kVerifyPairs; ++k) {
        if (r[k] != kNoHook && w[k] != static_cast<W>(r[k])) rec[at++] = r[k];
    }
    if (threadIdx.x == 63) lost_counts[blockIdx.x] = incl;
}

// Joins again the relations k_union_verify collected: x -> (y, p) is joined
// with compare-and-swap hooks (the slot has some smaller parent by now, so the
// relation is made between the two trees).  Each block takes the counts of
// kRejoinSpan check blocks of one XCD (check block b ran on XCD b % 8) and
// spreads their relations over its threads, one each, so that the walks run
// side by side in full waves.
constexpr uint32_t kRejoinSpan = 64;

__global__ __launch_bounds__(kBlock) void k_union_rejoin(
    const Seg* __restrict__ segs, uint32_t nsegs, uint32_t* __restrict__ uf,
    const NodeDesc* __restrict__ desc, const uint64_t* __restrict__ rec,
    const uint32_t* __restrict__ lost_counts, uint64_t vblocks, uint32_t group,
    DeferredEdge* __restrict__ deferred, unsigned long long* __restrict__ deferred_count,
    uint64_t defer_cap, unsigned long long* __restrict__ rejoined) {
    using W = uint32_t;
    __shared__ uint32_t start[kRejoinSpan + 1];
    // Check blocks x + kXcds * (kRejoinSpan * g + i), i < kRejoinSpan, for
    // this block's XCD x and group g.
    const uint64_t x = blockIdx.x % kXcds, g = blockIdx.x / kXcds;
    auto check_block = [&](uint32_t i) { return x + kXcds * (kRejoinSpan * g + i); };
    if (threadIdx.x < 64) {
        const uint64_t b = check_block(threadIdx.x);
        const uint32_t c = b < vblocks ? lost_counts[b] : 0u;
        start[threadIdx.x + 1] = static_cast<uint32_t>(wave_inclusive_sum(c));
        if (threadIdx.x == 0) start[0] = 0;
    }
    __syncthreads();
    const uint32_t total = start[kRejoinSpan];
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
        // The check block whose relations hold t: start[i] <= t < start[i + 1].
        uint32_t lo = 0, hi = kRejoinSpan;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (start[mid] <= t) lo = mid;
            else hi = mid;
        }
        const Seg* sgp;
        uint64_t node, chunk;
        // A check block that counted relations maps to a node.
        if (!level_block(segs, nsegs, check_block(lo), group, false, sgp, node, chunk)) continue;
        const NodeDesc& nd = desc[sgp->desc0 + node];
        const uint64_t off = nd.off;
        const uint64_t r = rec[((off + 2 * chunk * kBlock) >> 1) + (t - start[lo])];
        W* const u = uf + off;
        const W xs = static_cast<W>(r >> 32), word = static_cast<W>(r);
        // value(x) ^ value(y) == p for the relation x -> (y, p).
        if (uf_union_rem<DeviceForest>(u, xs, 0u, word >> 1, 0u, word & 1u, kWalkCap)) continue;
        const unsigned long long i = atomicAdd(deferred_count, 1ull);
        if (i < defer_cap) {
            deferred[i] = DeferredEdge{off, (word >> 1) | (static_cast<uint64_t>(word & 1u) << 63), xs};
        } else {
            uf_union_rem<DeviceForest>(u, xs, 0u, word >> 1, 0u, word & 1u, ~uint64_t{0});
        }
    }
    // Counted when profiling, into one of kRejoinCounters counters per depth.
    if (rejoined && threadIdx.x == 0 && total) {
        atomicAdd(rejoined + (blockIdx.x & (kRejoinCounters - 1)),
                  static_cast<unsigned long long>(total));
    }
}

// Runs only when k_union deferred an edge at this depth (`counts[0]`):
// every slot finds its root (halving, then pointing the slot at the root), as
// k_route_inputs does.  Run by all slots at once this is pointer jumping, so
// chains collapse in a logarithmic number of rounds instead of being walked
// end to end, provided each XCD sees the others' shortcuts: F is
// CoherentForest at depths whose nodes spread over several XCDs.  The last
// block to finish (`counts[1]` counts finished blocks)
// then joins the deferred edges, whose walks are now short.  The grid is
// small and strides over the level's blocks, so one launch that reads a zero
// count is all a level pays when nothing was deferred.
template <typename W, typename F>
__global__ __launch_bounds__(kBlock) void k_flatten(const Seg* __restrict__ segs, uint32_t nsegs,
                                                    uint64_t vblocks, W* __restrict__ uf,
                                                    const NodeDesc* __restrict__ desc,
                                                    const DeferredEdge* __restrict__ deferred,
                                                    unsigned long long* __restrict__ counts,
                                                    uint64_t defer_cap,
                                                    const int* __restrict__ errs) {
    const unsigned long long total = counts[0];
    if (total == 0) return;
    for (uint64_t b = blockIdx.x; b < vblocks; b += gridDim.x) {
        const Seg* sgp;
        uint64_t node, chunk;
        if (!level_block(segs, nsegs, b, 0u, false, sgp, node, chunk)) continue;
        const Seg& sg = *sgp;
        if (errs[sg.item]) continue;
        const NodeDesc& nd = desc[sg.desc0 + node];
        const uint64_t cnt = nd.cnt;
        const uint64_t p = chunk * kBlock + threadIdx.x;
        if (p >= (cnt >> 1)) continue;
        W* const u = uf + nd.off;
        const W x0 = static_cast<W>(p + (cnt & 1));
        uint32_t s;
        const W r = uf_find<F>(u, x0, s);
        if (r != x0) F::store(u + x0, static_cast<W>((r << 1) | static_cast<W>(s)));
    }
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&counts[1], 1ull) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const uint64_t n = total < defer_cap ? total : defer_cap;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const DeferredEdge e = deferred[i];
        const W lo = static_cast<W>(e.lo_label & ~(uint64_t{1} << 63));
        const uint32_t label = static_cast<uint32_t>(e.lo_label >> 63);
        uf_union_rem<F>(uf + e.off, static_cast<W>(e.hi), 0u, lo, 0u, label, ~uint64_t{0});
    }
}

// Entry i (node-local index `off + i` in shared terms) of the level's
// permutation: the item's own source at depth 0, the shared level array below.
__device__ __forceinline__ uint64_t load_perm(const ItemCtx& it, const uint32_t* lev, bool level0,
                                              uint64_t abs) {
    return level0 ? load_src(it, abs - it.base) : static_cast<uint64_t>(__builtin_nontemporal_load(lev + abs));
}

// A node of at most 2^16 entries (coloured by k_level_lds or k_subtree)
// keeps its level arrays as 16-bit entries: entry i of the node at offset
// `off` is 16-bit element 2 * off + i of the array, inside the node's own
// range of 4-byte entries, so that nodes of both widths (items of different
// sizes) share one array.  It halves what those levels stream.
__device__ __forceinline__ uint16_t* narrow_at(uint32_t* lev, uint64_t off) {
    return reinterpret_cast<uint16_t*>(lev) + 2 * off;
}
__device__ __forceinline__ const uint16_t* narrow_at(const uint32_t* lev, uint64_t off) {
    return reinterpret_cast<const uint16_t*>(lev) + 2 * off;
}

// Stores the two children's entries of pair p (and the odd tail entry,
// child 0's last) of the node at `off` with `lc` entries in child 0.
__device__ __forceinline__ void store_children(uint32_t* dst, bool narrow, uint64_t off,
                                               uint64_t lc, uint64_t p, uint32_t c0, uint32_t c1) {
    if (narrow) {
        __builtin_nontemporal_store(static_cast<uint16_t>(c0), narrow_at(dst, off) + p);
        __builtin_nontemporal_store(static_cast<uint16_t>(c1), narrow_at(dst, off + lc) + p);
    } else {
        __builtin_nontemporal_store(c0, dst + off + p);
        __builtin_nontemporal_store(c1, dst + off + lc + p);
    }
}
__device__ __forceinline__ void store_tail(uint32_t* dst, bool narrow, uint64_t off, uint64_t h,
                                           uint32_t v) {
    if (narrow) narrow_at(dst, off)[h] = static_cast<uint16_t>(v);
    else dst[off + h] = v;
}

// One thread per input pair: canonical switch value s_p (its slot's parity to
// the root), child routing of P, and s ballot-packed into the node's row.
// With a nonzero `group` (depths whose nodes keep to one XCD) each XCD takes
// its nodes `group` at a time, as the plain-store union pass does: the finds
// of small nodes then spread over several nodes' words, still L2-resident.
// When the first-hook pass counted more than `chain_threshold` hooks under
// the slot just below (chains of neighbouring slots) at a depth whose nodes
// spread over several XCDs, the walks read and write coherently across XCDs,
// so that each XCD's walks take the shortcuts the others wrote.
template <typename W>
__global__ __launch_bounds__(kBlock) void k_route_inputs(const ItemCtx* __restrict__ items,
                                                         const Seg* __restrict__ segs,
                                                         uint32_t nsegs, int level0,
                                                         const uint32_t* __restrict__ plev,
                                                         uint32_t* __restrict__ pdst,
                                                         W* __restrict__ uf,
                                                         const NodeDesc* __restrict__ desc,
                                                         uint64_t* __restrict__ srows,
                                                         const unsigned long long* __restrict__ adjacent,
                                                         unsigned long long chain_threshold,
                                                         uint32_t group,
                                                         const int* __restrict__ errs) {
    const Seg* sgp;
    uint64_t node, chunk;
    if (!level_block(segs, nsegs, blockIdx.x, group, false, sgp, node, chunk)) return;
    const Seg& sg = *sgp;
    if (errs[sg.item]) return;
    const ItemCtx it = items[sg.item];
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt, row = nd.row;
    const uint64_t h = cnt >> 1, lc = h + (cnt & 1);
    const uint64_t p = chunk * kBlock + threadIdx.x;
    uint32_t s = 0;
    if (p < h) {
        // The pair's entries are loaded first, so that their latency overlaps
        // the find's dependent loads.
        const uint64_t a0 = load_perm(it, plev, level0, off + 2 * p);
        const uint64_t a1 = load_perm(it, plev, level0, off + 2 * p + 1);
        // The walk halves its path; then the start slot is pointed straight
        // at its root.  Slots are taken in ascending order within a node and
        // parents are smaller, so later walks soon meet such words.
        W* const u = uf + off;
        const W x0 = static_cast<W>(p + (cnt & 1));
        if (*adjacent > chain_threshold) {
            const W r = uf_find<CoherentForest>(u, x0, s);
            if (r != x0) CoherentForest::store(u + x0, (r << 1) | static_cast<W>(s));
        } else {
            const W r = uf_find<DeviceForest>(u, x0, s);
            if (r != x0) DeviceForest::store(u + x0, (r << 1) | static_cast<W>(s));
        }
        store_children(pdst, sg.narrow_out != 0, off, lc, p,
                       static_cast<uint32_t>((s ? a1 : a0) >> 1),
                       static_cast<uint32_t>((s ? a0 : a1) >> 1));
    } else if ((cnt & 1) && p == h) {
        store_tail(pdst, sg.narrow_out != 0, off, h,
                   static_cast<uint32_t>(load_perm(it, plev, level0, off + cnt - 1) >> 1));
    }
    const uint64_t ms = __ballot(s != 0);
    const uint64_t first = p - lane_id();
    if (lane_id() == 0 && first < h) srows[row + (first >> 6)] = ms;
}

// One thread per output pair: t_q = colour(I[2q]) read from the input rows,
// child routing of I, and t ballot-packed into the node's row.
// With `next_uf` (the next depth colours these nodes' children in HBM), the
// children's union-find words are set to the root word here, where the
// parent's words are no longer read: thread q covers slot q of both
// children (a child of c entries has ceil(c/2) slots, at most the parent's
// pair count), as 64-bit words where the next depth uses them.  The next
// depth's first hooks then store only the words that have a parent.
template <typename VI>
__global__ __launch_bounds__(kBlock) void k_route_outputs(const Seg* __restrict__ segs,
                                                          uint32_t nsegs,
                                                          const VI* __restrict__ inv,
                                                          uint32_t* __restrict__ idst,
                                                          const NodeDesc* __restrict__ desc,
                                                          const uint64_t* __restrict__ srows,
                                                          uint64_t* __restrict__ trows,
                                                          void* __restrict__ next_uf,
                                                          uint32_t next_wide,
                                                          const int* __restrict__ errs) {
    const Seg* sgp;
    uint64_t node, chunk;
    if (!level_block(segs, nsegs, blockIdx.x, 0u, false, sgp, node, chunk)) return;
    const Seg& sg = *sgp;
    if (errs[sg.item]) return;
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt, row = nd.row;
    const uint64_t h = cnt >> 1, lc = h + (cnt & 1);
    const uint64_t q = chunk * kBlock + threadIdx.x;
    uint32_t t = 0;
    if (q < h) {
        const uint64_t a0 = static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q));
        const uint64_t a1 = static_cast<uint64_t>(__builtin_nontemporal_load(inv + off + 2 * q + 1));
        // colour(x) = s_{x>>1} ^ (x & 1); the odd tail entry has colour 0.
        if (!((cnt & 1) && a0 == cnt - 1)) t = row_bit(srows, row, a0 >> 1) ^ (a0 & 1);
        store_children(idst, sg.narrow_out != 0, off, lc, q,
                       static_cast<uint32_t>((t ? a1 : a0) >> 1),
                       static_cast<uint32_t>((t ? a0 : a1) >> 1));
        if (next_uf) {
            const uint64_t lslots = (lc + 1) >> 1, rslots = (h + 1) >> 1;
            if (next_wide) {
                uint64_t* const w = static_cast<uint64_t*>(next_uf);
                if (q < lslots) w[off + q] = UfWord<uint64_t>::root;
                if (q < rslots) w[off + lc + q] = UfWord<uint64_t>::root;
            } else {
                uint32_t* const w = static_cast<uint32_t*>(next_uf);
                if (q < lslots) w[off + q] = UfWord<uint32_t>::root;
                if (q < rslots) w[off + lc + q] = UfWord<uint32_t>::root;
            }
        }
    } else if ((cnt & 1) && q == h) {
        store_tail(idst, sg.narrow_out != 0, off, h,
                   static_cast<uint32_t>(static_cast<uint64_t>(inv[off + cnt - 1]) >> 1));
    }
    const uint64_t mt = __ballot(t != 0);
    const uint64_t first = q - lane_id();
    if (lane_id() == 0 && first < h) trows[row + (first >> 6)] = mt;
}

// ---------------------------------------------------------------------------
// Second colouring of the HBM levels, independent of the union-find: orbit
// minima by rulers.  It replaces k_hook_first / k_union / k_flatten /
// k_route_inputs when selected, and is there to check the union-find
// colouring byte for byte at sizes no CPU encoder reaches.
//
// Along next(x) = I[P[x] ^ 1] ^ 1 (through x's output pair to the other
// entry there, then to that entry's input-pair partner) the colour stays the
// same: each step crosses one output pair and one input pair, each of which
// joins entries of different colours.  A component of the colouring graph
// (benes.cpp:247-297) is therefore two orbits of next(), one per colour,
// holding one entry of each input pair of the component.  In a cycle the
// smallest entry is even (it is 2s for the smallest pair s, whose partner
// 2s + 1 is the smallest entry of the other orbit), and the reference gives
// it colour 0, so colour(x) = (smallest entry of x's orbit) & 1.  In an odd
// node the tail input t = c - 1 (no input partner) and the unpaired output
// c - 1 (no output partner) end two chains instead: the one from t, which the
// reference colours 0 first, stops where an entry reaches the unpaired
// output; the other, colour 1, stops where next() would reach t.
//
// Rulers sample the orbits: one entry per aligned 32 entries of the level
// array (its offset in the bucket hashed from the bucket), numbered by the
// bucket.  Each ruler walks forward to the next ruler or a chain end,
// recording the segment's smallest entry and marking the entries it passed
// as its own; pointer jumping over the rulers (each segment of a cycle holds
// at most one bucket's ruler, so ceil(log2(rulers per node)) rounds cover a
// whole cycle) gives each ruler its orbit's minimum or chain end.  An entry
// no ruler passed walks forward itself until it meets a ruler or a marked
// entry (whose colour it shares), a chain end, or itself again (an orbit
// without rulers: a short cycle, whose minimum it has then seen).
constexpr uint32_t kRulerLog2 = 5;
constexpr uint32_t kRulerEndA = 0xFFFFFFFFu;     // chain to the unpaired output: colour 0
constexpr uint32_t kRulerEndB = 0xFFFFFFFEu;     // chain to the tail input: colour 1
constexpr uint32_t kRulerUnwalked = 0xFFFFFFFDu; // bucket whose ruler no walk started
constexpr uint32_t kUnowned = 0xFFFFFFFFu;       // entry no ruler passed
constexpr uint64_t kOrbitEndA = ~uint64_t{0};
constexpr uint64_t kOrbitEndB = ~uint64_t{1};

__device__ __forceinline__ bool is_ruler(uint64_t abs) {
    uint64_t z = (abs >> kRulerLog2) * 0x9E3779B97F4A7C15ull;
    z ^= z >> 29;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 32;
    return (abs & ((1u << kRulerLog2) - 1)) == (static_cast<uint32_t>(z) & ((1u << kRulerLog2) - 1));
}

// next(x) within the node at `off` of `cnt` entries, or a chain end.
template <typename VI>
__device__ __forceinline__ uint64_t orbit_next(const ItemCtx& it, const uint32_t* plev, bool level0,
                                               const VI* inv, uint64_t off, uint64_t cnt,
                                               uint64_t x) {
    const uint64_t y = level0 ? load_src(it, off + x - it.base) : static_cast<uint64_t>(plev[off + x]);
    if ((cnt & 1) && y == cnt - 1) return kOrbitEndA;
    const uint64_t z = static_cast<uint64_t>(inv[off + (y ^ 1)]);
    if ((cnt & 1) && z == cnt - 1) return kOrbitEndB;
    return z ^ 1;
}

// Thread per input pair (both entries) and the tail: the walks of the rulers.
template <typename VI>
__global__ __launch_bounds__(kBlock) void k_ruler_walk(const ItemCtx* __restrict__ items,
                                                       const Seg* __restrict__ segs, uint32_t nsegs,
                                                       int level0, const uint32_t* __restrict__ plev,
                                                       const VI* __restrict__ inv,
                                                       const NodeDesc* __restrict__ desc,
                                                       uint32_t* __restrict__ owner,
                                                       uint32_t* __restrict__ rnext,
                                                       uint64_t* __restrict__ rmin,
                                                       const int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    uint64_t node, chunk;
    if (!block_to_node(sg, node, chunk, false)) return;
    const ItemCtx it = items[sg.item];
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt, h = cnt >> 1;
    const uint64_t q = chunk * kBlock + threadIdx.x;
    const uint32_t first = q < h ? 0u : ((cnt & 1) && q == h ? 2u : 3u);
    for (uint32_t k = first; k < 3; ++k) {
        const uint64_t x = k < 2 ? 2 * q + k : cnt - 1;
        const uint64_t abs = off + x;
        if (!is_ruler(abs)) {
            if (k == 1) break;
            continue;
        }
        const uint32_t id = static_cast<uint32_t>(abs >> kRulerLog2);
        uint64_t cur = x, m = x;
        uint32_t nxt;
        while (true) {
            const uint64_t v = orbit_next(it, plev, level0 != 0, inv, off, cnt, cur);
            if (v == kOrbitEndA) {
                nxt = kRulerEndA;
                break;
            }
            if (v == kOrbitEndB) {
                nxt = kRulerEndB;
                break;
            }
            if (is_ruler(off + v)) {
                nxt = static_cast<uint32_t>((off + v) >> kRulerLog2);
                break;
            }
            owner[off + v] = id;
            m = v < m ? v : m;
            cur = v;
        }
        owner[abs] = id;
        rnext[id] = nxt;
        rmin[id] = m;
        if (k == 1) break;
    }
}

// One pointer-jumping round over the rulers (thread per bucket).
__global__ __launch_bounds__(kBlock) void k_ruler_jump(uint64_t buckets,
                                                       const uint32_t* __restrict__ next_in,
                                                       const uint64_t* __restrict__ min_in,
                                                       uint32_t* __restrict__ next_out,
                                                       uint64_t* __restrict__ min_out) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < buckets;
         i += stride) {
        const uint32_t j = next_in[i];
        const uint64_t m = min_in[i];
        if (j >= kRulerUnwalked) {
            next_out[i] = j;
            min_out[i] = m;
            continue;
        }
        const uint64_t mj = min_in[j];
        next_out[i] = next_in[j];
        min_out[i] = mj < m ? mj : m;
    }
}

__device__ __forceinline__ uint32_t ruler_colour(const uint32_t* rnext, const uint64_t* rmin,
                                                 uint32_t id) {
    const uint32_t n = rnext[id];
    if (n == kRulerEndA) return 0;
    if (n == kRulerEndB) return 1;
    return static_cast<uint32_t>(rmin[id] & 1);
}

// Thread per input pair: s_p = colour(2p) from the rulers (an entry no ruler
// passed walks to one), child routing of P, s ballot-packed into the node's
// row, as k_route_inputs does from the union-find.
template <typename VI>
__global__ __launch_bounds__(kBlock) void k_ruler_route(const ItemCtx* __restrict__ items,
                                                        const Seg* __restrict__ segs, uint32_t nsegs,
                                                        int level0, const uint32_t* __restrict__ plev,
                                                        const VI* __restrict__ inv,
                                                        uint32_t* __restrict__ pdst,
                                                        const NodeDesc* __restrict__ desc,
                                                        const uint32_t* __restrict__ owner,
                                                        const uint32_t* __restrict__ rnext,
                                                        const uint64_t* __restrict__ rmin,
                                                        uint64_t* __restrict__ srows,
                                                        const int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    uint64_t node, chunk;
    if (!block_to_node(sg, node, chunk, false)) return;
    const ItemCtx it = items[sg.item];
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, cnt = nd.cnt, row = nd.row;
    const uint64_t h = cnt >> 1, lc = h + (cnt & 1);
    const uint64_t p = chunk * kBlock + threadIdx.x;
    uint32_t s = 0;
    if (p < h) {
        const uint64_t x = 2 * p;
        const uint32_t o = owner[off + x];
        if (is_ruler(off + x)) {
            s = ruler_colour(rnext, rmin, static_cast<uint32_t>((off + x) >> kRulerLog2));
        } else if (o != kUnowned) {
            s = ruler_colour(rnext, rmin, o);
        } else {
            uint64_t cur = x, m = x;
            while (true) {
                const uint64_t v = orbit_next(it, plev, level0 != 0, inv, off, cnt, cur);
                if (v == kOrbitEndA || v == kOrbitEndB) {
                    s = v == kOrbitEndA ? 0u : 1u;
                    break;
                }
                if (v == x) {
                    s = static_cast<uint32_t>(m & 1);
                    break;
                }
                if (is_ruler(off + v)) {
                    s = ruler_colour(rnext, rmin, static_cast<uint32_t>((off + v) >> kRulerLog2));
                    break;
                }
                const uint32_t ov = owner[off + v];
                if (ov != kUnowned) {
                    s = ruler_colour(rnext, rmin, ov);
                    break;
                }
                m = v < m ? v : m;
                cur = v;
            }
        }
        const uint64_t a0 = load_perm(it, plev, level0, off + 2 * p);
        const uint64_t a1 = load_perm(it, plev, level0, off + 2 * p + 1);
        store_children(pdst, sg.narrow_out != 0, off, lc, p,
                       static_cast<uint32_t>((s ? a1 : a0) >> 1),
                       static_cast<uint32_t>((s ? a0 : a1) >> 1));
    } else if ((cnt & 1) && p == h) {
        store_tail(pdst, sg.narrow_out != 0, off, h,
                   static_cast<uint32_t>(load_perm(it, plev, level0, off + cnt - 1) >> 1));
    }
    const uint64_t ms = __ballot(s != 0);
    const uint64_t first = p - lane_id();
    if (lane_id() == 0 && first < h) srows[row + (first >> 6)] = ms;
}

// One workgroup per node of at most kLevelLds entries: the same three steps
// as k_union / k_route_inputs / k_route_outputs, with the node's union-find
// words and its input switch values in LDS instead of HBM.  A node of 2^16
// entries has 2^15 input slots: 64 KiB of 16-bit parent words, so two
// 1024-thread workgroups share a CU.  (16-bit words do not pay in k_subtree,
// where the CU holds the same number of workgroups either way and the
// emulated 16-bit compare-and-swap costs more than it saves.)
constexpr uint32_t kLevelLdsLog2 = 16;
constexpr uint32_t kLevelLds = 1u << kLevelLdsLog2;
constexpr uint32_t kLevelThreads = 1024;

template <typename VI>
__global__ __launch_bounds__(kLevelThreads) void k_level_lds(
    const ItemCtx* __restrict__ items, const Seg* __restrict__ segs, uint32_t nsegs, int level0,
    const uint32_t* __restrict__ plev, const VI* __restrict__ isrc, uint32_t* __restrict__ pdst,
    uint32_t* __restrict__ idst, const NodeDesc* __restrict__ desc, uint64_t* __restrict__ srows,
    uint64_t* __restrict__ trows, const int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    // 16-bit parent words (slots below 2^15), so two workgroups fit a CU.
    static_assert(kLevelLds / 2 <= (1u << 15), "LDS-level slots must fit 16-bit words");
    __shared__ uint32_t uf_words[kLevelLds / 4 + 1];
    __shared__ uint64_t sbits[kLevelLds / 128 + 1];
    // Unions whose finds met a long path (chains of neighbouring slots from
    // structured permutations): (lo, hi, label) packed, joined after the
    // workgroup has flattened the forest.
    constexpr uint32_t kLdsDeferred = 64, kLdsWalkCap = 256;
    __shared__ uint32_t deferred[kLdsDeferred];
    __shared__ uint32_t ndeferred;
    uint16_t* const uf = reinterpret_cast<uint16_t*>(uf_words);
    const ItemCtx it = items[sg.item];
    const uint64_t node = blockIdx.x - sg.block0;
    const NodeDesc& nd = desc[sg.desc0 + node];
    const uint64_t off = nd.off, row = nd.row;
    const uint32_t cnt = static_cast<uint32_t>(nd.cnt);
    const uint32_t h = cnt >> 1, odd = cnt & 1, lc = h + odd;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nthreads = blockDim.x, nwaves = blockDim.x >> 6;
    for (uint32_t i = tid; i < (lc + 1) / 2; i += nthreads) uf_words[i] = 0xFFFFFFFFu;
    if (tid == 0) ndeferred = 0;
    __syncthreads();
    // Below depth 0 the node's P and I are 16-bit entries (narrow_at): a
    // pair is one 32-bit load.
    const uint32_t* const ipairs =
        reinterpret_cast<const uint32_t*>(narrow_at(reinterpret_cast<const uint32_t*>(isrc), off));
    const uint32_t* const ppairs = reinterpret_cast<const uint32_t*>(narrow_at(plev, off));
    auto ipair = [&](uint32_t q, uint32_t& a, uint32_t& b) {
        if (level0) {
            a = static_cast<uint32_t>(isrc[off + 2 * q]);
            b = static_cast<uint32_t>(isrc[off + 2 * q + 1]);
        } else {
            const uint32_t two = __builtin_nontemporal_load(ipairs + q);
            a = two & 0xFFFFu;
            b = two >> 16;
        }
    };
    for (uint32_t q = tid; q < h; q += nthreads) {
        uint32_t a, b;
        ipair(q, a, b);
        const uint16_t sa = static_cast<uint16_t>(in_slot(a, cnt));
        const uint16_t sb = static_cast<uint16_t>(in_slot(b, cnt));
        if (sa == sb) continue;
        const uint32_t label = 1u ^ ((a ^ b) & 1u);
        if (uf_union_slots_bounded<GroupForest>(uf, sa, sb, label, kLdsWalkCap)) continue;
        const uint32_t i = atomicAdd(&ndeferred, 1u);
        if (i < kLdsDeferred) {
            deferred[i] = (static_cast<uint32_t>(sa) << 16) | (static_cast<uint32_t>(sb) << 1) | label;
        } else {
            uf_union_slots<GroupForest>(uf, sa, sb, label);
        }
    }
    __syncthreads();
    if (ndeferred != 0) {
        // Every slot finds its root at once (pointer jumping), then the
        // deferred unions walk short paths.
        for (uint32_t p = tid; p < h; p += nthreads) {
            uint32_t s;
            const uint16_t x0 = static_cast<uint16_t>(p + odd);
            const uint16_t r = uf_find<GroupForest>(uf, x0, s);
            if (r != x0) GroupForest::store(uf + x0, static_cast<uint16_t>((r << 1) | s));
        }
        __syncthreads();
        const uint32_t n = ndeferred < kLdsDeferred ? ndeferred : kLdsDeferred;
        for (uint32_t i = tid; i < n; i += nthreads) {
            const uint32_t e = deferred[i];
            uf_union_slots<GroupForest>(uf, static_cast<uint16_t>(e >> 16),
                                        static_cast<uint16_t>((e >> 1) & 0x7FFFu), e & 1u);
        }
        __syncthreads();
    }
    // Input switch values: route P, keep s in LDS and in the node's row.
    for (uint32_t base = wave * 64; base < lc; base += nwaves * 64) {
        const uint32_t p = base + lane;
        uint32_t s = 0;
        if (p < h) {
            uint32_t a0, a1;
            if (level0) {
                a0 = static_cast<uint32_t>(load_src(it, off + 2 * p - it.base));
                a1 = static_cast<uint32_t>(load_src(it, off + 2 * p + 1 - it.base));
            } else {
                const uint32_t two = __builtin_nontemporal_load(ppairs + p);
                a0 = two & 0xFFFFu;
                a1 = two >> 16;
            }
            uf_find<GroupForest>(uf, static_cast<uint16_t>(p + odd), s);
            store_children(pdst, true, off, lc, p, (s ? a1 : a0) >> 1, (s ? a0 : a1) >> 1);
        } else if (odd && p == h) {
            const uint32_t v = level0 ? static_cast<uint32_t>(load_src(it, off + cnt - 1 - it.base))
                                      : narrow_at(plev, off)[cnt - 1];
            store_tail(pdst, true, off, h, v >> 1);
        }
        const uint64_t ms = __ballot(s != 0);
        if (lane == 0) {
            sbits[base >> 6] = ms;
            if (base < h) srows[row + (base >> 6)] = ms;
        }
    }
    __syncthreads();
    // Output switch values t_q = colour(I[2q]); route the inverse.
    for (uint32_t base = wave * 64; base < lc; base += nwaves * 64) {
        const uint32_t q = base + lane;
        uint32_t t = 0;
        if (q < h) {
            uint32_t a0, a1;
            ipair(q, a0, a1);
            if (!(odd && a0 == cnt - 1)) {
                const uint32_t pa = a0 >> 1;
                t = static_cast<uint32_t>((sbits[pa >> 6] >> (pa & 63)) & 1) ^ (a0 & 1);
            }
            store_children(idst, true, off, lc, q, (t ? a1 : a0) >> 1, (t ? a0 : a1) >> 1);
        } else if (odd && q == h) {
            const uint32_t v =
                level0 ? static_cast<uint32_t>(isrc[off + cnt - 1])
                       : narrow_at(reinterpret_cast<const uint32_t*>(isrc), off)[cnt - 1];
            store_tail(idst, true, off, h, v >> 1);
        }
        const uint64_t mt = __ballot(t != 0);
        if (lane == 0 && base < h) trows[row + (base >> 6)] = mt;
    }
}

// Levels where a packet gives each node at most 4 bits (span <= 4) in runs
// of at least 64 bits (group depth >= 7): their full-share packets are
// packed by k_pack_narrow from whole row words, k_pack keeps the rest.  (At
// wider spans a word gathers from few rows, and k_pack's per-word gather is
// the cheaper of the two.)
__host__ __device__ __forceinline__ bool pack_narrow(uint32_t group_depth, uint32_t span_log2) {
    return span_log2 <= 2 && group_depth >= 7;
}

// Words of the blob one packet's share of a level (2^(g-1) bits, g the
// cluster group's depth) can touch.
__host__ __device__ __forceinline__ uint64_t pack_words_per_job(uint32_t g) {
    return ((uint64_t{1} << (g - 1)) + 63) / 64 + 1;
}

// Jobs are (cluster group, packet) of one item at depth d on one side, and
// each job has pack_words_per_job(g) threads.  When every node of the group's
// level fills its share of the packet, the share is one run of 2^r slices of
// one span each (internal.hpp:604-649): each thread gathers one output word
// of it from the nodes' bit rows and stores it, OR-ing only words shared with
// neighbouring runs.  In the packet where node shares differ, the job's
// threads take the nodes in turn, each placing its node's share at the node's
// own address.
__global__ __launch_bounds__(kBlock) void k_pack(const ItemCtx* __restrict__ items,
                                                 const Seg* __restrict__ segs, uint32_t nsegs,
                                                 uint32_t d, const NodeDesc* __restrict__ desc,
                                                 const uint64_t* __restrict__ srows,
                                                 const uint64_t* __restrict__ trows,
                                                 uint64_t* __restrict__ blob_all,
                                                 const int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    const int side = static_cast<int>(sg.side);
    const uint64_t* const rows = side == 0 ? srows : trows;
    const ItemCtx it = items[sg.item];
    const Plan& P = *it.plan;
    uint64_t* blob = blob_all + it.blob;
    const NodeDesc* dsc = desc + sg.desc0;
    const uint64_t packets = sg.chunks;
    const SideDepth sd = P.side[side][d];
    const uint32_t r = sd.relative_level, g = sd.group_depth;
    // A launch has fewer than 2^32 work-items: 32-bit divisions.
    const uint32_t wpj = static_cast<uint32_t>(pack_words_per_job(g));
    const uint32_t t = static_cast<uint32_t>((blockIdx.x - sg.block0) * kBlock + threadIdx.x);
    const uint32_t job = t / wpj, wi = t - job * wpj;
    const uint32_t root_depth = d - r;
    const uint64_t groups = uint64_t{1} << root_depth;
    const uint64_t group = job / static_cast<uint32_t>(packets);
    const uint64_t packet = job % static_cast<uint32_t>(packets);
    if (group >= groups) return;
    const uint32_t span_log2 = g - r - 1;
    const uint64_t span = uint64_t{1} << span_log2;
    const uint64_t C = node_count(P.n, root_depth, group);
    const uint64_t first_pair = packet << span_log2;
    const uint64_t min_pairs = (C >> r) >> 1;
    const uint64_t b0 = group << r;
    const uint64_t nodes = uint64_t{1} << r;
    const uint64_t secbase = (side == 0 ? P.input_offset : P.output_offset) * 8;
    if ((C >> r) > P.leaf && min_pairs >= first_pair + span) {
        if (pack_narrow(g, span_log2)) return;  // k_pack_narrow's share
        const uint64_t len = nodes << span_log2;
        const uint64_t dst0 = secbase + bit_of(dsc[b0].a[side], first_pair);
        const uint64_t ow = (dst0 >> 6) + wi, wbit = ow << 6;
        if (wbit >= dst0 + len) return;
        const uint64_t o_lo = wbit > dst0 ? wbit - dst0 : 0;
        const uint64_t o_hi = min(len, wbit + 64 - dst0);
        uint32_t sh = static_cast<uint32_t>(dst0 + o_lo - wbit);
        uint64_t v = 0;
        for (uint64_t o = o_lo; o < o_hi;) {
            const uint64_t in_node = o & (span - 1);
            const uint32_t take = static_cast<uint32_t>(min(o_hi - o, span - in_node));
            const uint64_t row = sg.row0 + (b0 + (o >> span_log2)) * sg.row_words;
            v |= row_range(rows, row, first_pair + in_node, take) << sh;
            sh += take;
            o += take;
        }
        if (o_hi - o_lo == 64) {
            blob[ow] = v;
        } else if (v) {
            atomicOr(reinterpret_cast<unsigned long long*>(blob + ow),
                     static_cast<unsigned long long>(v));
        }
        return;
    }
    // The packet where node shares differ (or a group's last, partial
    // packet): each node's share goes to its own address.  With fewer nodes
    // than the job has threads (a tall cluster group near its root, where a
    // share runs to 2^(g-1) bits), the threads split each share by output
    // word; one thread per share would write up to 2^(g-7) words in a row
    // while the rest of the launch waits.
    const uint64_t tpn = nodes < wpj ? wpj / nodes : 1;  // threads per node
    const uint64_t first_node = tpn > 1 ? wi / tpn : wi;
    const uint64_t node_step = tpn > 1 ? nodes : wpj;    // tpn > 1: one node per thread
    const uint64_t k0 = tpn > 1 ? wi % tpn : 0;
    for (uint64_t pi = first_node; pi < nodes; pi += node_step) {
        const NodeDesc& nd = dsc[b0 + pi];
        const uint64_t len = packet_share(nd.cnt, packet, span_log2, P.leaf);
        if (len == 0) continue;
        const uint64_t dst = secbase + bit_of(nd.a[side], first_pair);
        const uint64_t w_end = ((dst + len - 1) >> 6) + 1;
        for (uint64_t ow = (dst >> 6) + k0; ow < w_end; ow += tpn) {
            const uint64_t lo = max(ow << 6, dst), hi = min((ow + 1) << 6, dst + len);
            const uint64_t v =
                row_range(rows, nd.row, first_pair + (lo - dst), static_cast<uint32_t>(hi - lo))
                << (lo & 63);
            if (hi - lo == 64) {
                blob[ow] = v;
            } else if (v) {
                atomicOr(reinterpret_cast<unsigned long long*>(blob + ow),
                         static_cast<unsigned long long>(v));
            }
        }
    }
}

// Full-share packets of the narrow levels (pack_narrow).  A packet's run is
// the concatenation, over the group's nodes in path order, of each node's
// span bits [first_pair, first_pair + span): the run's word k holds the
// 64/span nodes of node group k.  A block holds kNarrowGroups waves, one per
// (node group, row word = 64 pairs): lane L loads the row word of the group's
// node L / span, and one ballot per packet of those 64 pairs, over the bits
// (packet * span + L % span), is that packet's word for the group.  One load
// per lane gives 64/span words, where gathering each word from the rows costs
// 64/span loads of its own.  The words go through LDS so that consecutive
// threads store consecutive words of one run.  A group-root set with fewer
// than kNarrowGroups node groups (small items) gives its block's remaining
// waves further row words instead of leaving them idle (narrow_shape).
constexpr uint32_t kNarrowGroups = 16;
constexpr uint32_t kNarrowThreads = 64 * kNarrowGroups;

// A narrow block's shape for `node_groups` node groups per set (a power of
// two): `ngb` node groups of each of `wpb` row words, `chunks` blocks across
// the node groups.
struct NarrowShape {
    uint64_t ngb, wpb, chunks;
};
__host__ __device__ __forceinline__ NarrowShape narrow_shape(uint64_t node_groups) {
    const uint64_t ngb = node_groups < kNarrowGroups ? node_groups : kNarrowGroups;
    return NarrowShape{ngb, ngb ? kNarrowGroups / ngb : 0, ngb ? node_groups / ngb : 0};
}

__global__ __launch_bounds__(kNarrowThreads) void k_pack_narrow(
    const ItemCtx* __restrict__ items, const Seg* __restrict__ segs, uint32_t nsegs, uint32_t d,
    const NodeDesc* __restrict__ desc, const uint64_t* __restrict__ srows,
    const uint64_t* __restrict__ trows, uint64_t* __restrict__ blob_all,
    const int* __restrict__ errs) {
    __shared__ uint64_t tile[64][kNarrowGroups];
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    const int side = static_cast<int>(sg.side);
    const uint64_t* const rows = side == 0 ? srows : trows;
    const ItemCtx it = items[sg.item];
    const Plan& P = *it.plan;
    uint64_t* blob = blob_all + it.blob;
    const SideDepth sd = P.side[side][d];
    const uint32_t r = sd.relative_level, g = sd.group_depth;
    const uint32_t span_log2 = g - r - 1;
    const uint32_t span = 1u << span_log2;
    const uint32_t per_word = 64u >> span_log2;  // nodes of a group; packets of a row word
    const uint64_t node_groups = (uint64_t{1} << r) >> (6 - span_log2);
    const NarrowShape sh = narrow_shape(node_groups);
    if (sh.ngb == 0) return;
    const uint64_t rw = sg.row_words;
    const uint64_t wblocks = (rw + sh.wpb - 1) / sh.wpb;
    const uint64_t b = blockIdx.x - sg.block0;
    const uint32_t root_depth = d - r;
    const uint64_t set = b / (sh.chunks * wblocks);
    if (set >= (uint64_t{1} << root_depth)) return;
    const uint64_t rest = b - set * (sh.chunks * wblocks);
    const uint64_t chunk = rest / wblocks, wb = rest - chunk * wblocks;
    const uint64_t C = node_count(P.n, root_depth, set);
    if ((C >> r) <= P.leaf) return;
    const uint64_t min_pairs = (C >> r) >> 1;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t b0 = set << r;
    // Wave v: node group chunk * ngb + v % ngb of row word wb * wpb + v / ngb.
    const uint64_t w = wb * sh.wpb + wave / sh.ngb;
    const uint64_t ng = chunk * sh.ngb + wave % sh.ngb;
    if (w < rw && (w << 6) < min_pairs) {
        const uint64_t node = b0 + ng * per_word + (lane >> span_log2);
        const uint64_t word = rows[sg.row0 + node * rw + w];
        for (uint32_t pk = 0; pk < per_word; ++pk) {
            const uint64_t bits = __ballot(((word >> (pk * span + (lane & (span - 1)))) & 1) != 0);
            if (lane == pk) tile[pk][wave] = bits;
        }
    }
    __syncthreads();
    // Thread (pk, column): packet pk's word of the column's wave.
    const uint32_t pk = threadIdx.x / kNarrowGroups, col = threadIdx.x % kNarrowGroups;
    const uint64_t wc = wb * sh.wpb + col / sh.ngb;
    const uint64_t ngc = chunk * sh.ngb + col % sh.ngb;
    if (pk >= per_word || wc >= rw) return;
    const uint64_t first_pair = ((wc * per_word) + pk) << span_log2;
    if (first_pair + span > min_pairs) return;  // not a full share: k_pack places it
    const uint64_t secbase = (side == 0 ? P.input_offset : P.output_offset) * 8;
    const uint64_t dst = secbase + bit_of(desc[sg.desc0 + b0].a[side], first_pair) + ngc * 64;
    const uint64_t v = tile[pk][col];
    if ((dst & 63) == 0) {
        blob[dst >> 6] = v;  // inside the run: no other writer
    } else {
        or_word(blob, dst, v);
    }
}

// ---------------------------------------------------------------------------
// Lookup (the reference's lookup, internal.hpp:1036-1124): one thread per
// value.  Forward: descend through the input switches (the odd tail always
// goes left), decode the leaf's factoradic rank, climb back through the output
// switches; inverse swaps the two sides.  Addresses come from the same closed
// forms the encoder uses; each climb step recomputes its node's address, so a
// lookup costs O(depth^2) address steps and no per-thread stack.

__device__ __forceinline__ uint32_t blob_bit(const uint8_t* blob, uint64_t bit) {
    return (blob[bit >> 3] >> (bit & 7)) & 1u;
}

// Eight 16-bit limbs (least significant first) of the `bits` (<= 118) bits
// of the blob starting at bit `at`.
__device__ __forceinline__ void read_rank_limbs(const uint8_t* blob, uint64_t at, uint32_t bits,
                                                uint32_t limb[8]) {
    uint64_t lo = 0, hi = 0;
    if (bits == 0) {
#pragma unroll
        for (uint32_t l = 0; l < 8; ++l) limb[l] = 0;
        return;
    }
    const uint64_t byte0 = at >> 3;
    const uint32_t sh = at & 7;
    const uint32_t nbytes = (sh + bits + 7) >> 3;  // at most 16
    for (uint32_t b = 0; b < nbytes; ++b) {
        const uint64_t v = blob[byte0 + b];
        if (b < 8) lo |= v << (8 * b);
        else hi |= v << (8 * (b - 8));
    }
    // Shift the 128-bit (hi:lo) right by sh and keep `bits` bits.
    if (sh) {
        lo = (lo >> sh) | (hi << (64 - sh));
        hi >>= sh;
    }
    if (bits < 64) {
        lo &= bits ? (~uint64_t{0} >> (64 - bits)) : 0;
        hi = 0;
    } else if (bits < 128) {
        hi &= bits > 64 ? (~uint64_t{0} >> (128 - bits)) : 0;
    }
#pragma unroll
    for (uint32_t l = 0; l < 4; ++l) {
        limb[l] = static_cast<uint32_t>(lo >> (16 * l)) & 0xFFFFu;
        limb[l + 4] = static_cast<uint32_t>(hi >> (16 * l)) & 0xFFFFu;
    }
}

// Divides the limbs by k (<= 32) in place and returns the remainder.
__device__ __forceinline__ uint32_t divide_limbs(uint32_t limb[8], uint32_t k) {
    uint32_t rem = 0;
#pragma unroll
    for (int l = 7; l >= 0; --l) {
        const uint32_t cur = (rem << 16) | limb[l];
        limb[l] = cur / k;
        rem = cur - limb[l] * k;
    }
    return rem;
}

// k-th (0-based) set bit of m.
__device__ __forceinline__ uint32_t select_bit(uint32_t m, uint32_t k) {
    for (uint32_t i = 0; i < k; ++i) m &= m - 1;
    return static_cast<uint32_t>(__builtin_ctz(m));
}

template <typename VO>
__global__ __launch_bounds__(kBlock) void k_lookup(const Plan* __restrict__ plan,
                                                   const uint8_t* __restrict__ blob,
                                                   const uint64_t* __restrict__ values,
                                                   uint64_t count, int inverse,
                                                   VO* __restrict__ out, int* __restrict__ err) {
    const Plan& P = *plan;
    const int down = inverse ? 1 : 0, up = 1 - down;
    const uint64_t down_sec = (down ? P.output_offset : P.input_offset) * 8;
    const uint64_t up_sec = (up ? P.output_offset : P.input_offset) * 8;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < count;
         t += stride) {
        uint64_t index = values ? values[t] : t;
        if (index >= P.n) {
            err[0] = 1;
            out[t] = 0;
            continue;
        }
        // Descend.
        Walk w = walk_root(P, down);
        uint64_t cnt = P.n, path = 0, rmask = 0;
        uint32_t d = 0;
        while (cnt > P.leaf) {
            bool right = false;
            if (!((cnt & 1) && index == cnt - 1)) {
                const SideAddr a = address_at(P, w, d, path);
                right = blob_bit(blob, down_sec + bit_of(a, index >> 1)) != (index & 1);
            }
            rmask |= static_cast<uint64_t>(right) << d;
            const uint64_t left_cnt = cnt - (cnt >> 1);
            path = path * 2 + (right ? 1 : 0);
            ++d;
            w = walk_from(P, down, w, d - 1, d, path);
            index >>= 1;
            cnt = right ? (cnt >> 1) : left_cnt;
        }
        // The leaf: Lehmer digits from the top (d_i = rank mod (i+1), then
        // divide), entry i the d_i-th smallest value not taken by later ones.
        const uint32_t c = static_cast<uint32_t>(cnt);
        uint32_t limb[8];
        read_rank_limbs(blob, leaf_section_offset(P) * 8 + leaf_rank_bit(P, d, path),
                        rank_bits(c), limb);
        uint32_t avail = c >= 32 ? ~0u : ((1u << c) - 1u);
        uint64_t found = ~uint64_t{0};
        for (uint32_t i = c; i-- > 0;) {
            const uint32_t value = select_bit(avail, divide_limbs(limb, i + 1));
            avail &= ~(1u << value);
            if (!inverse && i == index) {
                found = value;
                break;
            }
            if (inverse && value == index) {
                found = i;
                break;
            }
        }
        index = found;
        // Climb.
        for (uint32_t k = d; k-- > 0;) {
            const uint64_t pk = path >> (d - k);
            const bool right = (rmask >> k) & 1;
            const uint64_t size = node_count(P.n, k, pk), pairs = size >> 1;
            if (!right && (size & 1) && index == pairs) {
                index = size - 1;
                continue;
            }
            const SideAddr a = node_address(P, up, k, pk);
            index = index * 2 + ((right ? 1u : 0u) ^ blob_bit(blob, up_sec + bit_of(a, index)));
        }
        out[t] = static_cast<VO>(index);
    }
}

// ---------------------------------------------------------------------------
// Subtrees in LDS
//
// One workgroup owns a subtree of at most 2^SubLog2 entries (a node at the fusion
// depth d0) and runs every remaining level in LDS: node-local permutation
// (u16, ping-pong; while a level is coloured, the other half holds the
// nodes' inverses), union-find words over input slots (u32), and the level's
// switch values as bit arrays indexed like the permutation.  Work is dealt as (node, 64-entry
// chunk) items to waves, so each wave stays inside one node and the node's
// offset comes from a per-level LDS table.

// Number of the first k paths at `level` whose node holds an extra entry
// (32-bit form of large_before for subtree-local counts).
__device__ __forceinline__ uint32_t large_before32(uint32_t k, uint32_t m, uint32_t level) {
    uint32_t included = 0;
    for (uint32_t bits = level; bits != 0; --bits) {
        const uint32_t half = 1u << (bits - 1);
        const uint32_t even = (k + 1) >> 1;
        if (m <= half) {
            k = even;
        } else {
            included += even;
            k >>= 1;
            m -= half;
        }
    }
    return included + ((k != 0 && m != 0) ? 1u : 0u);
}

__device__ __forceinline__ uint32_t rev32(uint32_t v, uint32_t bits) {
    return bits == 0 ? 0u : (__brev(v) >> (32 - bits));
}

// Count of node pi at relative level k below a root of n0 entries.
__device__ __forceinline__ uint32_t sub_count(uint32_t n0, uint32_t k, uint32_t pi) {
    const uint32_t m = n0 & ((1u << k) - 1);
    return (n0 >> k) + (rev32(pi, k) < m ? 1u : 0u);
}

// Factoradic rank (internal.hpp:148-167) of one leaf, OR'ed into the blob at
// bit `pos`.  Lehmer digit i counts earlier entries smaller than entry i,
// read off a bitmask of the values seen so far; Horner with radix i+1,
// i ascending.
__device__ __forceinline__ void rank_leaf(uint64_t pos, const uint16_t* vals, uint32_t c,
                                          uint64_t* blob) {
    uint32_t seen = 0;
    unsigned __int128 r = 0;
    for (uint32_t i = 0; i < c; ++i) {
        const uint32_t v = vals[i];
        const uint32_t digit = __popc(seen & ((1u << v) - 1u));
        seen |= 1u << v;
        r = r * (i + 1) + digit;
    }
    if (rank_bits(c) == 0) return;
    const uint64_t lo = static_cast<uint64_t>(r), hi = static_cast<uint64_t>(r >> 64);
    const uint64_t w = pos >> 6;
    const uint32_t sh = pos & 63;
    const uint64_t x0 = lo << sh;
    const uint64_t x1 = sh ? ((lo >> (64 - sh)) | (hi << sh)) : hi;
    const uint64_t x2 = sh ? (hi >> (64 - sh)) : 0;
    auto* b = reinterpret_cast<unsigned long long*>(blob);
    if (x0) atomicOr(b + w, static_cast<unsigned long long>(x0));
    if (x1) atomicOr(b + w + 1, static_cast<unsigned long long>(x1));
    if (x2) atomicOr(b + w + 2, static_cast<unsigned long long>(x2));
}

// Walk state at the subtree root, per side, plus the separate-placement rank
// offset there: every address below continues from it.
struct SubtreeRoot {
    Walk walk[2];
    uint64_t middle;
};

// Separate-placement rank offset of node (d, path), continuing from depth
// `from` (benes.cpp:317-319: a right child starts after its left sibling's
// ranks).
__device__ __forceinline__ uint64_t middle_from(const Plan& P, uint64_t mid, uint32_t from,
                                                uint32_t d, uint64_t path) {
    for (uint32_t i = from + 1; i <= d; ++i) {
        const uint64_t here = path >> (d - i);
        if (here & 1) mid += P.middle_bits[i][node_large(P.n, i, here - 1) ? 1 : 0];
    }
    return mid;
}

// Bits [pos, pos + len) (len <= 64) of an LDS bit array.
__device__ __forceinline__ uint64_t lds_bits(const uint64_t* bits, uint32_t pos, uint32_t len) {
    const uint32_t w = pos >> 6, sh = pos & 63;
    uint64_t v = bits[w] >> sh;
    if (sh + len > 64) v |= bits[w + 1] << (64 - sh);
    return len >= 64 ? v : v & ((uint64_t{1} << len) - 1);
}

// Writes the switch bits of every internal node at relative level k of a
// subtree (absolute depth d0 + k) into the blob.  Bits are produced packet by
// packet: within a cluster group, one packet's share of a level is the
// concatenation of that packet's pair range from each node in path order
// (internal.hpp:604-649), so a run of consecutive nodes is one contiguous
// stretch of the blob.  `sbits`/`tbits` hold the level's switch values at the
// position of each node's pair in LDS.
//
// Jobs are (set of nodes sharing a cluster group, packet).  When every node
// of the set fills its whole share of the packet (all but the last packet of
// a set, and sets without leaves), the run is `set_nodes` slices of one span
// each: one lane per output word gathers its word's slices from LDS and
// stores it, or ORs it in when the word is shared with a neighbouring run.
// The other jobs place nodes of differing shares with a prefix sum.
__device__ __forceinline__ void emit_level_runs(const Plan& P, const SubtreeRoot& root,
                                                const SideAddr (*first_addr)[kSubLog2Max + 1],
                                                uint32_t d0, uint64_t b0, uint32_t n0,
                                                uint32_t k, const uint64_t* sbits,
                                                const uint64_t* tbits, const uint16_t* noff,
                                                uint64_t* blob) {
    const uint32_t d = d0 + k;
    const uint32_t L = P.leaf;
    const uint32_t wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const uint32_t cmax = (n0 + ((1u << k) - 1)) >> k;
    const uint32_t maxpairs = cmax >> 1;
    for (int side = 0; side < 2; ++side) {
        const SideDepth& sd = P.side[side][d];
        const uint32_t r = sd.relative_level, g = sd.group_depth;
        const uint32_t R = d - r;
        const uint32_t span_log2 = g - r - 1;
        const uint32_t span = 1u << span_log2;
        const uint64_t* rows = side == 0 ? sbits : tbits;
        const uint64_t secbase = (side == 0 ? P.input_offset : P.output_offset) * 8;
        // Sets of consecutive nodes that share one cluster group.
        const bool one_set = R <= d0;
        const uint32_t nsets = one_set ? 1u : (1u << (R - d0));
        const uint32_t set_nodes = one_set ? (1u << k) : (1u << r);
        const uint32_t packets = (maxpairs + span - 1) >> span_log2;
        const uint32_t jobs = nsets * packets;
        const uint32_t len = set_nodes << span_log2;
        auto full_share = [&](uint32_t set, uint32_t first_pair) {
            const uint32_t minc = one_set ? (n0 >> k) : (sub_count(n0, R - d0, set) >> r);
            return minc > L && (minc >> 1) >= first_pair + span;
        };
        // Equal shares: one lane per output word.
        const uint32_t wpj = (len + 63) / 64 + 1;  // words one run can touch
        {
            uint32_t cached_set = ~0u;
            SideAddr a{};
            for (uint32_t t = threadIdx.x; t < jobs * wpj; t += blockDim.x) {
                const uint32_t job = t / wpj, wi = t - job * wpj;
                const uint32_t set = job / packets, pk = job - set * packets;
                const uint32_t first_pair = pk << span_log2;
                if (!full_share(set, first_pair)) continue;
                const uint32_t first = set * set_nodes;
                if (set != cached_set) {
                    const uint64_t path = (b0 << k) | first;
                    a = first == 0 ? first_addr[side][k]
                                   : address_at(P, walk_from(P, side, root.walk[side], d0, d, path),
                                                d, path);
                    cached_set = set;
                }
                const uint64_t dst0 = secbase + bit_of(a, first_pair);
                const uint64_t ow = (dst0 >> 6) + wi, wbit = ow << 6;
                if (wbit >= dst0 + len) continue;
                const uint32_t o_lo = wbit > dst0 ? static_cast<uint32_t>(wbit - dst0) : 0u;
                const uint32_t o_hi =
                    static_cast<uint32_t>(min(static_cast<uint64_t>(len), wbit + 64 - dst0));
                const uint32_t sh0 = static_cast<uint32_t>(dst0 + o_lo - wbit);
                // The word's slices are visited starting at a lane-dependent
                // one: lanes of a wave hold consecutive words, whose slices
                // lie one span of nodes apart in LDS, and in step they would
                // all hit the same bank.
                const uint32_t s_lo = o_lo >> span_log2, ns = ((o_hi - 1) >> span_log2) - s_lo + 1;
                uint32_t j = lane_id() % ns;
                uint64_t v = 0;
                for (uint32_t it = 0; it < ns; ++it) {
                    const uint32_t sl = s_lo + j;
                    const uint32_t lo = max(o_lo, sl << span_log2);
                    const uint32_t hi = min(o_hi, (sl + 1) << span_log2);
                    v |= lds_bits(rows, noff[first + sl] + first_pair + (lo & (span - 1)), hi - lo)
                         << (sh0 + lo - o_lo);
                    j = j + 1 == ns ? 0 : j + 1;
                }
                if (o_hi - o_lo == 64) {
                    blob[ow] = v;
                } else if (v) {
                    atomicOr(reinterpret_cast<unsigned long long*>(blob + ow),
                             static_cast<unsigned long long>(v));
                }
            }
        }
        if (len <= 64) {
            // Differing shares, short runs: one lane per (set, packet); the run
            // starts at the first internal node's address and each later node
            // continues where the previous one ended.
            for (uint32_t job = threadIdx.x; job < jobs; job += blockDim.x) {
                const uint32_t set = job / packets, pk = job - set * packets;
                const uint32_t first_pair = pk << span_log2;
                if (full_share(set, first_pair)) continue;
                const uint32_t first = set * set_nodes, end = first + set_nodes;
                uint32_t f = first;
                while (f < end && sub_count(n0, k, f) <= L) ++f;
                if (f == end) continue;
                // The first node's address is at hand when it is the
                // level's first node (its walk from the subtree root is a
                // chain of dependent table reads).
                const uint64_t path = (b0 << k) | f;
                const SideAddr a =
                    f == 0 ? first_addr[side][k]
                           : address_at(P, walk_from(P, side, root.walk[side], d0, d, path), d,
                                        path);
                // The run (at most 64 bits) is gathered in two words, then
                // ORed into the blob: two atomics per run, not one per node.
                const uint64_t dst0 = secbase + bit_of(a, first_pair);
                uint32_t at = static_cast<uint32_t>(dst0 & 63);
                uint64_t w0 = 0, w1 = 0;
                for (uint32_t pi = f; pi < end; ++pi) {
                    const uint32_t share = static_cast<uint32_t>(
                        packet_share(sub_count(n0, k, pi), pk, span_log2, L));
                    if (share) {
                        const uint64_t v = lds_bits(rows, noff[pi] + first_pair, share);
                        if (at < 64) {
                            w0 |= v << at;
                            if (at + share > 64) w1 |= v >> (64 - at);
                        } else {
                            w1 |= v << (at - 64);
                        }
                    }
                    at += share;
                }
                auto* const w = reinterpret_cast<unsigned long long*>(blob) + (dst0 >> 6);
                if (w0) atomicOr(w, static_cast<unsigned long long>(w0));
                if (w1) atomicOr(w + 1, static_cast<unsigned long long>(w1));
            }
            continue;
        }
        // Differing shares, long runs: lanes take nodes, a wave prefix sum
        // places them.
        for (uint32_t job = wave; job < jobs; job += nwaves) {
            const uint32_t set = job / packets, pk = job - set * packets;
            const uint32_t first_pair = pk << span_log2;
            if (full_share(set, first_pair)) continue;  // wave-uniform
            const uint32_t first = set * set_nodes, end = first + set_nodes;
            uint32_t f = first;
            while (f < end && sub_count(n0, k, f) <= L) ++f;  // wave-uniform
            if (f == end) continue;
            const uint64_t path = (b0 << k) | f;
            const SideAddr a =
                f == 0 ? first_addr[side][k]
                       : address_at(P, walk_from(P, side, root.walk[side], d0, d, path), d, path);
            const uint64_t run0 = secbase + bit_of(a, first_pair);
            uint64_t carry = 0;
            for (uint32_t base = f; base < end; base += 64) {
                const uint32_t pi = base + lane_id();
                const uint64_t share =
                    pi < end ? packet_share(sub_count(n0, k, pi), pk, span_log2, L) : 0;
                const uint64_t incl = wave_inclusive_sum(share);
                if (share) {
                    // The lane's share, up to 64 row bits at a time.
                    const uint32_t pos0 = noff[pi] + first_pair;
                    const uint64_t dst = run0 + carry + incl - share;
                    for (uint32_t b = 0; b < share; b += 64) {
                        const uint32_t cnt = min(64u, static_cast<uint32_t>(share) - b);
                        or_word(blob, dst + b, lds_bits(rows, pos0 + b, cnt));
                    }
                }
                carry += __shfl(incl, 63);
            }
        }
    }
}

// The subtree's switch rows of each level (both sides) leave the colouring
// kernel through HBM and are packed by k_subtree_emit: the packing's address
// arithmetic needs about twice the registers of the colouring, and kept in
// one kernel it held the colouring to 4 waves per SIMD.  Rows of subtree b at
// relative level k, side s: words [((b * levels + k) * 2 + s) * W, + W) of
// `lbits`, W = subtree_row_words(SubLog2).
__host__ __device__ constexpr uint32_t subtree_row_words(uint32_t sub_log2) {
    return (1u << sub_log2) / 64 + 2;
}

// Waves per SIMD the colouring kernel is compiled for.  2^12 x 512 (38 KB of
// LDS, four workgroups per CU) and 2^13 x 1024 (76 KB, two per CU) fill all
// 8 wave slots of a SIMD when the kernel keeps to 64 VGPRs; left to itself
// the compiler takes 71-73 and the CU runs 7 or 6 waves per SIMD (2^12: one
// workgroup fewer per CU).  The other shapes are held to 4 by their LDS.
__host__ __device__ constexpr uint32_t subtree_waves(uint32_t sub_log2, uint32_t threads) {
    return (sub_log2 == 12 && threads == 512) || (sub_log2 == 13 && threads == 1024) ? 8u : 4u;
}

// Unions each lane keeps in flight in the colouring kernel: four where
// registers allow, fewer where the 64-VGPR budget above applies (3 and 2 are
// the most that fit it without spills; below 8 waves the extra unions in
// flight hid less than the extra waves do).
__host__ __device__ constexpr uint32_t subtree_union_batch(uint32_t sub_log2, uint32_t threads) {
    return subtree_waves(sub_log2, threads) < 8 ? 4u : (sub_log2 == 12 ? 3u : 2u);
}

template <uint32_t SubLog2, uint32_t Threads>
__global__ __launch_bounds__(Threads) __attribute__((amdgpu_waves_per_eu(
    subtree_waves(SubLog2, Threads)))) void k_subtree(const ItemCtx* __restrict__ items,
                                                         const Seg* __restrict__ segs,
                                                         uint32_t nsegs,
                                                         const uint32_t* __restrict__ lev0,
                                                         const uint32_t* __restrict__ lev1,
                                                         uint64_t* __restrict__ lbits,
                                                         uint32_t levels,
                                                         uint64_t* __restrict__ blob_all,
                                                         const int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    constexpr uint32_t kSub = 1u << SubLog2;
    constexpr uint32_t kRowWords = subtree_row_words(SubLog2);
    __shared__ uint16_t buf[2][kSub];
    __shared__ uint32_t uf[kSub];
    __shared__ uint64_t sbits[kRowWords];
    __shared__ uint64_t tbits[kRowWords];
    __shared__ uint16_t noff[kSub / 2 + 1];
    __shared__ SubtreeRoot root;
    const ItemCtx it = items[sg.item];
    const Plan& P = *it.plan;
    uint64_t* blob = blob_all + it.blob;
    const uint32_t L = P.leaf;
    const uint32_t d0 = it.df;
    const uint64_t b0 = blockIdx.x - sg.block0;
    const uint32_t n0 = static_cast<uint32_t>(node_count(P.n, d0, b0));
    const uint64_t base = node_offset(P.n, d0, b0);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nthreads = blockDim.x, nwaves = blockDim.x >> 6;
    // Walks to the subtree root (for the leaves' rank addresses).  Each is a
    // chain of dependent table reads as long as the subtree is deep, which
    // the first barrier waits for: the two sides run in two lanes of one
    // branch and the middle-section walk in another wave, all at once (in
    // three branches of one wave they took 11.6 of the 13 us before it).
    if (tid < 2) root.walk[tid] = walk_to(P, static_cast<int>(tid), d0, b0);
    if (tid == 64) root.middle = middle_from(P, 0, 0, d0, b0);
    // Depth d0 - 1 wrote level (d0 - 1) & 1; depth 0 reads the item's source.
    // Every lane issues its loads before storing any (a loop that stored
    // each value before loading the next paid one memory latency per entry).
    const uint32_t* plev = ((d0 - 1) & 1) ? lev1 : lev0;
    {
        constexpr uint32_t kPer = kSub / Threads;
        static_assert(kPer * Threads == kSub, "a subtree's entries split evenly over the lanes");
        uint32_t v[kPer];
        if (d0 == 0) {
#pragma unroll
            for (uint32_t j = 0; j < kPer; ++j) {
                const uint32_t i = tid + j * Threads;
                v[j] = i < n0 ? static_cast<uint32_t>(load_src(it, base + i)) : 0u;
            }
        } else {
            // 16-bit entries (narrow_at), as every node of at most 2^16.
            const uint16_t* src = narrow_at(plev, it.base + base);
#pragma unroll
            for (uint32_t j = 0; j < kPer; ++j) {
                const uint32_t i = tid + j * Threads;
                v[j] = i < n0 ? static_cast<uint32_t>(__builtin_nontemporal_load(src + i)) : 0u;
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint32_t i = tid + j * Threads;
            if (i < n0) buf[0][i] = static_cast<uint16_t>(v[j]);
        }
    }
    __syncthreads();
    const int rank_side = P.placement == 1 ? 0 : 1;
    const uint64_t rank_section = leaf_section_offset(P) * 8;
    int cur = 0;
    for (uint32_t k = 0;; ++k) {
        const uint32_t d = d0 + k;
        uint16_t* perm = buf[cur];
        uint16_t* next = buf[cur ^ 1];
        const uint32_t nodes = 1u << k;
        const uint32_t q = n0 >> k, m = n0 & (nodes - 1);
        const uint32_t cmax = q + (m != 0 ? 1u : 0u);
        // Node offsets of this level (leaves included).
        for (uint32_t pi = tid; pi < nodes; pi += nthreads) {
            noff[pi] = static_cast<uint16_t>(pi * q + large_before32(pi, m, k));
        }
        __syncthreads();
        // Leaves whose parent was internal: rank them now.
        if (q <= L) {
            for (uint32_t pi = tid; pi < nodes; pi += nthreads) {
                const uint32_t c = sub_count(n0, k, pi);
                if (c > L) continue;
                if (k > 0 && sub_count(n0, k - 1, pi >> 1) <= L) continue;
                const uint64_t path = (b0 << k) | pi;
                uint64_t bit;
                if (P.placement == 0) {
                    bit = middle_from(P, root.middle, d0, d, path);
                } else {
                    bit = rank_bit_at(walk_from(P, rank_side, root.walk[rank_side], d0, d, path),
                                      L);
                }
                rank_leaf(rank_section + bit, perm + noff[pi], c, blob);
            }
        }
        if (cmax <= L) break;
        for (uint32_t i = tid; i < n0; i += nthreads) uf[i] = UfWord<uint32_t>::root;
        for (uint32_t i = tid; i < kSub / 64 + 2; i += nthreads) {
            sbits[i] = 0;
            tbits[i] = 0;
        }
        __syncthreads();
        // The node's inverse, held in the other half of the ping-pong buffer
        // until routing overwrites it.  Items are (node, 64-entry chunk).
        const uint32_t cpn = (cmax + 63) >> 6;
        for (uint32_t item = wave; item < nodes * cpn; item += nwaves) {
            const uint32_t pi = item / cpn, chunk = item - pi * cpn;
            const uint32_t c = sub_count(n0, k, pi);
            if (c <= L) continue;
            const uint32_t x = chunk * 64 + lane;
            if (x >= c) continue;
            const uint32_t off = noff[pi];
            next[off + perm[off + x]] = static_cast<uint16_t>(x);
        }
        __syncthreads();
        // Pair-indexed steps run over a flat index g = node * S + pair, S the
        // power of two covering a node's pairs and, in an odd node, its tail
        // slot (index c >> 1): (c + 1) >> 1 indices for c entries, at most
        // (cmax + 1) >> 1.  A wave covers 64 / S small nodes at once, or 64
        // pairs of one large node.  (Sizing S by the largest node's pairs
        // plus one whenever any node is odd doubled it below most subtrees
        // of sizes that are not powers of two, such as 4095 = 2048 + 2047.)
        const uint32_t s_log2 = 32u - __builtin_clz((((cmax + 1) >> 1) - 1) | 1u);
        const uint32_t S = 1u << s_log2, total = nodes << s_log2;
        // Union of the two input slots reaching each output pair (the same
        // input-slot forest as the HBM levels).  Each lane takes kUnionBatch
        // pairs at a time and issues their first hooks together (most settle
        // there), so the LDS round trips of independent unions overlap.
        constexpr uint32_t kUnionBatch = subtree_union_batch(SubLog2, Threads);
        for (uint32_t g0 = tid; g0 < total; g0 += kUnionBatch * nthreads) {
            uint32_t base[kUnionBatch], hi[kUnionBatch], lo[kUnionBatch], found[kUnionBatch];
            bool act[kUnionBatch];
#pragma unroll
            for (uint32_t u = 0; u < kUnionBatch; ++u) {
                const uint32_t g = g0 + u * nthreads;
                act[u] = false;
                if (g >= total) continue;
                const uint32_t pi = g >> s_log2, qq = g & (S - 1);
                const uint32_t c = sub_count(n0, k, pi);
                if (c <= L || qq >= (c >> 1)) continue;
                const uint32_t off = noff[pi];
                const uint32_t a = next[off + 2 * qq], b = next[off + 2 * qq + 1];
                const uint32_t sa = input_slot<uint32_t>(a, c), sb = input_slot<uint32_t>(b, c);
                if (sa == sb) continue;
                act[u] = true;
                base[u] = off;
                hi[u] = sa > sb ? sa : sb;
                lo[u] = ((sa > sb ? sb : sa) << 1) | (1u ^ ((a ^ b) & 1u));
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnionBatch; ++u) {
                if (act[u]) found[u] = GroupForest::hook(uf + base[u] + hi[u], lo[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnionBatch; ++u) {
                if (!act[u] || found[u] == UfWord<uint32_t>::root) continue;
                // hi already had a parent: join the two trees by their roots.
                uint32_t pl, ph;
                uint32_t* const f = uf + base[u];
                const uint32_t rl = uf_find<GroupForest>(f, lo[u] >> 1, pl);
                const uint32_t rh = uf_find<GroupForest>(f, found[u] >> 1, ph);
                uf_link<GroupForest>(f, rl, pl, rh, ph ^ (found[u] & 1u), lo[u] & 1u);
            }
        }
        __syncthreads();
        // ORs a wave's ballot of per-pair bits into an LDS bit array at each
        // node's pair positions: one 64-pair chunk of one node (S >= 64), or
        // 64 / S whole nodes, each by the lane holding its first pair.
        auto or_ballot = [&](uint64_t* bits, uint64_t mask, uint32_t pi, uint32_t qq) {
            uint64_t seg;
            uint32_t pos;
            if (S >= 64) {
                if (lane != 0) return;
                seg = mask;
                pos = noff[pi] + qq;
            } else {
                if (qq != 0 || pi >= nodes) return;
                seg = (mask >> lane) & ((uint64_t{1} << S) - 1);
                pos = noff[pi];
            }
            if (!seg) return;
            const uint32_t w = pos >> 6, sh = pos & 63;
            atomicOr(reinterpret_cast<unsigned long long*>(bits + w), seg << sh);
            if (sh && (seg >> (64 - sh))) {
                atomicOr(reinterpret_cast<unsigned long long*>(bits + w + 1), seg >> (64 - sh));
            }
        };
        // Input switch values: each pair's parity to its root.
        for (uint32_t gb = wave * 64; gb < total; gb += nthreads) {
            const uint32_t g = gb + lane;
            const uint32_t pi = g >> s_log2, qq = g & (S - 1);
            uint32_t sv = 0;
            if (g < total) {
                const uint32_t c = sub_count(n0, k, pi);
                if (c > L && qq < (c >> 1)) {
                    uf_find<GroupForest>(uf + noff[pi], input_pair_slot<uint32_t>(qq, c), sv);
                }
            }
            or_ballot(sbits, __ballot(sv != 0), pi, qq);
        }
        __syncthreads();
        // Output switch values t_q = colour(I[2q]), colour(x) = s_(x>>1) ^
        // (x & 1) and 0 for the odd tail; then the children of P.
        for (uint32_t gb = wave * 64; gb < total; gb += nthreads) {
            const uint32_t g = gb + lane;
            const uint32_t pi = g >> s_log2, qq = g & (S - 1);
            uint32_t t = 0;
            if (g < total) {
                const uint32_t c = sub_count(n0, k, pi);
                if (c > L && qq < (c >> 1)) {
                    const uint32_t off = noff[pi];
                    const uint32_t a0 = next[off + 2 * qq];
                    if (!((c & 1) && a0 == c - 1)) {
                        const uint32_t pos = off + (a0 >> 1);
                        t = static_cast<uint32_t>((sbits[pos >> 6] >> (pos & 63)) & 1) ^ (a0 & 1u);
                    }
                }
            }
            or_ballot(tbits, __ballot(t != 0), pi, qq);
        }
        __syncthreads();
        // The level's rows, for k_subtree_emit.
        uint64_t* const out = lbits + (static_cast<uint64_t>(blockIdx.x) * levels + k) * 2 * kRowWords;
        for (uint32_t i = tid; i < kRowWords; i += nthreads) {
            out[i] = sbits[i];
            out[kRowWords + i] = tbits[i];
        }
        for (uint32_t g = tid; g < total; g += nthreads) {
            const uint32_t pi = g >> s_log2, p = g & (S - 1);
            const uint32_t c = sub_count(n0, k, pi);
            if (c <= L) continue;
            const uint32_t off = noff[pi], h = c >> 1, lc = h + (c & 1);
            if (p < h) {
                const uint32_t pos = off + p;
                const bool sv = ((sbits[pos >> 6] >> (pos & 63)) & 1) != 0;
                const uint32_t a0 = perm[off + 2 * p], a1 = perm[off + 2 * p + 1];
                next[off + p] = static_cast<uint16_t>((sv ? a1 : a0) >> 1);
                next[off + lc + p] = static_cast<uint16_t>((sv ? a0 : a1) >> 1);
            } else if ((c & 1) && p == h) {
                next[off + h] = static_cast<uint16_t>(perm[off + c - 1] >> 1);
            }
        }
        __syncthreads();
        cur ^= 1;
    }
}

// Packs the switch bits k_subtree left in `lbits`, level by level, with the
// subtree's node offsets recomputed as there.  One workgroup per subtree, of
// one wave per 2^6 entries of subtree capacity (2^12: 64 threads, 2^13: 128,
// 2^14: 256): measured at 2^26, larger workgroups for the smaller subtrees
// left lanes idle (2^12 x 256 threads: 0.89 ms per compress, x 128: 0.56,
// x 64: 0.40) and smaller ones lengthened the larger subtrees' packing.
__host__ __device__ constexpr uint32_t emit_threads(uint32_t sub_log2) {
    return sub_log2 <= 12 ? 64u : (sub_log2 == 13 ? 128u : 256u);
}

template <uint32_t SubLog2>
__global__ __launch_bounds__(emit_threads(SubLog2)) void k_subtree_emit(const ItemCtx* __restrict__ items,
                                                               const Seg* __restrict__ segs,
                                                               uint32_t nsegs,
                                                               const uint64_t* __restrict__ lbits,
                                                               uint32_t levels,
                                                               uint64_t* __restrict__ blob_all,
                                                               const int* __restrict__ errs) {
    const Seg& sg = find_seg(segs, nsegs);
    if (errs[sg.item]) return;
    constexpr uint32_t kSub = 1u << SubLog2;
    constexpr uint32_t kRowWords = subtree_row_words(SubLog2);
    __shared__ uint64_t sbits[kRowWords];
    __shared__ uint64_t tbits[kRowWords];
    __shared__ uint16_t noff[kSub / 2 + 1];
    __shared__ SubtreeRoot root;
    __shared__ SideAddr first_addr[2][kSubLog2Max + 1];
    const ItemCtx it = items[sg.item];
    const Plan& P = *it.plan;
    uint64_t* blob = blob_all + it.blob;
    const uint32_t L = P.leaf;
    const uint32_t d0 = it.df;
    const uint64_t b0 = blockIdx.x - sg.block0;
    const uint32_t n0 = static_cast<uint32_t>(node_count(P.n, d0, b0));
    const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
    // Walks to the subtree root, and the address of the first node of every
    // internal level on both sides, computed by different lanes at once:
    // each is a chain of dependent table reads, and the emission of a level
    // would otherwise wait for one.
    // (Lanes of a second wave where there is one, else lanes 2 and up.)
    constexpr uint32_t kFirstAddrLane = emit_threads(SubLog2) >= 128 ? 64u : 2u;
    static_assert(kFirstAddrLane + 2 * (kSubLog2Max + 1) <= emit_threads(SubLog2),
                  "every first address needs a lane");
    // The root walks (lanes 0 and 1) and the first-node addresses take one
    // code path, so that lanes of one wave walk together rather than in turn.
    if (tid < 2 || (tid >= kFirstAddrLane && tid < kFirstAddrLane + 2 * (kSubLog2Max + 1))) {
        const bool is_root = tid < 2;
        const uint32_t t = is_root ? 0u : tid - kFirstAddrLane;
        const uint32_t side = is_root ? tid : (t & 1), k = is_root ? 0u : (t >> 1);
        if (is_root || ((n0 + (1u << k) - 1) >> k) > L) {
            const uint32_t d = d0 + k;
            const uint64_t path = b0 << k;
            const Walk w = walk_to(P, static_cast<int>(side), d, path);
            if (is_root) {
                root.walk[side] = w;
            } else {
                first_addr[side][k] = address_at(P, w, d, path);
            }
        }
    }
    for (uint32_t k = 0;; ++k) {
        const uint32_t nodes = 1u << k;
        const uint32_t q = n0 >> k, m = n0 & (nodes - 1);
        const uint32_t cmax = q + (m != 0 ? 1u : 0u);
        if (cmax <= L) break;
        for (uint32_t pi = tid; pi < nodes; pi += nthreads) {
            noff[pi] = static_cast<uint16_t>(pi * q + large_before32(pi, m, k));
        }
        const uint64_t* const in =
            lbits + (static_cast<uint64_t>(blockIdx.x) * levels + k) * 2 * kRowWords;
        for (uint32_t i = tid; i < kRowWords; i += nthreads) {
            sbits[i] = in[i];
            tbits[i] = in[kRowWords + i];
        }
        __syncthreads();
        emit_level_runs(P, root, first_addr, d0, b0, n0, k, sbits, tbits, noff, blob);
        __syncthreads();
    }
}

}  // namespace benes_hip
