// SPDX-License-Identifier: Apache-2.0
// gfx950 kernels of the compact-v1 compression path.
//
// Level array layout (HBM): at every depth d the nodes tile [0, n) in path
// order and each node holds its node-local permutation (values < its count);
// a node's two children occupy exactly its range (left = first ceil(c/2)).
// Levels ping-pong between two u32 arrays, so routing is a pure streaming
// pass; only the union-find touches memory at random.
//
// Per "global" level (nodes larger than one workgroup's LDS):
//   k_describe  one thread per node: offsets and both sides' SwitchBitAddress
//   k_union     one thread per entry: parity union-find in HBM
//   k_resolve   one thread per pair: canonical switch values, child routing,
//               switch values ballot-packed into node-major bit rows
//   k_pack      one wave per (cluster group, packet): gathers a packet's run
//               of switch bits from the rows and writes whole 64-bit words
// Below the fusion depth:
//   k_subtree   one workgroup per subtree: every remaining level in LDS, then
//               the factoradic ranks of its leaves.
#pragma once

#include <hip/hip_runtime.h>

#include "format.hpp"
#include "uf.hpp"

namespace benes_hip {

constexpr uint32_t kBlock = 256;          // threads per block of level kernels
constexpr uint32_t kSubLog2 = 14;          // largest subtree handled in LDS
constexpr uint32_t kSub = 1u << kSubLog2;
constexpr uint32_t kSubThreads = 1024;

struct NodeDesc {
    uint64_t off;
    uint64_t cnt;
    uint64_t row;  // first word of the node's bit rows
    SideAddr a[2];
};

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64)

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }

// ORs one bit per lane into the blob.  Addresses must be strictly increasing
// over the valid lanes, which then form a prefix of the wave.  Lanes hitting
// the same 64-bit word are merged first, so each word takes one atomic.
__device__ __forceinline__ void wave_or_bits(uint64_t* blob, uint64_t addr, bool bit,
                                             bool valid) {
    const uint32_t lane = lane_id();
    const uint64_t key = valid ? (addr >> 6) : ~uint64_t{0};
    uint64_t val = (valid && bit) ? (uint64_t{1} << (addr & 63)) : 0;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint64_t ov = __shfl_down(val, off);
        const uint64_t ok = __shfl_down(key, off);
        if (lane + off < 64 && ok == key) val |= ov;
    }
    const uint64_t prev = __shfl_up(key, 1);
    const bool head = lane == 0 || prev != key;
    if (head && valid && val) atomicOr(reinterpret_cast<unsigned long long*>(blob + key),
                                       static_cast<unsigned long long>(val));
}

// Copies `len` bits, bit k given by get(k), to blob bits [dst0, dst0 + len).
// Words wholly inside the run are stored; the (at most two) edge words are
// OR'ed atomically because neighbouring runs share them.
template <typename Get>
__device__ __forceinline__ void wave_write_run(uint64_t* blob, uint64_t dst0, uint64_t len,
                                               Get get) {
    if (len == 0) return;
    const uint32_t lane = lane_id();
    const uint64_t w_first = dst0 >> 6, w_last = (dst0 + len - 1) >> 6;
    for (uint64_t w = w_first; w <= w_last; ++w) {
        const uint64_t o = (w << 6) + lane;
        const bool valid = o >= dst0 && o < dst0 + len;
        const bool b = valid && get(o - dst0);
        const uint64_t val = __ballot(b);
        const uint64_t vm = __ballot(valid);
        if (lane == 0) {
            if (vm == ~uint64_t{0}) {
                blob[w] = val;
            } else if (val) {
                atomicOr(reinterpret_cast<unsigned long long*>(blob + w),
                         static_cast<unsigned long long>(val));
            }
        }
    }
}

__device__ __forceinline__ uint32_t row_bit(const uint64_t* rows, uint64_t row, uint64_t pair) {
    return static_cast<uint32_t>((rows[row + (pair >> 6)] >> (pair & 63)) & 1);
}

// ---------------------------------------------------------------------------
// Input validation: every value in [0, n) exactly once (benes.cpp:152-173).

template <typename V>
__global__ void k_validate(const V* __restrict__ perm, uint64_t n, uint32_t* __restrict__ seen,
                           int* __restrict__ err) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const uint64_t v = static_cast<uint64_t>(perm[i]);
        bool bad = v >= n;
        if (!bad) {
            const uint32_t bit = 1u << (v & 31);
            bad = (atomicOr(seen + (v >> 5), bit) & bit) != 0;
        }
        if (bad) atomicOr(err, 1);
    }
}

// ---------------------------------------------------------------------------
// Global levels

__global__ void k_describe(const Plan* __restrict__ plan, uint32_t d, uint64_t row_words,
                           NodeDesc* __restrict__ desc, const int* __restrict__ err) {
    if (*err) return;
    const Plan& P = *plan;
    const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= (uint64_t{1} << d)) return;
    NodeDesc nd;
    nd.off = node_offset(P.n, d, j);
    nd.cnt = node_count(P.n, d, j);
    nd.row = j * row_words;
    nd.a[0] = node_address(P, 0, d, j);
    nd.a[1] = node_address(P, 1, d, j);
    desc[j] = nd;
}

template <typename V, typename W>
__global__ __launch_bounds__(kBlock) void k_union(const V* __restrict__ src, W* __restrict__ uf,
                                                  const NodeDesc* __restrict__ desc,
                                                  uint64_t chunks, const int* __restrict__ err) {
    if (*err) return;
    const uint64_t node = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    const uint64_t off = desc[node].off, cnt = desc[node].cnt;
    const uint64_t x = chunk * kBlock + threadIdx.x;
    if (x >= cnt) return;
    const uint64_t v = static_cast<uint64_t>(src[off + x]);
    if ((cnt & 1) && v == cnt - 1) return;  // edge into the unpaired output
    uf_union<AgentScope>(uf + off, input_slot<W>(x, cnt), output_pair_slot<W>(v >> 1, cnt),
                         static_cast<uint32_t>((x ^ v) & 1));
}

template <typename V, typename W>
__global__ __launch_bounds__(kBlock) void k_resolve(const V* __restrict__ src,
                                                    uint32_t* __restrict__ dst, W* __restrict__ uf,
                                                    const NodeDesc* __restrict__ desc,
                                                    uint64_t chunks, uint64_t* __restrict__ srows,
                                                    uint64_t* __restrict__ trows,
                                                    const int* __restrict__ err) {
    if (*err) return;
    const uint64_t node = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    const uint64_t off = desc[node].off, cnt = desc[node].cnt, row = desc[node].row;
    const uint64_t h = cnt >> 1, lc = h + (cnt & 1);
    const uint64_t p = chunk * kBlock + threadIdx.x;
    uint32_t s = 0, t = 0;
    if (p < h) {
        W* u = uf + off;
        uf_find<AgentScope>(u, input_pair_slot<W>(p, cnt), s);
        uf_find<AgentScope>(u, output_pair_slot<W>(p, cnt), t);
        const uint64_t a0 = static_cast<uint64_t>(src[off + 2 * p]);
        const uint64_t a1 = static_cast<uint64_t>(src[off + 2 * p + 1]);
        dst[off + p] = static_cast<uint32_t>((s ? a1 : a0) >> 1);
        dst[off + lc + p] = static_cast<uint32_t>((s ? a0 : a1) >> 1);
    } else if ((cnt & 1) && p == h) {
        dst[off + h] = static_cast<uint32_t>(static_cast<uint64_t>(src[off + cnt - 1]) >> 1);
    }
    const uint64_t ms = __ballot(s != 0), mt = __ballot(t != 0);
    const uint64_t first = p - lane_id();
    if (lane_id() == 0 && first < h) {
        srows[row + (first >> 6)] = ms;
        trows[row + (first >> 6)] = mt;
    }
}

// One wave per (group, packet) at depth d on one side.
__global__ __launch_bounds__(kBlock) void k_pack(const Plan* __restrict__ plan,
                                                 const NodeDesc* __restrict__ desc, int side,
                                                 uint32_t d, uint64_t packets,
                                                 const uint64_t* __restrict__ rows,
                                                 uint64_t* __restrict__ blob,
                                                 const int* __restrict__ err) {
    if (*err) return;
    const Plan& P = *plan;
    const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    const SideDepth sd = P.side[side][d];
    const uint32_t r = sd.relative_level, g = sd.group_depth;
    const uint32_t root_depth = d - r;
    const uint64_t groups = uint64_t{1} << root_depth;
    const uint64_t group = gw / packets, packet = gw % packets;
    if (group >= groups) return;
    const uint32_t span_log2 = g - r - 1;
    const uint64_t C = node_count(P.n, root_depth, group);
    const uint64_t first_pair = packet << span_log2;
    const uint64_t min_pairs = (C >> r) >> 1;
    const uint64_t b0 = group << r;
    const uint64_t secbase = (side == 0 ? P.input_offset : P.output_offset) * 8;
    if (min_pairs >= first_pair + (uint64_t{1} << span_log2)) {
        // Every node of the level fills its share of this packet: one run.
        const uint64_t start = bit_of(desc[b0].a[side], first_pair);
        const uint64_t len = uint64_t{1} << (r + span_log2);
        const uint64_t mask = (uint64_t{1} << span_log2) - 1;
        wave_write_run(blob, secbase + start, len, [&](uint64_t k) {
            const NodeDesc& nd = desc[b0 + (k >> span_log2)];
            return row_bit(rows, nd.row, first_pair + (k & mask)) != 0;
        });
    } else {
        // The last packet: node shares differ, place each from its own address.
        for (uint64_t pi = 0; pi < (uint64_t{1} << r); ++pi) {
            const NodeDesc& nd = desc[b0 + pi];
            const uint64_t len = packet_share(nd.cnt, packet, span_log2, P.leaf);
            if (len == 0) continue;
            const uint64_t start = bit_of(nd.a[side], first_pair);
            wave_write_run(blob, secbase + start, len, [&](uint64_t k) {
                return row_bit(rows, nd.row, first_pair + k) != 0;
            });
        }
    }
}

// ---------------------------------------------------------------------------
// Subtrees in LDS

// Node of local entry i at relative level k below a subtree root of n0
// entries.  Returns false when an ancestor (or the node) is a leaf.
__device__ __forceinline__ bool locate(uint64_t n0, uint32_t k, uint32_t leaf, uint64_t i,
                                       uint64_t& pi, uint64_t& off, uint64_t& c) {
    c = n0;
    off = 0;
    pi = 0;
    for (uint32_t l = 0; l < k; ++l) {
        if (c <= leaf) return false;
        const uint64_t lc = (c + 1) >> 1;
        if (i - off < lc) {
            c = lc;
            pi = pi << 1;
        } else {
            off += lc;
            c >>= 1;
            pi = (pi << 1) | 1;
        }
    }
    return c > leaf;
}

// Factoradic rank (internal.hpp:148-167) of one leaf, OR'ed into the blob.
// Lehmer digit i counts earlier entries smaller than entry i, read off a
// bitmask of the values seen so far; Horner with radix i+1, i ascending.
__device__ __forceinline__ void rank_leaf(const Plan& P, uint32_t d, uint64_t path,
                                          const uint16_t* vals, uint32_t c, uint64_t* blob) {
    uint32_t seen = 0;
    unsigned __int128 r = 0;
    for (uint32_t i = 0; i < c; ++i) {
        const uint32_t v = vals[i];
        const uint32_t digit = __popc(seen & ((1u << v) - 1u));
        seen |= 1u << v;
        r = r * (i + 1) + digit;
    }
    if (rank_bits(c) == 0) return;
    const uint64_t pos = leaf_section_offset(P) * 8 + leaf_rank_bit(P, d, path);
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

// Writes the switch bits of every internal node at relative level k of a
// subtree (absolute depth d0 + k) into the blob.  Bits are produced packet by
// packet: within a cluster group, one packet's share of a level is the
// concatenation of that packet's pair range from each node in path order
// (internal.hpp:604-649), so a run of consecutive nodes is one contiguous
// stretch of the blob, written as whole words.  `sbits`/`tbits` hold the
// level's switch values at the position of each node's pair in LDS.
__device__ __forceinline__ void emit_level_runs(const Plan& P, uint32_t d0, uint64_t b0,
                                                uint64_t n0, uint32_t k,
                                                const uint64_t* sbits, const uint64_t* tbits,
                                                const uint16_t* noff, uint64_t* blob) {
    const uint32_t d = d0 + k;
    const uint32_t L = P.leaf;
    const uint32_t wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const uint64_t cmax = (n0 + ((uint64_t{1} << k) - 1)) >> k;
    const uint64_t maxpairs = cmax >> 1;
    for (int side = 0; side < 2; ++side) {
        const SideDepth& sd = P.side[side][d];
        const uint32_t r = sd.relative_level, g = sd.group_depth;
        const uint32_t R = d - r;
        const uint32_t span_log2 = g - r - 1;
        const uint64_t span = uint64_t{1} << span_log2;
        const uint64_t* rows = side == 0 ? sbits : tbits;
        const uint64_t secbase = (side == 0 ? P.input_offset : P.output_offset) * 8;
        // Sets of consecutive nodes that share one cluster group.
        const bool one_set = R <= d0;
        const uint64_t nsets = one_set ? 1 : (uint64_t{1} << (R - d0));
        const uint64_t set_nodes = one_set ? (uint64_t{1} << k) : (uint64_t{1} << r);
        const uint64_t packets = (maxpairs + span - 1) >> span_log2;
        for (uint64_t job = wave; job < nsets * packets; job += nwaves) {
            const uint64_t set = job / packets, pk = job % packets;
            const uint64_t first = set * set_nodes;
            const uint64_t minc = one_set ? (n0 >> k) : (node_count(n0, R - d0, set) >> r);
            const uint64_t first_pair = pk << span_log2;
            if (minc > L && (minc >> 1) >= first_pair + span) {
                // Every node fills its share of this packet: one run.
                const SideAddr a = node_address(P, side, d, (b0 << k) | first);
                wave_write_run(blob, secbase + bit_of(a, first_pair), set_nodes << span_log2,
                               [&](uint64_t q) {
                                   const uint64_t pos = noff[first + (q >> span_log2)] +
                                                        first_pair + (q & (span - 1));
                                   return ((rows[pos >> 6] >> (pos & 63)) & 1) != 0;
                               });
            } else {
                for (uint64_t pi = first; pi < first + set_nodes; ++pi) {
                    const uint64_t c = node_count(n0, k, pi);
                    if (c <= L) continue;
                    const uint64_t len = packet_share(c, pk, span_log2, L);
                    if (len == 0) continue;
                    const SideAddr a = node_address(P, side, d, (b0 << k) | pi);
                    const uint64_t base = noff[pi] + first_pair;
                    wave_write_run(blob, secbase + bit_of(a, first_pair), len, [&](uint64_t q) {
                        const uint64_t pos = base + q;
                        return ((rows[pos >> 6] >> (pos & 63)) & 1) != 0;
                    });
                }
            }
        }
    }
}

template <typename V>
__global__ __launch_bounds__(kSubThreads) void k_subtree(const V* __restrict__ src,
                                                         const Plan* __restrict__ plan,
                                                         uint32_t d0, uint64_t* __restrict__ blob,
                                                         const int* __restrict__ err) {
    if (*err) return;
    __shared__ uint16_t buf[2][kSub];
    __shared__ uint32_t uf[kSub];
    __shared__ uint64_t sbits[kSub / 64];
    __shared__ uint64_t tbits[kSub / 64];
    __shared__ uint16_t noff[kSub / 2];
    const Plan& P = *plan;
    const uint32_t L = P.leaf;
    const uint64_t b0 = blockIdx.x;
    const uint64_t n0 = node_count(P.n, d0, b0);
    const uint64_t base = node_offset(P.n, d0, b0);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nthreads = blockDim.x, nwaves = blockDim.x >> 6;
    for (uint64_t i = tid; i < n0; i += nthreads) buf[0][i] = static_cast<uint16_t>(src[base + i]);
    __syncthreads();
    int cur = 0;
    for (uint32_t k = 0;; ++k) {
        const uint32_t d = d0 + k;
        uint16_t* perm = buf[cur];
        uint16_t* next = buf[cur ^ 1];
        const uint64_t cmax = (n0 + ((uint64_t{1} << k) - 1)) >> k;
        const uint64_t cmin = n0 >> k;
        const uint64_t nodes = uint64_t{1} << k;
        if (cmin <= L) {
            for (uint64_t pi = tid; pi < nodes; pi += nthreads) {
                const uint64_t c = node_count(n0, k, pi);
                if (c > L) continue;
                if (k > 0 && node_count(n0, k - 1, pi >> 1) <= L) continue;
                const uint64_t off = node_offset(n0, k, pi);
                rank_leaf(P, d, (b0 << k) | pi, perm + off, static_cast<uint32_t>(c), blob);
            }
        }
        if (cmax <= L) break;
        for (uint64_t i = tid; i < n0; i += nthreads) uf[i] = UfWord<uint32_t>::root;
        __syncthreads();
        for (uint64_t i = tid; i < n0; i += nthreads) {
            uint64_t pi, off, c;
            if (!locate(n0, k, L, i, pi, off, c)) continue;
            const uint64_t x = i - off, v = perm[i];
            if ((c & 1) && v == c - 1) continue;
            uf_union<GroupScope>(uf + off, input_slot<uint32_t>(x, c),
                                 output_pair_slot<uint32_t>(v >> 1, c),
                                 static_cast<uint32_t>((x ^ v) & 1));
        }
        __syncthreads();
        for (uint64_t wbase = static_cast<uint64_t>(wave) * 64; wbase < n0; wbase += nwaves * 64) {
            const uint64_t i = wbase + lane;
            uint32_t s = 0, t = 0;
            uint64_t pi, off, c;
            if (i < n0 && locate(n0, k, L, i, pi, off, c)) {
                const uint64_t p = i - off, h = c >> 1;
                if (p < h) {
                    uint32_t* u = uf + off;
                    uf_find<GroupScope>(u, input_pair_slot<uint32_t>(p, c), s);
                    uf_find<GroupScope>(u, output_pair_slot<uint32_t>(p, c), t);
                    const uint32_t a0 = perm[off + 2 * p], a1 = perm[off + 2 * p + 1];
                    const uint64_t lc = h + (c & 1);
                    next[off + p] = static_cast<uint16_t>((s ? a1 : a0) >> 1);
                    next[off + lc + p] = static_cast<uint16_t>((s ? a0 : a1) >> 1);
                } else if ((c & 1) && p == h) {
                    next[off + h] = static_cast<uint16_t>(perm[off + c - 1] >> 1);
                }
            }
            const uint64_t ms = __ballot(s != 0), mt = __ballot(t != 0);
            if (lane == 0) {
                sbits[wbase >> 6] = ms;
                tbits[wbase >> 6] = mt;
            }
        }
        for (uint64_t pi = tid; pi < nodes; pi += nthreads) {
            noff[pi] = static_cast<uint16_t>(node_offset(n0, k, pi));
        }
        __syncthreads();
        emit_level_runs(P, d0, b0, n0, k, sbits, tbits, noff, blob);
        __syncthreads();
        cur ^= 1;
    }
}

}  // namespace benes_hip
