// SPDX-License-Identifier: Apache-2.0
// Parity union-find over the pair graph of one routing node.
//
// Vertices ("slots") are the node's input pairs and output pairs; entry x of
// a node with permutation P is an edge between input pair x>>1 and output pair
// P[x]>>1 whose label (x ^ P[x]) & 1 is the XOR of the two pairs' switch
// values.  A parent word holds (parent slot << 1) | parity-to-parent; the
// all-ones word marks a root.  Hooking always puts the larger root under the
// smaller one, so every tree is rooted at its minimum slot.  Input-pair slots
// are numbered below output-pair slots and the unpaired odd tail input gets
// slot 0, so each component's root is the vertex the reference colours 0 first
// (benes.cpp:278-297: the odd tail, else the component's smallest input), and a
// slot's parity to its root is exactly its canonical switch value.
//
// Path halving keeps every stored word a true parity relation to an ancestor:
// x -> gp with parity p(x->p) ^ p(p->gp), published by compare-and-swap so a
// concurrent hook or halving of x is never overwritten.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace benes_hip {

template <typename W>
struct UfWord;
template <>
struct UfWord<uint32_t> {
    static constexpr uint32_t root = 0xFFFFFFFFu;
};
template <>
struct UfWord<uint64_t> {
    static constexpr uint64_t root = ~uint64_t{0};
};

// Memory-scope policies: device-wide parent arrays (global memory shared by
// many workgroups, possibly on different XCDs) and workgroup-local ones (LDS).
struct AgentScope {
    template <typename W>
    static __device__ __forceinline__ W load(W* p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <typename W>
    static __device__ __forceinline__ bool cas(W* p, W expected, W desired) {
        return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

struct GroupScope {
    template <typename W>
    static __device__ __forceinline__ W load(W* p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <typename W>
    static __device__ __forceinline__ bool cas(W* p, W expected, W desired) {
        return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
};

// Root of x and the parity of x relative to it.
template <typename Scope, typename W>
__device__ __forceinline__ W uf_find(W* uf, W x, uint32_t& parity) {
    constexpr W kRoot = UfWord<W>::root;
    uint32_t par = 0;
    while (true) {
        const W w = Scope::load(uf + x);
        if (w == kRoot) break;
        const W p = w >> 1;
        const W w2 = Scope::load(uf + p);
        if (w2 == kRoot) {
            par ^= static_cast<uint32_t>(w & 1);
            x = p;
            break;
        }
        const W hop = (w ^ w2) & 1;
        Scope::cas(uf + x, w, (w2 & ~W{1}) | hop);
        par ^= static_cast<uint32_t>(hop);
        x = w2 >> 1;
    }
    parity = par;
    return x;
}

// Records value(u) ^ value(v) == label.
template <typename Scope, typename W>
__device__ __forceinline__ void uf_union(W* uf, W u, W v, uint32_t label) {
    constexpr W kRoot = UfWord<W>::root;
    while (true) {
        uint32_t pu, pv;
        const W ru = uf_find<Scope>(uf, u, pu);
        const W rv = uf_find<Scope>(uf, v, pv);
        if (ru == rv) return;
        const W rel = static_cast<W>((pu ^ pv ^ label) & 1);
        const W hi = ru > rv ? ru : rv;
        const W lo = ru > rv ? rv : ru;
        if (Scope::cas(uf + hi, kRoot, (lo << 1) | rel)) return;
    }
}

// Slot numbering inside a node of `count` entries (h = count/2 pairs).
template <typename W>
__device__ __forceinline__ W input_slot(uint64_t x, uint64_t count) {
    if (count & 1) return x == count - 1 ? W{0} : static_cast<W>((x >> 1) + 1);
    return static_cast<W>(x >> 1);
}

template <typename W>
__device__ __forceinline__ W input_pair_slot(uint64_t pair, uint64_t count) {
    return static_cast<W>(pair + (count & 1));
}

template <typename W>
__device__ __forceinline__ W output_pair_slot(uint64_t pair, uint64_t count) {
    return static_cast<W>((count >> 1) + (count & 1) + pair);
}

}  // namespace benes_hip
