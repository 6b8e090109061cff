// SPDX-License-Identifier: Apache-2.0
// Parity union-find over the input slots of one routing node.
//
// Slot s is input pair s (or, in an odd node, slot 0 is the unpaired tail
// input and pair p is slot p + 1).  The two entries I[2q], I[2q + 1] reaching
// output pair q must take different colours, which is an edge between their
// slots labelled 1 ^ (I[2q] & 1) ^ (I[2q + 1] & 1): the XOR of the two slots'
// switch values (a slot's colour is that of its even entry).  A parent word
// holds (parent slot << 1) | parity-to-parent; the all-ones word marks a
// root.  A slot is only ever hooked under a smaller slot (a root under a
// smaller root, or under any smaller vertex in the Rem walk), so parents are
// below children and every tree is rooted at its minimum slot, the tail being
// slot 0: each component's root is the vertex the reference colours 0 first
// (benes.cpp:278-297: the odd tail, else the component's smallest input), and
// a slot's parity to its root is exactly its canonical switch value.
//
// Why cached loads and plain stores are enough.  Every word ever stored is a
// true parity relation from a slot to a strictly smaller slot: a hook links a
// root (or, in the first pass over the edges, a slot no other thread writes)
// to a smaller slot, and path halving links a non-root to one of its
// ancestors.  A slot that has a parent never becomes a root again, so a load
// that returns an out-of-date word still returns a true relation to a smaller
// slot.  Therefore
//   * walks always terminate (indices strictly decrease) and their parities are
//     right, whatever mix of fresh and stale words they read;
//   * only the hook must be atomic: a compare-and-swap that expects the
//     all-ones word, at device scope, so exactly one of any competing hooks of
//     a root succeeds;
//   * a failed hook returns the root's true parent, from which the walk
//     resumes, so every retry moves to a strictly smaller slot;
//   * halving writes only to slots that are no longer roots, which no hook
//     touches, so a plain store is safe; losing one to a concurrent halving
//     only leaves a longer (still true) path.
// Two endpoints that reach the same apparent root are truly connected, since
// the relations they followed are facts.  After the last union the forest is
// static apart from halving, so every later walk reaches the true root.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace benes_hip {

template <typename W>
struct UfWord;
// 16-bit words hold slots below 2^15: a parent is at most the largest slot
// minus one, so no parent word reaches the all-ones root mark.
template <>
struct UfWord<uint16_t> {
    static constexpr uint16_t root = 0xFFFFu;
};
template <>
struct UfWord<uint32_t> {
    static constexpr uint32_t root = 0xFFFFFFFFu;
};
template <>
struct UfWord<uint64_t> {
    static constexpr uint64_t root = ~uint64_t{0};
};

// Parent arrays in HBM, shared by workgroups on every XCD.
struct DeviceForest {
    template <typename W>
    static __device__ __forceinline__ W load(W* p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <typename W>
    static __device__ __forceinline__ void store(W* p, W v) {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    // Returns the word found (the all-ones word when the hook succeeded).
    template <typename W>
    static __device__ __forceinline__ W hook(W* p, W desired) {
        W expected = UfWord<W>::root;
        __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return expected;
    }
};

// The same arrays read and written coherently across XCDs: loads and stores
// at agent scope bypass the XCD's own L2 copy (the L2s of different XCDs are
// not coherent within a kernel).  Used where halving walks of several XCDs
// cross the same long paths, which only pay off if each XCD sees the others'
// shortcuts.
struct CoherentForest {
    template <typename W>
    static __device__ __forceinline__ W load(W* p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <typename W>
    static __device__ __forceinline__ void store(W* p, W v) {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    template <typename W>
    static __device__ __forceinline__ W hook(W* p, W desired) {
        return DeviceForest::hook(p, desired);
    }
};

// Parent arrays in LDS, private to one workgroup.
struct GroupForest {
    template <typename W>
    static __device__ __forceinline__ W load(W* p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <typename W>
    static __device__ __forceinline__ void store(W* p, W v) {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    template <typename W>
    static __device__ __forceinline__ W hook(W* p, W desired) {
        W expected = UfWord<W>::root;
        __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return expected;
    }
    // A 16-bit word shares an aligned 32-bit LDS word with its neighbour:
    // the hook is a 32-bit compare-and-swap, retried while only the
    // neighbour's half changed (a halving store or a hook of the neighbour).
    static __device__ __forceinline__ uint16_t hook(uint16_t* p, uint16_t desired) {
        const uint32_t hi_half = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(p) >> 1) & 1u;
        uint32_t* const w = reinterpret_cast<uint32_t*>(p - hi_half);
        const uint32_t sh = hi_half * 16;
        uint32_t cur = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        while (true) {
            const uint16_t mine = static_cast<uint16_t>(cur >> sh);
            if (mine != UfWord<uint16_t>::root) return mine;
            const uint32_t want = (cur & ~(0xFFFFu << sh)) | (static_cast<uint32_t>(desired) << sh);
            uint32_t expected = cur;
            if (__hip_atomic_compare_exchange_strong(w, &expected, want, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP)) {
                return UfWord<uint16_t>::root;
            }
            cur = expected;
        }
    }
};

// Root of x and the parity of x relative to it (path halving on the way).
template <typename F, typename W>
__device__ __forceinline__ W uf_find(W* uf, W x, uint32_t& parity) {
    constexpr W kRoot = UfWord<W>::root;
    uint32_t par = 0;
    while (true) {
        const W w = F::load(uf + x);
        if (w == kRoot) break;
        const W p = w >> 1;
        const W w2 = F::load(uf + p);
        if (w2 == kRoot) {
            par ^= static_cast<uint32_t>(w & 1);
            x = p;
            break;
        }
        const W hop = static_cast<W>((w ^ w2) & 1);
        F::store(uf + x, static_cast<W>((w2 & ~W{1}) | hop));
        par ^= static_cast<uint32_t>(hop);
        x = w2 >> 1;
    }
    parity = par;
    return x;
}

// uf_find that reads at most 2 * max_steps parent words: false when it stops
// short of the root (the halving it did stays; x is then partway up).
template <typename F, typename W>
__device__ __forceinline__ bool uf_find_bounded(W* uf, W& x, uint32_t& parity,
                                                uint32_t max_steps) {
    constexpr W kRoot = UfWord<W>::root;
    uint32_t par = 0;
    for (uint32_t step = 0; step < max_steps; ++step) {
        const W w = F::load(uf + x);
        if (w == kRoot) {
            parity = par;
            return true;
        }
        const W p = w >> 1;
        const W w2 = F::load(uf + p);
        if (w2 == kRoot) {
            parity = par ^ static_cast<uint32_t>(w & 1);
            x = p;
            return true;
        }
        const W hop = static_cast<W>((w ^ w2) & 1);
        F::store(uf + x, static_cast<W>((w2 & ~W{1}) | hop));
        par ^= static_cast<uint32_t>(hop);
        x = w2 >> 1;
    }
    return false;
}

// Records value(u) ^ value(v) == label, given u's root (ru, parity pu) and
// v's root (rv, parity pv).
template <typename F, typename W>
__device__ __forceinline__ void uf_link(W* uf, W ru, uint32_t pu, W rv, uint32_t pv,
                                        uint32_t label) {
    constexpr W kRoot = UfWord<W>::root;
    while (ru != rv) {
        const bool u_high = ru > rv;
        const W hi = u_high ? ru : rv, lo = u_high ? rv : ru;
        const W rel = static_cast<W>((pu ^ pv ^ label) & 1);
        const W found = F::hook(uf + hi, static_cast<W>((lo << 1) | rel));
        if (found == kRoot) return;
        // hi already has a parent: continue from it.
        uint32_t q;
        const W r = uf_find<F>(uf, static_cast<W>(found >> 1), q);
        q ^= static_cast<uint32_t>(found & 1);
        if (u_high) {
            ru = r;
            pu ^= q;
        } else {
            rv = r;
            pv ^= q;
        }
    }
}

// Union of two input slots.  The larger slot is first hooked straight under
// the smaller one: while it is still a root that one CAS settles the edge
// (a root may hang under any smaller slot; the merged tree keeps its
// smallest slot as root), and neither endpoint's tree is walked.  Only when
// it already has a parent are both roots found and linked.
template <typename F, typename W>
__device__ __forceinline__ void uf_union_slots(W* uf, W a, W b, uint32_t label) {
    const W hi = a > b ? a : b, lo = a > b ? b : a;
    const W found = F::hook(uf + hi, static_cast<W>((lo << 1) | (label & 1)));
    if (found == UfWord<W>::root) return;
    uint32_t pl, ph;
    const W rl = uf_find<F>(uf, lo, pl);
    const W rh = uf_find<F>(uf, static_cast<W>(found >> 1), ph);
    uf_link<F>(uf, rl, pl, rh, ph ^ static_cast<uint32_t>(found & 1), label);
}

// uf_union_slots whose two root finds give up after max_steps halving steps
// each: false then, with nothing linked (the edge is to be joined later).
template <typename F, typename W>
__device__ __forceinline__ bool uf_union_slots_bounded(W* uf, W a, W b, uint32_t label,
                                                       uint32_t max_steps) {
    const W hi = a > b ? a : b, lo = a > b ? b : a;
    const W found = F::hook(uf + hi, static_cast<W>((lo << 1) | (label & 1)));
    if (found == UfWord<W>::root) return true;
    uint32_t pl, ph;
    W rl = lo, rh = static_cast<W>(found >> 1);
    if (!uf_find_bounded<F>(uf, rl, pl, max_steps) || !uf_find_bounded<F>(uf, rh, ph, max_steps)) {
        return false;
    }
    uf_link<F>(uf, rl, pl, rh, ph ^ static_cast<uint32_t>(found & 1), label);
    return true;
}

// Union in the manner of Rem's algorithm: records value(x0) ^ value(y0) ==
// label for two slots, starting from vertices x and y with parities px, py
// to them.  The walk always climbs from the larger of the two current
// vertices; it stops when they meet (already connected) or when the larger
// one is a root, which is then hooked under the smaller current vertex (not
// necessarily a root: any smaller vertex keeps parents below children, so
// each tree stays rooted at its minimum slot).  Neither walk has to reach the
// root of a large tree, so the words near big roots are read far less.  The
// walk writes nothing but the hook: halving here costs more stores than the
// later finds save.
//
// At most `max_steps` parent words are read; the walk returns false if it
// stops there, having written nothing, and true once the relation is recorded.
template <typename F, typename W>
__device__ __forceinline__ bool uf_union_rem(W* uf, W x, uint32_t px, W y, uint32_t py,
                                             uint32_t label, uint64_t max_steps) {
    constexpr W kRoot = UfWord<W>::root;
    for (uint64_t step = 0; x != y; ++step) {
        if (step == max_steps) return false;
        if (x < y) {
            const W t = x;
            x = y;
            y = t;
            const uint32_t tp = px;
            px = py;
            py = tp;
        }
        W w = F::load(uf + x);
        if (w == kRoot) {
            const W found = F::hook(uf + x, (y << 1) | static_cast<W>((px ^ py ^ label) & 1));
            if (found == kRoot) return true;
            w = found;
        }
        px ^= static_cast<uint32_t>(w & 1);
        x = w >> 1;
    }
    return true;
}

// uf_union_rem whose hook is a plain store instead of a compare-and-swap,
// for a union pass that checks its hooks afterwards.  Two walks that find
// the same root at once may both store to it, and the later store wins; the
// earlier relation is then missing from the forest.  Every stored word is
// still a true relation to a smaller slot, so the forest stays valid, and
// the caller keeps the hook it made (`hooked` slot, `word` stored) so that a
// later pass can find the ones that were overwritten and join them again.
// Returns 1 after a hook, 0 when the two vertices met (already joined) and
// -1 when it stopped after `max_steps` parent words without storing.
template <typename F, typename W>
__device__ __forceinline__ int uf_union_rem_plain(W* uf, W x, uint32_t px, W y, uint32_t py,
                                                  uint32_t label, uint32_t max_steps, W& hooked,
                                                  W& word) {
    constexpr W kRoot = UfWord<W>::root;
    for (uint32_t step = 0; x != y; ++step) {
        if (step == max_steps) return -1;
        if (x < y) {
            const W t = x;
            x = y;
            y = t;
            const uint32_t tp = px;
            px = py;
            py = tp;
        }
        const W w = F::load(uf + x);
        if (w == kRoot) {
            word = (y << 1) | static_cast<W>((px ^ py ^ label) & 1);
            F::store(uf + x, word);
            hooked = x;
            return 1;
        }
        px ^= static_cast<uint32_t>(w & 1);
        x = w >> 1;
    }
    return 0;
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

}  // namespace benes_hip
