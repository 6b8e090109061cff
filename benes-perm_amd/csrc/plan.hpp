// SPDX-License-Identifier: Apache-2.0
// Host-side construction of a Plan: format normalisation, layout and the
// per-depth metric tables the kernels read.
//
// Behaviour (and error classes/messages) follows the reference's
// normalize_format / make_layout / write_header (src/internal.hpp:315-454) and
// FormatSizeCache (src/internal.hpp:225-313); the metric recursion is
// evaluated bottom-up over depths instead of memoised per count.
#pragma once

#include "format.hpp"
#include "status.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

namespace benes_hip {

struct FormatArgs {
    uint32_t middle_group_log2 = 5;
    uint32_t placement = 0;
    std::vector<uint8_t> input_clusters;
    std::vector<uint8_t> output_clusters;
};

// Normalised format plus everything derived from it for one entry count.
struct Layout {
    Plan plan;
    std::vector<uint8_t> input_clusters;
    std::vector<uint8_t> output_clusters;
    uint64_t header_bytes = 0;
    uint64_t network_size = 0;
};

namespace detail {

struct Sat {
    uint64_t v = 0;
    bool overflow = false;
};

inline Sat sat_add(Sat a, Sat b) {
    Sat r;
    r.overflow = a.overflow || b.overflow || b.v > UINT64_MAX - a.v;
    r.v = r.overflow ? 0 : a.v + b.v;
    return r;
}

inline uint32_t routing_depth(uint64_t count, uint32_t g) {
    const uint64_t leaf = uint64_t{1} << g;
    uint32_t d = 0;
    while (count > leaf) {
        count = count / 2 + count % 2;
        ++d;
    }
    return d;
}

inline uint64_t bytes_of(uint64_t bits) { return bits / 8 + (bits % 8 != 0 ? 1 : 0); }

}  // namespace detail

// Builds the layout; returns false with `st` set on invalid formats or
// unrepresentable sizes.
inline bool build_layout(uint64_t n, const FormatArgs& fmt, Layout& out, Status& st) {
    using detail::Sat;
    using detail::sat_add;
    const uint32_t g = fmt.middle_group_log2;
    if (g == 0 || g > 5) {
        st.set(kInvalidArgument, "middle_group_log2 must be between 1 and 5");
        return false;
    }
    if (fmt.placement > 2) {
        st.set(kInvalidArgument, "invalid middle placement");
        return false;
    }
    const uint32_t D = detail::routing_depth(n, g);
    auto defaults = [&]() -> std::vector<uint8_t> {
        if (D == 0) return {};
        if (D == 1) return {1};
        return {static_cast<uint8_t>((D + 1) / 2), static_cast<uint8_t>(D / 2)};
    };
    out.input_clusters = fmt.input_clusters.empty() ? defaults() : fmt.input_clusters;
    out.output_clusters = fmt.output_clusters.empty() ? defaults() : fmt.output_clusters;
    for (const auto* list : {&out.input_clusters, &out.output_clusters}) {
        uint32_t sum = 0;
        for (uint8_t c : *list) {
            if (c == 0) {
                st.set(kInvalidArgument, "cluster depths must be positive");
                return false;
            }
            sum += c;
        }
        if (sum != D) {
            st.set(kInvalidArgument, "cluster depths do not span every routing level");
            return false;
        }
    }
    uint32_t placement = fmt.placement;
    if (D == 0) placement = 0;

    Plan& p = out.plan;
    std::memset(&p, 0, sizeof(p));
    p.n = n;
    p.leaf_log2 = g;
    p.leaf = 1u << g;
    p.depth = D;
    p.placement = placement;

    // Bottom-up metrics for the two counts present at each depth.  Below
    // depth D + 1 every count is at most a leaf (nodes at depth D hold at
    // most L entries, and q(D + 1) + 1 <= L / 2 + 1), so the recursion starts
    // there, and the switch bits of the top k levels of a depth-d subtree are
    // needed for k <= D - d only (every lookup below stays within the routing
    // depth, and so does the recursion from depth d + 1).  A table over all
    // kMaxDepth depths and levels cost about 20 us per call, most of a small
    // batch item's host time.
    constexpr int K = kMaxDepth;
    const uint64_t L = p.leaf;
    const int top = std::min(K, static_cast<int>(D) + 1);  // deepest depth recursed from
    const int KK = top;                                    // levels kept in the table
    std::array<Sat, (K + 1) * 2> S{}, M{};
    std::vector<Sat> PSB(static_cast<size_t>(top + 1) * 2 * (KK + 1));  // [d][x][k], zeroed
    auto q = [&](int d) { return d >= 64 ? uint64_t{0} : (n >> d); };
    auto idx = [&](int d, uint64_t c) { return static_cast<int>(c - q(d)); };
    auto psb = [&](int d, int x, int k) -> Sat& { return PSB[(d * 2 + x) * (KK + 1) + k]; };
    for (int d = K; d >= 0; --d) {
        for (int x = 0; x < 2; ++x) {
            const uint64_t c = q(d) + x;
            Sat& s = S[d * 2 + x];
            Sat& m = M[d * 2 + x];
            if (c <= L || d >= top) {
                s = Sat{};
                m = Sat{c <= 32 ? rank_bits(static_cast<uint32_t>(c)) : 0, c > L};
                continue;
            }
            const int a = idx(d + 1, c / 2 + c % 2), b = idx(d + 1, c / 2);
            s = sat_add(Sat{c / 2}, sat_add(S[(d + 1) * 2 + a], S[(d + 1) * 2 + b]));
            m = sat_add(M[(d + 1) * 2 + a], M[(d + 1) * 2 + b]);
            for (int k = 1; k <= KK && k <= static_cast<int>(D) - d; ++k) {
                psb(d, x, k) = sat_add(Sat{c / 2},
                                       sat_add(psb(d + 1, a, k - 1), psb(d + 1, b, k - 1)));
            }
        }
    }
    const Sat sw = S[0], mid = M[0];
    if (sw.overflow || mid.overflow) {
        st.set(kOverflow, "compressed permutation is too large");
        return false;
    }
    // Section bytes: internal.hpp:389-429.
    Sat header = sat_add(Sat{20}, Sat{out.input_clusters.size() + out.output_clusters.size()});
    Sat in_bytes = sat_add(Sat{detail::bytes_of(sw.v)},
                           Sat{placement == 1 ? detail::bytes_of(mid.v) : 0});
    Sat out_bytes = sat_add(Sat{detail::bytes_of(sw.v)},
                            Sat{placement == 2 ? detail::bytes_of(mid.v) : 0});
    Sat mid_bytes{placement == 0 ? detail::bytes_of(mid.v) : 0};
    if (n > (uint64_t{1} << 63)) {
        st.set(kOverflow, "logical network size is too large");
        return false;
    }
    Sat out_off = sat_add(header, in_bytes);
    Sat mid_off = sat_add(out_off, out_bytes);
    Sat end_off = sat_add(mid_off, mid_bytes);
    if (end_off.overflow) {
        st.set(kOverflow, "compressed permutation is too large");
        return false;
    }
    uint64_t ns = n > L ? n : L, p2 = 1;
    while (p2 < ns) p2 <<= 1;
    out.network_size = p2;
    out.header_bytes = header.v;
    p.input_offset = header.v;
    p.output_offset = out_off.v;
    p.middle_offset = mid_off.v;
    p.end_offset = end_off.v;
    p.switch_bits_total = sw.v;
    p.middle_bits_total = mid.v;
    for (int d = 0; d <= K; ++d) {
        for (int x = 0; x < 2; ++x) p.middle_bits[d][x] = M[d * 2 + x].v;
    }

    // Per-side group structure.
    for (int side = 0; side < 2; ++side) {
        const std::vector<uint8_t>& cl = side == 0 ? out.input_clusters : out.output_clusters;
        const bool embeds = (side == 0 && placement == 1) || (side == 1 && placement == 2);
        SideDepth* sd = p.side[side];
        // Group boundaries.
        std::vector<uint32_t> start(cl.size() + 1, 0);
        for (size_t k = 0; k < cl.size(); ++k) start[k + 1] = start[k] + cl[k];
        auto group_of = [&](uint32_t d) -> size_t {
            size_t k = 0;
            while (k + 1 < start.size() && start[k + 1] <= d) ++k;
            return k;  // may equal cl.size() when d == D
        };
        auto sec = [&](int d, int x) -> uint64_t {
            return S[d * 2 + x].v + (embeds ? M[d * 2 + x].v : 0);
        };
        for (uint32_t d = 0; d <= D && d <= static_cast<uint32_t>(K); ++d) {
            const size_t k = group_of(d);
            sd[d].group_root = (k < start.size() && start[k] == d) ? 1 : 0;
            sd[d].root_depth = static_cast<uint8_t>(k < start.size() ? start[k] : D);
            sd[d].group_depth = static_cast<uint8_t>(k < cl.size() ? cl[k] : 0);
            sd[d].relative_level = static_cast<uint8_t>(d - sd[d].root_depth);
            if (d >= 1) {
                const size_t kp = group_of(d - 1);
                const uint32_t end = kp < cl.size() ? start[kp + 1] : D;
                for (int x = 0; x < 2; ++x) {
                    sd[d].residual[x] = sec(d, x) - psb(d, x, static_cast<int>(end - d)).v;
                }
            }
            if (sd[d].group_root && k < cl.size()) {
                const uint32_t gk = cl[k];
                for (int x = 0; x < 2; ++x) {
                    const uint64_t C = q(d) + x;
                    sd[d].group_switch_bits[x] = psb(d, x, static_cast<int>(gk)).v;
                    uint64_t fpb = 0;
                    if (C > L) {
                        for (uint32_t l = 0; l < gk; ++l) {
                            fpb += level_prefix_bits(C, gk, l, 0, uint64_t{1} << l,
                                                     static_cast<uint32_t>(L));
                        }
                    }
                    sd[d].full_packet_bits[x] = fpb;
                }
            }
        }
    }
    return true;
}

// Writes the blob header (internal.hpp:431-454) into `dst`, which must hold
// layout.header_bytes bytes.
// Reads a blob's header (`avail` bytes of it at hdr) and rebuilds its layout,
// with the reference's checks and messages (benes.cpp:26-77); `blob_size` is
// the whole blob's size, which must equal the layout's end.
inline bool parse_header(const uint8_t* hdr, uint64_t avail, uint64_t blob_size, Layout& lay,
                         Status& st) {
    static const uint8_t magic[8] = {'B', 'E', 'N', 'E', 'S', '0', '1', 0};
    if (avail < 20) {
        st.set(kInvalidArgument, "compressed permutation header is truncated");
        return false;
    }
    if (std::memcmp(hdr, magic, 8) != 0) {
        st.set(kInvalidArgument, "compressed permutation has invalid magic");
        return false;
    }
    uint64_t n = 0;
    for (int i = 0; i < 8; ++i) n |= static_cast<uint64_t>(hdr[8 + i]) << (8 * i);
    const uint32_t placement = hdr[16], g = hdr[17], ni = hdr[18], no = hdr[19];
    const uint64_t header_bytes = 20 + ni + no;
    if (header_bytes > 20 + 64 || header_bytes > blob_size || avail < header_bytes) {
        st.set(kInvalidArgument, "unsupported compressed permutation header size");
        return false;
    }
    if (placement > 2) {
        st.set(kInvalidArgument, "compressed permutation has invalid format parameters");
        return false;
    }
    if (n == 0) {
        st.set(kInvalidArgument, "compressed permutation has invalid dimensions");
        return false;
    }
    FormatArgs fa;
    fa.middle_group_log2 = g;
    fa.placement = placement;
    fa.input_clusters.assign(hdr + 20, hdr + 20 + ni);
    fa.output_clusters.assign(hdr + 20 + ni, hdr + 20 + ni + no);
    if (!build_layout(n, fa, lay, st)) return false;
    if (lay.plan.end_offset != blob_size) {
        st.set(kInvalidArgument, "compressed permutation has invalid section bounds");
        return false;
    }
    return true;
}

inline void write_header(const Layout& lay, uint8_t* dst) {
    static const uint8_t magic[8] = {'B', 'E', 'N', 'E', 'S', '0', '1', 0};
    std::memcpy(dst, magic, 8);
    for (int i = 0; i < 8; ++i) dst[8 + i] = static_cast<uint8_t>(lay.plan.n >> (8 * i));
    dst[16] = static_cast<uint8_t>(lay.plan.placement);
    dst[17] = static_cast<uint8_t>(lay.plan.leaf_log2);
    dst[18] = static_cast<uint8_t>(lay.input_clusters.size());
    dst[19] = static_cast<uint8_t>(lay.output_clusters.size());
    std::memcpy(dst + 20, lay.input_clusters.data(), lay.input_clusters.size());
    std::memcpy(dst + 20 + lay.input_clusters.size(), lay.output_clusters.data(),
                lay.output_clusters.size());
}

}  // namespace benes_hip
