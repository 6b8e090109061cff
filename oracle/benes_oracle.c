/* Test infrastructure only -- the CPU oracle (see benes_oracle.h).
 *
 * Single-threaded restatement of the reference's canonical compact-v1 encoder.
 * Each function names the reference file:line it follows.  It is written for
 * clarity, not speed: it mirrors the reference's recursion, breadth-first
 * 2-colouring and per-bit ORs so that a mistake in the HIP path cannot be
 * "agreed with" by an oracle that shares its design.
 */
#include "benes_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Error context                                                             */

typedef struct {
    int status;
    char msg[256];
} Status;

static void fail(Status* st, int code, const char* msg) {
    if (st->status == ORC_OK) {
        st->status = code;
        strncpy(st->msg, msg, sizeof(st->msg) - 1);
        st->msg[sizeof(st->msg) - 1] = 0;
    }
}

static int report(const Status* st, char* err, uint64_t err_len) {
    if (st->status != ORC_OK && err && err_len) {
        strncpy(err, st->msg, (size_t)err_len - 1);
        err[err_len - 1] = 0;
    }
    return st->status;
}

static uint32_t bit_width64(uint64_t v) {
    return v ? 64u - (uint32_t)__builtin_clzll(v) : 0u;
}

/* internal.hpp:206-213 */
static uint64_t checked_add(Status* st, uint64_t a, uint64_t b) {
    if (b > UINT64_MAX - a) {
        fail(st, ORC_OVERFLOW, "compressed permutation is too large");
        return 0;
    }
    return a + b;
}

/* ------------------------------------------------------------------------ */
/* Factoradic code: internal.hpp:54-98, 148-204                              */

/* internal.hpp:173-183: bits of (size! - 1). */
uint32_t orc_factoradic_bits(uint32_t size) {
    u128 f = 1;
    for (uint32_t i = 2; i <= size; ++i) f *= i;
    f -= 1;
    const uint64_t hi = (uint64_t)(f >> 64);
    return hi ? 64u + bit_width64(hi) : bit_width64((uint64_t)f);
}

/* internal.hpp:148-167: right-to-left Lehmer digits, then Horner with
 * radix (i+1) for i ascending. */
void orc_encode_factoradic(const uint32_t* perm, uint32_t size, uint64_t out[2]) {
    uint32_t digits[32];
    memset(digits, 0, sizeof(digits));
    for (uint32_t i = 0; i < size; ++i) digits[i] = perm[i];
    for (uint32_t i = size; i-- > 1;) {
        for (uint32_t j = 0; j < i; ++j) {
            if (digits[j] > digits[i]) --digits[j];
        }
    }
    u128 r = 0;
    for (uint32_t i = 0; i < size; ++i) r = r * (u128)(i + 1) + digits[i];
    out[0] = (uint64_t)r;
    out[1] = (uint64_t)(r >> 64);
}

/* internal.hpp:189-204 */
void orc_decode_factoradic(uint64_t low, uint64_t high, uint32_t size,
                           uint32_t* out) {
    u128 v = ((u128)high << 64) | low;
    for (uint32_t i = size; i-- > 0;) {
        out[i] = (uint32_t)(v % (i + 1));
        v /= (i + 1);
    }
    for (uint32_t i = 0; i < size; ++i) {
        for (uint32_t j = 0; j < i; ++j) {
            if (out[j] >= out[i]) ++out[j];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FormatSizeCache: internal.hpp:225-313 (memoised recursion)               */

typedef struct {
    uint64_t count, sw, mid;
} MetricEntry;
typedef struct {
    uint64_t count;
    uint32_t depth;
    uint64_t bits;
} PacketEntry;

typedef struct {
    uint32_t leaf;
    Status* st;
    MetricEntry m[1024];
    uint32_t nm;
    PacketEntry p[4096];
    uint32_t np;
} Sizes;

static void sizes_init(Sizes* s, uint32_t g, Status* st) {
    s->leaf = 1u << g;
    s->st = st;
    s->nm = 0;
    s->np = 0;
}

/* internal.hpp:235-256 */
static void metrics(Sizes* s, uint64_t count, uint64_t* sw, uint64_t* mid) {
    if (count <= s->leaf) {
        *sw = 0;
        *mid = orc_factoradic_bits((uint32_t)count);
        return;
    }
    for (uint32_t i = 0; i < s->nm; ++i) {
        if (s->m[i].count == count) {
            *sw = s->m[i].sw;
            *mid = s->m[i].mid;
            return;
        }
    }
    uint64_t ls, lm, rs, rm;
    metrics(s, count / 2 + count % 2, &ls, &lm);
    metrics(s, count / 2, &rs, &rm);
    *sw = checked_add(s->st, count / 2, checked_add(s->st, ls, rs));
    *mid = checked_add(s->st, lm, rm);
    if (s->nm < sizeof(s->m) / sizeof(s->m[0])) {
        s->m[s->nm].count = count;
        s->m[s->nm].sw = *sw;
        s->m[s->nm].mid = *mid;
        ++s->nm;
    }
}

static uint64_t switch_bits(Sizes* s, uint64_t c) {
    uint64_t a, b;
    metrics(s, c, &a, &b);
    return a;
}

static uint64_t middle_bits(Sizes* s, uint64_t c) {
    uint64_t a, b;
    metrics(s, c, &a, &b);
    return b;
}

/* internal.hpp:271-291 */
static uint64_t packet_switch_bits(Sizes* s, uint64_t count, uint32_t depth) {
    if (count <= s->leaf || depth == 0) return 0;
    for (uint32_t i = 0; i < s->np; ++i) {
        if (s->p[i].count == count && s->p[i].depth == depth) return s->p[i].bits;
    }
    const uint64_t r = checked_add(s->st, count / 2,
        checked_add(s->st, packet_switch_bits(s, count / 2 + count % 2, depth - 1),
                    packet_switch_bits(s, count / 2, depth - 1)));
    if (s->np < sizeof(s->p) / sizeof(s->p[0])) {
        s->p[s->np].count = count;
        s->p[s->np].depth = depth;
        s->p[s->np].bits = r;
        ++s->np;
    }
    return r;
}

uint64_t orc_switch_bits(uint64_t count, uint32_t g) {
    Status st = {0, {0}};
    Sizes* s = malloc(sizeof(Sizes));
    sizes_init(s, g, &st);
    const uint64_t r = switch_bits(s, count);
    free(s);
    return r;
}

uint64_t orc_middle_bits(uint64_t count, uint32_t g) {
    Status st = {0, {0}};
    Sizes* s = malloc(sizeof(Sizes));
    sizes_init(s, g, &st);
    const uint64_t r = middle_bits(s, count);
    free(s);
    return r;
}

uint64_t orc_packet_switch_bits(uint64_t count, uint32_t g, uint32_t depth) {
    Status st = {0, {0}};
    Sizes* s = malloc(sizeof(Sizes));
    sizes_init(s, g, &st);
    const uint64_t r = packet_switch_bits(s, count, depth);
    free(s);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Format normalisation and layout: internal.hpp:315-454                    */

/* internal.hpp:315-326 */
static uint32_t routing_depth(uint64_t count, uint32_t g) {
    const uint64_t leaf = 1ull << g;
    uint32_t d = 0;
    while (count > leaf) {
        count = count / 2 + count % 2;
        ++d;
    }
    return d;
}

/* internal.hpp:328-338 */
static uint32_t default_clusters(uint32_t depth, uint8_t* out) {
    if (depth == 0) return 0;
    if (depth == 1) {
        out[0] = 1;
        return 1;
    }
    out[0] = (uint8_t)((depth + 1) / 2);
    out[1] = (uint8_t)(depth / 2);
    return 2;
}

/* internal.hpp:340-355 */
static void validate_clusters(Status* st, const uint8_t* c, uint32_t n,
                              uint32_t depth) {
    uint32_t sum = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (c[i] == 0) {
            fail(st, ORC_INVALID_ARGUMENT, "cluster depths must be positive");
            return;
        }
        sum += c[i];
    }
    if (sum != depth) {
        fail(st, ORC_INVALID_ARGUMENT,
             "cluster depths do not span every routing level");
    }
}

static uint64_t byte_count(uint64_t bits) {
    return bits / 8 + (uint64_t)(bits % 8 != 0);
}

/* normalize_format internal.hpp:357-383, then make_layout :389-429. */
static void make_layout(Status* st, uint64_t count, uint32_t g, uint32_t placement,
                        const uint8_t* in_c, uint32_t n_in,
                        const uint8_t* out_c, uint32_t n_out, orc_layout* L) {
    memset(L, 0, sizeof(*L));
    if (g == 0 || g > 5) {
        fail(st, ORC_INVALID_ARGUMENT, "middle_group_log2 must be between 1 and 5");
        return;
    }
    if (placement > 2) {
        fail(st, ORC_INVALID_ARGUMENT, "invalid middle placement");
        return;
    }
    if (n_in > 64 || n_out > 64) {
        /* A cluster list longer than 64 cannot sum to a routing depth. */
        fail(st, ORC_INVALID_ARGUMENT, "cluster depths do not span every routing level");
        return;
    }
    const uint32_t depth = routing_depth(count, g);
    uint8_t ic[64], oc[64];
    uint32_t ni = n_in, no = n_out;
    if (ni == 0) {
        ni = default_clusters(depth, ic);
    } else {
        memcpy(ic, in_c, ni);
    }
    if (no == 0) {
        no = default_clusters(depth, oc);
    } else {
        memcpy(oc, out_c, no);
    }
    validate_clusters(st, ic, ni, depth);
    if (st->status) return;
    validate_clusters(st, oc, no, depth);
    if (st->status) return;
    if (depth == 0 && placement != 0) placement = 0;

    Sizes* s = malloc(sizeof(Sizes));
    sizes_init(s, g, st);
    const uint64_t sw = switch_bits(s, count);
    const uint64_t mid = middle_bits(s, count);
    free(s);
    if (st->status) return;
    const uint64_t header = checked_add(st, 20, checked_add(st, ni, no));
    const uint64_t in_bytes = checked_add(st, byte_count(sw),
                                          placement == 1 ? byte_count(mid) : 0);
    const uint64_t out_bytes = checked_add(st, byte_count(sw),
                                           placement == 2 ? byte_count(mid) : 0);
    const uint64_t mid_bytes = placement == 0 ? byte_count(mid) : 0;
    if (st->status) return;
    L->size = count;
    if (count > (1ull << 63)) {
        fail(st, ORC_OVERFLOW, "logical network size is too large");
        return;
    }
    uint64_t ns = count > (1ull << g) ? count : (1ull << g);
    uint64_t p2 = 1;
    while (p2 < ns) p2 <<= 1;
    L->network_size = p2;
    L->middle_group_log2 = (uint8_t)g;
    L->middle_placement = (uint8_t)placement;
    L->n_input_clusters = (uint8_t)ni;
    L->n_output_clusters = (uint8_t)no;
    memcpy(L->input_clusters, ic, ni);
    memcpy(L->output_clusters, oc, no);
    L->input_offset = header;
    L->output_offset = checked_add(st, header, in_bytes);
    L->middle_offset = checked_add(st, L->output_offset, out_bytes);
    L->end_offset = checked_add(st, L->middle_offset, mid_bytes);
}

int orc_make_layout(uint64_t n, uint32_t g, uint32_t placement,
                    const uint8_t* in_c, uint32_t n_in,
                    const uint8_t* out_c, uint32_t n_out,
                    orc_layout* out, char* err, uint64_t err_len) {
    Status st = {0, {0}};
    make_layout(&st, n, g, placement, in_c, n_in, out_c, n_out, out);
    return report(&st, err, err_len);
}

/* internal.hpp:431-454 */
static void write_header(uint8_t* b, const orc_layout* L) {
    static const uint8_t magic[8] = {'B', 'E', 'N', 'E', 'S', '0', '1', 0};
    memcpy(b, magic, 8);
    for (int i = 0; i < 8; ++i) b[8 + i] = (uint8_t)(L->size >> (8 * i));
    b[16] = L->middle_placement;
    b[17] = L->middle_group_log2;
    b[18] = L->n_input_clusters;
    b[19] = L->n_output_clusters;
    memcpy(b + 20, L->input_clusters, L->n_input_clusters);
    memcpy(b + 20 + L->n_input_clusters, L->output_clusters, L->n_output_clusters);
}

/* ------------------------------------------------------------------------ */
/* SectionAddress: internal.hpp:456-718                                      */

typedef struct {
    const uint8_t* clusters;
    uint32_t n_clusters;
    int embeds;
    Sizes* sizes;
    uint32_t cluster_index;
    uint32_t group_depth;
    uint32_t relative_level;
    uint64_t path;
    uint64_t root_count;
    uint64_t node_count;
    uint64_t packet_base;
    uint64_t preceding_residual;
    uint64_t group_switch_bits;
    uint64_t full_packet_bits;
    uint64_t pair_span;
    uint64_t full_packet_prefix;
    uint64_t tail_packet;
    uint64_t tail_packet_prefix;
} Addr;

/* The `contribution` lambda of internal.hpp:617-629. */
static uint64_t contribution(const Addr* a, uint64_t count, uint64_t packet,
                             uint32_t level) {
    if (count <= a->sizes->leaf) return 0;
    const uint64_t span = 1ull << (a->group_depth - level - 1);
    const uint64_t first = packet * span;
    const uint64_t pairs = count / 2;
    if (first >= pairs) return 0;
    return pairs - first < span ? pairs - first : span;
}

/* internal.hpp:611-649 */
static uint64_t path_prefix_bits(const Addr* a, uint64_t packet, uint32_t level,
                                 uint64_t path_count) {
    const uint64_t nodes = 1ull << level;
    const uint64_t large = a->root_count % nodes;
    const uint64_t small = a->root_count / nodes;
    uint64_t included = 0, rp = path_count, rl = large;
    for (uint32_t bits = level; bits != 0; --bits) {
        const uint64_t half = 1ull << (bits - 1);
        if (rl <= half) {
            rp = (rp + 1) / 2;
        } else {
            included += (rp + 1) / 2;
            rp /= 2;
            rl -= half;
        }
    }
    if (rp != 0 && rl != 0) ++included;
    return included * contribution(a, small + 1, packet, level) +
           (path_count - included) * contribution(a, small, packet, level);
}

/* internal.hpp:604-609 */
static uint64_t packet_level_bits(const Addr* a, uint64_t packet, uint32_t level) {
    return path_prefix_bits(a, packet, level, 1ull << level);
}

/* internal.hpp:651-663 */
static void refresh_group(Addr* a) {
    a->group_switch_bits =
        packet_switch_bits(a->sizes, a->root_count, a->group_depth);
    uint64_t bits = 0;
    for (uint32_t l = 0; l < a->group_depth; ++l) bits += packet_level_bits(a, 0, l);
    a->full_packet_bits = bits;
}

/* internal.hpp:665-674 */
static uint64_t packet_prefix(const Addr* a, uint64_t packet) {
    uint64_t bits = 0;
    for (uint32_t l = 0; l < a->relative_level; ++l) bits += packet_level_bits(a, packet, l);
    return bits + path_prefix_bits(a, packet, a->relative_level, a->path);
}

/* internal.hpp:676-683 */
static void refresh_node(Addr* a) {
    a->pair_span = 1ull << (a->group_depth - a->relative_level - 1);
    a->full_packet_prefix = packet_prefix(a, 0);
    const uint64_t pairs = a->node_count / 2;
    a->tail_packet = pairs == 0 ? 0 : (pairs - 1) / a->pair_span;
    a->tail_packet_prefix = packet_prefix(a, a->tail_packet);
}

/* internal.hpp:479-494 */
static void addr_init(Addr* a, uint64_t count, const uint8_t* clusters,
                      uint32_t n_clusters, int embeds, Sizes* s) {
    memset(a, 0, sizeof(*a));
    a->clusters = clusters;
    a->n_clusters = n_clusters;
    a->embeds = embeds;
    a->sizes = s;
    a->root_count = count;
    a->node_count = count;
    if (count > s->leaf) {
        a->group_depth = clusters[0];
        refresh_group(a);
        refresh_node(a);
    }
}

/* internal.hpp:464-472 */
static uint64_t switch_bit(const Addr* a, uint64_t pair) {
    const uint64_t packet = pair / a->pair_span;
    return a->packet_base + packet * a->full_packet_bits +
           (packet == a->tail_packet ? a->tail_packet_prefix : a->full_packet_prefix) +
           pair % a->pair_span;
}

/* internal.hpp:689-699 */
static uint64_t section_bits(const Addr* a, uint64_t count) {
    return switch_bits(a->sizes, count) + (a->embeds ? middle_bits(a->sizes, count) : 0);
}

static uint64_t residual_bits(const Addr* a, uint64_t count, uint32_t consumed) {
    return section_bits(a, count) - packet_switch_bits(a->sizes, count, consumed);
}

/* internal.hpp:506-517 (callers guarantee embeds). */
static uint64_t middle_bit(const Addr* a) {
    if (a->root_count <= a->sizes->leaf) return a->packet_base;
    return a->packet_base + a->group_switch_bits + a->preceding_residual;
}

/* internal.hpp:553-601 */
static Addr addr_child(const Addr* a, int right, uint64_t count) {
    Addr r = *a;
    const uint64_t left_count = count / 2 + count % 2;
    const uint64_t child_count = right ? count / 2 : left_count;
    const uint32_t remaining = a->group_depth - a->relative_level;
    if (remaining > 1) {
        if (right) r.preceding_residual += residual_bits(a, left_count, remaining - 1);
        ++r.relative_level;
        r.path = a->path * 2 + (uint64_t)(right != 0);
        r.node_count = child_count;
        refresh_node(&r);
        return r;
    }
    uint64_t preceding = a->preceding_residual;
    if (right) preceding += section_bits(a, left_count);
    r.packet_base = a->packet_base + a->group_switch_bits + preceding;
    r.preceding_residual = 0;
    r.root_count = child_count;
    r.node_count = child_count;
    r.relative_level = 0;
    r.path = 0;
    if (child_count > a->sizes->leaf) {
        ++r.cluster_index;
        r.group_depth = a->clusters[r.cluster_index];
        refresh_group(&r);
        refresh_node(&r);
    } else {
        r.group_depth = 0;
        r.group_switch_bits = 0;
        r.full_packet_bits = 0;
        r.pair_span = 0;
        r.full_packet_prefix = 0;
        r.tail_packet = 0;
        r.tail_packet_prefix = 0;
    }
    return r;
}

int orc_node_address(const orc_layout* L, int side, uint32_t depth,
                     uint64_t path, orc_address* out) {
    Status st = {0, {0}};
    Sizes* s = malloc(sizeof(Sizes));
    sizes_init(s, L->middle_group_log2, &st);
    const int embeds = (side == 0 && L->middle_placement == 1) ||
                       (side == 1 && L->middle_placement == 2);
    Addr a;
    addr_init(&a, L->size,
              side == 0 ? L->input_clusters : L->output_clusters,
              side == 0 ? L->n_input_clusters : L->n_output_clusters, embeds, s);
    uint64_t count = L->size;
    uint64_t mid = 0;  /* separate-placement rank offset, benes.cpp:317-319 */
    for (uint32_t d = 0; d < depth; ++d) {
        if (count <= s->leaf) {
            free(s);
            return ORC_OUT_OF_RANGE;  /* path descends below a leaf */
        }
        const int right = (int)((path >> (depth - 1 - d)) & 1);
        const uint64_t left_count = count / 2 + count % 2;
        if (right) mid += middle_bits(s, left_count);
        a = addr_child(&a, right, count);
        count = right ? count / 2 : left_count;
    }
    memset(out, 0, sizeof(*out));
    out->count = count;
    out->packet_base = a.packet_base;
    out->full_packet_bits = a.full_packet_bits;
    out->full_packet_prefix = a.full_packet_prefix;
    out->tail_packet_prefix = a.tail_packet_prefix;
    out->pair_span = a.pair_span;
    out->tail_packet = a.tail_packet;
    out->group_switch_bits = a.group_switch_bits;
    out->relative_level = a.relative_level;
    out->group_depth = a.group_depth;
    out->is_leaf = count <= s->leaf;
    out->middle_bit = embeds ? middle_bit(&a) : mid;
    free(s);
    return st.status;
}

/* ------------------------------------------------------------------------ */
/* Canonical encoder: benes.cpp:152-382                                      */

typedef struct {
    uint8_t* bytes;
    const orc_layout* L;
    Sizes* sizes;
    Status* st;
} Enc;

/* benes.cpp:175-187 */
static void set_bit(uint8_t* bytes, uint64_t section, uint64_t bit, int value) {
    if (value) bytes[section + bit / 8] |= (uint8_t)(1u << (bit & 7));
}

/* benes.cpp:189-351, single-threaded (threads == 1). */
static void encode_node(Enc* e, const uint64_t* perm, uint64_t count,
                        const Addr* in_a, const Addr* out_a, uint64_t middle_off) {
    if (e->st->status) return;
    const orc_layout* L = e->L;
    if (count <= e->sizes->leaf) {
        uint32_t local[32];
        for (uint64_t i = 0; i < count; ++i) local[i] = (uint32_t)perm[i];
        uint64_t rank[2];
        orc_encode_factoradic(local, (uint32_t)count, rank);
        uint64_t section = L->middle_offset;
        if (L->middle_placement == 1) {
            section = L->input_offset;
            middle_off = middle_bit(in_a);
        } else if (L->middle_placement == 2) {
            section = L->output_offset;
            middle_off = middle_bit(out_a);
        }
        const uint32_t nbits = orc_factoradic_bits((uint32_t)count);
        for (uint32_t b = 0; b < nbits; ++b) {
            const uint64_t w = b < 64 ? rank[0] : rank[1];
            set_bit(e->bytes, section, middle_off + b, (int)((w >> (b & 63)) & 1));
        }
        return;
    }

    const uint64_t pairs = count / 2;
    const uint64_t nl = count / 2 + count % 2, nr = count / 2;
    uint64_t* left = malloc(sizeof(uint64_t) * nl);
    uint64_t* right = malloc(sizeof(uint64_t) * (nr ? nr : 1));
    {
        uint64_t* inverse = malloc(sizeof(uint64_t) * count);
        uint64_t* nb = malloc(sizeof(uint64_t) * 2 * count);
        uint8_t* degree = calloc(count, 1);
        int8_t* color = malloc(count);
        uint64_t* queue = malloc(sizeof(uint64_t) * count);
        for (uint64_t i = 0; i < count; ++i) inverse[perm[i]] = i;
#define CONNECT(a_, b_)                         \
    do {                                        \
        const uint64_t a0 = (a_), b0 = (b_);    \
        nb[2 * a0 + degree[a0]++] = b0;         \
        nb[2 * b0 + degree[b0]++] = a0;         \
    } while (0)
        for (uint64_t p = 0; p < pairs; ++p) {
            CONNECT(2 * p, 2 * p + 1);
            CONNECT(inverse[2 * p], inverse[2 * p + 1]);
        }
#undef CONNECT
        memset(color, -1, count);
        /* breadth-first colouring of one component, benes.cpp:255-277 */
#define COLOR_COMPONENT(first_)                                                \
    do {                                                                       \
        uint64_t head = 0, tail = 0;                                           \
        queue[tail++] = (first_);                                              \
        while (head != tail) {                                                 \
            const uint64_t v = queue[head++];                                  \
            for (uint8_t k = 0; k < degree[v]; ++k) {                          \
                const uint64_t nx = nb[2 * v + k];                             \
                const int8_t expected = (int8_t)(color[v] ^ 1);                \
                if (color[nx] < 0) {                                           \
                    color[nx] = expected;                                      \
                    queue[tail++] = nx;                                        \
                } else if (color[nx] != expected) {                            \
                    fail(e->st, ORC_LOGIC, "compact network coloring failed"); \
                }                                                              \
            }                                                                  \
        }                                                                      \
    } while (0)
        if (count % 2 == 1) {
            color[count - 1] = 0;
            COLOR_COMPONENT(count - 1);
            const uint64_t odd = inverse[count - 1];
            if (color[odd] < 0) {
                color[odd] = 0;
                COLOR_COMPONENT(odd);
            } else if (color[odd] != 0) {
                fail(e->st, ORC_LOGIC, "compact odd endpoints have incompatible colors");
            }
        }
        for (uint64_t v = 0; v < count; ++v) {
            if (color[v] < 0) {
                color[v] = 0;
                COLOR_COMPONENT(v);
            }
        }
#undef COLOR_COMPONENT
        for (uint64_t in = 0; in < count; ++in) {
            const int branch = color[in] != 0;
            if (in != count - 1 || count % 2 == 0) {
                set_bit(e->bytes, L->input_offset, switch_bit(in_a, in / 2),
                        branch != (int)(in & 1));
            }
            const uint64_t o = perm[in];
            if (o != count - 1 || count % 2 == 0) {
                set_bit(e->bytes, L->output_offset, switch_bit(out_a, o / 2),
                        branch != (int)(o & 1));
            }
            (branch ? right : left)[in / 2] = o / 2;
        }
        free(inverse);
        free(nb);
        free(degree);
        free(color);
        free(queue);
    }
    const uint64_t right_mid = middle_off + middle_bits(e->sizes, nl);
    Addr il = addr_child(in_a, 0, count), ol = addr_child(out_a, 0, count);
    encode_node(e, left, nl, &il, &ol, middle_off);
    free(left);
    Addr ir = addr_child(in_a, 1, count), orr = addr_child(out_a, 1, count);
    encode_node(e, right, nr, &ir, &orr, right_mid);
    free(right);
}

/* benes.cpp:152-173 */
static void validate_permutation(Status* st, const uint64_t* p, uint64_t n,
                                 uint32_t width) {
    if (n == 0) {
        fail(st, ORC_INVALID_ARGUMENT, "permutation size is out of range");
        return;
    }
    if (width == 4 && n > UINT32_MAX) {
        fail(st, ORC_INVALID_ARGUMENT, "permutation size requires 64-bit source values");
        return;
    }
    uint8_t* seen = calloc((size_t)((n + 7) / 8), 1);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t v = p[i];
        if (v >= n || (seen[v / 8] >> (v & 7)) & 1) {
            fail(st, ORC_INVALID_ARGUMENT,
                 "input must contain every value from zero through size minus one");
            break;
        }
        seen[v / 8] |= (uint8_t)(1u << (v & 7));
    }
    free(seen);
}

/* benes.cpp:355-382 with threads == 1. */
int orc_compress(const void* perm, uint64_t n, uint32_t width,
                 uint32_t g, uint32_t placement,
                 const uint8_t* in_c, uint32_t n_in,
                 const uint8_t* out_c, uint32_t n_out,
                 uint8_t* out, uint64_t cap, uint64_t* out_size,
                 char* err, uint64_t err_len) {
    Status st = {0, {0}};
    uint64_t* p = malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!p) {
        fail(&st, ORC_LENGTH, "allocation failed");
        return report(&st, err, err_len);
    }
    for (uint64_t i = 0; i < n; ++i) {
        p[i] = width == 4 ? ((const uint32_t*)perm)[i] : ((const uint64_t*)perm)[i];
    }
    validate_permutation(&st, p, n, width);
    orc_layout L;
    if (!st.status) make_layout(&st, n, g, placement, in_c, n_in, out_c, n_out, &L);
    if (!st.status) {
        *out_size = L.end_offset;
        if (!out || cap < L.end_offset) {
            /* size query: caller retries with a large enough buffer */
            free(p);
            return out ? ORC_LENGTH : ORC_OK;
        }
        memset(out, 0, L.end_offset);
        write_header(out, &L);
        Sizes* s = malloc(sizeof(Sizes));
        sizes_init(s, L.middle_group_log2, &st);
        Addr ia, oa;
        addr_init(&ia, n, L.input_clusters, L.n_input_clusters,
                  L.middle_placement == 1, s);
        addr_init(&oa, n, L.output_clusters, L.n_output_clusters,
                  L.middle_placement == 2, s);
        Enc e = {out, &L, s, &st};
        encode_node(&e, p, n, &ia, &oa, 0);
        free(s);
    }
    free(p);
    return report(&st, err, err_len);
}

/* ------------------------------------------------------------------------ */
/* Decoder: parse_header benes.cpp:26-77, validate_padding :79-115,           */
/* lookup internal.hpp:1035-1124.                                             */

static int read_bit(const uint8_t* b, uint64_t section, uint64_t bit) {
    return (b[section + bit / 8] >> (bit & 7)) & 1;
}

static void parse_layout(Status* st, const uint8_t* b, uint64_t size, orc_layout* L) {
    static const uint8_t magic[8] = {'B', 'E', 'N', 'E', 'S', '0', '1', 0};
    if (size < 20) {
        fail(st, ORC_INVALID_ARGUMENT, "compressed permutation header is truncated");
        return;
    }
    if (memcmp(b, magic, 8) != 0) {
        fail(st, ORC_INVALID_ARGUMENT, "compressed permutation has invalid magic");
        return;
    }
    uint64_t n = 0;
    for (int i = 0; i < 8; ++i) n |= (uint64_t)b[8 + i] << (8 * i);
    const uint32_t placement = b[16], g = b[17], ni = b[18], no = b[19];
    const uint64_t hdr = 20ull + ni + no;
    if (hdr > 20 + 64 || hdr > size) {
        fail(st, ORC_INVALID_ARGUMENT, "unsupported compressed permutation header size");
        return;
    }
    if (placement > 2) {
        fail(st, ORC_INVALID_ARGUMENT, "compressed permutation has invalid format parameters");
        return;
    }
    if (n == 0) {
        fail(st, ORC_INVALID_ARGUMENT, "compressed permutation has invalid dimensions");
        return;
    }
    make_layout(st, n, g, placement, b + 20, ni, b + 20 + ni, no, L);
    if (st->status) return;
    if (L->end_offset != size) {
        fail(st, ORC_INVALID_ARGUMENT, "compressed permutation has invalid section bounds");
        return;
    }
    Sizes* s = malloc(sizeof(Sizes));
    sizes_init(s, g, st);
    const uint64_t sw = switch_bits(s, n), mid = middle_bits(s, n);
    free(s);
    const uint64_t sec[3][3] = {
        {L->input_offset, L->output_offset, sw + (L->middle_placement == 1 ? mid : 0)},
        {L->output_offset, L->middle_offset, sw + (L->middle_placement == 2 ? mid : 0)},
        {L->middle_offset, L->end_offset, L->middle_placement == 0 ? mid : 0}};
    for (int k = 0; k < 3; ++k) {
        const uint64_t begin = sec[k][0], bytes = sec[k][1] - sec[k][0], used = sec[k][2];
        const uint64_t used_bytes = used / 8;
        if (used_bytes > bytes || (used_bytes == bytes && used % 8 != 0)) {
            fail(st, ORC_INVALID_ARGUMENT, "compressed permutation section is too small");
            return;
        }
        const uint64_t pad = (bytes - used_bytes) * 8 - used % 8;
        for (uint64_t i = 0; i < pad; ++i) {
            if (read_bit(b, begin, used + i)) {
                fail(st, ORC_INVALID_ARGUMENT,
                     "compressed permutation has nonzero section padding bits");
                return;
            }
        }
    }
}

static uint64_t lookup(Status* st, const uint8_t* b, const orc_layout* L, Sizes* s,
                       uint64_t value, int inverse) {
    if (value >= L->size) {
        fail(st, ORC_OUT_OF_RANGE, "permutation lookup is out of range");
        return 0;
    }
    struct Frame {
        uint64_t size;
        Addr in, out;
        int right;
    };
    struct Frame* frames = malloc(sizeof(struct Frame) * 64);
    uint32_t depth = 0;
    uint64_t count = L->size, index = value, mid = 0;
    Addr ia, oa;
    addr_init(&ia, count, L->input_clusters, L->n_input_clusters, L->middle_placement == 1, s);
    addr_init(&oa, count, L->output_clusters, L->n_output_clusters, L->middle_placement == 2, s);
    while (count > s->leaf) {
        const int odd_tail = count % 2 == 1 && index == count - 1;
        int right = 0;
        if (!odd_tail) {
            const uint64_t bit = switch_bit(inverse ? &oa : &ia, index / 2);
            const uint64_t section = inverse ? L->output_offset : L->input_offset;
            right = read_bit(b, section, bit) != (int)(index & 1);
        }
        frames[depth].size = count;
        frames[depth].in = ia;
        frames[depth].out = oa;
        frames[depth].right = right;
        ++depth;
        const uint64_t lc = count / 2 + count % 2;
        if (right) mid += middle_bits(s, lc);
        ia = addr_child(&ia, right, count);
        oa = addr_child(&oa, right, count);
        index /= 2;
        count = right ? count / 2 : lc;
    }
    uint64_t section = L->middle_offset, rbit = mid;
    if (L->middle_placement == 1) {
        section = L->input_offset;
        rbit = middle_bit(&ia);
    } else if (L->middle_placement == 2) {
        section = L->output_offset;
        rbit = middle_bit(&oa);
    }
    const uint32_t nbits = orc_factoradic_bits((uint32_t)count);
    uint64_t lo = 0, hi = 0;
    for (uint32_t k = 0; k < nbits; ++k) {
        if (!read_bit(b, section, rbit + k)) continue;
        if (k < 64) lo |= 1ull << k; else hi |= 1ull << (k - 64);
    }
    uint32_t leaf[32];
    orc_decode_factoradic(lo, hi, (uint32_t)count, leaf);
    if (!inverse) {
        index = leaf[index];
    } else {
        const uint32_t target = (uint32_t)index;
        for (index = 0; index < count && leaf[index] != target; ++index) {}
        if (index == count) {
            fail(st, ORC_RUNTIME, "invalid compact factorial code");
            free(frames);
            return 0;
        }
    }
    while (depth != 0) {
        const struct Frame* f = &frames[--depth];
        const uint64_t pairs = f->size / 2;
        if (!f->right && f->size % 2 == 1 && index == pairs) {
            index = f->size - 1;
            continue;
        }
        const uint64_t section2 = inverse ? L->input_offset : L->output_offset;
        const uint64_t bit = switch_bit(inverse ? &f->in : &f->out, index);
        const uint64_t parity = (uint64_t)f->right ^ (uint64_t)read_bit(b, section2, bit);
        index = index * 2 + parity;
    }
    free(frames);
    return index;
}

int orc_lookup(const uint8_t* blob, uint64_t size, uint64_t value, int inverse,
               uint64_t* result, char* err, uint64_t err_len) {
    Status st = {0, {0}};
    orc_layout L;
    parse_layout(&st, blob, size, &L);
    if (!st.status) {
        Sizes* s = malloc(sizeof(Sizes));
        sizes_init(s, L.middle_group_log2, &st);
        *result = lookup(&st, blob, &L, s, value, inverse);
        free(s);
    }
    return report(&st, err, err_len);
}

/* Forward (or inverse) images of `count` values, parsing the blob once: the
 * sampled round trip used where a whole decode would be too long. */
int orc_lookup_many(const uint8_t* blob, uint64_t size, const uint64_t* values,
                    uint64_t count, int inverse, uint64_t* results, char* err,
                    uint64_t err_len) {
    Status st = {0, {0}};
    orc_layout L;
    parse_layout(&st, blob, size, &L);
    if (!st.status) {
        Sizes* s = malloc(sizeof(Sizes));
        sizes_init(s, L.middle_group_log2, &st);
        for (uint64_t i = 0; i < count && !st.status; ++i) {
            results[i] = lookup(&st, blob, &L, s, values[i], inverse);
        }
        free(s);
    }
    return report(&st, err, err_len);
}

/* Decodes every forward image (and, when `inverse_out` is non-null, every
 * inverse image) of a blob: the round-trip verifier used by the tests. */
int orc_decode_all(const uint8_t* blob, uint64_t size, uint64_t* forward_out,
                   uint64_t* inverse_out, char* err, uint64_t err_len) {
    Status st = {0, {0}};
    orc_layout L;
    parse_layout(&st, blob, size, &L);
    if (!st.status) {
        Sizes* s = malloc(sizeof(Sizes));
        sizes_init(s, L.middle_group_log2, &st);
        for (uint64_t i = 0; i < L.size && !st.status; ++i) {
            forward_out[i] = lookup(&st, blob, &L, s, i, 0);
            if (inverse_out) inverse_out[i] = lookup(&st, blob, &L, s, i, 1);
        }
        free(s);
    }
    return report(&st, err, err_len);
}
