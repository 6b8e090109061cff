/* Test infrastructure only -- the CPU oracle.
 *
 * A single-threaded C restatement of the reference's canonical portable
 * encoder for the compact version-1 Benes format, plus its layout arithmetic,
 * section addressing and lookup decoder.  Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg load it (as liboracle.so via ctypes), and
 * only to check or time against; the product never links it.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * checks each function against something other than itself) is listed in
 * tests/test_oracle.py: the three reference golden blobs, the reference's exact
 * payload-size table, closed forms (identity / reverse), brute-force bijection
 * of the factoradic code, round trips through the restated decoder, and byte
 * equality with the reference library compiled from its own sources
 * (oracle/_ref).  No function is left "parity unpinned".
 */
#ifndef BENES_ORACLE_H
#define BENES_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror the reference's exception classes. */
enum {
    ORC_OK = 0,
    ORC_INVALID_ARGUMENT = 1,
    ORC_OVERFLOW = 2,
    ORC_LENGTH = 3,
    ORC_LOGIC = 4,
    ORC_RUNTIME = 5,
    ORC_OUT_OF_RANGE = 7
};

typedef struct {
    uint64_t size;
    uint64_t network_size;
    uint8_t middle_group_log2;
    uint8_t middle_placement;
    uint8_t n_input_clusters;
    uint8_t n_output_clusters;
    uint8_t input_clusters[64];
    uint8_t output_clusters[64];
    uint64_t input_offset;
    uint64_t output_offset;
    uint64_t middle_offset;
    uint64_t end_offset;
} orc_layout;

/* Address of one node on one side (0 input, 1 output), as the reference's
 * SectionAddress sees it after descending from the root along `path`. */
typedef struct {
    uint64_t count;
    uint64_t packet_base;
    uint64_t full_packet_bits;
    uint64_t full_packet_prefix;
    uint64_t tail_packet_prefix;
    uint64_t pair_span;
    uint64_t tail_packet;
    uint64_t middle_bit;       /* embedded rank offset, or separate offset */
    uint64_t group_switch_bits;
    uint32_t relative_level;
    uint32_t group_depth;
    uint32_t is_leaf;
} orc_address;

int orc_make_layout(uint64_t n, uint32_t middle_group_log2, uint32_t placement,
                    const uint8_t* in_clusters, uint32_t n_in,
                    const uint8_t* out_clusters, uint32_t n_out,
                    orc_layout* out, char* err, uint64_t err_len);

int orc_compress(const void* perm, uint64_t n, uint32_t width_bytes,
                 uint32_t middle_group_log2, uint32_t placement,
                 const uint8_t* in_clusters, uint32_t n_in,
                 const uint8_t* out_clusters, uint32_t n_out,
                 uint8_t* out, uint64_t out_capacity, uint64_t* out_size,
                 char* err, uint64_t err_len);

int orc_lookup(const uint8_t* blob, uint64_t size, uint64_t value,
               int inverse, uint64_t* result, char* err, uint64_t err_len);

int orc_lookup_many(const uint8_t* blob, uint64_t size, const uint64_t* values,
                    uint64_t count, int inverse, uint64_t* results, char* err,
                    uint64_t err_len);

int orc_decode_all(const uint8_t* blob, uint64_t size, uint64_t* forward_out,
                   uint64_t* inverse_out, char* err, uint64_t err_len);

int orc_node_address(const orc_layout* layout, int side, uint32_t depth,
                     uint64_t path, orc_address* out);

uint64_t orc_switch_bits(uint64_t count, uint32_t middle_group_log2);
uint64_t orc_middle_bits(uint64_t count, uint32_t middle_group_log2);
uint64_t orc_packet_switch_bits(uint64_t count, uint32_t middle_group_log2,
                                uint32_t depth);
uint32_t orc_factoradic_bits(uint32_t size);
void orc_encode_factoradic(const uint32_t* perm, uint32_t size, uint64_t out[2]);
void orc_decode_factoradic(uint64_t low, uint64_t high, uint32_t size,
                           uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif
