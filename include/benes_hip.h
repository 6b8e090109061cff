/* SPDX-License-Identifier: Apache-2.0
 *
 * benes_hip.h -- C ABI of the MI355X (gfx950) compression backend for the
 * compact version-1 Benes permutation format.
 *
 * This is the drop-in boundary for the reference's backend plug-in point:
 *
 *   bool detail::cuda_available() noexcept;                    src/internal.hpp:1191
 *   std::vector<std::byte> detail::compress_compact_cuda(
 *       std::span<const std::uint32_t>, FormatOptions);        src/internal.hpp:1193-1195
 *   std::vector<std::byte> detail::compress_compact_cuda(
 *       std::span<const std::uint64_t>, FormatOptions);        src/internal.hpp:1196-1198
 *
 * (the same slot the Metal backend fills at src/internal.hpp:1183-1190 and the
 * stubs at src/backends_stub.cpp:8-40 fill when a backend is not built).
 * Plain pointers and sizes only; no C++ or torch types.  Status codes map onto
 * the exception classes the reference throws for the same condition; the
 * C++ layer (benes-perm_amd/include/benes/benes.hpp) rethrows them.
 *
 * Every output blob is byte-identical to the reference's single-threaded
 * portable encoder (src/benes.cpp:189-382) for both input widths.
 */
#ifndef BENES_HIP_H
#define BENES_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define BENES_HIP_OK 0
#define BENES_HIP_INVALID_ARGUMENT 1 /* std::invalid_argument */
#define BENES_HIP_OVERFLOW 2         /* std::overflow_error   */
#define BENES_HIP_LENGTH 3           /* std::length_error     */
#define BENES_HIP_LOGIC 4            /* std::logic_error      */
#define BENES_HIP_RUNTIME 5          /* std::runtime_error (device error, unavailable) */
#define BENES_HIP_BUFFER_TOO_SMALL 6 /* out_capacity < needed; *out_size holds the need */
#define BENES_HIP_PENDING 7          /* benes_hip_server_wait timed out; request still running */
#define BENES_HIP_OUT_OF_RANGE 8     /* std::out_of_range (lookup of a value >= n) */

/* FormatOptions (reference include/benes/benes.hpp:32-39).  Empty cluster
 * lists (count 0) select the balanced default, as in the reference. */
typedef struct benes_hip_format {
    uint8_t middle_group_log2;     /* 1..5, default 5 */
    uint8_t middle_placement;      /* 0 separate, 1 embedded_input, 2 embedded_output */
    uint8_t n_input_clusters;
    uint8_t n_output_clusters;
    const uint8_t* input_clusters;
    const uint8_t* output_clusters;
} benes_hip_format;

/* Replaces detail::cuda_available() (src/internal.hpp:1191): non-zero when
 * a gfx950 device is present and the kernels in this library load on it. */
int benes_hip_available(void);

/* Replaces detail::compress_compact_cuda(span, FormatOptions)
 * (src/internal.hpp:1193-1198).  `perm` is a host array of n entries of
 * width_bytes (4 or 8).  Writes the whole blob (header included) to `out`.
 * When out_capacity is too small, returns BENES_HIP_BUFFER_TOO_SMALL with the
 * needed size in *out_size (call with out == NULL to query).
 * *device_seconds (optional) receives the device time of the call, the
 * reference's last_device_seconds() (src/benes.cpp:119-125). */
int benes_hip_compress(const void* perm, uint64_t n, uint32_t width_bytes,
                       const benes_hip_format* format,
                       uint8_t* out, uint64_t out_capacity, uint64_t* out_size,
                       double* device_seconds, char* err, size_t err_len);

/* Device-resident variant of benes_hip_compress: `d_perm` and `d_out` are
 * device pointers on the current device; work is queued on `stream`
 * (a hipStream_t, NULL for the default stream) and the call returns once the
 * blob is complete in d_out.  Same validation and error behaviour. */
int benes_hip_compress_device(const void* d_perm, uint64_t n, uint32_t width_bytes,
                              const benes_hip_format* format,
                              uint8_t* d_out, uint64_t out_capacity, uint64_t* out_size,
                              void* stream, char* err, size_t err_len);

/* Allocation-free variant of benes_hip_compress_device.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or other allocating routine and
 * creates no stream or event: every device buffer the kernels need beside the
 * input and output is a slice of one fixed array of
 * BENES_HIP_NOALLOC_SCRATCH_BYTES per device, which the device holds from the
 * time the library's code object is loaded on it until the process ends.
 * Work is queued on `stream` (NULL: the default stream) and the call returns
 * once the blob is complete.  Limits: n <= BENES_HIP_NOALLOC_MAX_ENTRIES
 * (BENES_HIP_LENGTH beyond), `d_out` 8-byte aligned with out_capacity at
 * least the blob size rounded up to a multiple of 8 (BENES_HIP_BUFFER_TOO_SMALL
 * otherwise, with the blob size in *out_size).  Every call on one device
 * shares that device's array and keeps its contents until the next call:
 * calls whose device work overlapped would race, so the library holds a
 * per-device lock from a call's first launch until its stream has drained,
 * and calls from several host threads run one after another.  Bytes are those
 * of benes_hip_compress_device; per-kernel profiling does not cover it. */
#define BENES_HIP_NOALLOC_MAX_ENTRIES (1ull << 20)
#define BENES_HIP_NOALLOC_SCRATCH_BYTES (36ull << 20)
int benes_hip_compress_device_noalloc(const void* d_perm, uint64_t n, uint32_t width_bytes,
                                      const benes_hip_format* format,
                                      uint8_t* d_out, uint64_t out_capacity, uint64_t* out_size,
                                      void* stream, char* err, size_t err_len);

/* ---- Batches: many permutations of uneven sizes and widths in one call. ----
 * Every launch of the pipeline covers all items of the batch at once; the
 * bytes of each blob are the same as those of a plain compress of the item.
 * Items failing validation get their own status and message; the others
 * still complete.  No counterpart in the reference, whose compress() takes one
 * permutation per call. */
typedef struct benes_hip_item {
    const void* perm;         /* host (compress_batch) or device (_device) pointer */
    uint64_t n;
    uint32_t width_bytes;     /* 4 or 8 */
    int32_t status;           /* [out] BENES_HIP_* of this item */
    benes_hip_format format;
    uint8_t* out;             /* compress_batch: host destination of the blob */
    uint64_t out_capacity;    /* compress_batch: its capacity in bytes */
    uint64_t out_offset;      /* [out] _device / batch_layout: byte offset in d_out */
    uint64_t out_size;        /* [out] blob bytes */
    char message[96];         /* [out] error text when status != OK */
} benes_hip_item;

/* Host-only: validates formats, fills out_size and out_offset of every item
 * and the device bytes a compress_batch_device output needs (*total_bytes). */
int benes_hip_batch_layout(benes_hip_item* items, uint64_t count, uint64_t* total_bytes);

/* Host inputs and outputs (one staging copy in, one copy out per item). */
int benes_hip_compress_batch(benes_hip_item* items, uint64_t count, double* device_seconds,
                             char* err, size_t err_len);

/* Device inputs; every blob is written into d_out (8-byte aligned, at least
 * benes_hip_batch_layout's total) at its item's out_offset. */
int benes_hip_compress_batch_device(benes_hip_item* items, uint64_t count, uint8_t* d_out,
                                    uint64_t out_capacity, void* stream, double* device_seconds,
                                    char* err, size_t err_len);

/* ---- Serving: a request loop with continuous batching. ----
 * No counterpart in the reference.  Requests (host arrays, any size and width,
 * any format) are queued on submit; a scheduler thread batches the queue's
 * oldest requests whenever one of `lanes` device streams is idle, admitting
 * only what fits that lane's share of hbm_budget_bytes, so batches on
 * different lanes overlap on the device and requests arriving meanwhile join
 * the next batch.  Each blob is identical to a plain compress of its request.
 * The caller keeps `perm` readable and `out` writable until the request
 * completes (benes_hip_server_wait).  A request larger than one lane's share
 * of the budget is refused with BENES_HIP_LENGTH. */
typedef struct benes_hip_server benes_hip_server;

typedef struct benes_hip_server_counters {
    uint64_t completed;              /* requests finished (any status) */
    uint64_t entries;                /* entries of requests finished OK */
    uint64_t batches;                /* batches launched */
    uint64_t peak_bytes;             /* largest admitted footprint in flight */
    uint64_t max_concurrent_batches; /* most lanes busy at once */
    uint64_t queued;                 /* requests waiting now */
    uint64_t budget_bytes;
} benes_hip_server_counters;

benes_hip_server* benes_hip_server_create(uint64_t hbm_budget_bytes, uint32_t max_batch,
                                          uint32_t lanes, char* err, size_t err_len);
/* Returns the request id (>= 1), or 0 with *status set when refused. */
uint64_t benes_hip_server_submit(benes_hip_server* server, const void* perm, uint64_t n,
                                 uint32_t width_bytes, const benes_hip_format* format,
                                 uint8_t* out, uint64_t out_capacity, int* status, char* err,
                                 size_t err_len);
/* Waits (timeout_ms < 0: without limit) and returns the request's status;
 * *latency_seconds is the time from submit to the blob being in `out`.
 * BENES_HIP_PENDING means the time ran out first; any other status releases
 * the id. */
int benes_hip_server_wait(benes_hip_server* server, uint64_t id, int64_t timeout_ms,
                          uint64_t* out_size, double* latency_seconds, char* err, size_t err_len);
int benes_hip_server_stats(benes_hip_server* server, benes_hip_server_counters* out);
/* Finishes queued and running requests, then stops the loop. */
void benes_hip_server_destroy(benes_hip_server* server);

/* ---- Files: ingest and egress. ----
 * Replaces compress_file (src/benes.cpp:498-553): raw little-endian values of
 * width_bytes (4 or 8) from `input_path`, the blob written to `output_path`.
 * The file is read whole into host memory and compressed as a host array
 * (as benes_hip_compress does: staged to HBM and validated there), and the
 * blob is written to the output file.  Errors and messages are the reference's, in
 * its order: "cannot open permutation input file" (BENES_HIP_RUNTIME),
 * "permutation input has an invalid element width or size"
 * (BENES_HIP_INVALID_ARGUMENT), "failed to read permutation input file"
 * (BENES_HIP_RUNTIME), the compress errors, then "failed to write compressed
 * permutation file" (BENES_HIP_RUNTIME).  The output is created or truncated
 * only after compression succeeded.  *out_size receives the blob's size. */
int benes_hip_compress_file(const char* input_path, const char* output_path,
                            uint32_t width_bytes, const benes_hip_format* format,
                            uint64_t* out_size, double* device_seconds, char* err,
                            size_t err_len);

/* ---- Lookup: the consumer side of the format. ----
 * Replaces the reference's forward_lookup / inverse_lookup
 * (src/internal.hpp:1036-1139), as used by CompressedPermutationView
 * (src/benes.cpp:581-700): images of `count` values under the permutation a
 * blob encodes (inverse != 0: under its inverse).  The header is checked as
 * the reference's parse_header does (src/benes.cpp:26-77); section padding is
 * not checked.  A value >= n gives BENES_HIP_OUT_OF_RANGE.  Results are
 * written as width_bytes (4 or 8) entries.
 * _device: blob, values and results are device memory, work on `stream`;
 * d_values == NULL looks up 0..count-1 (a whole decode when count == n).
 * Host form: host buffers, staged through the library's pinned ring. */
int benes_hip_lookup_device(const uint8_t* d_blob, uint64_t blob_size, const uint64_t* d_values,
                            uint64_t count, int inverse, void* d_out, uint32_t width_bytes,
                            void* stream, char* err, size_t err_len);
int benes_hip_lookup(const uint8_t* blob, uint64_t blob_size, const uint64_t* values,
                     uint64_t count, int inverse, void* out, uint32_t width_bytes, char* err,
                     size_t err_len);

/* Size of the blob for n entries (reference make_layout, src/internal.hpp:389-429). */
int benes_hip_encoded_size(uint64_t n, const benes_hip_format* format,
                           uint64_t* out_size, char* err, size_t err_len);

/* Device seconds of the last compress on this thread
 * (reference detail::last_device_seconds, src/benes.cpp:123-125). */
double benes_hip_last_device_seconds(void);

/* Per-kernel device timing (no counterpart in the reference, whose only timer
 * is last_device_seconds): while enabled, every kernel launch of later
 * compress calls is bracketed by a HIP event pair on the stream it runs on,
 * and the elapsed times are summed per kernel name.  Enabling clears the sums.
 * benes_hip_kernel_stats returns 0 once `index` is past the last kernel. */
void benes_hip_set_profiling(int enable);
int benes_hip_kernel_stats(int index, char* name, size_t name_len, double* total_ms,
                           uint64_t* launches);

/* ---- Host-side format arithmetic (no device needed), exported so that the
 * ---- layout and addressing can be checked against the oracle on a CPU. ---- */

typedef struct benes_hip_layout {
    uint64_t size;
    uint64_t network_size;
    uint64_t input_offset, output_offset, middle_offset, end_offset; /* bytes */
    uint64_t switch_bits, middle_bits;
    uint32_t routing_depth;
    uint8_t middle_group_log2, middle_placement, n_input_clusters, n_output_clusters;
    uint8_t input_clusters[64];
    uint8_t output_clusters[64];
} benes_hip_layout;

/* make_layout + normalize_format (src/internal.hpp:357-429). */
int benes_hip_layout_of(uint64_t n, const benes_hip_format* format,
                        benes_hip_layout* out, char* err, size_t err_len);

/* Switch-bit address of node (depth, path) on side 0 (input) or 1 (output):
 * the reference's SwitchBitAddress (src/internal.hpp:456-473) reached by
 * SectionAddress::child() along `path` (first branch = most significant of
 * `depth` bits).  Also the node's count and, when the node is a leaf, the bit
 * offset of its rank within its section (rank_bit). */
typedef struct benes_hip_address {
    uint64_t count;
    uint64_t packet_base;
    uint64_t full_packet_bits;
    uint64_t full_packet_prefix;
    uint64_t tail_packet_prefix;
    uint64_t pair_span;
    uint64_t tail_packet;
    uint64_t rank_bit;
    uint64_t offset; /* position of the node's entries in the level array */
} benes_hip_address;

int benes_hip_node_address(uint64_t n, const benes_hip_format* format, int side,
                           uint32_t depth, uint64_t path, benes_hip_address* out,
                           char* err, size_t err_len);

/* Bit offset of switch `pair` of a node (SwitchBitAddress::bit). */
uint64_t benes_hip_address_bit(const benes_hip_address* address, uint64_t pair);

/* ---- Synthetic seeded permutations (benchmark and test workloads). ---- */

#define BENES_HIP_PERM_FISHER_YATES 0 /* uniform: for i = n-1..1 swap(i, U[0,i]) */
#define BENES_HIP_PERM_SINGLE_CYCLE 1 /* Sattolo: for i = n-1..1 swap(i, U[0,i-1]) */
#define BENES_HIP_PERM_NEAR_IDENTITY 2 /* identity + floor(rate_ppm*n/1e6) swaps */
#define BENES_HIP_PERM_IDENTITY 3
#define BENES_HIP_PERM_REVERSE 4
#define BENES_HIP_PERM_KEYED 5 /* keyed Feistel bijection walked into [0, n); parallel */
#define BENES_HIP_PERM_UNIFORM_PARALLEL 6 /* uniform, drawn on all host threads: each index to
                                           * one of 2^k buckets by its own draw, buckets laid
                                           * out in order, each shuffled by Fisher-Yates
                                           * (Rao-Sandelius); host only */

/* Writes n entries of width_bytes (4 or 8) into `out` (host memory).  The
 * generator is splitmix64 with Lemire's bounded draw, so the sequence is fixed
 * by (kind, n, seed, rate_ppm) on every platform. */
int benes_hip_generate(uint32_t kind, uint64_t n, uint64_t seed, uint32_t rate_ppm,
                       void* out, uint32_t width_bytes);

/* The same permutations written straight into device memory `d_out` on
 * `stream` (returns once they are written), for the kinds computable entry by
 * entry: IDENTITY, REVERSE and KEYED (the others are sequential draws and
 * return BENES_HIP_INVALID_ARGUMENT).  Used for inputs too large to generate
 * on the host, such as 2^32 + 12345 entries. */
int benes_hip_generate_device(uint32_t kind, uint64_t n, uint64_t seed, void* d_out,
                              uint32_t width_bytes, void* stream);

/* The subtree shape (entries per LDS subtree, threads per workgroup) is
 * chosen per call by an online tuner: calls are classed by the log2 of their
 * largest item and of their total entries; the first
 * BENES_HIP_AUTOTUNE_TRIAL_CALLS calls of a class are trials that cycle twice
 * through the shapes (every shape gives the same bytes), and later calls of
 * the class use the shape whose fastest trial was fastest, without further
 * updates.  The environment variable BENES_HIP_SUBTREE=<log2>x<threads>
 * (14x1024, 13x1024, 13x512 or 12x512) pins one.  This writes a text summary
 * of the classes seen, the trial device times and the choices into `out`;
 * returns the summary's full length.  No counterpart in the reference. */
#define BENES_HIP_AUTOTUNE_TRIAL_CALLS 8
int benes_hip_autotune_report(char* out, size_t len);

/* Pins the subtree shape for later calls (one of the four above), or hands the
 * choice back to the tuner with (0, 0).  No counterpart in the reference. */
int benes_hip_set_subtree_shape(uint32_t log2_entries, uint32_t threads);

/* First-hook binning of the union-find (staging the random first-hook stores
 * of very large nodes in bins that one XCD then writes): -1 by size (the
 * default: nodes of at least 64 MB of words in levels of at least 2^31
 * entries), 0 never, 1 wherever a node's words exceed 4 MB.  The blob is the
 * same in every mode; tests use 1 to check the binned path at small sizes.
 * No counterpart in the reference. */
int benes_hip_set_hook_binning(int mode);

/* 64-bit union-find words and level-0 inverse: 0 by size (the default:
 * 64-bit words in nodes of 2^31 input slots or more, a 64-bit inverse from
 * 2^32 - 1 entries), 1 at every level of every call.  The blob is the same in
 * both modes; tests use 1 to check the 64-bit kernels at small sizes, as the
 * reference's Metal backend does with BENES_METAL_FORCE_PARENT64
 * (src/metal/benes_metal.mm:615-633). */
int benes_hip_set_wide_words(int mode);

/* Union pass of the HBM levels whose nodes each keep to one XCD: -1 by size
 * (the default: hooks by plain stores, checked afterwards and the overwritten
 * relations joined again, where a node's union-find words fit in an XCD's
 * L2), 0 compare-and-swap hooks everywhere, 1 plain-store hooks at every such
 * depth.  The blob is the same in every mode; tests use 1 and 0 to check both
 * paths at small sizes.  No counterpart in the reference. */
int benes_hip_set_plain_union(int mode);

/* Colouring of the levels coloured in HBM (nodes of more than 2^16 entries):
 * 0 the parity union-find (the default), 1 ruler orbit minima, a second
 * colouring that shares no code with the union-find: each entry's colour is
 * read off the smallest entry of its orbit under x -> I[P[x] ^ 1] ^ 1, found
 * by sampled rulers and pointer jumping.  Both give the reference's
 * canonical colours (src/benes.cpp:278-297), so the blob is the same; mode 1
 * is slower and exists to check the union-find byte for byte at sizes no CPU
 * encoder reaches.  No counterpart in the reference. */
int benes_hip_set_colouring(int mode);

/* Frees the device buffers the library keeps for the calling thread's
 * current device (they grow to the largest call so far and are otherwise
 * kept for reuse).  No counterpart in the reference. */
void benes_hip_release_workspace(void);

#ifdef __cplusplus
}
#endif

#endif /* BENES_HIP_H */
