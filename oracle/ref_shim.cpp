// Test infrastructure only: a C-ABI shim over the reference library's own
// benes::compress (compiled from the reference sources by oracle/Makefile into
// oracle/_ref/).  Tests and bench.py's CPU-baseline leg load it through ctypes
// to pin the oracle and to time the reference's multithreaded portable_cpu
// encoder.  Nothing in the product links or loads it.
//
// Bound interface: benes::compress(std::span<const std::uint32_t|uint64_t>,
// CompressOptions) -- include/benes/benes.hpp:47-52 of the reference.
#include <benes/benes.hpp>

#include <cstdint>
#include <cstring>
#include <exception>
#include <span>
#include <stdexcept>
#include <vector>

namespace {

int classify(const std::exception& e) {
    if (dynamic_cast<const std::invalid_argument*>(&e)) return 1;
    if (dynamic_cast<const std::overflow_error*>(&e)) return 2;
    if (dynamic_cast<const std::length_error*>(&e)) return 3;
    if (dynamic_cast<const std::logic_error*>(&e)) return 4;
    if (dynamic_cast<const std::runtime_error*>(&e)) return 5;
    return 6;
}

std::vector<std::byte> g_last;

}  // namespace

extern "C" {

// Compresses with the reference portable_cpu backend.  Returns 0 on success and
// the blob size in *out_size; the blob itself is fetched by ref_fetch().
// On failure returns the exception class (1 invalid_argument, 2 overflow_error,
// 3 length_error, 4 logic_error, 5 runtime_error, 6 other) and copies the
// message into err.
int ref_compress(const void* perm, std::uint64_t n, std::uint32_t width_bytes,
                 std::uint32_t middle_group_log2, std::uint32_t placement,
                 const std::uint8_t* in_clusters, std::uint32_t n_in,
                 const std::uint8_t* out_clusters, std::uint32_t n_out,
                 std::uint64_t threads, std::uint64_t* out_size,
                 char* err, std::uint64_t err_len) {
    try {
        benes::CompressOptions opt;
        opt.backend = benes::Backend::portable_cpu;
        opt.threads = static_cast<std::size_t>(threads);
        opt.format.middle_group_log2 = static_cast<std::uint8_t>(middle_group_log2);
        opt.format.input_clusters.assign(in_clusters, in_clusters + n_in);
        opt.format.output_clusters.assign(out_clusters, out_clusters + n_out);
        opt.format.middle_placement = static_cast<benes::MiddlePlacement>(placement);
        if (width_bytes == 4) {
            g_last = benes::compress(
                std::span<const std::uint32_t>(
                    static_cast<const std::uint32_t*>(perm), n), opt);
        } else {
            g_last = benes::compress(
                std::span<const std::uint64_t>(
                    static_cast<const std::uint64_t*>(perm), n), opt);
        }
        *out_size = g_last.size();
        return 0;
    } catch (const std::exception& e) {
        if (err && err_len) {
            std::strncpy(err, e.what(), err_len - 1);
            err[err_len - 1] = 0;
        }
        return classify(e);
    }
}

void ref_fetch(void* dst) {
    std::memcpy(dst, g_last.data(), g_last.size());
}

// Forward (inverse=0) or inverse (inverse=1) lookup on a blob with the
// reference's CompressedPermutationView.  Returns 0 and writes *result, or an
// exception class as above.
int ref_lookup(const void* blob, std::uint64_t size, std::uint64_t value,
               int inverse, std::uint64_t* result) {
    try {
        benes::CompressedPermutationView view(std::span<const std::byte>(
            static_cast<const std::byte*>(blob), size));
        *result = inverse ? view.inverse(value) : view.forward(value);
        return 0;
    } catch (const std::exception& e) {
        return classify(e);
    }
}

}  // extern "C"
