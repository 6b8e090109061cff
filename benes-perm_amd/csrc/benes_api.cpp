// SPDX-License-Identifier: Apache-2.0
// C++ layer over the C ABI: turns status codes back into the exception types
// the reference throws (src/benes.cpp, src/internal.hpp) and returns blobs by
// value, as detail::compress_compact_cuda does (src/internal.hpp:1193-1198).
#include <benes/benes.hpp>

#include <fstream>
#include <stdexcept>
#include <string>

#include "benes_hip.h"

namespace benes {
namespace detail {

[[noreturn]] void throw_status(int code, const char* message) {
    switch (code) {
    case BENES_HIP_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case BENES_HIP_OVERFLOW: throw std::overflow_error(message);
    case BENES_HIP_LENGTH: throw std::length_error(message);
    case BENES_HIP_LOGIC: throw std::logic_error(message);
    case BENES_HIP_OUT_OF_RANGE: throw std::out_of_range(message);
    default: throw std::runtime_error(message);
    }
}

}  // namespace detail

namespace {

[[noreturn]] void rethrow(int code, const char* message) { detail::throw_status(code, message); }

benes_hip_format c_format(const FormatOptions& format) {
    benes_hip_format f{};
    f.middle_group_log2 = format.middle_group_log2;
    f.middle_placement = static_cast<std::uint8_t>(format.middle_placement);
    if (format.input_clusters.size() > 255 || format.output_clusters.size() > 255) {
        throw std::invalid_argument("cluster depths do not span every routing level");
    }
    f.n_input_clusters = static_cast<std::uint8_t>(format.input_clusters.size());
    f.n_output_clusters = static_cast<std::uint8_t>(format.output_clusters.size());
    f.input_clusters = format.input_clusters.data();
    f.output_clusters = format.output_clusters.data();
    return f;
}

void require_hip(Backend backend) {
    if (backend == Backend::automatic) backend = Backend::hip;
    if (backend != Backend::hip || !detail::hip_available()) {
        throw std::runtime_error(std::string(backend_name(backend)) + " backend is unavailable");
    }
}

template <typename Value>
std::vector<std::byte> compress_hip(std::span<const Value> permutation,
                                    const FormatOptions& format) {
    const benes_hip_format f = c_format(format);
    char err[256] = {0};
    std::uint64_t size = 0;
    // A size query (no output buffer).  Errors come back in the reference's
    // order: the input's size, then for a rejected format the permutation
    // itself (checked on the device, src/benes.cpp:361-363), then the
    // format; an acceptable request answers BENES_HIP_BUFFER_TOO_SMALL with
    // the blob's size.
    int rc = benes_hip_compress(permutation.data(), permutation.size(), sizeof(Value), &f,
                                nullptr, 0, &size, nullptr, err, sizeof(err));
    if (rc != BENES_HIP_BUFFER_TOO_SMALL) {
        rethrow(rc == BENES_HIP_OK ? BENES_HIP_LOGIC : rc, rc == BENES_HIP_OK ? "size query failed" : err);
    }
    std::vector<std::byte> out(static_cast<std::size_t>(size));
    rc = benes_hip_compress(permutation.data(), permutation.size(), sizeof(Value), &f,
                            reinterpret_cast<std::uint8_t*>(out.data()), out.size(), &size,
                            nullptr, err, sizeof(err));
    if (rc != BENES_HIP_OK) rethrow(rc, err);
    return out;
}

template <typename Value>
std::vector<std::byte> dispatch(std::span<const Value> permutation, CompressOptions options) {
    require_hip(options.backend);
    return compress_hip(permutation, options.format);
}

}  // namespace

std::vector<std::byte> compress(std::span<const std::uint32_t> permutation,
                                CompressOptions options) {
    return dispatch(permutation, std::move(options));
}

std::vector<std::byte> compress(std::span<const std::uint64_t> permutation,
                                CompressOptions options) {
    return dispatch(permutation, std::move(options));
}

// The reference's compress_file (src/benes.cpp:498-553), in its order and with
// its messages: open and size the input, read it whole, decode the
// little-endian values, compress them (through benes_hip_compress, so every
// compress error comes as compress raises it), and only then create or
// truncate the output and write the blob.
void compress_file(const std::filesystem::path& input, const std::filesystem::path& output,
                   InputValueWidth input_width, CompressOptions options) {
    std::ifstream stream(input, std::ios::binary | std::ios::ate);
    if (!stream) throw std::runtime_error("cannot open permutation input file");
    const std::streamoff length = stream.tellg();
    const std::size_t value_bytes = static_cast<std::size_t>(input_width);
    if ((value_bytes != 4 && value_bytes != 8) || length <= 0 ||
        static_cast<std::uint64_t>(length) % value_bytes != 0) {
        throw std::invalid_argument("permutation input has an invalid element width or size");
    }
    stream.seekg(0);
    std::vector<unsigned char> raw(static_cast<std::size_t>(length));
    stream.read(reinterpret_cast<char*>(raw.data()), length);
    if (!stream) throw std::runtime_error("failed to read permutation input file");
    std::vector<std::byte> encoded;
    const std::size_t count = raw.size() / value_bytes;
    auto load = [&](std::size_t i) {
        std::uint64_t v = 0;
        for (std::size_t b = value_bytes; b-- > 0;) v = (v << 8) | raw[i * value_bytes + b];
        return v;
    };
    if (input_width == InputValueWidth::bits32) {
        std::vector<std::uint32_t> permutation(count);
        for (std::size_t i = 0; i < count; ++i) permutation[i] = static_cast<std::uint32_t>(load(i));
        raw = {};
        encoded = compress(std::span<const std::uint32_t>(permutation), std::move(options));
    } else {
        std::vector<std::uint64_t> permutation(count);
        for (std::size_t i = 0; i < count; ++i) permutation[i] = load(i);
        raw = {};
        encoded = compress(std::span<const std::uint64_t>(permutation), std::move(options));
    }
    std::ofstream out(output, std::ios::binary | std::ios::trunc);
    out.write(reinterpret_cast<const char*>(encoded.data()),
              static_cast<std::streamsize>(encoded.size()));
    if (!out) throw std::runtime_error("failed to write compressed permutation file");
}

bool backend_available(Backend backend) noexcept {
    return (backend == Backend::automatic || backend == Backend::hip) && detail::hip_available();
}

const char* backend_name(Backend backend) noexcept {
    switch (backend) {
    case Backend::automatic: return "automatic";
    case Backend::portable_cpu: return "portable-cpu";
    case Backend::metal: return "metal";
    case Backend::cuda: return "cuda";
    case Backend::avx512: return "avx512";
    case Backend::hip: return "hip";
    }
    return "unknown";
}

namespace detail {

bool hip_available() noexcept { return benes_hip_available() != 0; }

std::vector<std::byte> compress_compact_hip(std::span<const std::uint32_t> permutation,
                                            FormatOptions format) {
    return compress_hip(permutation, format);
}

std::vector<std::byte> compress_compact_hip(std::span<const std::uint64_t> permutation,
                                            FormatOptions format) {
    return compress_hip(permutation, format);
}

double last_device_seconds() noexcept { return benes_hip_last_device_seconds(); }

}  // namespace detail
}  // namespace benes
