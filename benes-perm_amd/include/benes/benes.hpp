// SPDX-License-Identifier: Apache-2.0
// C++ interface of the gfx950 compression backend.
//
// The names, defaults and argument meanings are the reference's public API for
// compression (include/benes/benes.hpp:13-61 of bramcohen/benes-perm) with one
// enumerator added, Backend::hip.  Only the compression path is provided here:
// Backend::automatic resolves to hip, and any other backend reports itself
// unavailable exactly as the reference does for a backend it was built
// without (src/benes.cpp:437-442).  compress_file is the reference's
// (src/benes.cpp:498-553), over compress().  Lookup, views and file readers stay with the
// reference library, which reads the byte-identical blobs this produces.
//
// detail::hip_available / detail::compress_compact_hip are the backend
// plug-in functions, the counterparts of detail::cuda_available /
// detail::compress_compact_cuda (src/internal.hpp:1191-1198), for the
// reference's own dispatcher to call.
#pragma once

#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <span>
#include <vector>

namespace benes {

enum class Backend {
    automatic,
    portable_cpu,
    metal,
    cuda,
    avx512,
    hip,
};

enum class MiddlePlacement : std::uint8_t {
    separate,
    embedded_input,
    embedded_output,
};

enum class InputValueWidth : std::uint8_t {
    bits32 = 4,
    bits64 = 8,
};

struct FormatOptions {
    // Middle groups contain up to 2^middle_group_log2 entries.
    std::uint8_t middle_group_log2 = 5;
    // Each value is a base-2 logarithmic cluster size.
    std::vector<std::uint8_t> input_clusters;
    std::vector<std::uint8_t> output_clusters;
    MiddlePlacement middle_placement = MiddlePlacement::separate;
};

struct CompressOptions {
    Backend backend = Backend::automatic;
    std::size_t threads = 0;  // ignored by the GPU backend, as by the reference's
    FormatOptions format;
};

[[nodiscard]] std::vector<std::byte> compress(std::span<const std::uint32_t> permutation,
                                              CompressOptions options = {});
[[nodiscard]] std::vector<std::byte> compress(std::span<const std::uint64_t> permutation,
                                              CompressOptions options = {});

// Raw little-endian values from `input`, the blob to `output`: the
// reference's compress_file (src/benes.cpp:498-553), reading the whole file,
// compressing it with compress() and writing the blob, with the reference's
// errors in its order.
void compress_file(const std::filesystem::path& input, const std::filesystem::path& output,
                   InputValueWidth input_width = InputValueWidth::bits64,
                   CompressOptions options = {});

[[nodiscard]] bool backend_available(Backend backend) noexcept;
[[nodiscard]] const char* backend_name(Backend backend) noexcept;

namespace detail {

bool hip_available() noexcept;
std::vector<std::byte> compress_compact_hip(std::span<const std::uint32_t> permutation,
                                            FormatOptions format);
std::vector<std::byte> compress_compact_hip(std::span<const std::uint64_t> permutation,
                                            FormatOptions format);
double last_device_seconds() noexcept;

// Throws the exception class the reference uses for a C ABI status code
// (BENES_HIP_INVALID_ARGUMENT -> std::invalid_argument, _OVERFLOW ->
// std::overflow_error, _LENGTH -> std::length_error, _LOGIC ->
// std::logic_error, _OUT_OF_RANGE -> std::out_of_range, anything else ->
// std::runtime_error) with `message`.
[[noreturn]] void throw_status(int code, const char* message);

}  // namespace detail

}  // namespace benes
