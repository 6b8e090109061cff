// SPDX-License-Identifier: Apache-2.0
// Status codes carried across the C ABI.  Each maps onto the exception class
// the reference throws for the same condition (std::invalid_argument,
// std::overflow_error, std::length_error, std::logic_error,
// std::runtime_error; see src/benes.cpp and src/internal.hpp).
#pragma once

#include <string>

#include "../../include/benes_hip.h"

namespace benes_hip {

enum Code : int {
    kOk = BENES_HIP_OK,
    kInvalidArgument = BENES_HIP_INVALID_ARGUMENT,
    kOverflow = BENES_HIP_OVERFLOW,
    kLength = BENES_HIP_LENGTH,
    kLogic = BENES_HIP_LOGIC,
    kRuntime = BENES_HIP_RUNTIME,
    kBufferTooSmall = BENES_HIP_BUFFER_TOO_SMALL,
    kPending = BENES_HIP_PENDING,
    kOutOfRange = BENES_HIP_OUT_OF_RANGE,
};

struct Status {
    int code = kOk;
    std::string message;
    void set(int c, const std::string& m) {
        if (code == kOk) {
            code = c;
            message = m;
        }
    }
    bool ok() const { return code == kOk; }
};

}  // namespace benes_hip
