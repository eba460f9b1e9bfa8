// Shared error handling for the loki_amd library.
//
// Errors are C++ exceptions inside the library (the reference's convention,
// lib/loki/exceptions.hpp:11-25 and cuda_utils.cuh:129-236); the C-ABI layer
// converts them to status codes plus a thread-local message.
#pragma once

#include <cstddef>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>

namespace loki_amd {

using SizeType = std::size_t;

/// Invalid argument / configuration (maps to LOKI_ERR_INVALID).
struct InvalidArgument : std::runtime_error {
    using std::runtime_error::runtime_error;
};

/// HIP runtime failure (maps to LOKI_ERR_HIP).
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] inline void throw_invalid(const std::string& msg, const char* file, int line) {
    std::ostringstream os;
    os << msg << " [" << file << ":" << line << "]";
    throw InvalidArgument(os.str());
}

}  // namespace loki_amd

#define LOKI_REQUIRE(cond, msg)                                                  \
    do {                                                                         \
        if (!(cond)) {                                                           \
            ::loki_amd::throw_invalid((msg), __FILE__, __LINE__);                \
        }                                                                        \
    } while (0)
