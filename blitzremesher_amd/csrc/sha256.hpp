// SHA-256 (FIPS 180-4), for the model files' checksum trailer and hash.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace blitz_hip {

// Lower-case hexadecimal SHA-256 of n bytes.
std::string sha256_hex(const void* data, size_t n);

}  // namespace blitz_hip
