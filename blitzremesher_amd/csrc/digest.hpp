// Identity digests of mesh streams. Two independent 64-bit digests (an
// FNV-1a style fold and a multiplicative mix, over 8-byte words, then the
// remaining bytes) of the streams blitz's identical() compares
// (src/evaluate.cpp:77-82). Equal digests are only a hint: the audit confirms
// identity by comparing the streams themselves before it skips any view.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace blitz_hip {

struct Digest {
    uint64_t a = 1469598103934665603ull, b = 0x9E3779B97F4A7C15ull;
    void word(uint64_t w) {
        a = (a ^ w) * 1099511628211ull;
        b = (b ^ w) * 0xff51afd7ed558ccdull;
        b ^= b >> 29;
    }
    void bytes(const void* p, size_t n) {
        const auto* c = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, c + i, 8);
            word(w);
        }
        for (; i < n; ++i) word(c[i]);
    }
    void value(uint64_t v) { word(v); }
};

// The identity digest of a mesh from its vertex-stream digest state and its
// triangle streams (the order host_mesh hashes them in).
inline void digest_triangles(Digest& d, const uint32_t* indices, size_t index_count, const uint16_t* materials,
                      size_t material_count, const uint8_t* double_sided, size_t double_sided_count) {
    d.value(index_count);
    d.bytes(indices, index_count * 4);
    d.value(material_count);
    d.bytes(materials, material_count * 2);
    d.value(double_sided_count);
    d.bytes(double_sided, double_sided_count);
}

}  // namespace blitz_hip
