// Classical quadric edge-collapse reducer (host C++); see simplify.cpp.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace blitz_hip {
// Returns the surviving triangles (3 indices each, into the input vertex
// array) after greedy quadric collapses down to at most `target` triangles,
// or fewer collapses when no legal one remains.
std::vector<uint32_t> reduce_reuse(const float* xyz, size_t vertices, const uint32_t* indices, size_t faces,
                                   size_t target, uint64_t* collapses);
}  // namespace blitz_hip
