// Local candidate patches (host C++): seeded breadth-first face regions of a
// mesh, each re-meshed by legal interior edge collapses onto existing
// vertices. A patch replaces its region's faces by the re-meshed faces; the
// candidate mesh of patch p is the source faces not in region p, in source
// order, followed by p's new faces (same vertex array).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace blitz_hip {

struct Patches {
    std::vector<uint32_t> removed_offsets{0}, removed;  // CSR over patches: source face ids, ascending
    std::vector<uint32_t> added_offsets{0}, added;      // CSR over patches: 3 vertex ids per new face
};

// `count` patches with region sizes uniform in [min_faces, max_faces];
// seeded (splitmix64) and independent of thread count.
Patches local_patches(const float* xyz, size_t vertices, const uint32_t* indices, size_t faces, uint32_t count,
                      uint64_t seed, uint32_t min_faces, uint32_t max_faces);

// Candidate index stream of patch p (source faces minus the region, then the
// new faces).
std::vector<uint32_t> patch_candidate(const uint32_t* indices, size_t faces, const uint32_t* removed,
                                      size_t removed_count, const uint32_t* added, size_t added_faces);

}  // namespace blitz_hip
