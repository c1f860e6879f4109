// Classical quadric edge collapse on the device (reduce.hip).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace blitz_hip {

struct HostMesh;

struct ReduceReport {
    uint32_t passes = 0;          // passes run (src/simplify.cpp:158)
    uint32_t rounds = 0;          // greedy rounds over all passes
    uint64_t collapses = 0;       // collapses over all passes
    uint32_t tail_collapses = 0;  // collapses taken after a budget skip
};

// blitz::reduce with OutputMode::Reuse and the default quadric objective
// (src/simplify.cpp:43-329, 337-367): the surviving triangles, as indices
// into the input's vertices, in the reference's order. Positions, indices,
// materials (material junctions lock) and vertex normals (the bend cost) are
// read; a mesh whose faces are all degenerate comes back unchanged.
std::vector<uint32_t> reduce_reuse_device(int device, const HostMesh& m, size_t target_triangles,
                                          ReduceReport* report);

}  // namespace blitz_hip
