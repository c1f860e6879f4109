// Scorer graph topology built on the device from the graph's face indices
// (graph_topology.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace blitz_hip {

struct Graph;
// From g.indices (g.faces x 3, resident) and g.vertices: offsets,
// neighbors, src, reverse, face_offsets, faces_of and boundary, equal to
// build_topology()'s arrays, plus g.actions and g.corners. Waits for
// `stream` once (the edge count sizes the action arrays).
void build_graph_topology(Graph& g, hipStream_t stream);

}  // namespace blitz_hip
