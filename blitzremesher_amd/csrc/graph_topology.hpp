// Scorer graph topology built on the device from the graph's face indices
// (graph_topology.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace blitz_hip {

struct Graph;
// From g.indices (g.faces x 3, resident) and g.vertices: the neighbour CSR
// (offsets, neighbors; rows ascending, degenerate face edges dropped), the
// source (src) and reverse entry of every directed action, the vertex ->
// face CSR in ascending face order (face_offsets, faces_of) and per-vertex
// boundary-edge counts (boundary), plus g.actions and g.corners. Waits for
// `stream` once (the edge count sizes the action arrays).
void build_graph_topology(Graph& g, hipStream_t stream);

}  // namespace blitz_hip
