// Scorer graph topology on the device, from the graph's uploaded face
// indices: the undirected neighbour CSR with ascending rows (degenerate
// face edges dropped), the source vertex and the reverse entry of every
// directed action, per-vertex counts of boundary edges (edges with exactly
// one face corner pair) and the vertex -> face CSR in ascending face order
// (a face listed once per corner of the vertex). The mesh states of an
// executor run are ranked from them without a host pass over the faces.
//
// Packed keys: with b = bits of (vertices - 1), an edge (min, max) is
// min << b | max and a directed entry (from, to) is from << b | to, so the
// sorts run over 2b bits; a degenerate corner edge gets the all-ones key of
// 2b bits, which sorts after every edge (min < max <= 2^b - 1 keeps valid
// keys below it).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "device_sort.hpp"
#include "graph_topology.hpp"
#include "scorer.hpp"

namespace blitz_hip {
namespace {

constexpr uint32_t kThreads = 256;

uint32_t blocks_for(size_t n) { return uint32_t(std::max<size_t>(1, (n + kThreads - 1) / kThreads)); }

__global__ void corner_edge_kernel(const uint32_t* __restrict__ idx, uint32_t corners, int bits,
                                   uint64_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= corners) return;
    const uint32_t a = idx[i], b = idx[i % 3 == 2 ? i - 2 : i + 1];
    const uint64_t none = (uint64_t(1) << (2 * bits)) - 1;
    keys[i] = a == b ? none : (uint64_t(min(a, b)) << bits) | max(a, b);
}

// Edges = runs minus the trailing run of degenerate corners, if any.
__global__ void edge_total_kernel(const uint64_t* __restrict__ unique, const uint32_t* __restrict__ runs, int bits,
                                  uint32_t* __restrict__ total) {
    if (threadIdx.x != 0) return;
    const uint32_t r = *runs;
    const uint64_t none = (uint64_t(1) << (2 * bits)) - 1;
    *total = r && unique[r - 1] == none ? r - 1 : r;
}

__global__ void boundary_kernel(const uint64_t* __restrict__ unique, const uint32_t* __restrict__ counts,
                                uint32_t edges, int bits, uint32_t* __restrict__ boundary) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= edges || counts[e] != 1) return;
    const uint64_t mask = (uint64_t(1) << bits) - 1;
    atomicAdd(&boundary[uint32_t(unique[e] >> bits)], 1u);
    atomicAdd(&boundary[uint32_t(unique[e] & mask)], 1u);
}

__global__ void directed_kernel(const uint64_t* __restrict__ unique, uint32_t edges, int bits,
                                uint64_t* __restrict__ keys) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= edges) return;
    const uint64_t mask = (uint64_t(1) << bits) - 1;
    const uint64_t a = unique[e] >> bits, b = unique[e] & mask;
    keys[2 * e] = unique[e];
    keys[2 * e + 1] = (b << bits) | a;
}

// Row starts of a list sorted by row: the entry that opens a row writes
// the start of every row from the previous entry's row (exclusive) to its
// own, so rows without entries get the next row's start; the last entry
// closes the remaining rows with the entry count.
__device__ void row_starts(uint32_t k, uint32_t n, uint32_t row, int64_t previous_row, uint32_t rows,
                           uint32_t* __restrict__ offsets) {
    for (int64_t v = previous_row + 1; v <= int64_t(row); ++v) offsets[v] = k;
    if (k == n - 1)
        for (uint32_t v = row + 1; v <= rows; ++v) offsets[v] = n;
}

__global__ void neighbor_csr_kernel(const uint64_t* __restrict__ sorted, uint32_t n, int bits, uint32_t vertices,
                                    uint32_t* __restrict__ offsets, uint32_t* __restrict__ neighbors,
                                    uint32_t* __restrict__ src) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint64_t mask = (uint64_t(1) << bits) - 1;
    const uint32_t from = uint32_t(sorted[k] >> bits);
    neighbors[k] = uint32_t(sorted[k] & mask);
    src[k] = from;
    const int64_t previous = k ? int64_t(sorted[k - 1] >> bits) : -1;
    if (previous != int64_t(from) || k == n - 1) row_starts(k, n, from, previous == int64_t(from) ? from : previous, vertices, offsets);
}

// reverse[k]: the entry to -> from, found by binary search in row `to`.
__global__ void reverse_kernel(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ neighbors,
                               const uint32_t* __restrict__ src, uint32_t n, uint32_t* __restrict__ reverse) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t to = neighbors[k], from = src[k];
    uint32_t lo = offsets[to], hi = offsets[to + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (neighbors[mid] < from)
            lo = mid + 1;
        else
            hi = mid;
    }
    reverse[k] = lo;
}

__global__ void face_csr_kernel(const uint32_t* __restrict__ sorted_vertex, const uint32_t* __restrict__ corner,
                                uint32_t n, uint32_t vertices, uint32_t* __restrict__ face_offsets,
                                uint32_t* __restrict__ faces_of) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = sorted_vertex[k];
    faces_of[k] = corner[k] / 3;
    const int64_t previous = k ? int64_t(sorted_vertex[k - 1]) : -1;
    if (previous != int64_t(v) || k == n - 1) row_starts(k, n, v, previous == int64_t(v) ? v : previous, vertices, face_offsets);
}

__global__ void zero_kernel(uint32_t* __restrict__ p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

void build_graph_topology(Graph& g, hipStream_t stream) {
    const uint32_t nv = g.vertices, corners = g.faces * 3;
    if (size_t(g.faces) * 3 > (size_t(1) << 31) || nv > (1u << 30)) throw std::invalid_argument("graph too large");
    int bits = 1;
    while (bits < 32 && (uint64_t(1) << bits) < nv) ++bits;
    const int key_bits = 2 * bits;
    g.corners = corners;
    g.offsets.resize((size_t(nv) + 1) * 4);
    g.face_offsets.resize((size_t(nv) + 1) * 4);
    g.boundary.resize(size_t(nv) * 4);
    g.faces_of.resize(size_t(corners) * 4);
    // Scratch: corner keys / directed keys, their sorted copies, unique
    // edges, run counts, corner positions, edge totals and sort workspace.
    const size_t entries = std::max<size_t>(1, corners);
    g.build_keys.resize(entries * 2 * 8);
    g.build_sorted.resize(entries * 2 * 8);
    g.build_unique.resize(entries * 8);
    g.build_counts.resize(entries * 4 + 16);
    g.build_values.resize(entries * 4);
    uint64_t* keys = g.build_keys.as<uint64_t>();
    uint64_t* sorted = g.build_sorted.as<uint64_t>();
    uint64_t* unique = g.build_unique.as<uint64_t>();
    uint32_t* counts = g.build_counts.as<uint32_t>();
    uint32_t* runs = counts + entries;
    uint32_t* total = runs + 1;
    size_t need = 0, bytes = 0;
    (void)radix_sort_pairs(nullptr, bytes, keys, sorted, nullptr, nullptr, uint32_t(2 * entries), 0, key_bits, stream);
    need = std::max(need, bytes);
    (void)run_length_encode(nullptr, bytes, sorted, unique, counts, runs, corners, stream);
    need = std::max(need, bytes);
    (void)radix_sort_pairs(nullptr, bytes, g.indices.as<uint32_t>(), reinterpret_cast<uint32_t*>(sorted), nullptr,
                           g.build_values.as<uint32_t>(), corners, 0, bits, stream);
    need = std::max(need, bytes);
    g.build_temp.resize(need);

    // Undirected edges and their corner counts.
    corner_edge_kernel<<<blocks_for(corners), kThreads, 0, stream>>>(g.indices.as<uint32_t>(), corners, bits, keys);
    bytes = g.build_temp.bytes;
    hip_check(radix_sort_pairs(g.build_temp.p, bytes, keys, sorted, nullptr, nullptr, corners, 0, key_bits, stream),
          "sort corner edges");
    bytes = g.build_temp.bytes;
    hip_check(run_length_encode(g.build_temp.p, bytes, sorted, unique, counts, runs, corners, stream), "edge runs");
    edge_total_kernel<<<1, 64, 0, stream>>>(unique, runs, bits, total);
    uint32_t edges = 0;
    hip_check(hipMemcpyAsync(&edges, total, 4, hipMemcpyDeviceToHost, stream), "download edge count");
    hip_check(hipStreamSynchronize(stream), "edge count");
    if (uint64_t(edges) * 2 > UINT32_MAX) throw std::invalid_argument("graph too large");
    const uint32_t n = 2 * edges;
    g.actions = n;
    g.neighbors.resize(size_t(n) * 4);
    g.src.resize(size_t(n) * 4);
    g.reverse.resize(size_t(n) * 4);

    zero_kernel<<<blocks_for(nv), kThreads, 0, stream>>>(g.boundary.as<uint32_t>(), nv);
    boundary_kernel<<<blocks_for(edges), kThreads, 0, stream>>>(unique, counts, edges, bits, g.boundary.as<uint32_t>());
    if (n) {
        directed_kernel<<<blocks_for(edges), kThreads, 0, stream>>>(unique, edges, bits, keys);
        bytes = g.build_temp.bytes;
        hip_check(radix_sort_pairs(g.build_temp.p, bytes, keys, sorted, nullptr, nullptr, n, 0, key_bits, stream),
              "sort directed entries");
        neighbor_csr_kernel<<<blocks_for(n), kThreads, 0, stream>>>(sorted, n, bits, nv, g.offsets.as<uint32_t>(),
                                                                    g.neighbors.as<uint32_t>(), g.src.as<uint32_t>());
        reverse_kernel<<<blocks_for(n), kThreads, 0, stream>>>(g.offsets.as<uint32_t>(), g.neighbors.as<uint32_t>(),
                                                               g.src.as<uint32_t>(), n, g.reverse.as<uint32_t>());
    } else {
        zero_kernel<<<blocks_for(nv + 1), kThreads, 0, stream>>>(g.offsets.as<uint32_t>(), nv + 1);
    }
    if (corners) {
        // Corners by vertex; the stable sort keeps ascending corner (and so
        // face) order within a vertex.
        uint32_t* sorted_vertex = reinterpret_cast<uint32_t*>(sorted);
        bytes = g.build_temp.bytes;
        hip_check(radix_sort_pairs(g.build_temp.p, bytes, g.indices.as<uint32_t>(), sorted_vertex, nullptr,
                               g.build_values.as<uint32_t>(), corners, 0, bits, stream),
              "sort corners");
        face_csr_kernel<<<blocks_for(corners), kThreads, 0, stream>>>(sorted_vertex, g.build_values.as<uint32_t>(),
                                                                      corners, nv, g.face_offsets.as<uint32_t>(),
                                                                      g.faces_of.as<uint32_t>());
    } else {
        zero_kernel<<<blocks_for(nv + 1), kThreads, 0, stream>>>(g.face_offsets.as<uint32_t>(), nv + 1);
    }
    hip_check(hipGetLastError(), "graph topology launch");
}

}  // namespace blitz_hip
