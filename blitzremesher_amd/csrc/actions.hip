// Collapse actions on gfx950: device topology, legality, greedy independent
// batches, trial edits and per-action patches, with the semantics of the
// reference's CPU ActionState (src/neural/action.cpp).
//
// Topology (src/neural/action.cpp:27-145). Vertices with bit-equal positions
// (-0 read as +0) weld to the smallest index: three stable radix passes sort
// (x, y, z, index) and an inclusive max-scan carries each run's head. Face
// corners give the per-vertex face lists (a stable sort by geometric vertex
// keeps faces ascending) and the undirected edge keys; a run-length pass
// counts each edge's corners (1: boundary, >2: non-manifold). Neighbour lists
// come from sorting both directions of every edge. A vertex whose link is not
// one connected fan is marked invalid by a per-vertex relaxation over its
// neighbour slots.
//
// Legality (src/neural/action.cpp:146-256), one thread per directed edge:
// the wedge mapping, boundary rule, link condition (common neighbours are
// exactly the opposite vertices), material junctions, duplicate and
// degenerate results, the 0.05 fold bound and UV orientation, with the same
// float/double arithmetic. Legal actions are compacted in enumeration order.
//
// Selection (src/neural/action.cpp:261-304) is inherently sequential (each
// pick blocks its footprint for later ones); one wavefront walks the rank
// order, testing and claiming a footprint (both endpoints and their
// neighbours) 64 neighbour slots at a time.
#include "actions.hpp"
#include "device_sort.hpp"
#include "vec_device.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <future>
#include <stdexcept>

namespace blitz_hip {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

using Topo = ActionTopo;

__global__ void weld_keys_kernel(const float4* __restrict__ p, uint32_t n, uint32_t* kx, uint32_t* ky, uint32_t* kz,
                                 uint32_t* ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = p[i];
    kx[i] = __float_as_uint(v.x == 0 ? 0.f : v.x);
    ky[i] = __float_as_uint(v.y == 0 ? 0.f : v.y);
    kz[i] = __float_as_uint(v.z == 0 ? 0.f : v.z);
    ids[i] = i;
}

__global__ void gather_kernel(const uint32_t* __restrict__ key, const uint32_t* __restrict__ order, uint32_t n,
                              uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = key[order[i]];
}

__global__ void weld_heads_kernel(const uint32_t* __restrict__ order, const uint32_t* kx, const uint32_t* ky,
                                  const uint32_t* kz, uint32_t n, uint32_t* head) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    bool same = false;
    if (k) {
        const uint32_t a = order[k], b = order[k - 1];
        same = kx[a] == kx[b] && ky[a] == ky[b] && kz[a] == kz[b];
    }
    head[k] = same ? 0 : k;
}

__global__ void weld_assign_kernel(const uint32_t* __restrict__ order, const uint32_t* __restrict__ head, uint32_t n,
                                   uint32_t* geometry) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) geometry[order[k]] = order[head[k]];
}

// ---- Topology by counting (src/neural/action.cpp:58-104 builds the same
// lists by sorting): per-vertex face and adjacency lists are filled by
// atomic cursors and then each list is sorted by its own thread, which
// gives the reference's ascending orders without a device-wide sort.

// Each corner adds its face to its geometric vertex's face list and, when
// its edge is not degenerate, each end to the other's adjacency list; a
// degenerate corner marks the face's vertices invalid (action.cpp:63-68).
__global__ void topo_count_kernel(const uint32_t* __restrict__ indices, const uint32_t* __restrict__ geometry,
                                  uint32_t corners, uint32_t* face_count, uint32_t* adj_count, uint8_t* invalid) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= corners) return;
    const uint32_t f = c / 3, j = c % 3;
    const uint32_t a = geometry[indices[c]], b = geometry[indices[3 * f + (j + 1) % 3]];
    atomicAdd(&face_count[a], 1u);
    if (a == b) {
        for (uint32_t t = 0; t < 3; ++t) invalid[geometry[indices[3 * f + t]]] = 1;
        return;
    }
    atomicAdd(&adj_count[a], 1u);
    atomicAdd(&adj_count[b], 1u);
}

__global__ void topo_fill_kernel(const uint32_t* __restrict__ indices, const uint32_t* __restrict__ geometry,
                                 uint32_t corners, const uint32_t* __restrict__ face_offsets,
                                 const uint32_t* __restrict__ adj_offsets, uint32_t* face_cursor, uint32_t* adj_cursor,
                                 uint32_t* faces, uint32_t* adj) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= corners) return;
    const uint32_t f = c / 3, j = c % 3;
    const uint32_t a = geometry[indices[c]], b = geometry[indices[3 * f + (j + 1) % 3]];
    faces[face_offsets[a] + atomicAdd(&face_cursor[a], 1u)] = f;
    if (a == b) return;
    adj[adj_offsets[a] + atomicAdd(&adj_cursor[a], 1u)] = b;
    adj[adj_offsets[b] + atomicAdd(&adj_cursor[b], 1u)] = a;
}

// Ascending in place: insertion sort for short lists, heap sort otherwise
// (a vertex shared by very many faces stays O(d log d)).
__device__ void sort_list(uint32_t* p, uint32_t n) {
    if (n <= 32) {
        for (uint32_t i = 1; i < n; ++i) {
            const uint32_t v = p[i];
            uint32_t k = i;
            for (; k > 0 && p[k - 1] > v; --k) p[k] = p[k - 1];
            p[k] = v;
        }
        return;
    }
    const auto sift = [&](uint32_t root, uint32_t end) {
        for (;;) {
            uint32_t child = 2 * root + 1;
            if (child >= end) return;
            if (child + 1 < end && p[child] < p[child + 1]) ++child;
            if (!(p[root] < p[child])) return;
            const uint32_t t = p[root];
            p[root] = p[child];
            p[child] = t;
            root = child;
        }
    };
    for (uint32_t start = n / 2; start-- > 0;) sift(start, n);
    for (uint32_t end = n - 1; end > 0; --end) {
        const uint32_t t = p[0];
        p[0] = p[end];
        p[end] = t;
        sift(0, end);
    }
}

// Per vertex u: its face list ascending (the reference fills it in face
// order), its adjacency sorted; each run of equal neighbours is one edge
// whose length is the edge's face count (action.cpp:77-92: one face marks
// both ends boundary, more than two marks them invalid; each end sets its
// own flag here). degree = distinct neighbours, up = those above u.
// Per-vertex lists longer than kShortList entries are sorted by
// long_sort_kernel, one wave per vertex in LDS (up to kLongList entries):
// a thread walking a high-valence vertex's list (the pole of a
// latitude-longitude sphere has one neighbour per longitude) through
// global memory would hold up the whole build.
constexpr uint32_t kShortList = 32;
constexpr uint32_t kLongList = 1024;

// Stable rank sort of the per-vertex lists with more than kShortList
// entries, one wave per vertex (four per workgroup): the wave copies the
// list into LDS and lane i writes entry i to the number of entries smaller
// than it or equal and earlier. Adjacency lists first, then face lists.
// Longer lists than kLongList are heap-sorted in place by the wave's first
// lane.
__global__ __launch_bounds__(256) void long_sort_kernel(uint32_t n, const uint32_t* __restrict__ face_offsets,
                                                        uint32_t* faces, const uint32_t* __restrict__ adj_offsets,
                                                        uint32_t* adj) {
    __shared__ uint32_t lds[4][kLongList];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t u = blockIdx.x * 4 + wave;
    for (int which = 0; which < 2; ++which) {
        uint32_t* list = nullptr;
        uint32_t m = 0;
        if (u < n) {
            const uint32_t* offsets = which ? face_offsets : adj_offsets;
            list = (which ? faces : adj) + offsets[u];
            m = offsets[u + 1] - offsets[u];
        }
        const bool mine = m > kShortList && m <= kLongList;
        if (mine)
            for (uint32_t i = lane; i < m; i += 64) lds[wave][i] = list[i];
        __syncthreads();
        if (mine) {
            for (uint32_t i = lane; i < m; i += 64) {
                const uint32_t v = lds[wave][i];
                uint32_t rank = 0;
                for (uint32_t j = 0; j < m; ++j) {
                    const uint32_t w = lds[wave][j];
                    rank += uint32_t(w < v) | uint32_t(w == v && j < i);
                }
                list[rank] = v;
            }
        } else if (m > kLongList && lane == 0) {
            sort_list(list, m);
        }
        __syncthreads();
    }
}

__global__ void topo_vertex_kernel(uint32_t n, const uint32_t* __restrict__ face_offsets, uint32_t* faces,
                                   const uint32_t* __restrict__ adj_offsets, uint32_t* adj, uint32_t* degree,
                                   uint32_t* up, uint8_t* boundary, uint8_t* invalid) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u == 0) {
        // The scans' last entries (offsets[n], edge_offsets[n] = totals).
        degree[n] = 0;
        up[n] = 0;
    }
    if (u >= n) return;
    // Longer lists are sorted already (long_sort_kernel).
    const uint32_t d = face_offsets[u + 1] - face_offsets[u];
    if (d <= kShortList) sort_list(faces + face_offsets[u], d);
    uint32_t* a = adj + adj_offsets[u];
    const uint32_t m = adj_offsets[u + 1] - adj_offsets[u];
    if (m <= kShortList) sort_list(a, m);
    uint32_t distinct = 0, above = 0;
    bool open = false, bad = false;
    for (uint32_t i = 0; i < m;) {
        uint32_t k = i + 1;
        while (k < m && a[k] == a[i]) ++k;
        ++distinct;
        above += a[i] > u;
        open |= k - i == 1;
        bad |= k - i > 2;
        i = k;
    }
    degree[u] = distinct;
    up[u] = above;
    if (open) boundary[u] = 1;
    if (bad) invalid[u] = 1;
}

// Distinct sorted neighbours into the CSR lists and the edges (u, v > u),
// ascending, as keys (u << 32) | v.
__global__ void topo_emit_kernel(uint32_t n, const uint32_t* __restrict__ adj_offsets, const uint32_t* __restrict__ adj,
                                 const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ edge_offsets,
                                 uint32_t* neighbors, unsigned long long* edges) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const uint32_t* a = adj + adj_offsets[u];
    const uint32_t m = adj_offsets[u + 1] - adj_offsets[u];
    uint32_t w = offsets[u], e = edge_offsets[u];
    for (uint32_t i = 0; i < m; ++i) {
        if (i && a[i] == a[i - 1]) continue;
        neighbors[w++] = a[i];
        if (a[i] > u) edges[e++] = (static_cast<unsigned long long>(u) << 32) | a[i];
    }
}


__device__ uint32_t find_neighbor(const Topo& t, uint32_t u, uint32_t v) {
    uint32_t lo = t.offsets[u], hi = t.offsets[u + 1];
    const uint32_t end = hi;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (t.neighbors[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < end && t.neighbors[lo] == v ? lo : kNone;
}

// The two vertices of face f other than u (the first two met, as
// src/neural/action.cpp:120-126); false when there are not two.
__device__ bool link_of(const Topo& t, uint32_t f, uint32_t u, uint32_t other[2]) {
    uint32_t count = 0;
    for (uint32_t j = 0; j < 3; ++j) {
        const uint32_t v = t.geometry[t.indices[3 * f + j]];
        if (v != u && count < 2) other[count++] = v;
    }
    return count == 2;
}

// Sorted geometric vertices of every face, read by the link and legality
// tests in place of three index and three weld lookups per face.
__global__ void face_geo_kernel(const uint32_t* __restrict__ indices, const uint32_t* __restrict__ geometry,
                                uint32_t faces, uint4* out) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces) return;
    uint32_t a = geometry[indices[3 * f]], b = geometry[indices[3 * f + 1]], c = geometry[indices[3 * f + 2]];
    if (a > b) { const uint32_t x = a; a = b; b = x; }
    if (b > c) { const uint32_t x = b; b = c; c = x; }
    if (a > b) { const uint32_t x = a; a = b; b = x; }
    out[f] = make_uint4(a, b, c, 0u);
}

// The two vertices of face f other than u from its sorted triple; false
// unless u occurs exactly once (link_of's two-others condition).
__device__ __forceinline__ bool link_from_geo(uint4 g, uint32_t u, uint32_t other[2]) {
    const uint32_t hits = (g.x == u) + (g.y == u) + (g.z == u);
    if (hits != 1) return false;
    other[0] = g.x == u ? g.y : g.x;
    other[1] = g.z == u ? g.y : g.z;
    return true;
}

// src/neural/action.cpp:107-144: the link of u must be one connected fan
// over all of u's neighbours.
constexpr uint32_t kFan = 32;

__global__ __launch_bounds__(256) void fans_kernel(Topo t, uint8_t* invalid, uint8_t* visited) {
    __shared__ uint8_t fan_slots[256][2 * kFan];
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= t.vertices) return;
    const uint32_t f0 = t.face_offsets[u], f1 = t.face_offsets[u + 1];
    if (f0 == f1 || invalid[u]) return;
    if (f1 - f0 <= kFan && t.offsets[u + 1] - t.offsets[u] <= kFan) {
        // Fans of up to 32 faces and neighbours: each face's link as a pair of
        // neighbour slots (0xFF: not a neighbour), found once and kept in
        // LDS; reach spreads over 32-bit slot masks.
        uint8_t* pa = fan_slots[threadIdx.x];
        uint8_t* pb = pa + kFan;
        const uint32_t d = f1 - f0, n0 = t.offsets[u];
        for (uint32_t k = 0; k < d; ++k) {
            uint32_t other[2];
            if (!link_from_geo(t.face_geo[t.faces[f0 + k]], u, other)) {
                invalid[u] = 1;
                return;
            }
            const uint32_t a = find_neighbor(t, u, other[0]), b = find_neighbor(t, u, other[1]);
            pa[k] = a == kNone ? uint8_t(0xFF) : uint8_t(a - n0);
            pb[k] = b == kNone ? uint8_t(0xFF) : uint8_t(b - n0);
        }
        if (pa[0] == 0xFF) {
            invalid[u] = 1;
            return;
        }
        uint32_t reached = 1u << pa[0];
        for (bool changed = true; changed;) {
            changed = false;
            for (uint32_t k = 0; k < d; ++k) {
                if (pa[k] == 0xFF || pb[k] == 0xFF) continue;
                const uint32_t m = (1u << pa[k]) | (1u << pb[k]);
                if ((reached & m) && (reached & m) != m) {
                    reached |= m;
                    changed = true;
                }
            }
        }
        const uint32_t slots = t.offsets[u + 1] - n0;
        if (reached != (slots == 32 ? ~0u : (1u << slots) - 1)) invalid[u] = 1;
        return;
    }
    // Longer fans within kLongList faces and neighbours: long_fans_kernel.
    if (f1 - f0 <= kLongList && t.offsets[u + 1] - t.offsets[u] <= kLongList) return;
    uint32_t other[2];
    for (uint32_t k = f0; k < f1; ++k)
        if (!link_of(t, t.faces[k], u, other)) {
            invalid[u] = 1;
            return;
        }
    const uint32_t n0 = t.offsets[u], n1 = t.offsets[u + 1];
    for (uint32_t k = n0; k < n1; ++k) visited[k] = 0;
    link_of(t, t.faces[f0], u, other);
    const uint32_t seed = find_neighbor(t, u, other[0]);
    if (seed == kNone) {
        invalid[u] = 1;
        return;
    }
    visited[seed] = 1;
    for (bool changed = true; changed;) {
        changed = false;
        for (uint32_t k = f0; k < f1; ++k) {
            link_of(t, t.faces[k], u, other);
            const uint32_t a = find_neighbor(t, u, other[0]), b = find_neighbor(t, u, other[1]);
            if (a == kNone || b == kNone) continue;
            if (visited[a] != visited[b]) {
                visited[a] = visited[b] = 1;
                changed = true;
            }
        }
    }
    for (uint32_t k = n0; k < n1; ++k)
        if (!visited[k]) {
            invalid[u] = 1;
            return;
        }
}

// The same fan test for vertices with more than kFan faces or neighbours
// (and at most kLongList of each), one wave per vertex: the lanes find each
// face's link as two neighbour slots (link_of, find_neighbor) into LDS, and
// the wave's first lane joins the slots of every link with union-find in LDS
// and checks that every neighbour lies in the first face's component -- the
// component the reference's flood from that face's first link vertex
// reaches.
__global__ __launch_bounds__(256) void long_fans_kernel(Topo t, uint8_t* invalid) {
    __shared__ uint32_t link_a[4][kLongList], link_b[4][kLongList], parent[4][kLongList];
    __shared__ uint32_t broken[4];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t u = blockIdx.x * 4 + wave;
    uint32_t f0 = 0, d = 0, n0 = 0, slots = 0;
    bool mine = false;
    if (u < t.vertices) {
        f0 = t.face_offsets[u];
        d = t.face_offsets[u + 1] - f0;
        n0 = t.offsets[u];
        slots = t.offsets[u + 1] - n0;
        mine = d && !invalid[u] && (d > kFan || slots > kFan) && d <= kLongList && slots <= kLongList;
    }
    if (lane == 0) broken[wave] = 0;
    __syncthreads();
    if (mine) {
        for (uint32_t k = lane; k < d; k += 64) {
            uint32_t other[2];
            uint32_t a = kNone, b = kNone;
            if (!link_of(t, t.faces[f0 + k], u, other)) {
                broken[wave] = 1;
            } else {
                a = find_neighbor(t, u, other[0]);
                b = find_neighbor(t, u, other[1]);
            }
            link_a[wave][k] = a == kNone ? kNone : a - n0;
            link_b[wave][k] = b == kNone ? kNone : b - n0;
        }
        for (uint32_t i = lane; i < slots; i += 64) parent[wave][i] = i;
    }
    __syncthreads();
    if (mine && lane == 0) {
        uint32_t* p = parent[wave];
        const auto find = [&](uint32_t x) {
            while (p[x] != x) {
                p[x] = p[p[x]];
                x = p[x];
            }
            return x;
        };
        bool bad = broken[wave] != 0 || link_a[wave][0] == kNone;
        if (!bad) {
            for (uint32_t k = 0; k < d; ++k) {
                const uint32_t a = link_a[wave][k], b = link_b[wave][k];
                if (a == kNone || b == kNone) continue;
                const uint32_t ra = find(a), rb = find(b);
                if (ra != rb) p[max(ra, rb)] = min(ra, rb);
            }
            const uint32_t root = find(link_a[wave][0]);
            for (uint32_t i = 0; i < slots && !bad; ++i) bad = find(i) != root;
        }
        if (bad) invalid[u] = 1;
    }
}

// include/blitz/mesh.hpp:22-26

__device__ __forceinline__ uint16_t material_of(const Topo& t, uint32_t f) {
    return t.materials ? t.materials[f] : uint16_t(0);
}

__device__ __forceinline__ void sort3(uint32_t v[3]) {
    uint32_t a = v[0], b = v[1], c = v[2];
    if (a > b) { const uint32_t s = a; a = b; b = s; }
    if (b > c) { const uint32_t s = b; b = c; c = s; }
    if (a > b) { const uint32_t s = a; a = b; b = s; }
    v[0] = a;
    v[1] = b;
    v[2] = c;
}

// src/neural/action.cpp:13-15
__device__ __forceinline__ double uv_area(const float2* uv, const uint32_t ids[3]) {
    const float2 a = uv[ids[0]], b = uv[ids[1]], c = uv[ids[2]];
    return (double(b.x) - a.x) * (double(c.y) - a.y) - (double(b.y) - a.y) * (double(c.x) - a.x);
}

__device__ __forceinline__ uint32_t map_id(const LegalAction& a, uint32_t id) {
    for (uint32_t w = 0; w < a.wedges; ++w)
        if (a.wedge_from[w] == id) return a.wedge_to[w];
    return id;
}

// Sorted triple of a face's geometric vertices with `from` replaced by `to`.
__device__ __forceinline__ void moved_triple(uint4 g, uint32_t from, uint32_t to, uint32_t out[3]) {
    out[0] = g.x == from ? to : g.x;
    out[1] = g.y == from ? to : g.y;
    out[2] = g.z == from ? to : g.z;
    sort3(out);
}

// src/neural/action.cpp:146-256
__device__ bool legal_action(const Topo& t, uint32_t from, uint32_t to, LegalAction& a) {
    a.from = from;
    a.to = to;
    a.wedges = 0;
    a.removed = 0;
    if (from == to || t.invalid[from] || t.invalid[to]) return false;
    if (find_neighbor(t, from, to) == kNone) return false;
    const uint32_t f0 = t.face_offsets[from], f1 = t.face_offsets[from + 1];
    const uint32_t g0 = t.face_offsets[to], g1 = t.face_offsets[to + 1];
    if (f0 == f1) return false;
    uint32_t opposite[2], pair_u[2], pair_v[2], wedge[2];
    uint32_t edge_faces = 0, wedges = 0;
    bool junction = false;
    const uint16_t first_material = material_of(t, t.faces[f0]);
    for (uint32_t k = f0; k < f1; ++k) {
        const uint32_t f = t.faces[k];
        junction |= material_of(t, f) != first_material;
        uint32_t u = kNone, v = kNone, w = kNone;
        for (uint32_t j = 0; j < 3; ++j) {
            const uint32_t id = t.indices[3 * f + j], g = t.geometry[id];
            if (g == from)
                u = id;
            else if (g == to)
                v = id;
            else
                w = g;
        }
        if (u == kNone) return false;
        // Distinct wedges of `from`; more than two can never match the at most
        // two mapped pairs.
        bool seen = false;
        for (uint32_t i = 0; i < wedges; ++i) seen |= wedge[i] == u;
        if (!seen) {
            if (wedges == 2) return false;
            wedge[wedges++] = u;
        }
        if (v != kNone) {
            if (w == kNone || edge_faces == 2) return false;
            opposite[edge_faces] = w;
            pair_u[edge_faces] = u;
            pair_v[edge_faces] = v;
            ++edge_faces;
        }
    }
    if (edge_faces < 1 || edge_faces >= t.triangles) return false;
    // A boundary vertex may only move along its boundary.
    if (t.boundary[from] && (!t.boundary[to] || edge_faces != 1)) return false;
    // Common neighbours must be exactly the (distinct) opposite vertices.
    uint32_t nopp = edge_faces;
    if (nopp == 2) {
        if (opposite[0] > opposite[1]) {
            const uint32_t s = opposite[0];
            opposite[0] = opposite[1];
            opposite[1] = s;
        }
        if (opposite[0] == opposite[1]) nopp = 1;
    }
    if (nopp != edge_faces) return false;
    {
        uint32_t i = t.offsets[from], i1 = t.offsets[from + 1], j = t.offsets[to], j1 = t.offsets[to + 1];
        uint32_t matched = 0;
        while (i < i1 && j < j1) {
            const uint32_t x = t.neighbors[i], y = t.neighbors[j];
            if (x < y)
                ++i;
            else if (y < x)
                ++j;
            else {
                if (matched >= nopp || opposite[matched] != x) return false;
                ++matched;
                ++i;
                ++j;
            }
        }
        if (matched != nopp) return false;
    }
    // Each wedge of `from` maps to exactly one destination wedge.
    uint32_t npairs = edge_faces;
    if (npairs == 2) {
        const bool swap = pair_u[0] > pair_u[1] || (pair_u[0] == pair_u[1] && pair_v[0] > pair_v[1]);
        if (swap) {
            uint32_t s = pair_u[0];
            pair_u[0] = pair_u[1];
            pair_u[1] = s;
            s = pair_v[0];
            pair_v[0] = pair_v[1];
            pair_v[1] = s;
        }
        if (pair_u[0] == pair_u[1] && pair_v[0] == pair_v[1]) npairs = 1;
    }
    if (wedges == 2 && wedge[0] > wedge[1]) {
        const uint32_t s = wedge[0];
        wedge[0] = wedge[1];
        wedge[1] = s;
    }
    if (npairs != wedges) return false;
    for (uint32_t i = 0; i < npairs; ++i)
        if (pair_u[i] != wedge[i]) return false;
    a.wedges = npairs;
    for (uint32_t i = 0; i < npairs; ++i) {
        a.wedge_from[i] = pair_u[i];
        a.wedge_to[i] = pair_v[i];
    }
    a.removed = edge_faces;
    // Material sets around both endpoints must agree at a junction.
    if (junction) {
        for (uint32_t k = f0; k < f1; ++k) {
            const uint16_t m = material_of(t, t.faces[k]);
            bool found = false;
            for (uint32_t l = g0; l < g1 && !found; ++l) found = material_of(t, t.faces[l]) == m;
            if (!found) return false;
        }
        for (uint32_t l = g0; l < g1; ++l) {
            const uint16_t m = material_of(t, t.faces[l]);
            bool found = false;
            for (uint32_t k = f0; k < f1 && !found; ++k) found = material_of(t, t.faces[k]) == m;
            if (!found) return false;
        }
    }
    // The moved faces: not degenerate, not duplicating a face around `to` or
    // each other, not folded, not flipped in UV. Every corner of `from` in
    // these faces is one of the mapped wedges (checked above), so a moved
    // face's geometric triple is its triple with `from` replaced by `to`.
    for (uint32_t k = f0; k < f1; ++k) {
        const uint32_t f = t.faces[k];
        const uint4 geo = t.face_geo[f];
        if (geo.x == to || geo.y == to || geo.z == to) continue;  // removed by the collapse
        uint32_t ids[3];
        moved_triple(geo, from, to, ids);
        if (ids[0] == ids[1] || ids[1] == ids[2]) return false;
        // Duplicates, searched among the faces of one of the moved face's
        // two other vertices (the one with fewer faces), which every face
        // that could match contains: a face around `to` with the moved
        // triple, or another moved face -- replacing `from` by `to` is
        // one-to-one on triples holding `from` once and not `to`, so two
        // moved faces coincide after the move exactly when their triples
        // coincide before it. (Scanning the faces around `to` and around
        // `from` instead is quadratic in the valence of a pole.)
        uint32_t o0 = kNone, o1 = kNone;
        for (uint32_t j = 0; j < 3; ++j)
            if (ids[j] != to) (o0 == kNone ? o0 : o1) = ids[j];
        const uint32_t c0 = t.face_offsets[o0 + 1] - t.face_offsets[o0];
        const uint32_t c1 = t.face_offsets[o1 + 1] - t.face_offsets[o1];
        const uint32_t pick = c0 <= c1 ? o0 : o1;
        for (uint32_t l = t.face_offsets[pick], l1 = t.face_offsets[pick + 1]; l < l1; ++l) {
            const uint32_t h = t.faces[l];
            if (h == f) continue;
            const uint4 g = t.face_geo[h];
            if ((g.x == ids[0] && g.y == ids[1] && g.z == ids[2]) || (g.x == geo.x && g.y == geo.y && g.z == geo.z))
                return false;
        }
        uint32_t old[3], next[3];
        for (uint32_t j = 0; j < 3; ++j) {
            old[j] = t.indices[3 * f + j];
            next[j] = map_id(a, old[j]);
        }
        const float4 p0 = t.positions[old[0]], p1 = t.positions[old[1]], p2 = t.positions[old[2]];
        const float4 q0 = t.positions[next[0]], q1 = t.positions[next[1]], q2 = t.positions[next[2]];
        const F3 before = cross3(sub3(p1, p0), sub3(p2, p0));
        const F3 after = cross3(sub3(q1, q0), sub3(q2, q0));
        if (len3(after) <= 1e-15 || dot3(before, after) <= .05 * len3(before) * len3(after)) return false;
        if (t.uv) {
            const double x = uv_area(t.uv, old), y = uv_area(t.uv, next);
            if (fabs(x) > 1e-20 && x * y <= 0) return false;
        }
    }
    return true;
}

__global__ void legal_kernel(Topo t, const unsigned long long* __restrict__ edges, uint32_t count, LegalAction* out,
                             uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) flags[2 * count] = 0;  // the scan's last entry (the legal count)
    if (i >= 2 * count) return;
    const unsigned long long key = edges[i / 2];
    const uint32_t a = uint32_t(key >> 32), b = uint32_t(key);
    LegalAction act;
    const bool ok = (i & 1) ? legal_action(t, b, a, act) : legal_action(t, a, b, act);
    out[i] = act;
    flags[i] = ok;
}

__global__ void compact_kernel(const LegalAction* __restrict__ in, const uint32_t* __restrict__ flags,
                               const uint32_t* __restrict__ slots, uint32_t n, LegalAction* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) out[slots[i]] = in[i];
}

__device__ __forceinline__ uint32_t load_used(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_used(uint32_t* p) {
    __hip_atomic_store(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// src/neural/action.cpp:261-304 on one wavefront, 64 candidates at a time:
// each lane tests one candidate's footprint (both endpoints and their
// neighbours) against the footprints already claimed, the chunk's footprints
// are intersected pairwise through per-vertex lane masks (`touch`, all zero
// between chunks), and the sequential greedy choice is replayed over those
// 64-bit masks in rank order, so the picks are exactly the sequential ones.
// out[0] = picked count, out[1] = 1 when a strict batch (a trial) is not
// independent, out[2] = the position in `rows` of the first pick, out[3..]
// the picked rows.
__global__ __launch_bounds__(64) void claim_kernel(Topo t, const LegalAction* __restrict__ legal,
                                                   const uint32_t* __restrict__ rows, uint32_t n, uint32_t maximum,
                                                   unsigned long long removable, uint32_t strict, uint32_t* used,
                                                   unsigned long long* touch, uint32_t* out) {
    const uint32_t lane = threadIdx.x;
    const unsigned long long bit = 1ull << lane;
    uint32_t picked = 0, error = 0, first = kNone;
    bool done = false;
    for (uint32_t k0 = 0; k0 < n && !done; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool valid = k < n;
        LegalAction a{};
        if (valid) a = legal[rows[k]];
        const auto footprint = [&](auto&& visit) {
            for (uint32_t e = 0; e < 2; ++e) {
                const uint32_t u = e ? a.to : a.from;
                visit(u);
                const uint32_t j1 = t.offsets[u + 1];
                for (uint32_t j = t.offsets[u]; j < j1; ++j) visit(t.neighbors[j]);
            }
        };
        bool hit = false;
        unsigned long long conflict = 0;
        if (valid) {
            footprint([&](uint32_t v) {
                hit |= load_used(&used[v]) != 0;
                atomicOr(&touch[v], bit);
            });
        }
        __threadfence();  // every lane's touches land before any is read back
        if (valid)
            footprint([&](uint32_t v) { conflict |= __hip_atomic_load(&touch[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); });
        conflict &= ~bit;
        // Sequential replay (wave-uniform): candidate j is taken when its
        // removal fits, it meets no earlier claim and no footprint taken
        // earlier in this chunk.
        const unsigned long long hits = __ballot(hit), valids = __ballot(valid);
        const uint32_t before = picked;
        unsigned long long taken = 0;
        for (uint32_t j = 0; j < 64; ++j) {
            if (!((valids >> j) & 1)) break;
            const uint32_t removed = uint32_t(__shfl(int(a.removed), int(j)));
            const unsigned long long meets = __shfl(conflict, int(j));
            if (removed > removable || ((hits >> j) & 1) || (meets & taken)) {
                error |= strict;
                continue;
            }
            taken |= 1ull << j;
            if (!picked) first = k0 + j;
            ++picked;
            removable -= removed;
            if (picked == maximum || !removable) {
                done = true;
                break;
            }
        }
        if (taken & bit) {
            out[3 + before + uint32_t(__popcll(taken & (bit - 1)))] = rows[k];
            footprint([&](uint32_t v) { store_used(&used[v]); });
        }
        if (valid) footprint([&](uint32_t v) { __hip_atomic_store(&touch[v], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); });
        __threadfence();  // claims and cleared masks are seen by the next chunk
    }
    if (lane == 0) {
        out[0] = picked;
        out[1] = error;
        out[2] = first;
    }
}

// claim_kernel with its per-vertex state in LDS hash tables instead of
// global arrays: the chunk's lane masks (`touch`, cleared per chunk) and the
// vertices claimed so far in this call. Insert attempts are counted ahead
// of each insert and capped at three quarters of a table, so probes always
// end; past the cap the kernel sets out[3 + 64] and stops, and the caller
// runs claim_kernel instead. The picks are claim_kernel's.
constexpr uint32_t kTouchSlots = 4096, kClaimSlots = 2048;

__device__ __forceinline__ uint32_t slot_of(uint32_t v, uint32_t slots) { return (v * 2654435761u) & (slots - 1); }

__global__ __launch_bounds__(64) void claim_lds_kernel(Topo t, const LegalAction* __restrict__ legal,
                                                       const uint32_t* __restrict__ rows, uint32_t n, uint32_t maximum,
                                                       unsigned long long removable, uint32_t strict, uint32_t* out) {
    __shared__ uint32_t touch_key[kTouchSlots];
    __shared__ unsigned long long touch_mask[kTouchSlots];
    __shared__ uint32_t claim_key[kClaimSlots];
    __shared__ uint32_t attempts[2];  // touch inserts this chunk, claim inserts this call
    __shared__ uint32_t overflow;
    const uint32_t lane = threadIdx.x;
    const unsigned long long bit = 1ull << lane;
    for (uint32_t i = lane; i < kTouchSlots; i += 64) {
        touch_key[i] = kNone;
        touch_mask[i] = 0;
    }
    for (uint32_t i = lane; i < kClaimSlots; i += 64) claim_key[i] = kNone;
    if (lane == 0) {
        attempts[0] = attempts[1] = 0;
        overflow = 0;
    }
    __syncthreads();
    const auto touch = [&](uint32_t v) {
        if (atomicAdd(&attempts[0], 1u) >= kTouchSlots / 4 * 3) {
            overflow = 1;
            return;
        }
        for (uint32_t h = slot_of(v, kTouchSlots);; h = (h + 1) & (kTouchSlots - 1)) {
            const uint32_t k = atomicCAS(&touch_key[h], kNone, v);
            if (k == kNone || k == v) {
                atomicOr(&touch_mask[h], bit);
                return;
            }
        }
    };
    const auto touched = [&](uint32_t v) -> unsigned long long {
        for (uint32_t h = slot_of(v, kTouchSlots);; h = (h + 1) & (kTouchSlots - 1)) {
            const uint32_t k = touch_key[h];
            if (k == v) return touch_mask[h];
            if (k == kNone) return 0;
        }
    };
    const auto claim = [&](uint32_t v) {
        if (atomicAdd(&attempts[1], 1u) >= kClaimSlots / 4 * 3) {
            overflow = 1;
            return;
        }
        for (uint32_t h = slot_of(v, kClaimSlots);; h = (h + 1) & (kClaimSlots - 1)) {
            const uint32_t k = atomicCAS(&claim_key[h], kNone, v);
            if (k == kNone || k == v) return;
        }
    };
    const auto claimed = [&](uint32_t v) {
        for (uint32_t h = slot_of(v, kClaimSlots);; h = (h + 1) & (kClaimSlots - 1)) {
            const uint32_t k = claim_key[h];
            if (k == v) return true;
            if (k == kNone) return false;
        }
    };
    uint32_t picked = 0, error = 0, first = kNone;
    bool done = false;
    for (uint32_t k0 = 0; k0 < n && !done; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool valid = k < n;
        LegalAction a{};
        if (valid) a = legal[rows[k]];
        const auto footprint = [&](auto&& visit) {
            for (uint32_t e = 0; e < 2; ++e) {
                const uint32_t u = e ? a.to : a.from;
                visit(u);
                const uint32_t j1 = t.offsets[u + 1];
                for (uint32_t j = t.offsets[u]; j < j1; ++j) visit(t.neighbors[j]);
            }
        };
        bool hit = false;
        if (valid)
            footprint([&](uint32_t v) {
                hit |= claimed(v);
                touch(v);
            });
        __syncthreads();
        if (overflow) break;
        unsigned long long conflict = 0;
        if (valid) footprint([&](uint32_t v) { conflict |= touched(v); });
        conflict &= ~bit;
        // Sequential replay, as in claim_kernel.
        const unsigned long long hits = __ballot(hit), valids = __ballot(valid);
        const uint32_t before = picked;
        unsigned long long taken = 0;
        for (uint32_t j = 0; j < 64; ++j) {
            if (!((valids >> j) & 1)) break;
            const uint32_t removed = uint32_t(__shfl(int(a.removed), int(j)));
            const unsigned long long meets = __shfl(conflict, int(j));
            if (removed > removable || ((hits >> j) & 1) || (meets & taken)) {
                error |= strict;
                continue;
            }
            taken |= 1ull << j;
            if (!picked) first = k0 + j;
            ++picked;
            removable -= removed;
            if (picked == maximum || !removable) {
                done = true;
                break;
            }
        }
        if (taken & bit) {
            out[3 + before + uint32_t(__popcll(taken & (bit - 1)))] = rows[k];
            footprint(claim);
        }
        __syncthreads();
        for (uint32_t i = lane; i < kTouchSlots; i += 64) {
            touch_key[i] = kNone;
            touch_mask[i] = 0;
        }
        if (lane == 0) attempts[0] = 0;
        __syncthreads();
        if (overflow) break;
    }
    if (lane == 0) {
        out[0] = picked;
        out[1] = error;
        out[2] = first;
        out[3 + 64] = overflow;
    }
}

// Ranking keys of the legal actions (src/neural/action_gpu.cu:809-844 with
// the scores of src/neural/action.cpp:355-370, tests/neural_action_gpu.cpp:
// 155-160): 1 constant, 3 shortest edge, 4 current plane, 0 learned (the
// logit of the scorer action from -> to, found in the scorer graph's sorted
// neighbour lists). A non-finite score is an error, as in
// src/neural/action.cpp:477-479. -0 sorts as +0 so
// that equal scores keep ascending action order, as a stable sort does.
// rank_kernel mode of per-action policy outputs (rank_policy).
constexpr uint32_t kPolicyScores = 5;

struct RankInput {
    const LegalAction* legal;
    uint32_t count;
    uint32_t mode;
    double diameter;
    const uint32_t* graph_offsets;
    const uint32_t* graph_neighbors;
    const float* outputs;
    uint32_t outputs_per_action;
    float* keys;
    uint32_t* order;
    uint32_t* bad;  // set when a score is not finite
};

__device__ __forceinline__ F3 f3(float4 p) { return {p.x, p.y, p.z}; }

__global__ void rank_kernel(Topo t, RankInput in) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.count) return;
    const LegalAction a = in.legal[i];
    const double scale = fmax(1e-20, in.diameter);
    float score = 0.f;
    if (in.mode == 3) {
        score = shortest_edge_score(t.positions, a.from, a.to, scale);
    } else if (in.mode == 4) {
        score = current_plane_score(t.positions, t.indices, t.face_offsets, t.faces, a.from, a.to, scale);
    } else if (in.mode == 0) {
        uint32_t lo = in.graph_offsets[a.from], hi = in.graph_offsets[a.from + 1];
        const uint32_t end = hi;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (in.graph_neighbors[mid] < a.to)
                lo = mid + 1;
            else
                hi = mid;
        }
        score = lo < end && in.graph_neighbors[lo] == a.to ? in.outputs[size_t(lo) * in.outputs_per_action] : NAN;
    } else if (in.mode == kPolicyScores) {
        // The policy's first output for this legal action
        // (rank_actions, src/neural/action_gpu.cu:821-822).
        score = in.outputs[size_t(i) * in.outputs_per_action];
    }
    if (!isfinite(score)) atomicExch(in.bad, 1u);
    in.keys[i] = score + 0.f;
    in.order[i] = i;
}

__global__ void remap_kernel(const LegalAction* __restrict__ legal, const uint32_t* __restrict__ rows, uint32_t n,
                             uint32_t* remap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LegalAction a = legal[rows[i]];
    for (uint32_t w = 0; w < a.wedges; ++w) remap[a.wedge_from[w]] = a.wedge_to[w];
}

// src/neural/action.cpp:322-340: remapped faces that became degenerate drop out.
__global__ void trial_keep_kernel(Topo t, const uint32_t* __restrict__ remap, uint32_t* next, uint32_t* keep) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= t.triangles) return;
    uint32_t ids[3];
    bool changed = false;
    for (uint32_t j = 0; j < 3; ++j) {
        ids[j] = t.indices[3 * f + j];
        if (remap[ids[j]] != kNone) {
            ids[j] = remap[ids[j]];
            changed = true;
        }
        next[3 * f + j] = ids[j];
    }
    const uint32_t a = t.geometry[ids[0]], b = t.geometry[ids[1]], c = t.geometry[ids[2]];
    keep[f] = !(changed && (a == b || b == c || a == c));
}

__global__ void trial_copy_kernel(const uint32_t* __restrict__ next, const uint32_t* __restrict__ keep,
                                  const uint32_t* __restrict__ slots, const uint16_t* __restrict__ materials,
                                  uint32_t faces, uint32_t* out, uint16_t* out_materials) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces || !keep[f]) return;
    for (uint32_t j = 0; j < 3; ++j) out[3 * slots[f] + j] = next[3 * f + j];
    if (materials) out_materials[slots[f]] = materials[f];
}

// Per patch: faces around `from` removed, the remapped ones not holding `to`
// added (a legal action never makes them degenerate).
__global__ void patch_count_kernel(Topo t, const LegalAction* __restrict__ legal, const uint32_t* __restrict__ rows,
                                   uint32_t n, uint32_t* removed_count, uint32_t* added_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LegalAction a = legal[rows[i]];
    const uint32_t len = t.face_offsets[a.from + 1] - t.face_offsets[a.from];
    removed_count[i] = len;
    added_count[i] = len - a.removed;
}

__global__ void patch_fill_kernel(Topo t, const LegalAction* __restrict__ legal, const uint32_t* __restrict__ rows,
                                  uint32_t n, const uint32_t* __restrict__ removed_at,
                                  const uint32_t* __restrict__ added_at, uint32_t* removed, uint32_t* added) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LegalAction a = legal[rows[i]];
    uint32_t r = removed_at[i], w = added_at[i];
    for (uint32_t k = t.face_offsets[a.from]; k < t.face_offsets[a.from + 1]; ++k) {
        const uint32_t f = t.faces[k];
        removed[r++] = f;
        bool holds_to = false;
        for (uint32_t j = 0; j < 3; ++j) holds_to |= t.geometry[t.indices[3 * f + j]] == a.to;
        if (holds_to) continue;
        for (uint32_t j = 0; j < 3; ++j) added[3 * w + j] = map_id(a, t.indices[3 * f + j]);
        ++w;
    }
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

uint32_t blocks_for(size_t n) { return uint32_t(std::max<size_t>(1, (n + 255) / 256)); }

template <class T> T* grow(DeviceBuffer& b, size_t count) {
    b.resize(std::max<size_t>(count, 1) * sizeof(T));
    return b.as<T>();
}

struct Clock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

}  // namespace

ActionSet::ActionSet(int device, const HostMesh& m, const float* uv) : device_(device) {
    DeviceGuard guard(device_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    build(m, uv);
}

void ActionSet::rebuild(const HostMesh& m, const float* uv) {
    DeviceGuard guard(device_);
    hip_check(hipStreamSynchronize(stream_), "action state");
    build(m, uv);
}

void ActionSet::build(const HostMesh& m, const float* uv) {
    Clock clock;
    const hipStream_t st = stream_;
    legal_count_ = edges_ = ranked_ = 0;
    touch_ready_ = false;
    const uint32_t previous_faces = faces_;
    vertices_ = uint32_t(m.positions.size() / 3);
    faces_ = uint32_t(m.indices.size() / 3);
    if (!vertices_ || !faces_) throw std::invalid_argument("action state needs a non-empty mesh");
    if (uint64_t(faces_) * 6 >= uint64_t(INT32_MAX)) throw std::invalid_argument("mesh too large for 32-bit corners");
    if (staging_ && faces_ > previous_faces) {
        (void)hipHostFree(staging_);
        staging_ = nullptr;
    }
    const uint32_t n = vertices_, corners = 3 * faces_;
    // The weld (geometric vertex of each index) depends on the positions
    // only: a rebuild over the same positions keeps it.
    const bool same_positions = weld_ready_ && weld_positions_ == m.positions;
    if (!same_positions) diameter_ = std::max(1e-20, bounds_of(m.positions.data(), n).radius * 2);
    // Every buffer of the build is sized here, before any launch, so nothing
    // is reallocated under queued work; the buffers persist across rebuilds.
    DeviceBuffer &weld = build_weld_, &weld2 = build_weld2_, &corner_buf = build_corners_,
                 &temp = build_temp_, &all_buf = build_all_, &flag_buf = build_flags_;
    uint32_t* kx = grow<uint32_t>(weld, 6 * size_t(n));
    uint32_t* ky = kx + n;
    uint32_t* kz = ky + n;
    uint32_t* ids = kz + n;
    uint32_t* key_tmp = ids + n;
    uint32_t* ord = key_tmp + n;
    uint32_t* ord2 = grow<uint32_t>(weld2, 2 * size_t(n));
    uint32_t* key_sorted = ord2 + n;
    // Counting topology: per-vertex counts, offsets and cursors, the face
    // and adjacency lists, distinct-neighbour counts and edge offsets.
    const size_t n1 = size_t(n) + 1;
    uint32_t* face_count = grow<uint32_t>(corner_buf, 8 * n1 + 2 * size_t(corners));
    uint32_t* adj_count = face_count + n1;
    uint32_t* adj_offsets = adj_count + n1;
    uint32_t* face_cursor = adj_offsets + n1;  // n
    uint32_t* adj_cursor = face_cursor + n1;   // n
    uint32_t* degree = adj_cursor + n1;
    uint32_t* up = degree + n1;
    uint32_t* edge_offsets = up + n1;
    uint32_t* adj = edge_offsets + n1;         // 2 * corners
    // At most one edge per corner: every per-edge buffer is sized from the
    // corner count, so nothing is reallocated under queued work.
    auto* unique = grow<unsigned long long>(edge_keys_, corners);
    uint32_t* geometry = grow<uint32_t>(geometry_, n);
    uint8_t* invalid = grow<uint8_t>(invalid_, n);
    uint8_t* boundary = grow<uint8_t>(boundary_, n);
    uint32_t* faces = grow<uint32_t>(face_lists_, corners);
    uint32_t* face_offsets = grow<uint32_t>(face_offsets_, n1);
    uint32_t* offsets = grow<uint32_t>(offsets_, n1);
    const size_t max_slots = 2 * size_t(corners);
    uint32_t* neighbors = grow<uint32_t>(neighbors_, max_slots);
    uint8_t* visited = grow<uint8_t>(used_, std::max<size_t>(max_slots, size_t(n) * 4));
    auto* all = grow<LegalAction>(all_buf, max_slots);
    uint32_t* flags = grow<uint32_t>(flag_buf, 2 * (max_slots + 1));
    uint32_t* where = flags + max_slots + 1;
    auto* legal = grow<LegalAction>(legal_, max_slots);
    uint4* face_geo = grow<uint4>(face_geo_, faces_);
    size_t bytes = 16;
    {
        // Workspace of the weld's sorts and scan (run when positions change).
        size_t need = 0;
        auto take = [&](hipError_t e) {
            hip_check(e, "workspace query");
            bytes = std::max(bytes, need);
        };
        take(radix_sort_pairs(nullptr, need, key_tmp, key_sorted, ids, ord, n, 0, 32, st));
        take(inclusive_max(nullptr, need, ord2, key_sorted, n, st));
        take(exclusive_sum(nullptr, need, face_count, face_offsets, uint32_t(n1), st));
        take(exclusive_sum(nullptr, need, flags, where, uint32_t(max_slots + 1), st));
    }
    void* work = grow<char>(temp, bytes);
    auto run = [&](const char* what, auto call) {
        size_t b = bytes;
        hip_check(call(b), what);
    };
    if (!same_positions) {
        std::vector<float4> pos(n);
        for (uint32_t i = 0; i < n; ++i)
            pos[i] = make_float4(m.positions[3 * i], m.positions[3 * i + 1], m.positions[3 * i + 2], 0.f);
        hip_check(hipMemcpyAsync(grow<float4>(positions_, n), pos.data(), n * sizeof(float4), hipMemcpyHostToDevice, st),
                  "upload positions");
        hip_check(hipStreamSynchronize(st), "upload");
    }
    hip_check(hipMemcpyAsync(grow<uint32_t>(indices_, corners), m.indices.data(), size_t(corners) * 4,
                             hipMemcpyHostToDevice, st),
              "upload indices");
    has_materials_ = !m.materials.empty();
    if (has_materials_)
        hip_check(hipMemcpyAsync(grow<uint16_t>(materials_, faces_), m.materials.data(), size_t(faces_) * 2,
                                 hipMemcpyHostToDevice, st),
                  "upload materials");
    if (uv)
        hip_check(hipMemcpyAsync(grow<float2>(uv_, n), uv, size_t(n) * 8, hipMemcpyHostToDevice, st), "upload uv");
    has_uv_ = uv != nullptr;
    if (!same_positions) {
        // ---- weld: stable radix passes on the z, y, x keys carry the index order.
        weld_keys_kernel<<<blocks_for(n), 256, 0, st>>>(positions_.as<float4>(), n, kx, ky, kz, ids);
        auto sort_pass = [&](const uint32_t* key, const uint32_t* order_in, uint32_t* order_out) {
            gather_kernel<<<blocks_for(n), 256, 0, st>>>(key, order_in, n, key_tmp);
            run("weld sort", [&](size_t& b) {
                return radix_sort_pairs(work, b, key_tmp, key_sorted, order_in, order_out, n, 0, 32, st);
            });
        };
        sort_pass(kz, ids, ord);
        sort_pass(ky, ord, ord2);
        sort_pass(kx, ord2, ord);
        weld_heads_kernel<<<blocks_for(n), 256, 0, st>>>(ord, kx, ky, kz, n, ord2);
        run("weld scan", [&](size_t& b) {
            return inclusive_max(work, b, ord2, key_sorted, n, st);
        });
        weld_assign_kernel<<<blocks_for(n), 256, 0, st>>>(ord, key_sorted, n, geometry);
        weld_positions_ = m.positions;
        weld_ready_ = true;
    }
    // ---- faces and edges around every geometric vertex, by counting.
    hip_check(hipMemsetAsync(invalid, 0, n, st), "clear");
    hip_check(hipMemsetAsync(boundary, 0, n, st), "clear");
    hip_check(hipMemsetAsync(face_count, 0, 5 * n1 * 4, st), "clear");  // counts and cursors
    const uint32_t* idx = indices_.as<uint32_t>();
    topo_count_kernel<<<blocks_for(corners), 256, 0, st>>>(idx, geometry, corners, face_count, adj_count, invalid);
    face_geo_kernel<<<blocks_for(faces_), 256, 0, st>>>(idx, geometry, faces_, face_geo);
    run("face scan", [&](size_t& b) { return exclusive_sum(work, b, face_count, face_offsets, uint32_t(n1), st); });
    run("adjacency scan", [&](size_t& b) { return exclusive_sum(work, b, adj_count, adj_offsets, uint32_t(n1), st); });
    topo_fill_kernel<<<blocks_for(corners), 256, 0, st>>>(idx, geometry, corners, face_offsets, adj_offsets,
                                                          face_cursor, adj_cursor, faces, adj);
    long_sort_kernel<<<(n + 3) / 4, 256, 0, st>>>(n, face_offsets, faces, adj_offsets, adj);
    topo_vertex_kernel<<<blocks_for(n), 256, 0, st>>>(n, face_offsets, faces, adj_offsets, adj, degree, up, boundary,
                                                      invalid);
    run("degree scan", [&](size_t& b) { return exclusive_sum(work, b, degree, offsets, uint32_t(n1), st); });
    run("edge scan", [&](size_t& b) { return exclusive_sum(work, b, up, edge_offsets, uint32_t(n1), st); });
    topo_emit_kernel<<<blocks_for(n), 256, 0, st>>>(n, adj_offsets, adj, offsets, edge_offsets, neighbors, unique);
    uint32_t edge_total = 0;
    hip_check(hipMemcpyAsync(&edge_total, edge_offsets + n, 4, hipMemcpyDeviceToHost, st), "download edges");
    hip_check(hipStreamSynchronize(st), "edges");
    edges_ = edge_total;
    const uint32_t slots = 2 * edges_;
    Topo t{positions_.as<float4>(), has_uv_ ? uv_.as<float2>() : nullptr, idx,
           m.materials.empty() ? nullptr : materials_.as<uint16_t>(), geometry, invalid, boundary, offsets,
           neighbors, face_offsets, faces, n, faces_, face_geo};
    fans_kernel<<<blocks_for(n), 256, 0, st>>>(t, invalid, visited);
    long_fans_kernel<<<(n + 3) / 4, 256, 0, st>>>(t, invalid);
    // ---- legality of every directed edge, compacted in enumeration order.
    legal_count_ = 0;
    if (slots) {
        legal_kernel<<<blocks_for(slots), 256, 0, st>>>(t, unique, edges_, all, flags);
        run("legal scan", [&](size_t& b) { return exclusive_sum(work, b, flags, where, slots + 1, st); });
        compact_kernel<<<blocks_for(slots), 256, 0, st>>>(all, flags, where, slots, legal);
        hip_check(hipMemcpyAsync(&legal_count_, where + slots, 4, hipMemcpyDeviceToHost, st), "download count");
    }
    hip_check(hipGetLastError(), "action topology launch");
    hip_check(hipStreamSynchronize(st), "action topology");
    build_ms_ = clock.ms();
}

ActionSet::~ActionSet() {
    DeviceGuard guard(device_);
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
    (void)hipHostFree(staging_);
}

void ActionSet::download_legal(uint32_t* pairs) const {
    if (!legal_count_) return;
    DeviceGuard guard(device_);
    std::vector<LegalAction> host(legal_count_);
    hip_check(hipMemcpy(host.data(), legal_.p, host.size() * sizeof(LegalAction), hipMemcpyDeviceToHost),
              "download legal actions");
    for (size_t i = 0; i < host.size(); ++i) {
        pairs[2 * i] = host[i].from;
        pairs[2 * i + 1] = host[i].to;
    }
}

ActionTopo ActionSet::topo() const {
    return {positions_.as<float4>(), has_uv_ ? uv_.as<float2>() : nullptr, indices_.as<uint32_t>(),
            has_materials_ ? materials_.as<uint16_t>() : nullptr, geometry_.as<uint32_t>(), invalid_.as<uint8_t>(),
            boundary_.as<uint8_t>(), offsets_.as<uint32_t>(), neighbors_.as<uint32_t>(),
            face_offsets_.as<uint32_t>(), face_lists_.as<uint32_t>(), vertices_, faces_, face_geo_.as<uint4>()};
}

uint32_t ActionSet::claim_device(const uint32_t* drows, size_t n, uint32_t maximum, uint64_t removable, bool strict,
                                 std::vector<uint32_t>& picked) {
    const hipStream_t st = stream_;
    uint32_t* out = grow<uint32_t>(claim_out_, 3 + 64 + 1);
    uint32_t host[3 + 64 + 1];
    // The LDS form first; the global form when a table would fill.
    claim_lds_kernel<<<1, 64, 0, st>>>(topo(), legal_.as<LegalAction>(), drows, uint32_t(n), maximum, removable,
                                       strict ? 1u : 0u, out);
    hip_check(hipGetLastError(), "claim launch");
    hip_check(hipMemcpyAsync(host, out, sizeof host, hipMemcpyDeviceToHost, st), "download selection");
    hip_check(hipStreamSynchronize(st), "claim");
    if (!host[3 + 64]) {
        if (strict && host[1]) throw std::invalid_argument("illegal, stale or overlapping endpoint actions");
        picked.assign(host + 3, host + 3 + host[0]);
        return host[2];
    }
    uint32_t* used = grow<uint32_t>(used_, vertices_);
    // Lane masks per vertex, zero between chunks; the claim kernel clears
    // what it sets, so they are zeroed once per state.
    if (!touch_ready_) {
        unsigned long long* touch = grow<unsigned long long>(touch_, vertices_);
        fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(2 * vertices_)), 256, 0, st>>>(
            reinterpret_cast<uint32_t*>(touch), 2 * vertices_, 0);
        touch_ready_ = true;
    }
    fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(vertices_)), 256, 0, st>>>(used, vertices_, 0);
    claim_kernel<<<1, 64, 0, st>>>(topo(), legal_.as<LegalAction>(), drows, uint32_t(n), maximum, removable,
                                   strict ? 1u : 0u, used, touch_.as<unsigned long long>(), out);
    hip_check(hipGetLastError(), "claim launch");
    hip_check(hipMemcpyAsync(host, out, (3 + 64) * 4, hipMemcpyDeviceToHost, st), "download selection");
    hip_check(hipStreamSynchronize(st), "claim");
    if (strict && host[1]) throw std::invalid_argument("illegal, stale or overlapping endpoint actions");
    picked.assign(host + 3, host + 3 + host[0]);
    return host[2];
}

void ActionSet::claim(const uint32_t* rows, size_t n, uint32_t maximum, uint64_t removable, bool strict,
                      std::vector<uint32_t>& picked) {
    DeviceGuard guard(device_);
    for (size_t i = 0; i < n; ++i)
        if (rows[i] >= legal_count_) throw std::invalid_argument("action order index outside the legal actions");
    uint32_t* drows = grow<uint32_t>(selection_, n);
    if (n) hip_check(hipMemcpyAsync(drows, rows, n * 4, hipMemcpyHostToDevice, stream_), "upload order");
    claim_device(drows, n, maximum, removable, strict, picked);
}

void ActionSet::rank(uint32_t mode, double diameter, const Graph* graph, uint32_t* order_out) {
    if (mode == 2) throw Unavailable("shuffled ranking is not implemented by the HIP action state");
    if (mode > 4) throw std::invalid_argument("unknown ranking");
    if (mode == 0 && (!graph || !graph->has_outputs)) throw std::invalid_argument("learned ranking needs scorer outputs");
    DeviceGuard guard(device_);
    const uint32_t n = legal_count_;
    ranked_ = 0;
    if (!n) return;
    RankInput in{};
    in.legal = legal_.as<LegalAction>();
    in.count = n;
    in.mode = mode;
    in.diameter = diameter;
    if (mode == 0) {
        in.graph_offsets = graph->offsets.as<uint32_t>();
        in.graph_neighbors = graph->neighbors.as<uint32_t>();
        in.outputs = graph->outputs.as<float>();
        in.outputs_per_action = kOutputs;
        if (graph->vertices != vertices_) throw std::invalid_argument("scorer graph of another mesh");
    }
    run_rank(&in, order_out);
}

void ActionSet::rank_scores(const float* scores, uint32_t stride, uint32_t* order_out) {
    ranked_ = 0;
    if (!legal_count_) return;
    RankInput in{};
    in.legal = legal_.as<LegalAction>();
    in.count = legal_count_;
    in.mode = kPolicyScores;
    in.diameter = diameter_;
    in.outputs = scores;
    in.outputs_per_action = stride;
    run_rank(&in, order_out);
}

// Scores every legal action with `in` (mode, inputs), sorts the order by
// descending score (stable: ties keep enumeration order) and keeps it on the
// device; a score that is not finite is an error.
void ActionSet::run_rank(const void* input, uint32_t* order_out) {
    RankInput in = *static_cast<const RankInput*>(input);
    const hipStream_t st = stream_;
    const uint32_t n = legal_count_;
    float* keys = grow<float>(rank_keys_, 2 * size_t(n));
    uint32_t* idx = grow<uint32_t>(rank_idx_, n);
    uint32_t* order = grow<uint32_t>(order_, n);
    uint32_t* bad = grow<uint32_t>(claim_out_, 3 + 64);
    hip_check(hipMemsetAsync(bad, 0, 4, st), "clear");
    in.keys = keys;
    in.order = idx;
    in.bad = bad;
    rank_kernel<<<blocks_for(n), 256, 0, st>>>(topo(), in);
    size_t need = 0;
    hip_check(radix_sort_pairs_descending(nullptr, need, keys, keys + n, idx, order, n, st),
              "sort workspace");
    void* temp = grow<char>(temp_, std::max(need, temp_.bytes));
    size_t b = temp_.bytes;
    hip_check(radix_sort_pairs_descending(temp, b, keys, keys + n, idx, order, n, st),
              "rank sort");
    uint32_t flag = 0;
    hip_check(hipMemcpyAsync(&flag, bad, 4, hipMemcpyDeviceToHost, st), "download rank flag");
    if (order_out) hip_check(hipMemcpyAsync(order_out, order, size_t(n) * 4, hipMemcpyDeviceToHost, st), "download order");
    hip_check(hipGetLastError(), "rank launch");
    hip_check(hipStreamSynchronize(st), "rank");
    if (flag) throw std::invalid_argument("nonfinite action rank");
    ranked_ = n;
}

std::vector<uint32_t> ActionSet::independent_ranked(size_t position, uint32_t maximum, uint64_t removable,
                                                    size_t* first) {
    if (!maximum || maximum > 64) throw std::invalid_argument("action batch outside 1..64");
    if (ranked_ != legal_count_) throw std::invalid_argument("rank() has not run");
    std::vector<uint32_t> picked;
    *first = size_t(-1);
    if (position >= legal_count_ || !removable) return picked;
    DeviceGuard guard(device_);
    const uint32_t k = claim_device(order_.as<uint32_t>() + position, legal_count_ - position, maximum, removable,
                                    false, picked);
    if (!picked.empty()) *first = position + k;
    return picked;
}

std::vector<uint32_t> ActionSet::independent(const uint32_t* order, size_t n, uint32_t maximum, uint64_t removable) {
    if (!maximum || maximum > 64) throw std::invalid_argument("action batch outside 1..64");
    std::vector<uint32_t> picked;
    if (!n || !removable) return picked;
    claim(order, n, maximum, removable, false, picked);
    return picked;
}

std::vector<uint32_t> ActionSet::trial(const uint32_t* rows, size_t n, std::vector<uint16_t>* materials) {
    if (!n || n > 64) throw std::invalid_argument("invalid action batch");
    std::vector<uint32_t> picked;
    claim(rows, n, uint32_t(n), ~uint64_t(0), true, picked);
    DeviceGuard guard(device_);
    const hipStream_t st = stream_;
    uint32_t* remap = grow<uint32_t>(scratch_a_, vertices_ + 2 * size_t(faces_) + 1 + 3 * size_t(faces_));
    uint32_t* keep = remap + vertices_;
    uint32_t* slot = keep + faces_ + 1;
    uint32_t* next = slot + faces_;
    fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(vertices_)), 256, 0, st>>>(remap, vertices_, kNone);
    const uint32_t* drows = selection_.as<uint32_t>();
    remap_kernel<<<1, 64, 0, st>>>(legal_.as<LegalAction>(), drows, uint32_t(n), remap);
    const Topo t = topo();
    trial_keep_kernel<<<blocks_for(faces_), 256, 0, st>>>(t, remap, next, keep);
    size_t need = 0;
    hip_check(exclusive_sum(nullptr, need, keep, slot, faces_, st), "scan workspace");
    void* temp = grow<char>(temp_, std::max(need, temp_.bytes));
    size_t b = temp_.bytes;
    hip_check(exclusive_sum(temp, b, keep, slot, faces_, st), "trial slots");
    uint32_t tail[2] = {0, 0};
    hip_check(hipMemcpyAsync(&tail[0], slot + faces_ - 1, 4, hipMemcpyDeviceToHost, st), "download count");
    hip_check(hipMemcpyAsync(&tail[1], keep + faces_ - 1, 4, hipMemcpyDeviceToHost, st), "download count");
    hip_check(hipStreamSynchronize(st), "trial count");
    const uint32_t kept = tail[0] + tail[1];
    if (!kept) throw std::invalid_argument("batch would remove the entire mesh");
    uint32_t* out = grow<uint32_t>(scratch_b_, 4 * size_t(kept));
    uint16_t* out_materials = reinterpret_cast<uint16_t*>(out + 3 * size_t(kept));
    trial_copy_kernel<<<blocks_for(faces_), 256, 0, st>>>(next, keep, slot,
                                                          has_materials_ ? materials_.as<uint16_t>() : nullptr,
                                                          faces_, out, out_materials);
    // One copy of indices and materials through pinned staging memory (sized
    // for the state's whole mesh once).
    const size_t bytes = size_t(kept) * (12 + (has_materials_ ? 2 : 0));
    if (!staging_) hip_check(hipHostMalloc(&staging_, size_t(faces_) * 14, hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipMemcpyAsync(staging_, out, bytes, hipMemcpyDeviceToHost, st), "download trial");
    hip_check(hipStreamSynchronize(st), "trial");
    const auto* staged = static_cast<const uint32_t*>(staging_);
    std::vector<uint32_t> host(staged, staged + 3 * size_t(kept));
    if (materials) {
        materials->assign(has_materials_ ? kept : 0, 0);
        if (has_materials_)
            std::memcpy(materials->data(), staged + 3 * size_t(kept), size_t(kept) * 2);
    }
    return host;
}

// src/neural/action.cpp:453-525 with the gate supplied by the caller.
std::vector<uint32_t> execute_actions(int device, const HostMesh& input, const float* uv, uint64_t target,
                                      uint32_t budget, uint32_t ranking, uint32_t batch, const ActionGate& gate,
                                      ExecuteStats& stats, std::vector<uint16_t>* materials_out, Scorer* scorer,
                                      const Bounds* bounds, ExecutorCache* cache, const ActionLadderGate* ladder) {
    if (!batch || batch > 64) throw std::invalid_argument("action batch outside 1..64");
    if (ranking == 2) throw Unavailable("shuffled ranking is not implemented by the HIP executor");
    if (ranking > 4) throw std::invalid_argument("unknown ranking");
    if (ranking == 0 && (!scorer || !bounds)) throw std::invalid_argument("learned ranking needs a scorer");
    if (ranking == 0 && scorer->device() != device) throw std::invalid_argument("scorer on another device");
    const auto clock0 = std::chrono::steady_clock::now();
    stats = {};
    const double diameter = std::max(1e-20, bounds_of(input.positions.data(), input.positions.size() / 3).radius * 2);
    HostMesh state = input;
    target = std::max<uint64_t>(1, target);
    auto since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    bool no_legal = false;
    ExecutorCache local;
    ExecutorCache& c = cache ? *cache : local;
    const std::vector<float> uv_values = uv ? std::vector<float>(uv, uv + input.positions.size() / 3 * 2)
                                            : std::vector<float>{};
    if (c.input_state && (c.input_state->device() != device || c.input_indices != input.indices ||
                          c.input_materials != input.materials || c.input_uv != uv_values))
        c = ExecutorCache{};
    bool first = true;
    // The next state, built and ranked from a trial's candidate while the
    // gate audits it (the gate runs on another thread); kept when the gate
    // accepts, dropped when it rejects. Positions do not change within a
    // call, so the speculative mesh takes them once.
    HostMesh next_mesh;
    next_mesh.positions = state.positions;
    bool next_ready = false;
    // One scorer graph for every state (rank() waits for its sort, so the
    // next state's build may overwrite the buffers); the states share the
    // input's vertex stream, uploaded once.
    std::unique_ptr<Graph> graph;
    // The scorer's graph of a mesh state and its BF16 logits (learned
    // ranking only); independent of the action state, so it can run beside
    // that state's build.
    const auto score_state = [&](const HostMesh& mesh) {
        const bool same_positions = graph != nullptr;
        if (!graph) graph = std::make_unique<Graph>();
        scorer->rebuild(*graph, mesh, *bounds, same_positions);
        scorer->forward(*graph, Precision::Bf16);
    };
    const auto rank_state = [&](ActionSet* s, const HostMesh& mesh) {
        if (ranking == 0) score_state(mesh);
        s->rank(ranking, diameter, ranking == 0 ? graph.get() : nullptr, nullptr);
    };
    while (state.indices.size() / 3 > target && stats.trials < budget) {
        auto t = std::chrono::steady_clock::now();
        // The input's state (ranked once per ranking order that does not
        // depend on a scorer), then reused states for the later meshes.
        const bool cached_rank = next_ready || (first && ranking != 0 && c.input_state && c.input_ranking == int(ranking));
        ActionSet* set;
        if (next_ready) {
            std::swap(c.work, c.spec);
            set = c.work.get();
            next_ready = false;
        } else if (first) {
            if (!c.input_state) {
                c.input_state = std::make_unique<ActionSet>(device, state, uv);
                c.input_indices = input.indices;
                c.input_materials = input.materials;
                c.input_uv = uv_values;
                c.input_ranking = -1;
                ++stats.states;
            }
            set = c.input_state.get();
        } else {
            if (c.work)
                c.work->rebuild(state, uv);
            else
                c.work = std::make_unique<ActionSet>(device, state, uv);
            ++stats.states;
            set = c.work.get();
        }
        first = false;
        stats.build_ms += since(t);
        const size_t n = set->legal_count();
        if (!n) {
            no_legal = true;
            break;
        }
        t = std::chrono::steady_clock::now();
        if (!cached_rank) {
            rank_state(set, state);
            if (set == c.input_state.get()) c.input_ranking = ranking == 0 ? -1 : int(ranking);
        }
        stats.rank_ms += since(t);
        stats.ranked += n;
        const uint64_t faces = state.indices.size() / 3;
        bool accepted = false;
        for (size_t position = 0; position < n && stats.trials < budget && !accepted;) {
            t = std::chrono::steady_clock::now();
            size_t first_pick = 0;
            std::vector<uint32_t> picked = set->independent_ranked(position, batch, faces - target, &first_pick);
            stats.select_ms += since(t);
            if (picked.empty()) break;
            for (;;) {
                if (stats.trials == budget) break;
                std::vector<uint16_t> materials;
                t = std::chrono::steady_clock::now();
                std::vector<uint32_t> candidate = set->trial(picked.data(), picked.size(), &materials);
                stats.trial_ms += since(t);
                ++stats.trials;
                t = std::chrono::steady_clock::now();
                auto verdict = std::async(std::launch::async, [&] { return gate(candidate, materials); });
                // Speculate only when an acceptance would continue the loop.
                bool built = false;
                if (candidate.size() / 3 > target && stats.trials < budget) {
                    // With learned ranking the scorer's graph and logits of
                    // the next state are computed on another thread (the
                    // scorer's own stream) while its action state builds.
                    std::future<void> scored;
                    try {
                        const auto tb = std::chrono::steady_clock::now();
                        next_mesh.indices = candidate;
                        next_mesh.materials = materials;
                        if (ranking == 0) scored = std::async(std::launch::async, [&] { score_state(next_mesh); });
                        if (c.spec)
                            c.spec->rebuild(next_mesh, uv);
                        else
                            c.spec = std::make_unique<ActionSet>(device, next_mesh, uv);
                        stats.speculative_build_ms += since(tb);
                        if (scored.valid()) scored.get();
                        if (c.spec->legal_count())
                            c.spec->rank(ranking, diameter, ranking == 0 ? graph.get() : nullptr, nullptr);
                        stats.speculative_ms += since(tb);
                        built = true;
                    } catch (...) {
                        if (scored.valid()) scored.wait();
                        verdict.wait();
                        throw;
                    }
                }
                const bool pass = verdict.get();
                stats.gate_ms += since(t);
                if (pass) {
                    state.indices = std::move(candidate);
                    state.materials = std::move(materials);
                    stats.accepted += picked.size();
                    ++stats.accepted_batches;
                    accepted = true;
                    if (built) {
                        next_ready = true;
                        ++stats.states;
                    }
                    break;
                }
                stats.rejected += picked.size();
                if (picked.size() == 1) break;
                picked.resize(picked.size() / 2);
                if (ladder) {
                    // The rest of the ladder in one call: every shorter
                    // prefix the trial budget leaves, trialled, then audited
                    // together. The first accepted one is the one the
                    // one-by-one loop keeps; trials and rejections are
                    // counted up to it, as that loop counts them.
                    std::vector<std::vector<uint32_t>> candidates;
                    std::vector<std::vector<uint16_t>> kept;
                    std::vector<size_t> sizes;
                    t = std::chrono::steady_clock::now();
                    for (size_t s = picked.size(); s >= 1 && stats.trials + candidates.size() < budget; s /= 2) {
                        std::vector<uint16_t> m;
                        candidates.push_back(set->trial(picked.data(), s, &m));
                        kept.push_back(std::move(m));
                        sizes.push_back(s);
                    }
                    stats.trial_ms += since(t);
                    if (candidates.empty()) break;
                    t = std::chrono::steady_clock::now();
                    const int p = (*ladder)(candidates, kept);
                    stats.gate_ms += since(t);
                    if (p >= int(candidates.size())) throw std::logic_error("ladder gate index out of range");
                    const size_t counted = p >= 0 ? size_t(p) + 1 : candidates.size();
                    stats.trials += counted;
                    for (size_t i = 0; i < counted; ++i)
                        if (int(i) != p) stats.rejected += sizes[i];
                    if (p >= 0) {
                        state.indices = std::move(candidates[size_t(p)]);
                        state.materials = std::move(kept[size_t(p)]);
                        stats.accepted += sizes[size_t(p)];
                        ++stats.accepted_batches;
                        accepted = true;
                    }
                    break;
                }
            }
            // A rejected singleton is not retried: continue after the first pick.
            position = first_pick + 1;
        }
        if (!accepted) break;
    }
    stats.stop_reason = state.indices.size() / 3 <= target ? 0 : stats.trials >= budget ? 1 : no_legal ? 2 : 3;
    stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - clock0).count();
    if (materials_out) *materials_out = state.materials;
    return state.indices;
}

void ActionSet::patches(const uint32_t* rows, size_t n, std::vector<uint32_t>& removed_offsets,
                        std::vector<uint32_t>& removed, std::vector<uint32_t>& added_offsets,
                        std::vector<uint32_t>& added) {
    removed_offsets.assign(n + 1, 0);
    added_offsets.assign(n + 1, 0);
    removed.clear();
    added.clear();
    if (!n) return;
    for (size_t i = 0; i < n; ++i)
        if (rows[i] >= legal_count_) throw std::invalid_argument("patch row outside the legal actions");
    DeviceGuard guard(device_);
    const hipStream_t st = stream_;
    uint32_t* drows = grow<uint32_t>(selection_, 5 * n + 2);
    uint32_t* rc = drows + n;           // n + 1
    uint32_t* ac = rc + n + 1;          // n + 1
    uint32_t* ro = ac + n + 1;          // n (exclusive offsets)
    uint32_t* ao = ro + n;              // n
    hip_check(hipMemcpyAsync(drows, rows, n * 4, hipMemcpyHostToDevice, st), "upload rows");
    const Topo t = topo();
    patch_count_kernel<<<blocks_for(n), 256, 0, st>>>(t, legal_.as<LegalAction>(), drows, uint32_t(n), rc, ac);
    std::vector<uint32_t> hrc(n), hac(n);
    hip_check(hipMemcpyAsync(hrc.data(), rc, n * 4, hipMemcpyDeviceToHost, st), "download counts");
    hip_check(hipMemcpyAsync(hac.data(), ac, n * 4, hipMemcpyDeviceToHost, st), "download counts");
    hip_check(hipStreamSynchronize(st), "patch counts");
    for (size_t i = 0; i < n; ++i) {
        removed_offsets[i + 1] = removed_offsets[i] + hrc[i];
        added_offsets[i + 1] = added_offsets[i] + hac[i];
    }
    hip_check(hipMemcpyAsync(ro, removed_offsets.data(), n * 4, hipMemcpyHostToDevice, st), "upload offsets");
    hip_check(hipMemcpyAsync(ao, added_offsets.data(), n * 4, hipMemcpyHostToDevice, st), "upload offsets");
    const size_t nr = removed_offsets[n], na = added_offsets[n];
    uint32_t* dr = grow<uint32_t>(scratch_b_, nr + 3 * na);
    uint32_t* da = dr + nr;
    patch_fill_kernel<<<blocks_for(n), 256, 0, st>>>(t, legal_.as<LegalAction>(), drows, uint32_t(n), ro, ao, dr, da);
    removed.resize(nr);
    added.resize(3 * na);
    if (nr) hip_check(hipMemcpyAsync(removed.data(), dr, nr * 4, hipMemcpyDeviceToHost, st), "download removed");
    if (na) hip_check(hipMemcpyAsync(added.data(), da, 3 * na * 4, hipMemcpyDeviceToHost, st), "download added");
    hip_check(hipGetLastError(), "patch launch");
    hip_check(hipStreamSynchronize(st), "patches");
}

}  // namespace blitz_hip
