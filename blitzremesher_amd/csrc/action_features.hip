// Action features on the device: the reference's vertex graph features of an
// action state's current mesh (blitz::neural::graph, src/neural/graph.cpp:
// 35-187, 24 channels per vertex) and the feature rows of its legal actions
// (ActionState::actions, src/neural/action.cpp:371-450; the GPU rows of
// fill_action_features, src/neural/action_gpu.cu:644-730, add the second
// wedge's vertex features for 128-wide placement rows).
//
// Every value is the reference's own arithmetic: float Vec3 differences,
// cross products and sums, double dot products and lengths, float(...)
// conversions where the reference converts, and sums in the reference's
// order (a vertex's face normals in (face, corner) order, its neighbours'
// distances in ascending neighbour order). The library is compiled with
// -ffp-contract=off, so no product is fused into an add.
//
// Layout: one pass per stage over the whole mesh (vertex, face, corner or
// edge per thread); the only sorts are of 32-bit corner keys and 64-bit edge
// keys (radix, stable), as in the action topology build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "actions.hpp"
#include "device_sort.hpp"
#include "policy.hpp"
#include "vec_device.hpp"

namespace blitz_hip {
namespace {

// The 64-bit edge keys as the sort primitives take them (same width).
static_assert(sizeof(unsigned long long) == sizeof(uint64_t));
inline uint64_t* as_u64(unsigned long long* p) { return reinterpret_cast<uint64_t*>(p); }
inline const uint64_t* as_u64(const unsigned long long* p) { return reinterpret_cast<const uint64_t*>(p); }

constexpr uint32_t kChannels = 24;
// Graph flags (src/neural/internal.hpp:14).
constexpr uint32_t kLocked = 1, kBoundary = 2, kSeam = 4, kMaterial = 8, kUsed = 16;
constexpr unsigned long long kNoEdge = ~0ull;

uint32_t blocks(size_t n) { return uint32_t(std::max<size_t>(1, (n + 255) / 256)); }

// The attribute streams of the state's input (null when absent).
struct Attributes {
    const float* normals;    // xyz per vertex, as given
    const float2* uv;
    const uint32_t* colors;  // RGBA8 bytes of one vertex as one word
    const float4* tangents;
};

// graph()'s position(): the vertex relative to the bounds centre, over the
// bounds diameter, each coordinate float((double(p) - centre) / scale).
struct Frame {
    double cx, cy, cz, scale;
};
__device__ __forceinline__ F3 frame_position(const float4* positions, uint32_t i, const Frame& f) {
    const float4 p = positions[i];
    return {float((double(p.x) - f.cx) / f.scale), float((double(p.y) - f.cy) / f.scale),
            float((double(p.z) - f.cz) / f.scale)};
}

// The compare() chain of src/neural/graph.cpp:42-52 is zero exactly when the
// bytes of every present stream agree.
__device__ bool same_attributes(const Attributes& a, uint32_t i, uint32_t j) {
    if (a.normals)
        for (uint32_t k = 0; k < 3; ++k)
            if (__float_as_uint(a.normals[3 * i + k]) != __float_as_uint(a.normals[3 * j + k])) return false;
    if (a.uv && (__float_as_uint(a.uv[i].x) != __float_as_uint(a.uv[j].x) ||
                 __float_as_uint(a.uv[i].y) != __float_as_uint(a.uv[j].y)))
        return false;
    if (a.colors && a.colors[i] != a.colors[j]) return false;
    if (a.tangents) {
        const float4 s = a.tangents[i], t = a.tangents[j];
        if (__float_as_uint(s.x) != __float_as_uint(t.x) || __float_as_uint(s.y) != __float_as_uint(t.y) ||
            __float_as_uint(s.z) != __float_as_uint(t.z) || __float_as_uint(s.w) != __float_as_uint(t.w))
            return false;
    }
    return true;
}

__global__ void group_count_kernel(const uint32_t* __restrict__ geometry, uint32_t n, uint32_t* count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&count[geometry[i]], 1u);
}

__global__ void group_fill_kernel(const uint32_t* __restrict__ geometry, uint32_t n, const uint32_t* __restrict__ offsets,
                                  uint32_t* cursor, uint32_t* members) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = geometry[i];
    members[offsets[g] + atomicAdd(&cursor[g], 1u)] = i;
}

// src/neural/graph.cpp:53-74. The reference sorts by (position key,
// attribute bytes, index) and gives each run of equal attributes within a
// position its first vertex: the smallest index with the same position key
// and the same attribute bytes. A position holding more than one attribute
// class is a seam (Locked | Seam on all its vertices). Both facts are
// independent of the order the group's members are visited in.
__global__ void canonical_kernel(Attributes a, const uint32_t* __restrict__ geometry,
                                 const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ members, uint32_t n,
                                 uint32_t* canonical, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = geometry[i];
    uint32_t c = i;
    bool seam = false;
    for (uint32_t k = offsets[g]; k < offsets[g + 1]; ++k) {
        const uint32_t j = members[k];
        if (same_attributes(a, i, j))
            c = min(c, j);
        else
            seam = true;
    }
    canonical[i] = c;
    flags[i] = seam ? kLocked | kSeam : 0u;
}

// src/neural/graph.cpp:91-112, per face: the unit normal of its canonical
// corners, the corners keyed by canonical vertex for the per-vertex pass and
// the canonical edges (u < v; none where a corner repeats).
__global__ void graph_faces_kernel(const float4* __restrict__ positions, const uint32_t* __restrict__ indices,
                                   const uint32_t* __restrict__ canonical, uint32_t faces, Frame frame,
                                   float4* face_normal, uint32_t* corner_key, uint32_t* corner_id,
                                   unsigned long long* edge_key) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces) return;
    uint32_t id[3];
    for (uint32_t k = 0; k < 3; ++k) id[k] = canonical[indices[3 * f + k]];
    const F3 p0 = frame_position(positions, id[0], frame), p1 = frame_position(positions, id[1], frame),
             p2 = frame_position(positions, id[2], frame);
    const F3 n = norm3(cross3(sub3(p1, p0), sub3(p2, p0)));
    face_normal[f] = make_float4(n.x, n.y, n.z, 0.f);
    for (uint32_t k = 0; k < 3; ++k) {
        const uint32_t c = 3 * f + k, u = id[k], v = id[(k + 1) % 3];
        corner_key[c] = u;
        corner_id[c] = c;
        edge_key[c] = u != v ? (static_cast<unsigned long long>(min(u, v)) << 32) | max(u, v) : kNoEdge;
    }
}

__global__ void count_kernel(const uint32_t* __restrict__ key, uint32_t n, uint32_t* count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&count[key[i]], 1u);
}

// src/neural/graph.cpp:97-104, per canonical vertex over its corners in
// (face, corner) order (a stable sort of the corners by vertex): Used, the
// Material flag (the reference sets it when a corner's face material differs
// from the previous corner's) and the float sum of the faces' unit normals.
__global__ void graph_vertices_kernel(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ corners,
                                      const float4* __restrict__ face_normal, const uint16_t* __restrict__ materials,
                                      uint32_t n, uint32_t* flags, float4* normal_sum) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    F3 sum{0.f, 0.f, 0.f};
    uint32_t add = 0;
    uint16_t previous = 0;
    for (uint32_t k = offsets[u]; k < offsets[u + 1]; ++k) {
        const uint32_t f = corners[k] / 3;
        const uint16_t material = materials ? materials[f] : uint16_t(0);
        if (k > offsets[u] && material != previous) add |= kLocked | kMaterial;
        previous = material;
        add |= kUsed;
        sum = add3(sum, xyz(face_normal[f]));
    }
    flags[u] |= add;
    normal_sum[u] = make_float4(sum.x, sum.y, sum.z, 0.f);
}

// src/neural/graph.cpp:116-138, per unique canonical edge: boundary (one
// face) and non-manifold (more than two) flags, degrees, both directions.
__global__ void graph_edges_kernel(const unsigned long long* __restrict__ unique, const uint32_t* __restrict__ counts,
                                   uint32_t edges, uint32_t* flags, uint32_t* degree, unsigned long long* directed) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= edges) return;
    const unsigned long long key = unique[e];
    const uint32_t u = uint32_t(key >> 32), v = uint32_t(key);
    const uint32_t c = counts[e];
    if (c == 1) {
        atomicOr(&flags[u], kLocked | kBoundary);
        atomicOr(&flags[v], kLocked | kBoundary);
    }
    if (c > 2) {
        atomicOr(&flags[u], kLocked);
        atomicOr(&flags[v], kLocked);
    }
    atomicAdd(&degree[u], 1u);
    atomicAdd(&degree[v], 1u);
    directed[2 * e] = key;
    directed[2 * e + 1] = (static_cast<unsigned long long>(v) << 32) | u;
}

__global__ void low_words_kernel(const unsigned long long* __restrict__ keys, uint32_t n, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = uint32_t(keys[i]);
}

__device__ __forceinline__ float clamp16(float v) { return (v < -16.f ? -16.f : (16.f < v ? 16.f : v)) / 16; }

// linear_color (include/blitz/mesh.hpp:20): each byte / 255.f.
__device__ __forceinline__ float4 linear(uint32_t rgba) {
    return make_float4(float(rgba & 0xFFu) / 255.f, float((rgba >> 8) & 0xFFu) / 255.f,
                       float((rgba >> 16) & 0xFFu) / 255.f, float(rgba >> 24) / 255.f);
}

// src/neural/graph.cpp:145-185: the 24 channels of every vertex.
__global__ void graph_features_kernel(const float4* __restrict__ positions, Attributes a,
                                      const float4* __restrict__ normal_sum, const uint32_t* __restrict__ flags,
                                      const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ neighbors,
                                      uint32_t n, Frame frame, float* x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const F3 p = frame_position(positions, i, frame);
    const F3 normal = norm3(a.normals ? F3{a.normals[3 * i], a.normals[3 * i + 1], a.normals[3 * i + 2]}
                                      : xyz(normal_sum[i]));
    const float4 color = a.colors ? linear(a.colors[i]) : make_float4(1.f, 1.f, 1.f, 1.f);
    const float2 uv = a.uv ? a.uv[i] : make_float2(0.f, 0.f);
    const float4 t = a.tangents ? a.tangents[i] : make_float4(0.f, 0.f, 0.f, 1.f);
    const F3 direction = norm3(F3{t.x, t.y, t.z});
    double edge = 0;
    for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k)
        edge += len3(sub3(frame_position(positions, neighbors[k], frame), p));
    const uint32_t degree = offsets[i + 1] - offsets[i];
    edge /= double(max(1u, degree));
    const uint32_t fl = flags[i];
    float* o = x + size_t(i) * kChannels;
    o[0] = p.x;
    o[1] = p.y;
    o[2] = p.z;
    o[3] = normal.x;
    o[4] = normal.y;
    o[5] = normal.z;
    o[6] = color.x;
    o[7] = color.y;
    o[8] = color.z;
    o[9] = clamp16(uv.x);
    o[10] = clamp16(uv.y);
    o[11] = direction.x;
    o[12] = direction.y;
    o[13] = direction.z;
    o[14] = t.w;
    o[15] = a.normals ? 1.f : 0.f;
    o[16] = a.uv ? 1.f : 0.f;
    o[17] = a.colors ? 1.f : 0.f;
    o[18] = a.tangents ? 1.f : 0.f;
    o[19] = (fl & kSeam) ? 1.f : 0.f;
    o[20] = (fl & kBoundary) ? 1.f : 0.f;
    o[21] = (fl & kMaterial) ? 1.f : 0.f;
    o[22] = float(log2(double(1u + degree)) / 8);
    o[23] = float(log2(1 + edge * 1024) / 10);
}

struct RowInput {
    const LegalAction* legal;
    uint32_t count, width;
    const float* vertex;  // 24 channels per vertex
    ActionTopo t;
    const uint32_t* colors;
    float conditions[8];
    double scale;       // the state's diameter
    float face_ratio;   // float(current faces / input faces)
    float x79;          // < 0: the target's channel 23 (else this value)
    float* rows;
};

// src/neural/action.cpp:383-449 for legal action i (rows in legal order).
// `from`/`to` are geometric vertices; the first wedge pair (source, target)
// is the smallest (from-index, to-index) pair of the action's mapping, as
// the reference's sorted remap. A legal action's edge has as many faces as
// the collapse removes (one or two), the reference's edge.count.
__global__ void action_rows_kernel(RowInput in) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.count) return;
    float* x = in.rows + size_t(i) * in.width;
    for (uint32_t j = 0; j < in.width; ++j) x[j] = 0.f;
    const LegalAction a = in.legal[i];
    const ActionTopo& t = in.t;
    const uint32_t u = a.from, v = a.to, s = a.wedge_from[0], z = a.wedge_to[0];
    const float* xs = in.vertex + size_t(s) * kChannels;
    const float* xz = in.vertex + size_t(z) * kChannels;
    for (uint32_t j = 0; j < kChannels; ++j) {
        x[j] = xs[j];
        x[kChannels + j] = xz[j];
    }
    for (uint32_t j = 0; j < 8; ++j) x[48 + j] = in.conditions[j];
    const F3 d = scale3(sub3(t.positions[v], t.positions[u]), 1 / in.scale);
    x[56] = d.x;
    x[57] = d.y;
    x[58] = d.z;
    x[59] = float(len3(d));
    float along = 0.f;
    for (uint32_t j = 0; j < 3; ++j) along += x[3 + j] * x[27 + j];
    x[60] = along;
    if (t.uv) {
        x[61] = clamp16(t.uv[z].x - t.uv[s].x);
        x[62] = clamp16(t.uv[z].y - t.uv[s].y);
    }
    x[63] = float(t.offsets[u + 1] - t.offsets[u]) / 64;
    x[64] = float(t.offsets[v + 1] - t.offsets[v]) / 64;
    x[65] = float(a.removed) / 2.f;
    x[66] = in.face_ratio;
    x[67] = t.boundary[u];
    x[68] = t.boundary[v];
    x[69] = float(a.wedges) / 8;
    x[70] = float(t.face_offsets[u + 1] - t.face_offsets[u]) / 64;
    x[71] = float(t.face_offsets[v + 1] - t.face_offsets[v]) / 64;
    const auto face_cross = [&](uint32_t f) {
        const float4 p = t.positions[t.indices[3 * f]], q = t.positions[t.indices[3 * f + 1]],
                     r = t.positions[t.indices[3 * f + 2]];
        return cross3(sub3(q, p), sub3(r, p));
    };
    F3 mean{0.f, 0.f, 0.f}, target_mean{0.f, 0.f, 0.f};
    for (uint32_t k = t.face_offsets[u]; k < t.face_offsets[u + 1]; ++k)
        mean = add3(mean, norm3(face_cross(t.faces[k])));
    for (uint32_t k = t.face_offsets[v]; k < t.face_offsets[v + 1]; ++k)
        target_mean = add3(target_mean, norm3(face_cross(t.faces[k])));
    mean = norm3(mean);
    target_mean = norm3(target_mean);
    // std::min / std::max: the first argument unless the second compares
    // smaller (larger).
    double minimum = 1, maximum = 0;
    const double area_scale = in.scale * in.scale;
    for (uint32_t k = t.face_offsets[u]; k < t.face_offsets[u + 1]; ++k) {
        const F3 normal = face_cross(t.faces[k]);
        const double c = dot3(mean, norm3(normal));
        if (c < minimum) minimum = c;
        const double area = len3(normal) / area_scale;
        if (maximum < area) maximum = area;
    }
    x[72] = float(minimum);
    x[73] = float(dot3(mean, target_mean));
    x[74] = float(maximum);
    x[75] = a.wedges > 1 ? 1.f : 0.f;
    x[76] = t.materials ? 1.f : 0.f;
    if (in.colors) {
        const float4 p = linear(in.colors[s]), q = linear(in.colors[z]);
        x[77] = float(len3(F3{p.x - q.x, p.y - q.y, p.z - q.z}));
    }
    x[78] = xs[23];
    // The conditioned placement schema stores preserve_uv here instead
    // (src/neural/action_gpu.cu:721-722).
    x[79] = in.x79 < 0 ? xz[23] : in.x79;
    if (in.width == 128 && a.wedges == 2) {
        const float* s1 = in.vertex + size_t(a.wedge_from[1]) * kChannels;
        const float* z1 = in.vertex + size_t(a.wedge_to[1]) * kChannels;
        for (uint32_t j = 0; j < kChannels; ++j) {
            x[80 + j] = s1[j];
            x[104 + j] = z1[j];
        }
    }
}

// src/neural/vertex_storage.hpp:24-32: unorm16 position storage of a
// quantization box (zero extent encodes as zero).
__device__ __forceinline__ uint16_t pack_unorm16(float value, float low, float extent) {
    if (extent == 0) return 0;
    const double q = (double(value) - low) / extent * 65535.;
    return uint16_t(q <= 0 ? 0 : q >= 65535 ? 65535 : floor(q + .5));
}
__device__ __forceinline__ float unpack_unorm16(uint16_t value, float low, float extent) {
    return low + extent * (float(value) / 65535.f);
}

struct PlacementInput {
    const LegalAction* legal;
    uint32_t count;
    const float4* positions;
    const float* normals;     // xyz per vertex, or null
    const float* prediction;  // `outputs` per action, or null (default placements)
    uint32_t outputs;
    bool supervise;           // add the predicted normal residuals
    bool quantized;
    float low[3], extent[3];
    const uint32_t* exact;    // per-vertex bitmap, or null
    float* out;               // position, normal 0, normal 1 per action
    uint32_t* bad;
};

__device__ __forceinline__ bool exact_vertex(const uint32_t* bits, uint32_t i) {
    return bits && ((bits[i / 32] >> (i % 32)) & 1u);
}

// decode_placements / default_placements (src/neural/action_gpu.cu:
// 1157-1200) with grid_position and exact_edit (374-390): the edge midpoint
// plus the clamped offset in edge lengths, snapped to the quantization grid
// unless an endpoint wedge is exact; each destination wedge's unit normal,
// plus the predicted residual when supervised. Without a prediction the
// position is the destination vertex's.
__global__ void placements_kernel(PlacementInput in) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.count) return;
    const LegalAction a = in.legal[i];
    float* o = in.out + size_t(i) * 9;
    for (uint32_t j = 0; j < 9; ++j) o[j] = 0.f;
    F3 position = xyz(in.positions[a.to]);
    const float* y = in.prediction ? in.prediction + size_t(i) * in.outputs : nullptr;
    if (y) {
        for (uint32_t j = 0; j < in.outputs; ++j)
            if (!isfinite(y[j])) {
                atomicExch(in.bad, 1u);
                return;
            }
        const F3 from = xyz(in.positions[a.from]), to = xyz(in.positions[a.to]);
        const F3 middle = scale3(add3(from, to), .5);
        const double length = len3(sub3(from, to));
        const auto unit = [](float v) { return fminf(1.f, fmaxf(-1.f, v)); };
        position = add3(middle, scale3(F3{unit(y[3]), unit(y[4]), unit(y[5])}, length));
        bool exact = false;
        for (uint32_t j = 0; j < a.wedges; ++j)
            exact |= exact_vertex(in.exact, a.wedge_from[j]) || exact_vertex(in.exact, a.wedge_to[j]);
        if (!exact && in.quantized) {
            const bool inside = !(position.x < in.low[0] || position.y < in.low[1] || position.z < in.low[2] ||
                                  position.x > in.low[0] + in.extent[0] || position.y > in.low[1] + in.extent[1] ||
                                  position.z > in.low[2] + in.extent[2]);
            if (inside)
                position = {unpack_unorm16(pack_unorm16(position.x, in.low[0], in.extent[0]), in.low[0], in.extent[0]),
                            unpack_unorm16(pack_unorm16(position.y, in.low[1], in.extent[1]), in.low[1], in.extent[1]),
                            unpack_unorm16(pack_unorm16(position.z, in.low[2], in.extent[2]), in.low[2], in.extent[2])};
        }
    }
    o[0] = position.x;
    o[1] = position.y;
    o[2] = position.z;
    if (in.normals)
        for (uint32_t j = 0; j < a.wedges; ++j) {
            const float* n = in.normals + 3 * size_t(a.wedge_to[j]);
            F3 normal = norm3(F3{n[0], n[1], n[2]});
            if (y && in.supervise) normal = norm3(add3(normal, F3{y[6 + 3 * j], y[7 + 3 * j], y[8 + 3 * j]}));
            o[3 + 3 * j] = normal.x;
            o[4 + 3 * j] = normal.y;
            o[5 + 3 * j] = normal.z;
        }
}

__global__ void mapping_kernel(const LegalAction* __restrict__ legal, uint32_t count, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const LegalAction a = legal[i];
    uint32_t* o = out + size_t(i) * 5;
    o[0] = a.wedges;
    o[1] = a.wedge_from[0];
    o[2] = a.wedge_to[0];
    o[3] = a.wedges > 1 ? a.wedge_from[1] : 0xFFFFFFFFu;
    o[4] = a.wedges > 1 ? a.wedge_to[1] : 0xFFFFFFFFu;
}

// Sub-allocations of one workspace, each 256-byte aligned. With a null base
// it only measures the layout.
struct Carve {
    char* base;
    size_t used = 0;
    template <class T> T* take(size_t count) {
        const size_t at = used;
        used += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
        return base ? reinterpret_cast<T*>(base + at) : nullptr;
    }
};

}  // namespace

void ActionSet::upload_attributes(const VertexStreams& streams) {
    const hipStream_t st = stream_;
    const uint32_t n = vertices_;
    const size_t attr_bytes = (streams.normals ? 12 * size_t(n) : 0) + (streams.colors ? 4 * size_t(n) : 0) +
                              (streams.tangents ? 16 * size_t(n) : 0);
    feature_attr_.resize(std::max<size_t>(attr_bytes, 64));
    char* attr = static_cast<char*>(feature_attr_.p);
    feature_tangents_ = nullptr;
    feature_normals_ = nullptr;
    feature_colors_ = nullptr;
    size_t at = 0;
    if (streams.tangents) {
        hip_check(hipMemcpyAsync(attr + at, streams.tangents, 16 * size_t(n), hipMemcpyHostToDevice, st), "upload tangents");
        feature_tangents_ = reinterpret_cast<const float*>(attr + at);
        at += 16 * size_t(n);
    }
    if (streams.normals) {
        hip_check(hipMemcpyAsync(attr + at, streams.normals, 12 * size_t(n), hipMemcpyHostToDevice, st), "upload normals");
        feature_normals_ = reinterpret_cast<const float*>(attr + at);
        at += 12 * size_t(n);
    }
    if (streams.colors) {
        hip_check(hipMemcpyAsync(attr + at, streams.colors, 4 * size_t(n), hipMemcpyHostToDevice, st), "upload colors");
        feature_colors_ = reinterpret_cast<const uint32_t*>(attr + at);
    }
}

void ActionSet::graph_features_device(const VertexStreams& streams, float* device_out) {
    const hipStream_t st = stream_;
    const uint32_t n = vertices_, faces = faces_, corners = 3 * faces_;
    upload_attributes(streams);
    const Attributes a{feature_normals_, has_uv_ ? uv_.as<float2>() : nullptr, feature_colors_,
                       reinterpret_cast<const float4*>(feature_tangents_)};
    const Bounds b = bounds_of(weld_positions_.data(), n);
    const Frame frame{double(b.center.x), double(b.center.y), double(b.center.z), b.radius * 2};

    // Workspace: every array sized before the first launch.
    size_t temp_bytes = 16;
    {
        size_t need = 0;
        auto take = [&](hipError_t e) {
            hip_check(e, "workspace query");
            temp_bytes = std::max(temp_bytes, need);
        };
        uint32_t* k32 = nullptr;
        uint64_t* k64 = nullptr;
        take(exclusive_sum(nullptr, need, k32, k32, n + 1, st));
        take(radix_sort_pairs(nullptr, need, k32, k32, k32, k32, corners, 0, 32, st));
        take(radix_sort_pairs(nullptr, need, k64, k64, nullptr, nullptr, corners, 0, 64, st));
        take(run_length_encode(nullptr, need, k64, k64, k32, k32, corners, st));
        take(radix_sort_pairs(nullptr, need, k64, k64, nullptr, nullptr, 2 * corners, 0, 64, st));
    }
    const size_t n1 = size_t(n) + 1, nc = corners;
    uint32_t *group_count, *group_offsets, *cursor, *members, *canonical, *flags, *corner_key, *corner_id,
        *key_sorted, *corner_sorted, *vertex_count, *vertex_offsets, *counts, *degree, *neighbor_offsets, *neighbors,
        *runs;
    unsigned long long *edge_key, *edge_sorted, *unique, *directed, *directed_sorted;
    float4 *face_normal, *normal_sum;
    void* temp;
    const auto layout = [&](Carve& w) {
        group_count = w.take<uint32_t>(n1);
        group_offsets = w.take<uint32_t>(n1);
        cursor = w.take<uint32_t>(n);
        members = w.take<uint32_t>(n);
        canonical = w.take<uint32_t>(n);
        flags = w.take<uint32_t>(n1);
        corner_key = w.take<uint32_t>(nc);
        corner_id = w.take<uint32_t>(nc);
        key_sorted = w.take<uint32_t>(nc);
        corner_sorted = w.take<uint32_t>(nc);
        vertex_count = w.take<uint32_t>(n1);
        vertex_offsets = w.take<uint32_t>(n1);
        counts = w.take<uint32_t>(nc);
        degree = w.take<uint32_t>(n1);
        neighbor_offsets = w.take<uint32_t>(n1);
        neighbors = w.take<uint32_t>(2 * nc);
        edge_key = w.take<unsigned long long>(nc);
        edge_sorted = w.take<unsigned long long>(nc);
        unique = w.take<unsigned long long>(nc);
        directed = w.take<unsigned long long>(2 * nc);
        directed_sorted = w.take<unsigned long long>(2 * nc);
        face_normal = w.take<float4>(faces);
        normal_sum = w.take<float4>(n);
        temp = w.take<char>(temp_bytes);
        runs = w.take<uint32_t>(1);
    };
    Carve measure{nullptr};
    layout(measure);
    feature_work_.resize(measure.used);
    Carve w{static_cast<char*>(feature_work_.p)};
    layout(w);
    auto run = [&](const char* what, auto call) {
        size_t bytes = temp_bytes;
        hip_check(call(bytes), what);
    };

    // ---- canonical vertices and seams (position groups of the weld).
    hip_check(hipMemsetAsync(group_count, 0, 4 * n1, st), "clear");
    hip_check(hipMemsetAsync(cursor, 0, 4 * size_t(n), st), "clear");
    group_count_kernel<<<blocks(n), 256, 0, st>>>(geometry_.as<uint32_t>(), n, group_count);
    run("group offsets", [&](size_t& bt) {
        return exclusive_sum(temp, bt, group_count, group_offsets, n + 1, st);
    });
    group_fill_kernel<<<blocks(n), 256, 0, st>>>(geometry_.as<uint32_t>(), n, group_offsets, cursor, members);
    canonical_kernel<<<blocks(n), 256, 0, st>>>(a, geometry_.as<uint32_t>(), group_offsets, members, n, canonical,
                                                flags);
    // ---- faces: normals, corners by canonical vertex, canonical edges.
    graph_faces_kernel<<<blocks(faces), 256, 0, st>>>(positions_.as<float4>(), indices_.as<uint32_t>(), canonical,
                                                      faces, frame, face_normal, corner_key, corner_id, edge_key);
    run("corner sort", [&](size_t& bt) {
        return radix_sort_pairs(temp, bt, corner_key, key_sorted, corner_id, corner_sorted, corners, 0, 32, st);
    });
    hip_check(hipMemsetAsync(vertex_count, 0, 4 * n1, st), "clear");
    count_kernel<<<blocks(corners), 256, 0, st>>>(corner_key, corners, vertex_count);
    run("corner offsets", [&](size_t& bt) {
        return exclusive_sum(temp, bt, vertex_count, vertex_offsets, n + 1, st);
    });
    graph_vertices_kernel<<<blocks(n), 256, 0, st>>>(vertex_offsets, corner_sorted, face_normal,
                                                     has_materials_ ? materials_.as<uint16_t>() : nullptr, n, flags,
                                                     normal_sum);
    // ---- unique edges: flags, degrees, sorted neighbour lists.
    run("edge sort", [&](size_t& bt) {
        return radix_sort_pairs(temp, bt, as_u64(edge_key), as_u64(edge_sorted), nullptr, nullptr, corners, 0, 64, st);
    });
    run("edge runs", [&](size_t& bt) {
        return run_length_encode(temp, bt, as_u64(edge_sorted), as_u64(unique), counts, runs, corners, st);
    });
    uint32_t nruns = 0;
    hip_check(hipMemcpyAsync(&nruns, runs, 4, hipMemcpyDeviceToHost, st), "download runs");
    hip_check(hipStreamSynchronize(st), "graph edges");
    if (nruns) {
        unsigned long long last = 0;
        hip_check(hipMemcpy(&last, unique + nruns - 1, 8, hipMemcpyDeviceToHost), "download last edge");
        if (last == kNoEdge) --nruns;
    }
    hip_check(hipMemsetAsync(degree, 0, 4 * n1, st), "clear");
    if (nruns) {
        graph_edges_kernel<<<blocks(nruns), 256, 0, st>>>(unique, counts, nruns, flags, degree, directed);
        run("neighbour sort", [&](size_t& bt) {
            return radix_sort_pairs(temp, bt, as_u64(directed), as_u64(directed_sorted), nullptr, nullptr, 2 * nruns, 0,
                                    64, st);
        });
        low_words_kernel<<<blocks(2 * nruns), 256, 0, st>>>(directed_sorted, 2 * nruns, neighbors);
    }
    run("neighbour offsets", [&](size_t& bt) {
        return exclusive_sum(temp, bt, degree, neighbor_offsets, n + 1, st);
    });
    // ---- the 24 channels.
    graph_features_kernel<<<blocks(n), 256, 0, st>>>(positions_.as<float4>(), a, normal_sum, flags, neighbor_offsets,
                                                     neighbors, n, frame, device_out);
    hip_check(hipGetLastError(), "graph feature launch");
}

void ActionSet::vertex_features(const VertexStreams& streams, std::vector<float>& out) {
    DeviceGuard guard(device_);
    float* x = grow_vertex_features();
    graph_features_device(streams, x);
    out.resize(size_t(vertices_) * kChannels);
    hip_check(hipMemcpyAsync(out.data(), x, out.size() * 4, hipMemcpyDeviceToHost, stream_), "download features");
    hip_check(hipStreamSynchronize(stream_), "graph features");
}

float* ActionSet::grow_vertex_features() {
    feature_vertex_.resize(std::max<size_t>(size_t(vertices_) * kChannels * 4, 4));
    return feature_vertex_.as<float>();
}

float* ActionSet::action_rows_device(const VertexStreams& streams, const float* conditions, uint32_t width,
                                     uint64_t input_faces, float x79) {
    if (width != 80 && width != 128) throw std::invalid_argument("action feature width must be 80 or 128");
    for (uint32_t j = 0; j < 8; ++j)
        if (!std::isfinite(conditions[j])) throw std::invalid_argument("nonfinite action condition");
    if (!input_faces) input_faces = faces_;
    float* x = grow_vertex_features();
    graph_features_device(streams, x);
    feature_rows_.resize(std::max<size_t>(size_t(legal_count_) * width * 4, 4));
    if (!legal_count_) return feature_rows_.as<float>();
    RowInput in{};
    in.legal = legal_.as<LegalAction>();
    in.count = legal_count_;
    in.width = width;
    in.vertex = x;
    in.t = topo();
    in.colors = feature_colors_;
    for (uint32_t j = 0; j < 8; ++j) in.conditions[j] = conditions[j];
    in.scale = diameter_;
    in.face_ratio = float(double(faces_) / double(input_faces));
    in.x79 = x79;
    in.rows = feature_rows_.as<float>();
    action_rows_kernel<<<blocks(legal_count_), 256, 0, stream_>>>(in);
    hip_check(hipGetLastError(), "action feature launch");
    return in.rows;
}

void ActionSet::action_features(const VertexStreams& streams, const float* conditions, uint32_t width,
                                uint64_t input_faces, std::vector<float>& rows) {
    DeviceGuard guard(device_);
    const float* d = action_rows_device(streams, conditions, width, input_faces, -1.f);
    rows.assign(size_t(legal_count_) * width, 0.f);
    if (!rows.empty())
        hip_check(hipMemcpyAsync(rows.data(), d, rows.size() * 4, hipMemcpyDeviceToHost, stream_), "download rows");
    hip_check(hipStreamSynchronize(stream_), "action features");
}

void ActionSet::rank_policy(Policy& policy, const VertexStreams& streams, const float* conditions,
                            uint64_t input_faces, bool preserve_uv, uint32_t* order_out) {
    if (policy.device() != device_) throw std::invalid_argument("policy on another device");
    DeviceGuard guard(device_);
    const float* rows = action_rows_device(streams, conditions, policy.inputs(), input_faces,
                                           policy.architecture() == 4 ? (preserve_uv ? 1.f : 0.f) : -1.f);
    ranked_ = 0;
    if (!legal_count_) return;
    policy_out_.resize(size_t(legal_count_) * policy.outputs() * 4);
    policy.predict_device(rows, policy_out_.as<float>(), legal_count_, stream_);
    rank_scores(policy_out_.as<float>(), policy.outputs(), order_out);
}

void ActionSet::placements(Policy* policy, const VertexStreams& streams, const float* conditions,
                           uint64_t input_faces, const PlacementOptions& options, std::vector<float>& out) {
    if (policy && policy->device() != device_) throw std::invalid_argument("policy on another device");
    if (policy && policy->architecture() != 3 && policy->architecture() != 4)
        throw std::invalid_argument("placements need a placement policy (architecture 3 or 4)");
    if (options.quantized)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(options.low[k]) || !std::isfinite(options.extent[k]) || options.extent[k] < 0 ||
                !std::isfinite(options.low[k] + options.extent[k]))
                throw std::invalid_argument("invalid quantization box");
    if (options.exact && options.exact_words < (size_t(vertices_) + 31) / 32)
        throw std::invalid_argument("exact position bitmap shorter than the vertex count");
    DeviceGuard guard(device_);
    const hipStream_t st = stream_;
    out.assign(size_t(legal_count_) * 9, 0.f);
    PlacementInput in{};
    if (policy) {
        const float* rows = action_rows_device(streams, conditions, policy->inputs(), input_faces,
                                               policy->architecture() == 4 ? (options.preserve_uv ? 1.f : 0.f) : -1.f);
        policy_out_.resize(std::max<size_t>(size_t(legal_count_) * policy->outputs() * 4, 4));
        if (legal_count_) policy->predict_device(rows, policy_out_.as<float>(), legal_count_, st);
        in.prediction = policy_out_.as<float>();
        in.outputs = policy->outputs();
        // The reference supervises normals when any appearance condition is set
        // (src/neural/action_gpu.cu:1780-1782).
        in.supervise = conditions[3] != 0 || conditions[4] != 0 || conditions[5] != 0;
    } else {
        upload_attributes(streams);
    }
    if (!legal_count_) {
        hip_check(hipStreamSynchronize(st), "placements");
        return;
    }
    // Outputs, the exact bitmap and the error flag in one buffer.
    const size_t exact_bytes = options.exact ? 4 * ((size_t(vertices_) + 31) / 32) : 0;
    const size_t out_bytes = (out.size() * 4 + 255) & ~size_t(255);
    placement_work_.resize(out_bytes + exact_bytes + 4);
    char* base = static_cast<char*>(placement_work_.p);
    in.out = reinterpret_cast<float*>(base);
    uint32_t* exact = reinterpret_cast<uint32_t*>(base + out_bytes);
    uint32_t* bad = reinterpret_cast<uint32_t*>(base + out_bytes + exact_bytes);
    if (options.exact)
        hip_check(hipMemcpyAsync(exact, options.exact, exact_bytes, hipMemcpyHostToDevice, st), "upload exact bits");
    hip_check(hipMemsetAsync(bad, 0, 4, st), "clear");
    in.legal = legal_.as<LegalAction>();
    in.count = legal_count_;
    in.positions = positions_.as<float4>();
    in.normals = feature_normals_;
    in.quantized = options.quantized;
    for (int k = 0; k < 3; ++k) {
        in.low[k] = options.low[k];
        in.extent[k] = options.extent[k];
    }
    in.exact = options.exact ? exact : nullptr;
    in.bad = bad;
    placements_kernel<<<blocks(legal_count_), 256, 0, st>>>(in);
    hip_check(hipGetLastError(), "placement launch");
    uint32_t flag = 0;
    hip_check(hipMemcpyAsync(&flag, bad, 4, hipMemcpyDeviceToHost, st), "download flag");
    hip_check(hipMemcpyAsync(out.data(), in.out, out.size() * 4, hipMemcpyDeviceToHost, st), "download placements");
    hip_check(hipStreamSynchronize(st), "placements");
    if (flag) throw std::invalid_argument("nonfinite neural prediction");
}

void ActionSet::mapping(std::vector<uint32_t>& out) {
    DeviceGuard guard(device_);
    out.assign(size_t(legal_count_) * 5, 0);
    if (!legal_count_) return;
    placement_work_.resize(out.size() * 4);
    mapping_kernel<<<blocks(legal_count_), 256, 0, stream_>>>(legal_.as<LegalAction>(), legal_count_,
                                                             placement_work_.as<uint32_t>());
    hip_check(hipGetLastError(), "mapping launch");
    hip_check(hipMemcpyAsync(out.data(), placement_work_.p, out.size() * 4, hipMemcpyDeviceToHost, stream_),
              "download mapping");
    hip_check(hipStreamSynchronize(stream_), "mapping");
}

}  // namespace blitz_hip
