// Classical quadric edge collapse on the device: the reference's reduce()
// with endpoint (Reuse) output and the default quadric objective
// (src/simplify.cpp:43-329 reduce_impl, :337-367 reduce), the classical
// proposer of the LOD chain and the source of the raster-audit candidates.
//
// The reference works in passes. Each pass rebuilds the faces around every
// vertex and the sorted unique edges (boundary and non-manifold flags),
// makes one candidate per unlocked edge (summed quadrics, the cheaper
// endpoint), sorts the candidates by (cost, u, v) and walks them once: a
// candidate whose endpoints are not yet marked used, that passes the link
// test and the fold test and fits the triangle budget collapses, and marks
// every vertex of the faces around both endpoints used. The faces are then
// remapped and the degenerate ones dropped, in order.
//
// Here every step of a pass but the walk is data-parallel on its own. The
// walk is a greedy independent set in rank order, and two facts make it
// parallel too. (1) A candidate's link test, fold test and removed-face
// count depend only on the faces around its endpoints, and an earlier
// collapse that touched any of those faces would have marked one of its
// endpoints used; so for every candidate the walk can still take, the tests
// give the same answer on the pass's starting mesh, and they are evaluated
// for all candidates at once. (2) Candidate B is skipped because of an
// earlier accepted A exactly when an endpoint of A lies among the vertices of
// the faces around B's endpoints (the relation is symmetric). The walk's
// accepted set is therefore the fixed point of rounds in which a candidate
// is accepted once no earlier undecided candidate has an endpoint in its
// footprint and rejected once an accepted one has (the deterministic
// parallel greedy independent set). The budget only binds at the end of the
// last pass: the accepted set is cut where the reference's running count
// would stop or skip, and when it would skip (one triangle of room left), the
// first later candidate that removes a single face and is not blocked by the
// kept prefix is taken, as the walk would.
//
// The starting state (normalized positions, quadrics accumulated per vertex
// in face order, wedge and material locks, degenerate faces dropped) is made
// on the host in the reference's arithmetic; every pass runs on the device.
#include "reduce.hpp"

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "audit.hpp"
#include "audit_kernels.hpp"
#include "device_buffer.hpp"
#include "device_sort.hpp"

namespace blitz_hip {
namespace {

constexpr uint32_t kLocked = 1, kBoundary = 2;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kNoEdge = ~0ull;
constexpr uint64_t kAbsent = ~0ull;  // sorts after every non-negative double's bits
constexpr int kMaxPasses = 128;      // src/simplify.cpp:158
constexpr uint8_t kUndecided = 0, kAccepted = 1, kRejected = 2;

uint32_t blocks_for(size_t n) { return uint32_t((n + 255) / 256); }

// ---- the reference's Vec3 (include/blitz/mesh.hpp:21-27) and Quadric
// (src/simplify.cpp:10-32) arithmetic, host and device alike.
struct V3 {
    float x, y, z;
};
__host__ __device__ inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double vdot(V3 a, V3 b) {
    return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z;
}
__host__ __device__ inline double vlen(V3 a) { return sqrt(vdot(a, a)); }
__host__ __device__ inline V3 vscale(V3 a, double s) { return {float(a.x * s), float(a.y * s), float(a.z * s)}; }
__host__ __device__ inline V3 vnorm(V3 a) {
    const double l = vlen(a);
    return l > 0 ? vscale(a, 1 / l) : V3{0, 0, 0};
}
// std::max(0.0, c): c unless 0 < c fails.
__host__ __device__ inline double max0(double c) { return 0.0 < c ? c : 0.0; }

__host__ __device__ inline double quadric_cost(const double* a, V3 p) {
    const double x = p.x, y = p.y, z = p.z;
    return max0(a[0] * x * x + 2.0 * a[1] * x * y + 2.0 * a[2] * x * z + 2.0 * a[3] * x + a[4] * y * y +
                2.0 * a[5] * y * z + 2.0 * a[6] * y + a[7] * z * z + 2.0 * a[8] * z + a[9]);
}

void add_plane(double* a, double x, double y, double z, double w, double weight) {
    const double v[4] = {x, y, z, w};
    int k = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = i; j < 4; ++j) a[k++] += v[i] * v[j] * weight;
}

// ---- per-pass kernels ------------------------------------------------------

__global__ void count_faces_kernel(const uint32_t* __restrict__ idx, uint32_t corners, uint32_t* count) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < corners) atomicAdd(&count[idx[c]], 1u);
}

__global__ void fill_faces_kernel(const uint32_t* __restrict__ idx, uint32_t corners,
                                  const uint32_t* __restrict__ offsets, uint32_t* cursor, uint32_t* adj) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= corners) return;
    const uint32_t v = idx[c];
    adj[offsets[v] + atomicAdd(&cursor[v], 1u)] = c / 3;
}

__global__ void edge_keys_kernel(const uint32_t* __restrict__ idx, uint32_t faces, uint64_t* keys) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= 3 * faces) return;
    const uint32_t f = c / 3, k = c % 3;
    uint32_t u = idx[3 * f + k], v = idx[3 * f + (k + 1) % 3];
    if (u > v) {
        const uint32_t t = u;
        u = v;
        v = t;
    }
    keys[c] = u != v ? (uint64_t(u) << 32) | v : kNoEdge;
}

__global__ void keep_locks_kernel(uint32_t* flags, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] &= kLocked;
}

// src/simplify.cpp:172-179: a single face marks a boundary, more than two
// lock both endpoints.
__global__ void edge_flags_kernel(const uint64_t* __restrict__ unique, const uint32_t* __restrict__ mult,
                                  uint32_t edges, uint32_t* flags) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= edges) return;
    const uint32_t u = uint32_t(unique[e] >> 32), v = uint32_t(unique[e]);
    if (mult[e] == 1) {
        atomicOr(&flags[u], kBoundary);
        atomicOr(&flags[v], kBoundary);
    }
    if (mult[e] > 2) {
        atomicOr(&flags[u], kLocked);
        atomicOr(&flags[v], kLocked);
    }
}

struct CandidateInput {
    const uint64_t* unique;
    uint32_t edges;
    const uint32_t* flags;
    const float4* p;
    const double* q;         // 10 per vertex
    const float4* normals;   // null without vertex normals
    double normal_weight;
};

// src/simplify.cpp:181-223 for endpoint output: the cheaper endpoint of the
// summed quadric (a boundary endpoint kept on the boundary), plus the normal
// bend term when the mesh has normals.
__global__ void candidates_kernel(CandidateInput in, uint64_t* cost_key, uint64_t* pair_key, float4* point) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= in.edges) return;
    uint32_t u = uint32_t(in.unique[e] >> 32), v = uint32_t(in.unique[e]);
    const uint32_t fu = in.flags[u], fv = in.flags[v];
    if ((fu | fv) & kLocked) {
        cost_key[e] = kAbsent;
        pair_key[e] = in.unique[e];
        return;
    }
    double sum[10];
    for (int i = 0; i < 10; ++i) sum[i] = in.q[10 * size_t(u) + i] + in.q[10 * size_t(v) + i];
    if ((fv & kBoundary) && !(fu & kBoundary)) {
        const uint32_t t = u;
        u = v;
        v = t;
    }
    const float4 pu4 = in.p[u], pv4 = in.p[v];
    V3 pu{pu4.x, pu4.y, pu4.z}, pv{pv4.x, pv4.y, pv4.z};
    V3 pt = pu;
    double cost = quadric_cost(sum, pt);
    if (!(in.flags[u] & kBoundary) && quadric_cost(sum, pv) < cost) {
        const uint32_t t = u;
        u = v;
        v = t;
        const V3 s = pu;
        pu = pv;
        pv = s;
        pt = pu;
        cost = quadric_cost(sum, pt);
    }
    if (in.normals) {
        const float4 nu4 = in.normals[u], nv4 = in.normals[v];
        const double bend = max0(1 - vdot(vnorm({nu4.x, nu4.y, nu4.z}), vnorm({nv4.x, nv4.y, nv4.z})));
        const V3 d = vsub(pu, pv);
        cost += in.normal_weight * bend * vdot(d, d) * 1e-3;
    }
    uint64_t bits;
    memcpy(&bits, &cost, 8);
    cost_key[e] = bits;
    pair_key[e] = (uint64_t(u) << 32) | v;
    point[e] = make_float4(pt.x, pt.y, pt.z, 0.f);
}

__global__ void gather_u64_kernel(const uint64_t* __restrict__ in, const uint32_t* __restrict__ order, uint32_t n,
                                  uint64_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[order[i]];
}

struct Topology {
    const uint32_t* idx;
    const uint32_t* offsets;  // faces around each vertex (any order)
    const uint32_t* adj;
    const float4* p;
    const uint32_t* flags;
};

__device__ inline V3 pos3(const float4* p, uint32_t i) {
    const float4 v = p[i];
    return {v.x, v.y, v.z};
}

// Per rank: endpoints, the link and fold tests and the removed-face count of
// src/simplify.cpp:231-262 on the pass's starting mesh; kRejected for absent
// and failing candidates.
__global__ void tests_kernel(Topology t, const uint32_t* __restrict__ order, const uint64_t* __restrict__ cost_sorted,
                             const uint64_t* __restrict__ pair_key, const float4* __restrict__ point, uint32_t count,
                             uint32_t* ru, uint32_t* rv, uint32_t* removed_out, uint8_t* state, uint8_t* valid_out) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    const uint32_t e = order[r];
    const uint32_t u = uint32_t(pair_key[e] >> 32), v = uint32_t(pair_key[e]);
    ru[r] = u;
    rv[r] = v;
    bool ok = cost_sorted[r] != kAbsent;
    uint32_t removed = 0;
    if (ok) {
        // Link condition: every vertex next to both u and v lies on a face
        // holding the edge, unless the edge joins two boundary vertices
        // through an interior (more than one face).
        uint32_t edge_faces = 0;
        for (uint32_t k = t.offsets[u]; k < t.offsets[u + 1]; ++k) {
            const uint32_t f = t.adj[k];
            edge_faces += (t.idx[3 * f] == v) | (t.idx[3 * f + 1] == v) | (t.idx[3 * f + 2] == v);
        }
        bool link_ok = !(t.flags[u] & kBoundary) || !(t.flags[v] & kBoundary) || edge_faces == 1;
        for (uint32_t k = t.offsets[v]; k < t.offsets[v + 1] && link_ok; ++k)
            for (int j = 0; j < 3 && link_ok; ++j) {
                const uint32_t a = t.idx[3 * t.adj[k] + j];
                if (a == u || a == v) continue;
                bool neighbor = false, opposite = false;
                for (uint32_t m = t.offsets[u]; m < t.offsets[u + 1]; ++m) {
                    const uint32_t g = t.adj[m];
                    const uint32_t g0 = t.idx[3 * g], g1 = t.idx[3 * g + 1], g2 = t.idx[3 * g + 2];
                    if (g0 == a || g1 == a || g2 == a) {
                        neighbor = true;
                        opposite |= g0 == v || g1 == v || g2 == v;
                    }
                }
                if (neighbor && !opposite) link_ok = false;
            }
        ok = link_ok;
        // Fold test of every face around u and v that survives.
        const float4 c4 = point[e];
        const V3 c{c4.x, c4.y, c4.z};
        for (int side = 0; side < 2 && ok; ++side) {
            const uint32_t vertex = side ? v : u;
            for (uint32_t k = t.offsets[vertex]; k < t.offsets[vertex + 1] && ok; ++k) {
                const uint32_t f = t.adj[k];
                const uint32_t a[3] = {t.idx[3 * f], t.idx[3 * f + 1], t.idx[3 * f + 2]};
                const bool hasu = a[0] == u || a[1] == u || a[2] == u, hasv = a[0] == v || a[1] == v || a[2] == v;
                if (hasu && hasv) {
                    if (vertex == u) ++removed;
                    continue;
                }
                const V3 p0 = pos3(t.p, a[0]), p1 = pos3(t.p, a[1]), p2 = pos3(t.p, a[2]);
                const V3 oldn = vcross(vsub(p1, p0), vsub(p2, p0));
                V3 np[3] = {p0, p1, p2};
                for (int j = 0; j < 3; ++j)
                    if (a[j] == u || a[j] == v) np[j] = c;
                const V3 newn = vcross(vsub(np[1], np[0]), vsub(np[2], np[0]));
                if (vlen(newn) < 1e-15 || vdot(oldn, newn) <= .05 * vlen(oldn) * vlen(newn)) ok = false;
            }
        }
        ok = ok && removed > 0;
    }
    removed_out[r] = removed;
    valid_out[r] = ok;
    state[r] = ok ? kUndecided : kRejected;
}

__global__ void mis_mark_kernel(const uint8_t* __restrict__ state, const uint32_t* __restrict__ ru,
                                const uint32_t* __restrict__ rv, uint32_t count, uint32_t* min_end) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count || state[r] != kUndecided) return;
    atomicMin(&min_end[ru[r]], r);
    atomicMin(&min_end[rv[r]], r);
}

// One round of the greedy independent set: a candidate with an accepted
// endpoint in its footprint is rejected; one with no earlier undecided
// endpoint in its footprint is accepted (its own endpoints marked).
__global__ void mis_decide_kernel(Topology t, uint8_t* state, const uint32_t* __restrict__ ru,
                                  const uint32_t* __restrict__ rv, uint32_t count, const uint32_t* __restrict__ min_end,
                                  uint32_t* accepted_end, uint32_t* undecided) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count || state[r] != kUndecided) return;
    const uint32_t ends[2] = {ru[r], rv[r]};
    bool wait = false, blocked = false;
    for (int s = 0; s < 2 && !blocked; ++s)
        for (uint32_t k = t.offsets[ends[s]]; k < t.offsets[ends[s] + 1] && !blocked; ++k)
            for (int j = 0; j < 3; ++j) {
                const uint32_t y = t.idx[3 * t.adj[k] + j];
                if (__hip_atomic_load(&accepted_end[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    blocked = true;
                    break;
                }
                wait |= min_end[y] < r;
            }
    if (blocked) {
        state[r] = kRejected;
    } else if (wait) {
        atomicAdd(undecided, 1u);
    } else {
        state[r] = kAccepted;
        __hip_atomic_store(&accepted_end[ends[0]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&accepted_end[ends[1]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void accepted_flags_kernel(const uint8_t* __restrict__ state, uint32_t count, uint32_t* flag) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < count) flag[r] = state[r] == kAccepted;
}

// Accepted ranks in rank order with their removed-face counts.
__global__ void accepted_list_kernel(const uint8_t* __restrict__ state, const uint32_t* __restrict__ slot,
                                     const uint32_t* __restrict__ removed, uint32_t count, uint32_t* ranks,
                                     uint32_t* removed_list) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count || state[r] != kAccepted) return;
    ranks[slot[r]] = r;
    removed_list[slot[r]] = removed[r];
}

// Every vertex of the faces around the endpoints of ranks[0..n).
__global__ void mark_used_kernel(Topology t, const uint32_t* __restrict__ ranks, uint32_t n,
                                 const uint32_t* __restrict__ ru, const uint32_t* __restrict__ rv, uint32_t* used) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t ends[2] = {ru[ranks[i]], rv[ranks[i]]};
    for (int s = 0; s < 2; ++s)
        for (uint32_t k = t.offsets[ends[s]]; k < t.offsets[ends[s] + 1]; ++k)
            for (int j = 0; j < 3; ++j) used[t.idx[3 * t.adj[k] + j]] = 1u;
}

// The first rank after `after` that the budget still admits (one removed
// face) and that is valid and not blocked.
__global__ void tail_kernel(const uint8_t* __restrict__ valid, const uint32_t* __restrict__ removed,
                            const uint32_t* __restrict__ ru, const uint32_t* __restrict__ rv,
                            const uint32_t* __restrict__ used, uint32_t after, uint32_t count, uint32_t room,
                            uint32_t* first) {
    const uint32_t r = after + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    if (valid[r] && removed[r] <= room && !used[ru[r]] && !used[rv[r]]) atomicMin(first, r);
}

__global__ void apply_kernel(const uint32_t* __restrict__ ranks, uint32_t n, const uint32_t* __restrict__ order,
                             const uint32_t* __restrict__ ru, const uint32_t* __restrict__ rv,
                             const float4* __restrict__ point, uint32_t* map, double* q, float4* p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = ranks[i], u = ru[r], v = rv[r];
    map[v] = u;
    for (int k = 0; k < 10; ++k) q[10 * size_t(u) + k] += q[10 * size_t(v) + k];
    p[u] = point[order[r]];
}

__global__ void iota_kernel(uint32_t* a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void fill_u32_kernel(uint32_t* a, uint32_t n, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = value;
}

// src/simplify.cpp:311-320: remapped faces, degenerate ones dropped, order kept.
__global__ void remap_faces_kernel(const uint32_t* __restrict__ idx, uint32_t faces, const uint32_t* __restrict__ map,
                                   uint32_t* mapped, uint32_t* keep) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces) return;
    const uint32_t a = map[idx[3 * f]], b = map[idx[3 * f + 1]], c = map[idx[3 * f + 2]];
    mapped[3 * f] = a;
    mapped[3 * f + 1] = b;
    mapped[3 * f + 2] = c;
    keep[f] = !(a == b || a == c || b == c);
}

__global__ void compact_faces_kernel(const uint32_t* __restrict__ mapped, const uint16_t* __restrict__ mat_in,
                                     const uint32_t* __restrict__ keep, const uint32_t* __restrict__ slot,
                                     uint32_t faces, uint32_t* idx, uint16_t* mat_out) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= faces || !keep[f]) return;
    const uint32_t s = slot[f];
    idx[3 * s] = mapped[3 * f];
    idx[3 * s + 1] = mapped[3 * f + 1];
    idx[3 * s + 2] = mapped[3 * f + 2];
    if (mat_in) mat_out[s] = mat_in[f];
}

template <class T> T* take(DeviceBuffer& b, size_t n) {
    b.resize(std::max<size_t>(n, 1) * sizeof(T));
    return b.as<T>();
}

}  // namespace

std::vector<uint32_t> reduce_reuse_device(int device, const HostMesh& m, size_t target_triangles,
                                          ReduceReport* report) {
    const size_t nv = m.positions.size() / 3, nf0 = m.indices.size() / 3;
    if (m.indices.size() % 3) throw std::invalid_argument("indices are not whole triangles");
    if (nv >= kNone || m.indices.size() >= kNone) throw std::invalid_argument("mesh exceeds 32-bit counts");
    for (uint32_t i : m.indices)
        if (i >= nv) throw std::invalid_argument("index outside vertex stream");
    if (!m.materials.empty() && m.materials.size() != nf0) throw std::invalid_argument("material count");
    const bool with_normals = !m.normals.empty();
    if (with_normals && m.normals.size() != m.positions.size()) throw std::invalid_argument("normal count");
    ReduceReport local{};
    ReduceReport& rep = report ? *report : local;
    rep = {};
    // ---- starting state, reference arithmetic (src/simplify.cpp:61-131).
    const Bounds b = bounds_of(m.positions.data(), nv);
    const double scale = b.radius * 2;
    std::vector<float4> p(nv);
    std::vector<uint32_t> flags(nv, 0);
    std::vector<double> q(10 * nv, 0.0);
    struct Key {
        uint32_t x, y, z, index;
    };
    std::vector<Key> keys(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        const float x = m.positions[3 * i], y = m.positions[3 * i + 1], z = m.positions[3 * i + 2];
        p[i] = make_float4(float((double(x) - b.center.x) / scale), float((double(y) - b.center.y) / scale),
                           float((double(z) - b.center.z) / scale), 0.f);
        keys[i] = {std::bit_cast<uint32_t>(x == 0 ? 0.f : x), std::bit_cast<uint32_t>(y == 0 ? 0.f : y),
                   std::bit_cast<uint32_t>(z == 0 ? 0.f : z), i};
    }
    std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& c) {
        return std::array{a.x, a.y, a.z} < std::array{c.x, c.y, c.z};
    });
    for (size_t i = 1; i < keys.size(); ++i)
        if (keys[i - 1].x == keys[i].x && keys[i - 1].y == keys[i].y && keys[i - 1].z == keys[i].z)
            flags[keys[i - 1].index] = flags[keys[i].index] = kLocked;
    std::vector<uint16_t> mat(nv, 0);
    std::vector<uint8_t> seen(nv, 0);
    const auto pv = [&](uint32_t i) { return V3{p[i].x, p[i].y, p[i].z}; };
    for (size_t f = 0; f < nf0; ++f) {
        const uint32_t ids[3] = {m.indices[3 * f], m.indices[3 * f + 1], m.indices[3 * f + 2]};
        V3 normal = vcross(vsub(pv(ids[1]), pv(ids[0])), vsub(pv(ids[2]), pv(ids[0])));
        const double area = vlen(normal);
        if (area <= 1e-24) continue;
        normal = vscale(normal, 1 / area);
        const double w = -vdot(normal, pv(ids[0]));
        for (uint32_t i : ids) {
            add_plane(&q[10 * size_t(i)], normal.x, normal.y, normal.z, w, area);
            const uint16_t material = m.materials.empty() ? 0 : m.materials[f];
            if (seen[i] && mat[i] != material) flags[i] |= kLocked;
            mat[i] = material;
            seen[i] = 1;
        }
    }
    std::vector<uint32_t> faces(m.indices);
    std::vector<uint16_t> materials(m.materials);
    if (nf0 > target_triangles) {
        size_t dst = 0;
        for (size_t f = 0; f < nf0; ++f) {
            const uint32_t a = faces[3 * f], c1 = faces[3 * f + 1], c = faces[3 * f + 2];
            if (a == c1 || a == c || c1 == c || vlen(vcross(vsub(pv(c1), pv(a)), vsub(pv(c), pv(a)))) == 0) continue;
            for (int k = 0; k < 3; ++k) faces[3 * dst + k] = faces[3 * f + k];
            if (!materials.empty()) materials[dst] = materials[f];
            ++dst;
        }
        if (!dst) return m.indices;  // no surface: the input as the only incumbent
        faces.resize(3 * dst);
        if (!materials.empty()) materials.resize(dst);
    }
    const size_t target = std::max<size_t>(1, target_triangles);
    uint32_t nf = uint32_t(faces.size() / 3);
    if (nf <= target) return faces;

    // ---- device passes.
    DeviceGuard guard(device);
    hipStream_t st{};
    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    struct StreamOwner {
        hipStream_t s;
        ~StreamOwner() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } owner{st};
    DeviceBuffer d_p, d_q, d_flags, d_idx, d_mat, d_mapped, d_mat2, d_normals, d_count, d_offsets, d_cursor, d_adj,
        d_keys, d_sorted, d_unique, d_mult, d_runs, d_cost, d_pair, d_point, d_order1, d_order, d_costg, d_costs,
        d_ru, d_rv, d_removed, d_state, d_valid, d_min, d_acc, d_und, d_flag, d_slot, d_ranks, d_rlist, d_used,
        d_first, d_map, d_keep, d_kslot, d_temp;
    const uint32_t n = uint32_t(nv);
    float4* P = take<float4>(d_p, n);
    double* Q = take<double>(d_q, 10 * size_t(n));
    uint32_t* F = take<uint32_t>(d_flags, n);
    hip_check(hipMemcpyAsync(P, p.data(), n * sizeof(float4), hipMemcpyHostToDevice, st), "upload positions");
    hip_check(hipMemcpyAsync(Q, q.data(), q.size() * 8, hipMemcpyHostToDevice, st), "upload quadrics");
    hip_check(hipMemcpyAsync(F, flags.data(), n * 4, hipMemcpyHostToDevice, st), "upload flags");
    uint32_t* I = take<uint32_t>(d_idx, faces.size());
    hip_check(hipMemcpyAsync(I, faces.data(), faces.size() * 4, hipMemcpyHostToDevice, st), "upload faces");
    uint16_t* M = nullptr;
    if (!materials.empty()) {
        M = take<uint16_t>(d_mat, materials.size());
        hip_check(hipMemcpyAsync(M, materials.data(), materials.size() * 2, hipMemcpyHostToDevice, st),
                  "upload materials");
    }
    float4* N = nullptr;
    if (with_normals) {
        std::vector<float4> nn(n);
        for (uint32_t i = 0; i < n; ++i)
            nn[i] = make_float4(m.normals[3 * i], m.normals[3 * i + 1], m.normals[3 * i + 2], 0.f);
        N = take<float4>(d_normals, n);
        hip_check(hipMemcpyAsync(N, nn.data(), n * sizeof(float4), hipMemcpyHostToDevice, st), "upload normals");
    }
    size_t temp_bytes = 0;
    const auto run = [&](const char* what, auto call) {
        size_t need = 0;
        hip_check(call(nullptr, need), what);
        if (need > temp_bytes) {
            hip_check(hipStreamSynchronize(st), what);
            temp_bytes = need;
        }
        void* temp = take<char>(d_temp, temp_bytes);
        size_t have = temp_bytes;
        hip_check(call(temp, have), what);
    };
    size_t remaining = nf;
    for (int pass = 0; pass < kMaxPasses && remaining > target; ++pass) {
        ++rep.passes;
        const uint32_t corners = 3 * nf;
        // Faces around every vertex.
        uint32_t* cnt = take<uint32_t>(d_count, n + 1);
        uint32_t* off = take<uint32_t>(d_offsets, n + 1);
        uint32_t* cur = take<uint32_t>(d_cursor, n);
        uint32_t* adj = take<uint32_t>(d_adj, corners);
        fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(n + 1)), 256, 0, st>>>(cnt, n + 1, 0);
        fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(n)), 256, 0, st>>>(cur, n, 0);
        count_faces_kernel<<<blocks_for(corners), 256, 0, st>>>(I, corners, cnt);
        run("face offsets", [&](void* t, size_t& bt) { return exclusive_sum(t, bt, cnt, off, n + 1, st); });
        fill_faces_kernel<<<blocks_for(corners), 256, 0, st>>>(I, corners, off, cur, adj);
        // Sorted unique edges and their face counts.
        uint64_t* keys = take<uint64_t>(d_keys, corners);
        uint64_t* sorted = take<uint64_t>(d_sorted, corners);
        uint64_t* unique = take<uint64_t>(d_unique, corners);
        uint32_t* mult = take<uint32_t>(d_mult, corners);
        uint32_t* runs = take<uint32_t>(d_runs, 1);
        edge_keys_kernel<<<blocks_for(corners), 256, 0, st>>>(I, nf, keys);
        run("edge sort", [&](void* t, size_t& bt) {
            return radix_sort_pairs(t, bt, keys, sorted, nullptr, nullptr, corners, 0, 64, st);
        });
        run("edge runs", [&](void* t, size_t& bt) {
            return run_length_encode(t, bt, sorted, unique, mult, runs, corners, st);
        });
        uint32_t edges = 0;
        uint64_t last = 0;
        hip_check(hipMemcpyAsync(&edges, runs, 4, hipMemcpyDeviceToHost, st), "download edges");
        hip_check(hipStreamSynchronize(st), "edges");
        if (edges) {
            hip_check(hipMemcpyAsync(&last, unique + edges - 1, 8, hipMemcpyDeviceToHost, st), "download edge");
            hip_check(hipStreamSynchronize(st), "edges");
            if (last == kNoEdge) --edges;
        }
        keep_locks_kernel<<<blocks_for(n), 256, 0, st>>>(F, n);
        if (edges) edge_flags_kernel<<<blocks_for(edges), 256, 0, st>>>(unique, mult, edges, F);
        if (!edges) break;
        // Candidates, ranked by (cost, u, v).
        uint64_t* cost = take<uint64_t>(d_cost, edges);
        uint64_t* pair = take<uint64_t>(d_pair, edges);
        float4* point = take<float4>(d_point, edges);
        CandidateInput ci{unique, edges, F, P, Q, N, 1.0};
        candidates_kernel<<<blocks_for(edges), 256, 0, st>>>(ci, cost, pair, point);
        uint32_t* order1 = take<uint32_t>(d_order1, edges);
        uint32_t* order = take<uint32_t>(d_order, edges);
        uint64_t* pair_sorted = take<uint64_t>(d_sorted, edges);  // reuses the edge-sort buffer
        run("pair sort", [&](void* t, size_t& bt) {
            return radix_sort_pairs(t, bt, pair, pair_sorted, nullptr, order1, edges, 0, 64, st);
        });
        uint64_t* cost_g = take<uint64_t>(d_costg, edges);
        uint64_t* cost_s = take<uint64_t>(d_costs, edges);
        gather_u64_kernel<<<blocks_for(edges), 256, 0, st>>>(cost, order1, edges, cost_g);
        run("cost sort", [&](void* t, size_t& bt) {
            return radix_sort_pairs(t, bt, cost_g, cost_s, order1, order, edges, 0, 64, st);
        });
        // Tests on the starting mesh, then the greedy independent set.
        const Topology topo{I, off, adj, P, F};
        uint32_t* ru = take<uint32_t>(d_ru, edges);
        uint32_t* rv = take<uint32_t>(d_rv, edges);
        uint32_t* removed = take<uint32_t>(d_removed, edges);
        uint8_t* state = take<uint8_t>(d_state, edges);
        uint8_t* valid = take<uint8_t>(d_valid, edges);
        tests_kernel<<<blocks_for(edges), 256, 0, st>>>(topo, order, cost_s, pair, point, edges, ru, rv, removed,
                                                         state, valid);
        uint32_t* min_end = take<uint32_t>(d_min, n);
        uint32_t* acc_end = take<uint32_t>(d_acc, n);
        uint32_t* und = take<uint32_t>(d_und, 1);
        fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(n)), 256, 0, st>>>(acc_end, n, 0);
        for (uint32_t round = 0;; ++round) {
            fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(n)), 256, 0, st>>>(min_end, n, kNone);
            fill_u32_kernel<<<1, 1, 0, st>>>(und, 1, 0);
            mis_mark_kernel<<<blocks_for(edges), 256, 0, st>>>(state, ru, rv, edges, min_end);
            mis_decide_kernel<<<blocks_for(edges), 256, 0, st>>>(topo, state, ru, rv, edges, min_end, acc_end, und);
            uint32_t left = 0;
            hip_check(hipMemcpyAsync(&left, und, 4, hipMemcpyDeviceToHost, st), "download undecided");
            hip_check(hipStreamSynchronize(st), "greedy round");
            ++rep.rounds;
            if (!left) break;
            if (round > edges) throw std::runtime_error("greedy rounds did not converge");
        }
        // Accepted ranks in order, cut where the running count stops or skips.
        uint32_t* flag = take<uint32_t>(d_flag, edges + 1);
        uint32_t* slot = take<uint32_t>(d_slot, edges + 1);
        accepted_flags_kernel<<<blocks_for(edges), 256, 0, st>>>(state, edges, flag);
        fill_u32_kernel<<<1, 1, 0, st>>>(flag + edges, 1, 0);
        run("accepted slots", [&](void* t, size_t& bt) { return exclusive_sum(t, bt, flag, slot, edges + 1, st); });
        uint32_t accepted = 0;
        hip_check(hipMemcpyAsync(&accepted, slot + edges, 4, hipMemcpyDeviceToHost, st), "download accepted");
        hip_check(hipStreamSynchronize(st), "accepted");
        uint32_t* ranks = take<uint32_t>(d_ranks, accepted + 1);
        uint32_t* rlist = take<uint32_t>(d_rlist, accepted + 1);
        accepted_list_kernel<<<blocks_for(edges), 256, 0, st>>>(state, slot, removed, edges, ranks, rlist);
        std::vector<uint32_t> host_removed(accepted), host_ranks(accepted);
        if (accepted) {
            hip_check(hipMemcpyAsync(host_removed.data(), rlist, accepted * 4, hipMemcpyDeviceToHost, st),
                      "download removed");
            hip_check(hipMemcpyAsync(host_ranks.data(), ranks, accepted * 4, hipMemcpyDeviceToHost, st),
                      "download ranks");
            hip_check(hipStreamSynchronize(st), "accepted list");
        }
        uint32_t keep_count = accepted;
        bool skip_at_cut = false;
        size_t left = remaining;
        for (uint32_t k = 0; k < accepted; ++k) {
            if (left <= target) {
                keep_count = k;
                break;
            }
            if (left < target + host_removed[k]) {
                keep_count = k;
                skip_at_cut = true;
                break;
            }
            left -= host_removed[k];
        }
        uint32_t collapsed = keep_count;
        if (skip_at_cut && left > target) {
            // The walk skips the candidate at the cut and goes on: the first
            // later candidate that fits the room and is not blocked by the
            // kept prefix collapses, after which the budget is met.
            uint32_t* used = take<uint32_t>(d_used, n);
            uint32_t* first = take<uint32_t>(d_first, 1);
            fill_u32_kernel<<<std::min<uint32_t>(1024, blocks_for(n)), 256, 0, st>>>(used, n, 0);
            fill_u32_kernel<<<1, 1, 0, st>>>(first, 1, kNone);
            if (keep_count) mark_used_kernel<<<blocks_for(keep_count), 256, 0, st>>>(topo, ranks, keep_count, ru, rv, used);
            const uint32_t after = host_ranks[keep_count];
            if (after + 1 < edges)
                tail_kernel<<<blocks_for(edges - after - 1), 256, 0, st>>>(valid, removed, ru, rv, used, after, edges,
                                                                            uint32_t(left - target), first);
            uint32_t found = kNone;
            hip_check(hipMemcpyAsync(&found, first, 4, hipMemcpyDeviceToHost, st), "download tail");
            hip_check(hipStreamSynchronize(st), "tail");
            if (found != kNone) {
                hip_check(hipMemcpyAsync(ranks + keep_count, &found, 4, hipMemcpyHostToDevice, st), "upload tail");
                uint32_t rm = 0;
                hip_check(hipMemcpyAsync(&rm, removed + found, 4, hipMemcpyDeviceToHost, st), "download tail");
                hip_check(hipStreamSynchronize(st), "tail");
                left -= rm;
                ++collapsed;
                ++rep.tail_collapses;
            }
        }
        if (!collapsed) break;
        rep.collapses += collapsed;
        // Collapse, remap and drop degenerate faces in order.
        uint32_t* map = take<uint32_t>(d_map, n);
        iota_kernel<<<blocks_for(n), 256, 0, st>>>(map, n);
        apply_kernel<<<blocks_for(collapsed), 256, 0, st>>>(ranks, collapsed, order, ru, rv, point, map, Q, P);
        uint32_t* mapped = take<uint32_t>(d_mapped, corners);
        uint32_t* keep = take<uint32_t>(d_keep, nf + 1);
        uint32_t* kslot = take<uint32_t>(d_kslot, nf + 1);
        remap_faces_kernel<<<blocks_for(nf), 256, 0, st>>>(I, nf, map, mapped, keep);
        fill_u32_kernel<<<1, 1, 0, st>>>(keep + nf, 1, 0);
        run("face slots", [&](void* t, size_t& bt) { return exclusive_sum(t, bt, keep, kslot, nf + 1, st); });
        uint16_t* mat_out = M ? take<uint16_t>(d_mat2, nf) : nullptr;
        compact_faces_kernel<<<blocks_for(nf), 256, 0, st>>>(mapped, M, keep, kslot, nf, I, mat_out);
        if (M) {
            std::swap(d_mat.p, d_mat2.p);
            std::swap(d_mat.bytes, d_mat2.bytes);
            M = d_mat.as<uint16_t>();
        }
        uint32_t kept = 0;
        hip_check(hipMemcpyAsync(&kept, kslot + nf, 4, hipMemcpyDeviceToHost, st), "download faces");
        hip_check(hipGetLastError(), "reduce launch");
        hip_check(hipStreamSynchronize(st), "reduce pass");
        nf = kept;
        remaining = nf;
    }
    std::vector<uint32_t> out(3 * size_t(nf));
    if (nf) hip_check(hipMemcpyAsync(out.data(), I, out.size() * 4, hipMemcpyDeviceToHost, st), "download faces");
    hip_check(hipStreamSynchronize(st), "reduce");
    return out;
}

}  // namespace blitz_hip
