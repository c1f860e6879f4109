// Local candidate patches: breadth-first face regions re-meshed by legal
// interior collapses (see patches.hpp). Host C++, FP64 geometry.
#include "patches.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

namespace blitz_hip {
namespace {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t state;
    uint64_t next() { return splitmix64(state++); }
    uint32_t below(uint32_t n) { return uint32_t(next() % n); }
};

struct V3 {
    double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Topology {
    std::vector<uint32_t> vf_offsets, vf;  // vertex -> incident faces
    std::vector<uint32_t> ff_offsets, ff;  // face -> faces sharing an edge
    std::vector<uint8_t> boundary;         // vertex on an edge with one face
};

Topology build(const uint32_t* idx, size_t faces, size_t vertices) {
    Topology t;
    t.vf_offsets.assign(vertices + 1, 0);
    for (size_t i = 0; i < faces * 3; ++i) ++t.vf_offsets[idx[i] + 1];
    for (size_t v = 0; v < vertices; ++v) t.vf_offsets[v + 1] += t.vf_offsets[v];
    t.vf.resize(faces * 3);
    std::vector<uint32_t> cursor(t.vf_offsets.begin(), t.vf_offsets.end() - 1);
    for (size_t f = 0; f < faces; ++f)
        for (int k = 0; k < 3; ++k) t.vf[cursor[idx[3 * f + k]]++] = uint32_t(f);
    // Edges as (min, max, face), sorted: faces sharing an edge are adjacent.
    std::vector<std::array<uint32_t, 3>> edges(faces * 3);
    for (size_t f = 0; f < faces; ++f)
        for (int k = 0; k < 3; ++k) {
            const uint32_t a = idx[3 * f + k], b = idx[3 * f + (k + 1) % 3];
            edges[3 * f + k] = {std::min(a, b), std::max(a, b), uint32_t(f)};
        }
    std::sort(edges.begin(), edges.end());
    std::vector<std::vector<uint32_t>> adj(faces);
    t.boundary.assign(vertices, 0);
    for (size_t i = 0; i < edges.size();) {
        size_t j = i;
        while (j < edges.size() && edges[j][0] == edges[i][0] && edges[j][1] == edges[i][1]) ++j;
        if (j - i == 1) t.boundary[edges[i][0]] = t.boundary[edges[i][1]] = 1;
        for (size_t a = i; a < j; ++a)
            for (size_t b = i; b < j; ++b)
                if (a != b) adj[edges[a][2]].push_back(edges[b][2]);
        i = j;
    }
    t.ff_offsets.assign(faces + 1, 0);
    for (size_t f = 0; f < faces; ++f) {
        auto& a = adj[f];
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        t.ff_offsets[f + 1] = t.ff_offsets[f] + uint32_t(a.size());
    }
    t.ff.reserve(t.ff_offsets.back());
    for (auto& a : adj) t.ff.insert(t.ff.end(), a.begin(), a.end());
    return t;
}

using Face = std::array<uint32_t, 3>;

bool has(const Face& f, uint32_t v) { return f[0] == v || f[1] == v || f[2] == v; }

}  // namespace

Patches local_patches(const float* xyz, size_t vertices, const uint32_t* idx, size_t faces, uint32_t count,
                      uint64_t seed, uint32_t min_faces, uint32_t max_faces) {
    if (!faces || min_faces < 1 || max_faces < min_faces) throw std::invalid_argument("invalid patch sizes");
    const Topology t = build(idx, faces, vertices);
    auto pos = [&](uint32_t v) { return V3{double(xyz[3 * v]), double(xyz[3 * v + 1]), double(xyz[3 * v + 2])}; };
    auto normal = [&](const Face& f) { return cross(sub(pos(f[1]), pos(f[0])), sub(pos(f[2]), pos(f[0]))); };
    Patches out;
    std::vector<uint32_t> stamp(faces, 0);
    std::vector<uint32_t> vstamp(vertices, 0);
    for (uint32_t p = 0; p < count; ++p) {
        Rng rng{splitmix64(seed ^ (uint64_t(p) * 0x9E3779B97F4A7C15ull))};
        const uint32_t target = min_faces + rng.below(max_faces - min_faces + 1);
        const uint32_t mark = p + 1;
        // Breadth-first region from a random seed face.
        std::vector<uint32_t> region{rng.below(uint32_t(faces))};
        stamp[region[0]] = mark;
        for (size_t head = 0; head < region.size() && region.size() < target; ++head) {
            const uint32_t f = region[head];
            for (uint32_t k = t.ff_offsets[f]; k < t.ff_offsets[f + 1] && region.size() < target; ++k) {
                const uint32_t g = t.ff[k];
                if (stamp[g] != mark) {
                    stamp[g] = mark;
                    region.push_back(g);
                }
            }
        }
        std::vector<Face> local(region.size());
        std::vector<uint32_t> interior;
        for (size_t i = 0; i < region.size(); ++i) {
            const uint32_t f = region[i];
            local[i] = {idx[3 * f], idx[3 * f + 1], idx[3 * f + 2]};
            for (int k = 0; k < 3; ++k) {
                const uint32_t v = idx[3 * f + k];
                if (vstamp[v] == mark) continue;
                vstamp[v] = mark;
                bool inside = !t.boundary[v];
                for (uint32_t j = t.vf_offsets[v]; inside && j < t.vf_offsets[v + 1]; ++j)
                    inside = stamp[t.vf[j]] == mark;
                if (inside) interior.push_back(v);
            }
        }
        for (size_t i = interior.size(); i > 1; --i) std::swap(interior[i - 1], interior[rng.below(uint32_t(i))]);
        std::vector<uint8_t> alive(local.size(), 1);
        // Neighbours of v in the current candidate: region faces (local) and
        // the source faces outside the region.
        auto neighbours = [&](uint32_t v) {
            std::vector<uint32_t> n;
            for (size_t i = 0; i < local.size(); ++i)
                if (alive[i] && has(local[i], v))
                    for (uint32_t w : local[i])
                        if (w != v) n.push_back(w);
            for (uint32_t j = t.vf_offsets[v]; j < t.vf_offsets[v + 1]; ++j) {
                const uint32_t f = t.vf[j];
                if (stamp[f] == mark) continue;
                for (int k = 0; k < 3; ++k)
                    if (idx[3 * f + k] != v) n.push_back(idx[3 * f + k]);
            }
            std::sort(n.begin(), n.end());
            n.erase(std::unique(n.begin(), n.end()), n.end());
            return n;
        };
        for (const uint32_t u : interior) {
            const std::vector<uint32_t> nu = neighbours(u);
            if (nu.empty()) continue;
            std::vector<uint32_t> order = nu;
            const V3 pu = pos(u);
            std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                const V3 da = sub(pos(a), pu), db = sub(pos(b), pu);
                return dot(da, da) < dot(db, db) || (dot(da, da) == dot(db, db) && a < b);
            });
            for (const uint32_t v : order) {
                // Link condition: common neighbours are exactly the apexes of
                // the two faces on edge (u, v).
                std::vector<uint32_t> apex;
                for (size_t i = 0; i < local.size(); ++i)
                    if (alive[i] && has(local[i], u) && has(local[i], v))
                        for (uint32_t w : local[i])
                            if (w != u && w != v) apex.push_back(w);
                if (apex.size() != 2) continue;
                const std::vector<uint32_t> nv = neighbours(v);
                std::vector<uint32_t> common;
                std::set_intersection(nu.begin(), nu.end(), nv.begin(), nv.end(), std::back_inserter(common));
                std::sort(apex.begin(), apex.end());
                if (common != apex) continue;
                // No folded or degenerate face after moving u onto v.
                bool legal = true;
                for (size_t i = 0; legal && i < local.size(); ++i) {
                    if (!alive[i] || !has(local[i], u) || has(local[i], v)) continue;
                    Face g = local[i];
                    for (auto& w : g)
                        if (w == u) w = v;
                    const V3 before = normal(local[i]), after = normal(g);
                    const double lb = std::sqrt(dot(before, before)), la = std::sqrt(dot(after, after));
                    legal = la > 1e-12 * std::max(1.0, lb) && dot(before, after) > 0.2 * lb * la;
                }
                if (!legal) continue;
                for (size_t i = 0; i < local.size(); ++i) {
                    if (!alive[i] || !has(local[i], u)) continue;
                    if (has(local[i], v))
                        alive[i] = 0;
                    else
                        for (auto& w : local[i])
                            if (w == u) w = v;
                }
                break;
            }
        }
        std::sort(region.begin(), region.end());
        out.removed.insert(out.removed.end(), region.begin(), region.end());
        out.removed_offsets.push_back(uint32_t(out.removed.size()));
        for (size_t i = 0; i < local.size(); ++i)
            if (alive[i]) out.added.insert(out.added.end(), local[i].begin(), local[i].end());
        out.added_offsets.push_back(uint32_t(out.added.size() / 3));
    }
    return out;
}

std::vector<uint32_t> patch_candidate(const uint32_t* idx, size_t faces, const uint32_t* removed, size_t removed_count,
                                      const uint32_t* added, size_t added_faces) {
    std::vector<uint32_t> out;
    out.reserve((faces - std::min(faces, removed_count) + added_faces) * 3);
    size_t r = 0;
    for (size_t f = 0; f < faces; ++f) {
        while (r < removed_count && removed[r] < f) ++r;
        if (r < removed_count && removed[r] == f) continue;
        out.insert(out.end(), idx + 3 * f, idx + 3 * f + 3);
    }
    out.insert(out.end(), added, added + 3 * added_faces);
    return out;
}

}  // namespace blitz_hip
