// CPU restatement of the reference's collapse-action state
// (src/neural/action.cpp, blitz::neural::ActionState): geometric welding,
// edge topology and vertex validity, the legality of a directed half-edge
// collapse with its wedge mapping, greedy independent batches and trial
// edits.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE. Only tests/ and bench.py's CPU
// baseline load it (oracle/liboracle.so, built by oracle/Makefile). It is
// pinned by tests/test_actions.py against a build of the reference's own
// ActionState (oracle/_ref, where present), against golden fixtures written
// by that build (tests/golden/action_golden.json), and against invariants a
// legal collapse must keep.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <utility>
#include <vector>

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

struct V3 {
    float x, y, z;
};
// include/blitz/mesh.hpp:22-26: subtraction and cross product in float, dot
// in double, length through the double dot.
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z; }
double length(V3 a) { return std::sqrt(dot(a, a)); }

class Actions {
  public:
    Actions(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat, uint64_t nm,
            const float* uv)
        : n_(uint32_t(nv)) {
        if (!nv || !ni || ni % 3) throw std::invalid_argument("empty mesh");
        positions_.resize(nv);
        std::memcpy(positions_.data(), pos, nv * 12);
        if (uv) uv_.assign(uv, uv + 2 * nv);
        indices_.assign(idx, idx + ni);
        if (nm) materials_.assign(mat, mat + nm);
        weld();
        rebuild();
    }
    uint32_t triangles() const { return uint32_t(indices_.size() / 3); }
    uint16_t material(uint32_t f) const { return materials_.empty() ? 0 : materials_[f]; }

    // src/neural/action.cpp:371-383: every edge, a->b then b->a, kept when
    // legal.
    std::vector<std::pair<uint32_t, uint32_t>> legal_actions() const {
        std::vector<std::pair<uint32_t, uint32_t>> out;
        std::vector<std::pair<uint32_t, uint32_t>> map;
        for (const Edge& e : edges_)
            for (auto [u, v] : {std::pair{e.a, e.b}, std::pair{e.b, e.a}})
                if (mapping(u, v, map)) out.emplace_back(u, v);
        return out;
    }

    // src/neural/action.cpp:280-304 over directed actions in rank order.
    std::vector<uint32_t> independent(const std::vector<std::pair<uint32_t, uint32_t>>& rows, const uint32_t* order,
                                      size_t n, uint32_t maximum, size_t removable) const {
        if (!maximum || maximum > 64) throw std::invalid_argument("action batch outside 1..64");
        std::vector<uint8_t> used(n_, 0);
        std::vector<uint32_t> picked;
        for (size_t k = 0; k < n; ++k) {
            const auto [u, v] = rows.at(order[k]);
            const uint32_t count = edge_count(u, v);
            if (!count) throw std::invalid_argument("stale batch edge");
            if (count > removable || !claim(u, v, used)) continue;
            picked.push_back(order[k]);
            removable -= count;
            if (picked.size() == maximum || !removable) break;
        }
        return picked;
    }

    // src/neural/action.cpp:308-344
    std::vector<uint32_t> trial(const std::vector<std::pair<uint32_t, uint32_t>>& batch) const {
        if (batch.empty() || batch.size() > 64) throw std::invalid_argument("invalid action batch");
        std::vector<uint32_t> remap(n_, kNone);
        std::vector<uint8_t> used(n_, 0);
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (auto [u, v] : batch) {
            if (!mapping(u, v, pairs) || !claim(u, v, used))
                throw std::invalid_argument("illegal, stale or overlapping endpoint actions");
            for (auto [a, b] : pairs) remap[a] = b;
        }
        std::vector<uint32_t> out;
        out.reserve(indices_.size());
        for (uint32_t f = 0; f < triangles(); ++f) {
            uint32_t ids[3];
            bool changed = false;
            for (unsigned j = 0; j < 3; ++j) {
                ids[j] = indices_[3 * f + j];
                if (remap[ids[j]] != kNone) {
                    ids[j] = remap[ids[j]];
                    changed = true;
                }
            }
            if (changed && (geometry_[ids[0]] == geometry_[ids[1]] || geometry_[ids[1]] == geometry_[ids[2]] ||
                            geometry_[ids[0]] == geometry_[ids[2]]))
                continue;
            out.insert(out.end(), ids, ids + 3);
        }
        if (out.empty()) throw std::invalid_argument("batch would remove the entire mesh");
        return out;
    }

    // Ranking scores of the legal actions (src/neural/action.cpp:355-370 and
    // tests/neural_action_gpu.cpp:155-160): 1 constant (0), 3 shortest edge
    // (-float(length((to - from) * (1 / diameter)))), 4 current plane
    // (-float(teacher_cost)).
    std::vector<float> scores(const std::vector<std::pair<uint32_t, uint32_t>>& rows, int mode, double diameter) const {
        std::vector<float> out(rows.size(), 0.f);
        const double scale = std::max(1e-20, diameter);
        for (size_t i = 0; i < rows.size(); ++i) {
            const auto [u, v] = rows[i];
            if (mode == 3) {
                const V3 d = sub(positions_[v], positions_[u]);
                const double s = 1 / scale;
                const V3 e{float(d.x * s), float(d.y * s), float(d.z * s)};
                out[i] = -float(length(e));
            } else if (mode == 4) {
                double cost = 0;
                for (uint32_t k = face_offsets_[u]; k < face_offsets_[u + 1]; ++k) {
                    const uint32_t f = faces_[k];
                    const V3 p = positions_[indices_[3 * f]], q = positions_[indices_[3 * f + 1]],
                             r = positions_[indices_[3 * f + 2]];
                    const V3 normal = cross(sub(q, p), sub(r, p));
                    const double weight = length(normal);
                    const double l = length(normal);
                    const V3 unit = l > 0 ? V3{float(normal.x * (1 / l)), float(normal.y * (1 / l)), float(normal.z * (1 / l))}
                                          : V3{0, 0, 0};
                    const double distance = dot(unit, sub(positions_[v], p)) / scale;
                    cost += distance * distance * weight / (scale * scale);
                }
                out[i] = float(-cost);
            }
        }
        return out;
    }

  private:
    struct Edge {
        uint32_t a, b, count;  // count of incident face corners, capped at 3
    };
    uint32_t n_;
    std::vector<V3> positions_;
    std::vector<float> uv_;
    std::vector<uint32_t> indices_;
    std::vector<uint16_t> materials_;
    std::vector<uint32_t> geometry_, offsets_, neighbors_, face_offsets_, faces_;
    std::vector<Edge> edges_;
    std::vector<uint8_t> boundary_, invalid_;

    // src/neural/action.cpp:8-12,27-47: vertices with bit-equal positions
    // (-0 read as +0) share the smallest index among them.
    void weld() {
        auto key = [&](uint32_t i) {
            const V3 p = positions_[i];
            auto bits = [](float x) {
                uint32_t b;
                const float y = x == 0 ? 0.f : x;
                std::memcpy(&b, &y, 4);
                return b;
            };
            return std::array<uint32_t, 3>{bits(p.x), bits(p.y), bits(p.z)};
        };
        std::vector<uint32_t> order(n_);
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const auto ka = key(a), kb = key(b);
            return ka != kb ? ka < kb : a < b;
        });
        geometry_.resize(n_);
        uint32_t root = order[0];
        for (uint32_t i : order) {
            if (key(i) != key(root)) root = i;
            geometry_[i] = root;
        }
    }

    // src/neural/action.cpp:49-145
    void rebuild() {
        offsets_.assign(n_ + 1, 0);
        face_offsets_.assign(n_ + 1, 0);
        boundary_.assign(n_, 0);
        invalid_.assign(n_, 0);
        edges_.clear();
        std::vector<uint64_t> pairs;
        for (uint32_t f = 0; f < triangles(); ++f) {
            uint32_t v[3];
            for (unsigned j = 0; j < 3; ++j) v[j] = geometry_[indices_[3 * f + j]];
            for (unsigned j = 0; j < 3; ++j) {
                uint32_t a = v[j], b = v[(j + 1) % 3];
                ++face_offsets_[a + 1];
                if (a == b) {
                    for (uint32_t i : v) invalid_[i] = 1;
                    continue;
                }
                if (a > b) std::swap(a, b);
                pairs.push_back(uint64_t(a) << 32 | b);
            }
        }
        std::sort(pairs.begin(), pairs.end());
        for (size_t i = 0; i < pairs.size();) {
            size_t end = i + 1;
            while (end < pairs.size() && pairs[end] == pairs[i]) ++end;
            const uint32_t a = uint32_t(pairs[i] >> 32), b = uint32_t(pairs[i]);
            const size_t count = end - i;
            if (count == 1) boundary_[a] = boundary_[b] = 1;
            if (count > 2) invalid_[a] = invalid_[b] = 1;
            edges_.push_back({a, b, uint32_t(std::min<size_t>(3, count))});
            ++offsets_[a + 1];
            ++offsets_[b + 1];
            i = end;
        }
        std::partial_sum(offsets_.begin(), offsets_.end(), offsets_.begin());
        std::partial_sum(face_offsets_.begin(), face_offsets_.end(), face_offsets_.begin());
        neighbors_.resize(edges_.size() * 2);
        faces_.resize(indices_.size());
        auto next = offsets_, face_next = face_offsets_;
        for (const Edge& e : edges_) {
            neighbors_[next[e.a]++] = e.b;
            neighbors_[next[e.b]++] = e.a;
        }
        for (uint32_t f = 0; f < triangles(); ++f)
            for (unsigned j = 0; j < 3; ++j) faces_[face_next[geometry_[indices_[3 * f + j]]]++] = f;
        for (uint32_t i = 0; i < n_; ++i)
            std::sort(neighbors_.begin() + offsets_[i], neighbors_.begin() + offsets_[i + 1]);
        // A vertex whose link (the opposite edges of its faces) is not one
        // connected fan over all of its neighbours is ambiguous.
        std::vector<std::pair<uint32_t, uint32_t>> links;
        std::vector<uint32_t> reached;
        for (uint32_t u = 0; u < n_; ++u) {
            if (face_offsets_[u] == face_offsets_[u + 1] || invalid_[u]) continue;
            links.clear();
            for (uint32_t k = face_offsets_[u]; k < face_offsets_[u + 1]; ++k) {
                const uint32_t f = faces_[k];
                uint32_t other[2], count = 0;
                for (unsigned j = 0; j < 3; ++j) {
                    const uint32_t v = geometry_[indices_[3 * f + j]];
                    if (v != u && count < 2) other[count++] = v;
                }
                if (count != 2) {
                    invalid_[u] = 1;
                    break;
                }
                links.emplace_back(other[0], other[1]);
            }
            if (invalid_[u]) continue;
            reached.assign(1, links[0].first);
            for (size_t i = 0; i < reached.size(); ++i)
                for (auto [a, b] : links) {
                    const uint32_t v = a == reached[i] ? b : b == reached[i] ? a : kNone;
                    if (v != kNone && std::find(reached.begin(), reached.end(), v) == reached.end())
                        reached.push_back(v);
                }
            if (reached.size() != offsets_[u + 1] - offsets_[u]) invalid_[u] = 1;
        }
    }

    bool neighbour(uint32_t u, uint32_t v) const {
        return std::binary_search(neighbors_.begin() + offsets_[u], neighbors_.begin() + offsets_[u + 1], v);
    }
    uint32_t edge_count(uint32_t u, uint32_t v) const {
        const uint32_t lo = std::min(u, v), hi = std::max(u, v);
        auto it = std::lower_bound(edges_.begin(), edges_.end(), std::pair{lo, hi},
                                   [](const Edge& e, std::pair<uint32_t, uint32_t> p) {
                                       return std::pair{e.a, e.b} < p;
                                   });
        return it != edges_.end() && it->a == lo && it->b == hi ? it->count : 0;
    }

    // src/neural/action.cpp:261-279: the footprint is both endpoints and
    // all their neighbours.
    bool claim(uint32_t u0, uint32_t v0, std::vector<uint8_t>& used) const {
        for (uint32_t u : {u0, v0}) {
            if (used[u]) return false;
            for (uint32_t k = offsets_[u]; k < offsets_[u + 1]; ++k)
                if (used[neighbors_[k]]) return false;
        }
        for (uint32_t u : {u0, v0}) {
            used[u] = 1;
            for (uint32_t k = offsets_[u]; k < offsets_[u + 1]; ++k) used[neighbors_[k]] = 1;
        }
        return true;
    }

    // src/neural/action.cpp:146-256: whether collapsing `from` onto `to` is
    // legal, and the wedge mapping (index of `from` -> index of `to` in each
    // face holding both).
    bool mapping(uint32_t from, uint32_t to, std::vector<std::pair<uint32_t, uint32_t>>& remap) const {
        remap.clear();
        if (from >= n_ || to >= n_ || from == to || invalid_[from] || invalid_[to]) return false;
        if (!neighbour(from, to)) return false;
        std::vector<uint32_t> opposite, common, from_wedges;
        std::vector<uint16_t> from_materials, to_materials;
        uint32_t edge_faces = 0;
        for (uint32_t k = face_offsets_[from]; k < face_offsets_[from + 1]; ++k) {
            const uint32_t f = faces_[k];
            uint32_t u = kNone, v = kNone, w = kNone;
            from_materials.push_back(material(f));
            for (unsigned j = 0; j < 3; ++j) {
                const uint32_t id = indices_[3 * f + j], g = geometry_[id];
                if (g == from)
                    u = id;
                else if (g == to)
                    v = id;
                else
                    w = g;
            }
            if (u == kNone) return false;
            from_wedges.push_back(u);
            if (v != kNone) {
                ++edge_faces;
                if (w == kNone) return false;
                opposite.push_back(w);
                remap.emplace_back(u, v);
            }
        }
        if (edge_faces < 1 || edge_faces > 2 || edge_faces >= triangles()) return false;
        // A boundary vertex may only move along its boundary.
        if (boundary_[from] && (!boundary_[to] || edge_faces != 1)) return false;
        std::set_intersection(neighbors_.begin() + offsets_[from], neighbors_.begin() + offsets_[from + 1],
                              neighbors_.begin() + offsets_[to], neighbors_.begin() + offsets_[to + 1],
                              std::back_inserter(common));
        auto unique = [](auto& a) {
            std::sort(a.begin(), a.end());
            a.erase(std::unique(a.begin(), a.end()), a.end());
        };
        unique(opposite);
        if (common != opposite || opposite.size() != edge_faces) return false;
        unique(remap);
        unique(from_wedges);
        // Each wedge of `from` must have exactly one destination wedge.
        if (remap.size() != from_wedges.size()) return false;
        for (size_t i = 0; i < remap.size(); ++i)
            if (remap[i].first != from_wedges[i]) return false;
        for (uint32_t k = face_offsets_[to]; k < face_offsets_[to + 1]; ++k) to_materials.push_back(material(faces_[k]));
        unique(from_materials);
        unique(to_materials);
        if (from_materials.size() > 1 && from_materials != to_materials) return false;
        auto map = [&](uint32_t id) {
            for (auto [a, b] : remap)
                if (id == a) return b;
            return id;
        };
        std::vector<std::array<uint32_t, 3>> destination;
        for (uint32_t k = face_offsets_[to]; k < face_offsets_[to + 1]; ++k) {
            const uint32_t f = faces_[k];
            std::array<uint32_t, 3> ids{geometry_[indices_[3 * f]], geometry_[indices_[3 * f + 1]],
                                        geometry_[indices_[3 * f + 2]]};
            std::sort(ids.begin(), ids.end());
            destination.push_back(ids);
        }
        for (uint32_t k = face_offsets_[from]; k < face_offsets_[from + 1]; ++k) {
            const uint32_t f = faces_[k];
            uint32_t old[3], next[3];
            bool removed = false;
            for (unsigned j = 0; j < 3; ++j) {
                old[j] = indices_[3 * f + j];
                next[j] = map(old[j]);
                removed |= geometry_[old[j]] == to;
            }
            if (removed) continue;
            std::array<uint32_t, 3> ids{geometry_[next[0]], geometry_[next[1]], geometry_[next[2]]};
            std::sort(ids.begin(), ids.end());
            if (ids[0] == ids[1] || ids[1] == ids[2] ||
                std::find(destination.begin(), destination.end(), ids) != destination.end())
                return false;
            destination.push_back(ids);
            // No fold: the face normal may turn by less than acos(0.05).
            const V3 before = cross(sub(positions_[old[1]], positions_[old[0]]), sub(positions_[old[2]], positions_[old[0]]));
            const V3 after = cross(sub(positions_[next[1]], positions_[next[0]]), sub(positions_[next[2]], positions_[next[0]]));
            if (length(after) <= 1e-15 || dot(before, after) <= .05 * length(before) * length(after)) return false;
            if (!uv_.empty()) {  // src/neural/action.cpp:13-15,248-253: no UV flip
                auto area = [&](const uint32_t* t) {
                    const float* a = &uv_[2 * t[0]];
                    const float* b = &uv_[2 * t[1]];
                    const float* c = &uv_[2 * t[2]];
                    return (double(b[0]) - a[0]) * (double(c[1]) - a[1]) - (double(b[1]) - a[1]) * (double(c[0]) - a[0]);
                };
                const double x = area(old), y = area(next);
                if (std::abs(x) > 1e-20 && x * y <= 0) return false;
            }
        }
        return true;
    }
};

struct Handle {
    std::unique_ptr<Actions> state;
    std::vector<std::pair<uint32_t, uint32_t>> rows;
};

}  // namespace

extern "C" {
void* oracle_actions_create(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat,
                            uint64_t nm, const uint8_t*, uint64_t, const float* uv) {
    try {
        auto h = std::make_unique<Handle>();
        h->state = std::make_unique<Actions>(pos, nv, idx, ni, mat, nm, uv);
        h->rows = h->state->legal_actions();
        return h.release();
    } catch (...) {
        return nullptr;
    }
}
void oracle_actions_destroy(void* h) { delete static_cast<Handle*>(h); }
int64_t oracle_actions_legal(void* p, uint32_t* pairs, uint64_t capacity) {
    auto* h = static_cast<Handle*>(p);
    if (h->rows.size() > capacity) return -1;
    for (size_t i = 0; i < h->rows.size(); ++i) {
        pairs[2 * i] = h->rows[i].first;
        pairs[2 * i + 1] = h->rows[i].second;
    }
    return int64_t(h->rows.size());
}
int64_t oracle_actions_independent(void* p, const uint32_t* order, uint64_t n, uint32_t maximum,
                                   uint64_t maximum_removed, uint32_t* selected) {
    try {
        auto* h = static_cast<Handle*>(p);
        const auto picked = h->state->independent(h->rows, order, size_t(n), maximum, size_t(maximum_removed));
        std::copy(picked.begin(), picked.end(), selected);
        return int64_t(picked.size());
    } catch (...) {
        return -1;
    }
}
int oracle_actions_scores(void* p, int mode, double diameter, float* out) {
    auto* h = static_cast<Handle*>(p);
    const auto s = h->state->scores(h->rows, mode, diameter);
    std::copy(s.begin(), s.end(), out);
    return 0;
}
int64_t oracle_actions_trial(void* p, const uint32_t* rows, uint64_t n, uint32_t* indices, uint64_t capacity) {
    try {
        auto* h = static_cast<Handle*>(p);
        std::vector<std::pair<uint32_t, uint32_t>> batch;
        for (uint64_t i = 0; i < n; ++i) batch.push_back(h->rows.at(rows[i]));
        const auto out = h->state->trial(batch);
        if (out.size() > capacity) return -1;
        std::copy(out.begin(), out.end(), indices);
        return int64_t(out.size());
    } catch (...) {
        return -1;
    }
}
}
