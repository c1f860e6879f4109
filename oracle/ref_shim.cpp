// C entry points over a build of the reference CPU sources.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE. oracle/Makefile compiles this file
// together with the reference's own CPU sources (read in place under the
// reference checkout, never copied) into oracle/_ref/libblitz_ref.so. It is used
// only to write the golden fixtures in tests/golden/, to cross-check the oracle
// restatement, and as the "reference" CPU baseline in bench.py.
//
// Signatures mirror oracle/audit_oracle.cpp so tests can swap one for the other.
#include "blitz/evaluate.hpp"
#include "blitz/remesher.hpp"
#include "chain_hooks.hpp"
#include "coverage.hpp"
#include "neural/action.hpp"
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <span>
#include <memory>
#include <stdexcept>
#include <vector>

using namespace blitz;

namespace {
struct Measure {
    double error, coverage, coverage_upper, changed_area, normal_degrees;
    uint32_t worst_view, changed_area_worst_view, views_evaluated;
    uint8_t supersample, complete, passed, resource_limited, cancelled;
};
struct OracleSettings {
    double screen_size, limit, max_changed_area;
    uint32_t rotation_seed;
    uint16_t orthographic, perspective;
    uint8_t supersample, max_supersample, force_two_sided, conservative_screen;
    int32_t profile;
    double normal_weight, color_weight, material_weight;
};
MeshView view_of(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat, uint64_t nm,
                 const uint8_t* ds, uint64_t nds) {
    MeshView v;
    v.positions = Stream<Vec3>(reinterpret_cast<const Vec3*>(pos), nv);
    v.indices = {idx, size_t(ni)};
    v.materials = {mat, size_t(nm)};
    v.double_sided = {ds, size_t(nds)};
    return v;
}
Measure out_of(const Measurement& m) {
    return {m.error, m.coverage, m.coverage_upper, m.changed_area, m.normal_degrees, m.worst_view,
            m.changed_area_worst_view, m.views_evaluated, m.supersample, m.complete, m.passed,
            m.resource_limited, m.cancelled};
}
}  // namespace

extern "C" {
int ref_bounds(const float* pos, uint64_t nv, float* center, double* radius) {
    MeshView v;
    v.positions = Stream<Vec3>(reinterpret_cast<const Vec3*>(pos), nv);
    auto b = bounds(v);
    center[0] = b.center.x;
    center[1] = b.center.y;
    center[2] = b.center.z;
    *radius = b.radius;
    return 0;
}

int ref_cameras(const float* center, double radius, double screen, uint16_t ortho, uint16_t persp, uint32_t seed,
                Camera* out, uint64_t capacity) {
    Bounds b{{center[0], center[1], center[2]}, radius};
    auto cams = cameras(b, screen, {ortho, persp, seed});
    if (cams.size() > capacity) return 1;
    for (size_t i = 0; i < cams.size(); ++i) out[i] = cams[i];
    return 0;
}

int ref_raster(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat, uint64_t nm,
               const uint8_t* ds, uint64_t nds, const float* center, double radius, const Camera* cam, double screen,
               uint8_t ss, uint8_t force_two, uint64_t* words, uint64_t capacity, uint8_t* clipped,
               uint32_t* size_out) {
    try {
        auto r = detail::rasterize_coverage(view_of(pos, nv, idx, ni, mat, nm, ds, nds),
                                            {{center[0], center[1], center[2]}, radius}, *cam, screen, ss, force_two);
        *size_out = r.width;
        *clipped = r.clipped;
        if (r.words.size() > capacity) return 1;
        std::memcpy(words, r.words.data(), r.words.size() * 8);
        return 0;
    } catch (const std::length_error&) {
        return 2;
    } catch (...) {
        return 1;
    }
}

int ref_coverage_distance(const uint64_t* a, const uint64_t* b, uint32_t size, uint8_t ss, double* out) {
    try {
        detail::CoverageRaster x, y;
        x.width = x.height = y.width = y.height = size;
        size_t n = (size_t(size) * size + 63) / 64;
        x.words.assign(a, a + n);
        y.words.assign(b, b + n);
        *out = detail::coverage_distance(x, y, ss, true);
        return 0;
    } catch (...) {
        return 1;
    }
}

int ref_evaluate(const float* rpos, uint64_t rnv, const uint32_t* ridx, uint64_t rni, const uint16_t* rmat,
                 uint64_t rnm, const uint8_t* rds, uint64_t rnds, const float* cpos, uint64_t cnv,
                 const uint32_t* cidx, uint64_t cni, const uint16_t* cmat, uint64_t cnm, const uint8_t* cds,
                 uint64_t cnds, const float* center, double radius, const OracleSettings* s, Measure* out,
                 int (*cancelled)(void*), void* user) {
    try {
        EvalSettings e;
        e.profile = Profile(s->profile);
        e.screen_size = s->screen_size;
        e.limit = s->limit;
        e.max_changed_area = s->max_changed_area;
        e.views = {s->orthographic, s->perspective, s->rotation_seed};
        e.supersample = s->supersample;
        e.max_supersample = s->max_supersample;
        e.force_two_sided = s->force_two_sided;
        e.conservative_screen = s->conservative_screen;
        e.weights = {s->normal_weight, s->color_weight, s->material_weight};
        e.force_scalar = true;
        if (cancelled) e.cancelled = [=] { return cancelled(user) != 0; };
        auto m = evaluate(view_of(rpos, rnv, ridx, rni, rmat, rnm, rds, rnds),
                          view_of(cpos, cnv, cidx, cni, cmat, cnm, cds, cnds),
                          {{center[0], center[1], center[2]}, radius}, e);
        *out = out_of(m);
        return 0;
    } catch (...) {
        return 1;
    }
}

// blitz::evaluate with normal and colour streams (null when absent).
int ref_evaluate_streams(const float* rpos, uint64_t rnv, const uint32_t* ridx, uint64_t rni, const uint16_t* rmat,
                         uint64_t rnm, const uint8_t* rds, uint64_t rnds, const float* rnormals,
                         const uint8_t* rcolors, const float* cpos, uint64_t cnv, const uint32_t* cidx, uint64_t cni,
                         const uint16_t* cmat, uint64_t cnm, const uint8_t* cds, uint64_t cnds,
                         const float* cnormals, const uint8_t* ccolors, const float* center, double radius,
                         const OracleSettings* s, Measure* out, int (*cancelled)(void*), void* user) {
    try {
        EvalSettings e;
        e.profile = Profile(s->profile);
        e.screen_size = s->screen_size;
        e.limit = s->limit;
        e.max_changed_area = s->max_changed_area;
        e.views = {s->orthographic, s->perspective, s->rotation_seed};
        e.supersample = s->supersample;
        e.max_supersample = s->max_supersample;
        e.force_two_sided = s->force_two_sided;
        e.conservative_screen = s->conservative_screen;
        e.weights = {s->normal_weight, s->color_weight, s->material_weight};
        e.force_scalar = true;
        if (cancelled) e.cancelled = [=] { return cancelled(user) != 0; };
        MeshView r = view_of(rpos, rnv, ridx, rni, rmat, rnm, rds, rnds);
        MeshView c = view_of(cpos, cnv, cidx, cni, cmat, cnm, cds, cnds);
        if (rnormals) r.normals = Stream<Vec3>(reinterpret_cast<const Vec3*>(rnormals), rnv);
        if (cnormals) c.normals = Stream<Vec3>(reinterpret_cast<const Vec3*>(cnormals), cnv);
        if (rcolors) r.colors = Stream<ColorRGBA8>(reinterpret_cast<const ColorRGBA8*>(rcolors), rnv);
        if (ccolors) c.colors = Stream<ColorRGBA8>(reinterpret_cast<const ColorRGBA8*>(ccolors), cnv);
        *out = out_of(evaluate(r, c, {{center[0], center[1], center[2]}, radius}, e));
        return 0;
    } catch (...) {
        return 1;
    }
}

// blitz::rasterize (attribute raster) of one view as row-major arrays.
int ref_raster_pixels(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat,
                      uint64_t nm, const uint8_t* ds, uint64_t nds, const float* normals, const uint8_t* colors,
                      const float* center, double radius, const Camera* cam, double screen, uint8_t ss,
                      uint8_t force_two, uint8_t* covered, uint8_t* visible, float* depth, float* normal, float* color,
                      uint16_t* material, uint64_t capacity, uint8_t* clipped, uint32_t* size_out) {
    try {
        MeshView m = view_of(pos, nv, idx, ni, mat, nm, ds, nds);
        if (normals) m.normals = Stream<Vec3>(reinterpret_cast<const Vec3*>(normals), nv);
        if (colors) m.colors = Stream<ColorRGBA8>(reinterpret_cast<const ColorRGBA8*>(colors), nv);
        const Raster r = rasterize(m, {{center[0], center[1], center[2]}, radius}, *cam, screen, ss, force_two);
        *size_out = r.width;
        *clipped = r.clipped;
        if (r.pixels.size() > capacity) return 1;
        for (size_t i = 0; i < r.pixels.size(); ++i) {
            const Pixel& p = r.pixels[i];
            covered[i] = p.covered;
            visible[i] = p.visible;
            depth[i] = p.depth;
            normal[3 * i] = p.normal.x;
            normal[3 * i + 1] = p.normal.y;
            normal[3 * i + 2] = p.normal.z;
            color[4 * i] = p.color.x;
            color[4 * i + 1] = p.color.y;
            color[4 * i + 2] = p.color.z;
            color[4 * i + 3] = p.color.w;
            material[i] = p.material;
        }
        return 0;
    } catch (const std::length_error&) {
        return 2;
    } catch (...) {
        return 1;
    }
}

// Classical quadric reduction (src/simplify.cpp reduce): Reuse output keeps the
// source vertex IDs; returns the new index count or -1.
int64_t ref_reduce_reuse(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, uint64_t target,
                         uint32_t* out, uint64_t capacity) {
    try {
        ReduceSettings rs;
        rs.output = OutputMode::Reuse;
        rs.target_triangles = target;
        auto lod = reduce(view_of(pos, nv, idx, ni, nullptr, 0, nullptr, 0), rs);
        if (!lod.shared_vertices || lod.data.indices.size() > capacity) return -1;
        std::memcpy(out, lod.data.indices.data(), lod.data.indices.size() * 4);
        return int64_t(lod.data.indices.size());
    } catch (...) {
        return -1;
    }
}

// The same reduction with per-face materials and per-vertex normals (null
// when absent): material junctions lock, normals add the bend cost.
int64_t ref_reduce_reuse_streams(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni,
                                 const uint16_t* mat, uint64_t nm, const float* normals, uint64_t target,
                                 uint32_t* out, uint64_t capacity) {
    try {
        ReduceSettings rs;
        rs.output = OutputMode::Reuse;
        rs.target_triangles = target;
        MeshView v = view_of(pos, nv, idx, ni, mat, nm, nullptr, 0);
        if (normals) v.normals = Stream<Vec3>(reinterpret_cast<const Vec3*>(normals), nv);
        auto lod = reduce(v, rs);
        if (!lod.shared_vertices || lod.data.indices.size() > capacity) return -1;
        std::memcpy(out, lod.data.indices.data(), lod.data.indices.size() * 4);
        return int64_t(lod.data.indices.size());
    } catch (...) {
        return -1;
    }
}

// Classical CPU generator (src/chain.cpp generate) on the coverage profile.
// Writes per-level triangle counts and the adjacent/source audit maxima.
struct ChainSettings {
    uint8_t levels, beam_width, search_supersample, audit_supersample, max_supersample, pad[3];
    uint16_t candidate_budget, search_ortho, search_persp, audit_ortho, audit_persp, pad2[3];
    uint32_t search_seed, audit_seed;
    double base_pixels, last_pixels, max_lod0_delta_px, transition0, transition1, max_changed_area;
};
int ref_generate(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const ChainSettings* c,
                 uint32_t* triangles, double* source_error, double* adjacent_error, uint64_t* audits,
                 uint32_t* status) {
    try {
        Settings s;
        s.levels = c->levels;
        s.profile = Profile::Coverage;
        s.beam_width = c->beam_width;
        s.search_supersample = c->search_supersample;
        s.audit_supersample = c->audit_supersample;
        s.max_supersample = c->max_supersample;
        s.candidate_budget = c->candidate_budget;
        s.search_views = {c->search_ortho, c->search_persp, c->search_seed};
        s.audit_views = {c->audit_ortho, c->audit_persp, c->audit_seed};
        s.base_pixels = c->base_pixels;
        s.last_pixels = c->last_pixels;
        if (c->max_lod0_delta_px > 0) s.max_lod0_delta_px = c->max_lod0_delta_px;
        s.transition.points = {{0, float(c->transition0)}, {1, float(c->transition1)}};
        s.max_changed_area = c->max_changed_area;
        auto r = generate(view_of(pos, nv, idx, ni, nullptr, 0, nullptr, 0), s);
        for (size_t i = 0; i < r.lods.size(); ++i) {
            triangles[i] = uint32_t(r.lods[i].data.indices.size() / 3);
            source_error[i] = r.lods[i].source_error.error;
            adjacent_error[i] = r.lods[i].adjacent.error;
        }
        uint64_t n = 0;
        for (auto a : r.audit_evaluations) n += a;
        *audits = n;
        *status = uint32_t(r.status);
        return int(r.lods.size());
    } catch (...) {
        return -1;
    }
}

// ---- ActionState (src/neural/action.cpp): collapse legality, independent
// batches and trial edits. The state owns copies of the streams it borrows.
struct RefActions {
    std::vector<Vec3> positions, normals;
    std::vector<Vec2> uv;
    std::vector<ColorRGBA8> colors;
    std::vector<Vec4> tangents;
    std::vector<uint32_t> indices;
    std::vector<uint16_t> materials;
    std::vector<uint8_t> double_sided;
    std::unique_ptr<neural::ActionState> state;
    std::vector<neural::ActionRecord> rows;
};
void* ref_actions_create(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat,
                         uint64_t nm, const uint8_t* ds, uint64_t nds, const float* uv) {
    try {
        auto r = std::make_unique<RefActions>();
        r->positions.assign(reinterpret_cast<const Vec3*>(pos), reinterpret_cast<const Vec3*>(pos) + nv);
        if (uv) r->uv.assign(reinterpret_cast<const Vec2*>(uv), reinterpret_cast<const Vec2*>(uv) + nv);
        r->indices.assign(idx, idx + ni);
        if (nm) r->materials.assign(mat, mat + nm);
        if (nds) r->double_sided.assign(ds, ds + nds);
        MeshView v;
        v.positions = Stream<Vec3>(r->positions.data(), nv);
        if (uv) v.uv = Stream<Vec2>(r->uv.data(), nv);
        v.indices = {r->indices.data(), r->indices.size()};
        v.materials = {r->materials.data(), r->materials.size()};
        v.double_sided = {r->double_sided.data(), r->double_sided.size()};
        r->state = std::make_unique<neural::ActionState>(v);
        r->rows = r->state->actions({});
        return r.release();
    } catch (...) {
        return nullptr;
    }
}
// ActionState over a mesh with normal, colour and tangent streams too (null
// when absent): its graph features (src/neural/graph.cpp) see seams,
// colours and tangents.
void* ref_actions_create_streams(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat,
                                 uint64_t nm, const float* uv, const float* normals, const uint8_t* colors,
                                 const float* tangents) {
    try {
        auto r = std::make_unique<RefActions>();
        r->positions.assign(reinterpret_cast<const Vec3*>(pos), reinterpret_cast<const Vec3*>(pos) + nv);
        if (uv) r->uv.assign(reinterpret_cast<const Vec2*>(uv), reinterpret_cast<const Vec2*>(uv) + nv);
        if (normals) r->normals.assign(reinterpret_cast<const Vec3*>(normals), reinterpret_cast<const Vec3*>(normals) + nv);
        if (colors)
            r->colors.assign(reinterpret_cast<const ColorRGBA8*>(colors), reinterpret_cast<const ColorRGBA8*>(colors) + nv);
        if (tangents)
            r->tangents.assign(reinterpret_cast<const Vec4*>(tangents), reinterpret_cast<const Vec4*>(tangents) + nv);
        r->indices.assign(idx, idx + ni);
        if (nm) r->materials.assign(mat, mat + nm);
        MeshView v;
        v.positions = Stream<Vec3>(r->positions.data(), nv);
        if (uv) v.uv = Stream<Vec2>(r->uv.data(), nv);
        if (normals) v.normals = Stream<Vec3>(r->normals.data(), nv);
        if (colors) v.colors = Stream<ColorRGBA8>(r->colors.data(), nv);
        if (tangents) v.tangents = Stream<Vec4>(r->tangents.data(), nv);
        v.indices = {r->indices.data(), r->indices.size()};
        v.materials = {r->materials.data(), r->materials.size()};
        r->state = std::make_unique<neural::ActionState>(v);
        r->rows = r->state->actions({});
        return r.release();
    } catch (...) {
        return nullptr;
    }
}

// neural::graph (src/neural/graph.cpp:35-187): 24 channels per vertex.
int ref_graph_features(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint16_t* mat,
                       uint64_t nm, const float* uv, const float* normals, const uint8_t* colors,
                       const float* tangents, float* out) {
    try {
        MeshView v = view_of(pos, nv, idx, ni, mat, nm, nullptr, 0);
        if (uv) v.uv = Stream<Vec2>(reinterpret_cast<const Vec2*>(uv), nv);
        if (normals) v.normals = Stream<Vec3>(reinterpret_cast<const Vec3*>(normals), nv);
        if (colors) v.colors = Stream<ColorRGBA8>(reinterpret_cast<const ColorRGBA8*>(colors), nv);
        if (tangents) v.tangents = Stream<Vec4>(reinterpret_cast<const Vec4*>(tangents), nv);
        const neural::Graph g = neural::graph(v);
        std::memcpy(out, g.x.data(), g.x.size() * 4);
        return 0;
    } catch (...) {
        return 1;
    }
}

void ref_actions_destroy(void* h) { delete static_cast<RefActions*>(h); }
// Legal directed actions in the state's enumeration order (edges ascending,
// a->b then b->a): from/to pairs.
int64_t ref_actions_legal(void* h, uint32_t* pairs, uint64_t capacity) {
    auto* r = static_cast<RefActions*>(h);
    if (r->rows.size() > capacity) return -1;
    for (size_t i = 0; i < r->rows.size(); ++i) {
        pairs[2 * i] = r->rows[i].action.from;
        pairs[2 * i + 1] = r->rows[i].action.to;
    }
    return int64_t(r->rows.size());
}
// ActionState::independent over the legal rows in `order` (indices into the
// legal list); writes the selected rows' indices.
int64_t ref_actions_independent(void* h, const uint32_t* order, uint64_t n, uint32_t maximum,
                                uint64_t maximum_removed, uint32_t* selected) {
    try {
        auto* r = static_cast<RefActions*>(h);
        auto picked = r->state->independent(r->rows, {order, size_t(n)}, maximum, size_t(maximum_removed));
        for (size_t i = 0; i < picked.size(); ++i) {
            uint32_t k = 0;
            while (!(r->rows[k].action == picked[i])) ++k;
            selected[i] = k;
        }
        return int64_t(picked.size());
    } catch (...) {
        return -1;
    }
}
// ActionState::trial of the legal rows `rows` (indices into the legal list):
// the candidate's index buffer.
int64_t ref_actions_trial(void* h, const uint32_t* rows, uint64_t n, uint32_t* indices, uint64_t capacity) {
    try {
        auto* r = static_cast<RefActions*>(h);
        std::vector<neural::Action> batch;
        for (uint64_t i = 0; i < n; ++i) batch.push_back(r->rows[rows[i]].action);
        const Lod lod = r->state->trial(batch);
        if (lod.data.indices.size() > capacity) return -1;
        std::memcpy(indices, lod.data.indices.data(), lod.data.indices.size() * 4);
        return int64_t(lod.data.indices.size());
    } catch (...) {
        return -1;
    }
}

// Ranking scores of the legal rows in enumeration order: 3 shortest edge
// (-x[59] of the row), 4 current plane (-teacher_cost), as
// tests/neural_action_gpu.cpp:155-160 ranks them.
int64_t ref_actions_scores(void* h, int mode, float* out) {
    try {
        auto* r = static_cast<RefActions*>(h);
        for (size_t i = 0; i < r->rows.size(); ++i)
            out[i] = mode == 3 ? -r->rows[i].x[59] : float(-r->state->teacher_cost(r->rows[i].action));
        return int64_t(r->rows.size());
    } catch (...) {
        return -1;
    }
}
// The 80 action features of every legal row (ActionRecord::x,
// src/neural/action.cpp:371-450) with the 8 conditions `c`.
int64_t ref_actions_features(void* h, const float* c, float* out) {
    try {
        auto* r = static_cast<RefActions*>(h);
        std::array<float, neural::conditions> cond{};
        for (size_t i = 0; i < cond.size(); ++i) cond[i] = c[i];
        auto rows = r->state->actions(cond);
        for (size_t i = 0; i < rows.size(); ++i)
            std::memcpy(out + i * neural::action_features, rows[i].x.data(), neural::action_features * 4);
        return int64_t(rows.size());
    } catch (...) {
        return -1;
    }
}

// neural::execute_actions (src/neural/action.cpp:453-525) with the rankings of
// tests/neural_action_gpu.cpp:155-160 (1 constant, 3 shortest edge, 4 current
// plane) and, as the gate, blitz::evaluate of the candidate against the input
// under `s` passing. Writes the result's index buffer and the work counters
// ranked, trials, accepted, rejected.
int64_t ref_execute_actions(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const uint8_t* ds,
                            uint64_t nds, uint64_t target, uint32_t budget, int ranking, uint32_t batch,
                            const float* center, double radius, const OracleSettings* s, uint32_t* out,
                            uint64_t capacity, uint64_t* stats) {
    try {
        MeshView input = view_of(pos, nv, idx, ni, nullptr, 0, ds, nds);
        EvalSettings e;
        e.profile = Profile(s->profile);
        e.screen_size = s->screen_size;
        e.limit = s->limit;
        e.max_changed_area = s->max_changed_area;
        e.views = {s->orthographic, s->perspective, s->rotation_seed};
        e.supersample = s->supersample;
        e.max_supersample = s->max_supersample;
        e.force_two_sided = s->force_two_sided;
        e.conservative_screen = s->conservative_screen;
        e.weights = {s->normal_weight, s->color_weight, s->material_weight};
        e.force_scalar = true;
        const Bounds b{{center[0], center[1], center[2]}, radius};
        auto rank = [&](const neural::ActionState& state, std::span<const neural::ActionRecord> rows) {
            std::vector<float> scores(rows.size());
            for (size_t i = 0; i < rows.size(); ++i)
                if (ranking == 3)
                    scores[i] = -rows[i].x[59];
                else if (ranking == 4)
                    scores[i] = float(-state.teacher_cost(rows[i].action));
            return scores;
        };
        auto gate = [&](MeshView candidate) { return evaluate(input, candidate, b, e).passed; };
        neural::ActionStats st;
        const Lod lod = neural::execute_actions(input, {}, size_t(target), budget, rank, gate, &st, {}, batch);
        if (lod.data.indices.size() > capacity) return -1;
        std::memcpy(out, lod.data.indices.data(), lod.data.indices.size() * 4);
        stats[0] = st.ranked;
        stats[1] = st.trials;
        stats[2] = st.accepted;
        stats[3] = st.rejected;
        return int64_t(lod.data.indices.size());
    } catch (...) {
        return -1;
    }
}

// detail::generate_with_hooks (src/chain.cpp:1175) driven by CPU hooks that
// follow src/neural/generate.cpp for the explicit-reuse, source-origin,
// action-policy configuration: hooks.evaluate is blitz::evaluate with the
// refinement bounded by neural::bounded_refinement (:215-224); the guarded
// proposal is neural::execute_actions from the input towards the requested
// target, ranked by shortest edge (3) or current plane (4) as in
// ref_execute_actions, each trial passing four bounded evaluates (source
// search, adjacent search, source audit, adjacent audit; :352-388); confirm
// is the CPU confirmation of :456-548; no overdraw tie-break.
struct HookSettings {
    uint8_t levels, beam_width, search_supersample, audit_supersample, max_supersample, pad[3];
    uint16_t candidate_budget, search_ortho, search_persp, audit_ortho, audit_persp, pad2[3];
    uint32_t search_seed, audit_seed, action_trials, action_batch, ranking, transition_count;
    double base_pixels, last_pixels, max_lod0_delta_px, max_changed_area;
    const float* transition;  // x, y pairs
};
// counters: status, candidate_evaluations, audit_evaluations[4],
// rejected_gates[4], legal_collapses, rejected_collapses, action_trials,
// action_ranked, decoded, finalists, fallback_levels (17 values).
int ref_generate_hooks(const float* pos, uint64_t nv, const uint32_t* idx, uint64_t ni, const HookSettings* c,
                       uint32_t* indices_out, uint64_t capacity, uint64_t* offsets, Measure* source_out,
                       Measure* adjacent_out, double* schedule_out, uint64_t* counters, uint32_t* finalists,
                       uint64_t finalist_capacity) {
    try {
        MeshView source = view_of(pos, nv, idx, ni, nullptr, 0, nullptr, 0);
        Settings s;
        s.levels = c->levels;
        s.profile = Profile::Coverage;
        s.beam_width = c->beam_width;
        s.search_supersample = c->search_supersample;
        s.audit_supersample = c->audit_supersample;
        s.max_supersample = c->max_supersample;
        s.candidate_budget = c->candidate_budget;
        s.search_views = {c->search_ortho, c->search_persp, c->search_seed};
        s.audit_views = {c->audit_ortho, c->audit_persp, c->audit_seed};
        s.base_pixels = c->base_pixels;
        s.last_pixels = c->last_pixels;
        if (c->max_lod0_delta_px > 0) s.max_lod0_delta_px = c->max_lod0_delta_px;
        s.transition.points.clear();
        for (uint32_t i = 0; i < c->transition_count; ++i)
            s.transition.points.push_back({c->transition[2 * i], c->transition[2 * i + 1]});
        s.max_changed_area = c->max_changed_area;
        s.research.output = OutputMode::Reuse;
        s.research.chain = ChainMode::Direct;
        uint64_t legal = 0, rejected = 0, trials = 0, ranked = 0, decoded = 0;
        auto bounded = [](const EvalSettings& e) {
            auto b = e;
            b.max_supersample = neural::bounded_refinement(e);
            return b;
        };
        detail::GenerationHooks hooks;
        hooks.evaluate = [&](MeshView a, MeshView b, const Bounds& bounds, const EvalSettings& e) {
            return evaluate(a, b, bounds, bounded(e));
        };
        hooks.propose_guarded = [&](MeshView input, MeshView fixed_source, MeshView previous, const Bounds& bounds,
                                    const ReduceSettings& rs, const EvalSettings& source_eval,
                                    const EvalSettings& adjacent_eval, const EvalSettings& source_search,
                                    const EvalSettings& adjacent_search) {
            auto pass = [&](MeshView reference, MeshView candidate, const EvalSettings& e) {
                return evaluate(reference, candidate, bounds, bounded(e)).passed;
            };
            auto gate = [&](MeshView candidate) {
                return pass(fixed_source, candidate, source_search) && pass(previous, candidate, adjacent_search) &&
                       pass(fixed_source, candidate, source_eval) && pass(previous, candidate, adjacent_eval);
            };
            auto rank = [&](const neural::ActionState& state, std::span<const neural::ActionRecord> rows) {
                std::vector<float> scores(rows.size());
                for (size_t i = 0; i < rows.size(); ++i)
                    if (c->ranking == 3)
                        scores[i] = -rows[i].x[59];
                    else if (c->ranking == 4)
                        scores[i] = float(-state.teacher_cost(rows[i].action));
                return scores;
            };
            neural::ActionStats st;
            Lod lod = neural::execute_actions(input, {}, rs.target_triangles, c->action_trials, rank, gate, &st, {},
                                              c->action_batch);
            ++decoded;
            legal += st.accepted;
            rejected += st.rejected;
            trials += st.trials;
            ranked += st.ranked;
            return lod;
        };
        hooks.confirm = [&](Result& r) {
            bool good = true;
            auto confirm = [&](MeshView reference, MeshView candidate, const EvalSettings& e, Measurement& output) {
                output = evaluate(reference, candidate, r.reference_bounds, e);
                bool nonfinite = std::isnan(output.error) || std::isnan(output.coverage) ||
                                 std::isnan(output.coverage_upper) || !std::isfinite(output.changed_area) ||
                                 !std::isfinite(output.normal_degrees) ||
                                 (output.passed && (!std::isfinite(output.error) || !std::isfinite(output.coverage) ||
                                                    !std::isfinite(output.coverage_upper)));
                if (!output.cancelled && !output.resource_limited && !nonfinite && output.complete && output.passed)
                    return true;
                output.complete = false;
                output.passed = false;
                return false;
            };
            for (size_t i = 1; i < r.lods.size() && good; ++i) {
                auto& lod = r.lods[i];
                EvalSettings e;
                e.profile = s.profile;
                e.weights = s.weights;
                e.views = s.audit_views;
                e.supersample = s.audit_supersample;
                e.max_supersample = s.max_supersample;
                e.screen_size = lod.schedule.pixels;
                e.max_changed_area = s.max_changed_area;
                double t = s.levels == 2 ? 0 : double(i - 1) / (s.levels - 2);
                e.weights.normal *= s.normal_importance.at(t);
                e.weights.color *= s.attribute_importance.at(t);
                e.weights.material *= s.attribute_importance.at(t);
                e.limit = lod.schedule.source;
                good = confirm(source, lod.view(source), e, lod.source_error);
                if (good) {
                    e.limit = lod.schedule.transition;
                    good = confirm(r.lods[i - 1].view(source), lod.view(source), e, lod.adjacent);
                }
            }
            return good;
        };
        auto r = detail::generate_with_hooks(source, s, {}, &hooks);
        uint64_t total = 0;
        offsets[0] = 0;
        for (size_t i = 0; i < r.lods.size(); ++i) {
            auto v = r.lods[i].view(source);
            if (r.lods[i].shared_vertices == false) return -2;
            if (total + v.indices.size() > capacity) return -3;
            std::memcpy(indices_out + total, v.indices.data(), v.indices.size() * 4);
            total += v.indices.size();
            offsets[i + 1] = total;
            source_out[i] = out_of(r.lods[i].source_error);
            adjacent_out[i] = out_of(r.lods[i].adjacent);
            schedule_out[3 * i] = r.lods[i].schedule.pixels;
            schedule_out[3 * i + 1] = r.lods[i].schedule.transition;
            schedule_out[3 * i + 2] = r.lods[i].schedule.source;
        }
        uint32_t fallback_levels = 0;
        for (size_t i = 1; i < r.lods.size(); ++i)
            fallback_levels += same_mesh_data(r.lods[i].view(source), r.lods[i - 1].view(source));
        uint64_t k = 0;
        counters[k++] = uint64_t(r.status);
        counters[k++] = r.candidate_evaluations;
        for (auto a : r.audit_evaluations) counters[k++] = a;
        for (auto a : r.rejected_gates) counters[k++] = a;
        counters[k++] = legal;
        counters[k++] = rejected;
        counters[k++] = trials;
        counters[k++] = ranked;
        counters[k++] = decoded;
        counters[k++] = r.candidates.size();
        counters[k++] = fallback_levels;
        uint64_t w = 0;
        for (auto& cand : r.candidates)
            for (auto t : cand.triangles)
                if (w < finalist_capacity) finalists[w++] = t;
        return int(r.lods.size());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_generate_hooks: %s\n", e.what());
        return -1;
    }
}

// blitz::schedule (src/chain.cpp:101-118) of the settings of ref_generate_hooks
// plus pixels_per_meter / meters_per_unit (base_pixels 0: unset).
int ref_schedule(const HookSettings* c, double pixels_per_meter, double meters_per_unit, const float* center,
                 double radius, double* out) {
    try {
        Settings s;
        s.levels = c->levels;
        s.pixels_per_meter = pixels_per_meter;
        s.meters_per_unit = meters_per_unit;
        if (c->base_pixels > 0) s.base_pixels = c->base_pixels;
        s.last_pixels = c->last_pixels;
        if (c->max_lod0_delta_px > 0) s.max_lod0_delta_px = c->max_lod0_delta_px;
        s.transition.points.clear();
        for (uint32_t i = 0; i < c->transition_count; ++i)
            s.transition.points.push_back({c->transition[2 * i], c->transition[2 * i + 1]});
        auto steps = schedule({{center[0], center[1], center[2]}, radius}, s);
        for (size_t i = 0; i < steps.size(); ++i) {
            out[3 * i] = steps[i].pixels;
            out[3 * i + 1] = steps[i].transition;
            out[3 * i + 2] = steps[i].source;
        }
        return int(steps.size());
    } catch (...) {
        return -1;
    }
}
}
