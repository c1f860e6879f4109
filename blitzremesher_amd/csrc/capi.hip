// extern "C" layer of include/blitz_hip.h over the C++ audit controller.
// Mirrors the reference C shim: descriptor checks (src/c_api.cpp:131-139),
// exception -> status mapping (src/c_api.cpp:187-202) and truncated error
// messages (src/c_api.cpp:26-32). No exception leaves this file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/blitz_hip.h"
#include "actions.hpp"
#include "model_io.hpp"
#include "policy.hpp"
#include "chain.hpp"
#include "digest.hpp"
#include "noalloc.hpp"
#include "audit.hpp"
#include "device_sort.hpp"
#include "reduce.hpp"
#include "patch_audit.hpp"
#include "patches.hpp"
#include "scorer.hpp"
#include "simplify.hpp"

using namespace blitz_hip;

struct blitz_hip_audit {
    std::unique_ptr<Audit> impl;
};
struct blitz_hip_device_mesh {
    blitz_hip_audit* owner;
    std::unique_ptr<DeviceMesh> mesh;
};
struct blitz_hip_scorer {
    std::unique_ptr<Scorer> impl;
};
struct blitz_hip_binding {
    blitz_hip_audit* owner;
    std::unique_ptr<Binding> impl;
};
struct blitz_hip_actions {
    std::unique_ptr<ActionSet> impl;
    // Host copies of the input's attribute streams for the graph features.
    std::vector<float> normals, tangents;
    std::vector<uint8_t> colors;
    VertexStreams streams() const {
        return {normals.empty() ? nullptr : normals.data(), colors.empty() ? nullptr : colors.data(),
                tangents.empty() ? nullptr : tangents.data()};
    }
};
struct blitz_hip_policy {
    std::unique_ptr<Policy> impl;
};
struct blitz_hip_model {
    ModelData data;
};
struct blitz_hip_chain {
    ChainResult result;
};
struct blitz_hip_graph {
    blitz_hip_scorer* owner;
    std::unique_ptr<Graph> graph;
    std::vector<uint32_t> src, dst;  // host copy of the actions
};

namespace {
void message(char* out, size_t n, const char* text) {
    if (out && n) {
        const size_t k = std::min(n - 1, std::strlen(text));
        std::memcpy(out, text, k);
        out[k] = 0;
    }
}

template <class F> blitz_status guarded(char* error, size_t capacity, F&& body) {
    message(error, capacity, "");
    try {
        body();
        return BLITZ_OK;
    } catch (const Unavailable& e) {
        message(error, capacity, e.what());
        return BLITZ_UNAVAILABLE;
    } catch (const std::invalid_argument& e) {
        message(error, capacity, e.what());
        return BLITZ_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        message(error, capacity, "allocation failed");
        return BLITZ_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        message(error, capacity, e.what());
        return BLITZ_INTERNAL_ERROR;
    } catch (...) {
        message(error, capacity, "unknown exception");
        return BLITZ_INTERNAL_ERROR;
    }
}

bool valid_stream(const blitz_stream& s, size_t n, size_t element, bool required) {
    if (!s.count) return !required;
    if (!s.data) return false;
    return s.count == n && s.stride >= element && (!n || (n - 1) <= (SIZE_MAX - element) / s.stride);
}

void digest_stream(Digest& d, const blitz_stream& s, size_t element) {
    d.value(s.count);
    const auto* base = static_cast<const unsigned char*>(s.data);
    if (s.stride == element) {
        d.bytes(base, s.count * element);
        return;
    }
    for (size_t i = 0; i < s.count; ++i) d.bytes(base + i * s.stride, element);
}

HostMesh host_mesh(const blitz_mesh* m) {
    if (!m || m->struct_size != sizeof(blitz_mesh) || m->abi_version != BLITZ_ABI_VERSION)
        throw std::invalid_argument("null pointer, descriptor size or ABI version mismatch");
    if ((!m->indices && m->index_count) || (!m->materials && m->material_count) ||
        (!m->double_sided && m->double_sided_count))
        throw std::invalid_argument("null index or material pointer");
    if ((reinterpret_cast<uintptr_t>(m->indices) % alignof(uint32_t)) ||
        (reinterpret_cast<uintptr_t>(m->materials) % alignof(uint16_t)))
        throw std::invalid_argument("unaligned index or material stream");
    const size_t n = m->positions.count;
    if (n >= UINT32_MAX || m->index_count % 3 || m->index_count >= UINT32_MAX)
        throw std::invalid_argument("mesh requires 32-bit representable counts and whole triangles");
    if (!valid_stream(m->positions, n, 12, false) || !valid_stream(m->normals, n, 12, false) ||
        !valid_stream(m->uv, n, 8, false) || !valid_stream(m->colors, n, 4, false) ||
        !valid_stream(m->tangents, n, 16, false))
        throw std::invalid_argument("invalid stream count, pointer or stride");
    if (m->material_count && m->material_count != m->index_count / 3)
        throw std::invalid_argument("material count differs from triangle count");
    for (size_t i = 0; i < m->index_count; ++i)
        if (m->indices[i] >= n) throw std::invalid_argument("index outside vertex stream");
    HostMesh h;
    h.positions.resize(n * 3);
    const auto* base = static_cast<const unsigned char*>(m->positions.data);
    for (size_t i = 0; i < n; ++i) std::memcpy(&h.positions[3 * i], base + i * m->positions.stride, 12);
    h.indices.assign(m->indices, m->indices + m->index_count);
    if (m->material_count) h.materials.assign(m->materials, m->materials + m->material_count);
    if (m->double_sided_count) h.double_sided.assign(m->double_sided, m->double_sided + m->double_sided_count);
    if (m->normals.count) {
        h.normals.resize(n * 3);
        const auto* nb = static_cast<const unsigned char*>(m->normals.data);
        for (size_t i = 0; i < n; ++i) std::memcpy(&h.normals[3 * i], nb + i * m->normals.stride, 12);
    }
    if (m->colors.count) {
        h.colors.resize(n * 4);
        const auto* cb = static_cast<const unsigned char*>(m->colors.data);
        for (size_t i = 0; i < n; ++i) std::memcpy(&h.colors[4 * i], cb + i * m->colors.stride, 4);
    }
    Digest d;
    digest_stream(d, m->positions, 12);
    digest_stream(d, m->normals, 12);
    digest_stream(d, m->colors, 4);
    h.vertex_digest[0] = d.a;
    h.vertex_digest[1] = d.b;
    digest_triangles(d, m->indices, m->index_count, m->materials, m->material_count, m->double_sided,
                     m->double_sided_count);
    h.digest[0] = d.a;
    h.digest[1] = d.b;
    Digest pd;
    digest_stream(pd, m->positions, 12);
    h.pos_digest[0] = pd.a;
    h.pos_digest[1] = pd.b;
    return h;
}

Settings settings_of(const blitz_hip_eval_settings* s) {
    if (!s || s->struct_size != sizeof(blitz_hip_eval_settings) || s->version != BLITZ_HIP_API_VERSION)
        throw std::invalid_argument("null pointer, settings size or version mismatch");
    Settings o;
    o.profile = s->profile;
    o.supersample = s->supersample;
    o.max_supersample = s->max_supersample;
    o.force_two_sided = s->force_two_sided != 0;
    o.conservative_screen = s->conservative_screen != 0;
    o.orthographic = s->views_orthographic;
    o.perspective = s->views_perspective;
    o.rotation_seed = s->views_rotation_seed;
    o.screen_size = s->screen_size;
    o.limit = s->limit;
    o.max_changed_area = s->max_changed_area;
    o.normal_weight = s->normal_weight;
    o.color_weight = s->color_weight;
    o.material_weight = s->material_weight;
    if (s->cancelled) {
        auto fn = s->cancelled;
        void* user = s->user_data;
        o.cancelled = [fn, user] { return fn(user) != 0; };
    }
    return o;
}

Bounds bounds_from(const blitz_hip_bounds* b) {
    if (!b) throw std::invalid_argument("null bounds");
    return {{b->center[0], b->center[1], b->center[2]}, b->radius};
}

void to_c(const Measurement& m, blitz_hip_measurement* out) {
    *out = {};
    out->error = m.error;
    out->coverage = m.coverage;
    out->coverage_upper = m.coverage_upper;
    out->changed_area = m.changed_area;
    out->normal_degrees = m.normal_degrees;
    out->worst_view = m.worst_view;
    out->changed_area_worst_view = m.changed_area_worst_view;
    out->views_evaluated = m.views_evaluated;
    out->supersample = m.supersample;
    out->complete = m.complete;
    out->passed = m.passed;
    out->resource_limited = m.resource_limited;
    out->cancelled = m.cancelled;
}

Audit& audit_of(blitz_hip_audit* a) {
    if (!a || !a->impl) throw std::invalid_argument("null audit");
    return *a->impl;
}

const DeviceMesh& mesh_of(const blitz_hip_audit* a, const blitz_hip_device_mesh* m) {
    if (!m || !m->mesh) throw std::invalid_argument("null device mesh");
    if (m->owner != a) throw std::invalid_argument("device mesh belongs to another audit");
    return *m->mesh;
}

ActionSet& actions_of(blitz_hip_actions* a) {
    if (!a || !a->impl) throw std::invalid_argument("null action state");
    return *a->impl;
}
// blitz::Curve from a C curve (src/c_api.cpp:33-42): an empty curve keeps
// the default.
Curve curve_of(const blitz_curve_point* p, size_t n, Curve fallback) {
    if (!n) return fallback;
    if (!p || n > 1024) throw std::invalid_argument("invalid curve");
    fallback.points.clear();
    for (size_t i = 0; i < n; ++i) fallback.points.push_back({p[i].x, p[i].y});
    return fallback;
}

// blitz::Settings from the ABI-5 descriptor (src/c_api.cpp:141-178), the
// fields neural generation reads.
ChainSettings chain_settings_of(const blitz_settings* c) {
    if (!c || c->struct_size != sizeof(blitz_settings) || c->abi_version != BLITZ_ABI_VERSION)
        throw std::invalid_argument("null pointer, descriptor size or ABI version mismatch");
    ChainSettings s;
    s.levels = c->levels;
    s.triangle_overhead_bps = c->triangle_overhead_bps;
    s.profile = c->profile;
    s.objective = c->objective;
    s.beam_width = c->beam_width;
    s.candidate_budget = c->candidate_budget;
    s.search_supersample = c->search_supersample;
    s.audit_supersample = c->audit_supersample;
    s.max_supersample = c->max_supersample;
    s.search_orthographic = c->search_ortho;
    s.search_perspective = c->search_perspective;
    s.search_seed = c->search_seed;
    s.audit_orthographic = c->audit_ortho;
    s.audit_perspective = c->audit_perspective;
    s.audit_seed = c->audit_seed;
    s.pixels_per_meter = c->pixels_per_meter;
    s.meters_per_unit = c->meters_per_unit;
    s.last_pixels = c->last_pixels;
    s.base_pixels = c->base_pixels;
    s.max_lod0_delta_px = c->max_lod0_delta_px;
    s.normal_weight = c->normal_weight;
    s.color_weight = c->color_weight;
    s.material_weight = c->material_weight;
    s.max_changed_area = c->max_changed_area;
    s.transition = curve_of(c->transition, c->transition_count, s.transition);
    s.normal_importance = curve_of(c->normal_importance, c->normal_importance_count, s.normal_importance);
    s.attribute_importance = curve_of(c->attribute_importance, c->attribute_importance_count, s.attribute_importance);
    if (c->cancelled) {
        auto fn = c->cancelled;
        void* user = c->user_data;
        s.cancelled = [fn, user] { return fn(user) != 0; };
    }
    return s;
}

const ChainResult& chain_of(const blitz_hip_chain* c) {
    if (!c) throw std::invalid_argument("null chain");
    return c->result;
}
}  // namespace

extern "C" {
uint32_t blitz_hip_api_version(void) { return BLITZ_HIP_API_VERSION; }

int blitz_hip_available(int32_t device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

blitz_status blitz_hip_device_synchronize(int32_t device) {
    return guarded(nullptr, 0, [&] {
        if (!blitz_hip_available(device)) throw Unavailable("no gfx950 HIP device at the requested index");
        DeviceGuard guard(device);
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

blitz_status blitz_hip_audit_options_init(blitz_hip_audit_options* o, size_t n) {
    if (!o || n != sizeof(*o)) return BLITZ_INVALID_ARGUMENT;
    *o = {};
    o->struct_size = sizeof(*o);
    o->version = BLITZ_HIP_API_VERSION;
    o->device = 0;
    o->memory_mib = 0;
    o->cache_rasters = 0;
    o->autotune = 1;
    return BLITZ_OK;
}

blitz_status blitz_hip_eval_settings_init(blitz_hip_eval_settings* s, size_t n) {
    if (!s || n != sizeof(*s)) return BLITZ_INVALID_ARGUMENT;
    const Settings d;
    *s = {};
    s->struct_size = sizeof(*s);
    s->version = BLITZ_HIP_API_VERSION;
    s->profile = d.profile;
    s->supersample = d.supersample;
    s->max_supersample = d.max_supersample;
    s->views_orthographic = d.orthographic;
    s->views_perspective = d.perspective;
    s->views_rotation_seed = d.rotation_seed;
    s->screen_size = d.screen_size;
    s->limit = d.limit;
    s->max_changed_area = d.max_changed_area;
    s->normal_weight = d.normal_weight;
    s->color_weight = d.color_weight;
    s->material_weight = d.material_weight;
    return BLITZ_OK;
}

blitz_status blitz_hip_bounds_of(const blitz_mesh* m, blitz_hip_bounds* out, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null bounds");
        const HostMesh h = host_mesh(m);
        const Bounds b = bounds_of(h.positions.data(), h.positions.size() / 3);
        *out = {};
        out->center[0] = b.center.x;
        out->center[1] = b.center.y;
        out->center[2] = b.center.z;
        out->radius = b.radius;
    });
}

blitz_status blitz_hip_cameras(const blitz_hip_bounds* b, double screen, uint16_t ortho, uint16_t persp, uint32_t seed,
                               blitz_hip_camera* out, size_t capacity) {
    return guarded(nullptr, 0, [&] {
        const auto cams = cameras(bounds_from(b), screen, ortho, persp, seed);
        if (!out || cams.size() > capacity) throw std::invalid_argument("camera capacity");
        for (size_t i = 0; i < cams.size(); ++i) {
            blitz_hip_camera& c = out[i];
            c = {};
            std::memcpy(c.right, &cams[i].right, 12);
            std::memcpy(c.up, &cams[i].up, 12);
            std::memcpy(c.forward, &cams[i].forward, 12);
            c.distance = cams[i].distance;
            c.focal = cams[i].focal;
            c.scale = cams[i].scale;
            c.perspective = cams[i].perspective;
        }
    });
}

blitz_status blitz_hip_audit_create(const blitz_hip_audit_options* o, blitz_hip_audit** out, char* error,
                                    size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!o || !out || o->struct_size != sizeof(*o) || o->version != BLITZ_HIP_API_VERSION)
            throw std::invalid_argument("null pointer, options size or version mismatch");
        if (!blitz_hip_available(o->device)) throw Unavailable("no gfx950 HIP device at the requested index");
        auto a = std::make_unique<blitz_hip_audit>();
        if (o->face_split > 2) throw std::invalid_argument("face_split must be 0, 1 or 2");
        a->impl = std::make_unique<Audit>(o->device, uint64_t(o->memory_mib) << 20, o->cache_rasters != 0,
                                          o->face_split, o->autotune != 0);
        *out = a.release();
    });
}

void blitz_hip_audit_destroy(blitz_hip_audit* a) { delete a; }

blitz_status blitz_hip_mesh_upload(blitz_hip_audit* a, const blitz_mesh* m, blitz_hip_device_mesh** out, char* error,
                                   size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null mesh destination");
        Audit& audit = audit_of(a);
        const HostMesh h = host_mesh(m);
        auto d = std::make_unique<blitz_hip_device_mesh>();
        d->owner = a;
        d->mesh = audit.upload(h);
        *out = d.release();
    });
}

void blitz_hip_mesh_release(blitz_hip_device_mesh* m) {
    if (!m) return;
    if (m->owner && m->owner->impl) {
        DeviceGuard guard(m->owner->impl->device());
        (void)hipStreamSynchronize(m->owner->impl->stream());
        m->mesh.reset();
    }
    delete m;
}

blitz_status blitz_hip_audit_evaluate(blitz_hip_audit* a, const blitz_mesh* reference, const blitz_mesh* candidate,
                                      const blitz_hip_bounds* b, const blitz_hip_eval_settings* s,
                                      blitz_hip_measurement* out, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null measurement");
        Audit& audit = audit_of(a);
        const Settings settings = settings_of(s);
        const Bounds bounds = bounds_from(b);
        const HostMesh hr = host_mesh(reference), hc = host_mesh(candidate);
        validate_settings(bounds, settings);
        auto dr = audit.upload(hr);
        auto dc = audit.upload(hc);
        to_c(audit.evaluate(*dr, *dc, bounds, settings), out);
    });
}

blitz_status blitz_hip_audit_evaluate_device(blitz_hip_audit* a, const blitz_hip_device_mesh* reference,
                                             const blitz_hip_device_mesh* candidate, const blitz_hip_bounds* b,
                                             const blitz_hip_eval_settings* s, blitz_hip_measurement* out, char* error,
                                             size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null measurement");
        Audit& audit = audit_of(a);
        to_c(audit.evaluate(mesh_of(a, reference), mesh_of(a, candidate), bounds_from(b), settings_of(s)), out);
    });
}

blitz_status blitz_hip_audit_evaluate_batch(blitz_hip_audit* a, const blitz_hip_device_mesh* reference,
                                            const blitz_hip_device_mesh* const* candidates, size_t count,
                                            const blitz_hip_bounds* b, const blitz_hip_eval_settings* s,
                                            blitz_hip_measurement* out, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out || (!candidates && count)) throw std::invalid_argument("null candidate list or output");
        Audit& audit = audit_of(a);
        std::vector<const DeviceMesh*> list(count);
        for (size_t i = 0; i < count; ++i) list[i] = &mesh_of(a, candidates[i]);
        if (!count) return;
        const auto results = audit.evaluate_batch(mesh_of(a, reference), list, bounds_from(b), settings_of(s));
        for (size_t i = 0; i < count; ++i) to_c(results[i], out + i);
    });
}

blitz_status blitz_hip_audit_bind(blitz_hip_audit* a, const blitz_mesh* reference, const blitz_hip_bounds* b,
                                  const blitz_hip_eval_settings* s, blitz_hip_binding** out, char* error,
                                  size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null binding output");
        *out = nullptr;
        Audit& audit = audit_of(a);
        const HostMesh h = host_mesh(reference);
        auto binding = std::make_unique<blitz_hip_binding>();
        binding->owner = a;
        binding->impl = audit.bind(h, bounds_from(b), settings_of(s));
        *out = binding.release();
    });
}

void blitz_hip_binding_release(blitz_hip_binding* b) {
    if (!b) return;
    if (b->owner && b->owner->impl) {
        DeviceGuard guard(b->owner->impl->device());
        b->impl.reset();
    }
    delete b;
}

blitz_status blitz_hip_audit_evaluate_patches(blitz_hip_audit* a, blitz_hip_binding* binding, size_t count,
                                              const uint32_t* removed_offsets, const uint32_t* removed,
                                              const uint32_t* added_offsets, const uint32_t* added,
                                              const blitz_hip_eval_settings* s, blitz_hip_measurement* out,
                                              uint32_t* fallbacks, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!binding || !binding->impl || binding->owner != a) throw std::invalid_argument("binding of another audit");
        if (count && !out) throw std::invalid_argument("null measurement output");
        Audit& audit = audit_of(a);
        PatchSet set;
        set.removed_offsets = removed_offsets;
        set.removed = removed;
        set.added_offsets = added_offsets;
        set.added = added;
        set.count = count;
        const auto results = audit.evaluate_patches(*binding->impl, set, settings_of(s), fallbacks);
        for (size_t i = 0; i < count; ++i) to_c(results[i], out + i);
    });
}

blitz_status blitz_hip_local_patches(const blitz_mesh* m, uint32_t count, uint64_t seed, uint32_t min_faces,
                                     uint32_t max_faces, uint32_t* removed_offsets, uint32_t* removed,
                                     size_t removed_capacity, uint32_t* added_offsets, uint32_t* added,
                                     size_t added_capacity, size_t* removed_count, size_t* added_count, char* error,
                                     size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!removed_count || !added_count) throw std::invalid_argument("null size outputs");
        const HostMesh h = host_mesh(m);
        const Patches p = local_patches(h.positions.data(), h.positions.size() / 3, h.indices.data(),
                                        h.indices.size() / 3, count, seed, min_faces, max_faces);
        *removed_count = p.removed.size();
        *added_count = p.added.size() / 3;
        if (!removed_offsets && !removed && !added_offsets && !added) return;
        if (!removed_offsets || !added_offsets || (!removed && !p.removed.empty()) || (!added && !p.added.empty()))
            throw std::invalid_argument("null patch output arrays");
        if (removed_capacity < p.removed.size() || added_capacity < p.added.size() / 3)
            throw std::invalid_argument("patch output capacity too small");
        std::memcpy(removed_offsets, p.removed_offsets.data(), p.removed_offsets.size() * 4);
        std::memcpy(added_offsets, p.added_offsets.data(), p.added_offsets.size() * 4);
        if (!p.removed.empty()) std::memcpy(removed, p.removed.data(), p.removed.size() * 4);
        if (!p.added.empty()) std::memcpy(added, p.added.data(), p.added.size() * 4);
    });
}

// ---- collapse actions
blitz_status blitz_hip_actions_create(int32_t device, const blitz_mesh* m, blitz_hip_actions** out, char* error,
                                      size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null action state output");
        if (!blitz_hip_available(device)) throw Unavailable("no gfx950 HIP device at the requested index");
        const HostMesh h = host_mesh(m);
        std::vector<float> uv;
        if (m->uv.count) {
            const size_t n = h.positions.size() / 3;
            uv.resize(2 * n);
            const auto* base = static_cast<const unsigned char*>(m->uv.data);
            for (size_t i = 0; i < n; ++i) std::memcpy(&uv[2 * i], base + i * m->uv.stride, 8);
        }
        auto a = std::make_unique<blitz_hip_actions>();
        a->impl = std::make_unique<ActionSet>(device, h, uv.empty() ? nullptr : uv.data());
        a->normals = h.normals;
        a->colors = h.colors;
        if (m->tangents.count) {
            const size_t n = h.positions.size() / 3;
            a->tangents.resize(4 * n);
            const auto* base = static_cast<const unsigned char*>(m->tangents.data);
            for (size_t i = 0; i < n; ++i) std::memcpy(&a->tangents[4 * i], base + i * m->tangents.stride, 16);
        }
        *out = a.release();
    });
}

void blitz_hip_actions_destroy(blitz_hip_actions* a) { delete a; }

blitz_status blitz_hip_actions_info(const blitz_hip_actions* a, uint32_t* edges, uint32_t* legal, double* build_ms) {
    if (!a || !a->impl) return BLITZ_INVALID_ARGUMENT;
    if (edges) *edges = a->impl->edge_count();
    if (legal) *legal = a->impl->legal_count();
    if (build_ms) *build_ms = a->impl->build_ms();
    return BLITZ_OK;
}

blitz_status blitz_hip_actions_legal(blitz_hip_actions* a, uint32_t* pairs, size_t capacity, char* error,
                                     size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        ActionSet& s = actions_of(a);
        if (!pairs || capacity < s.legal_count()) throw std::invalid_argument("legal action capacity too small");
        s.download_legal(pairs);
    });
}

blitz_status blitz_hip_actions_vertex_features(blitz_hip_actions* a, float* out, size_t capacity, char* error,
                                               size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        ActionSet& s = actions_of(a);
        std::vector<float> x;
        s.vertex_features(a->streams(), x);
        if (!out || capacity < x.size()) throw std::invalid_argument("vertex feature capacity too small");
        std::copy(x.begin(), x.end(), out);
    });
}

blitz_status blitz_hip_actions_features(blitz_hip_actions* a, const float* conditions, uint32_t width,
                                        uint64_t input_triangles, float* rows, size_t capacity, char* error,
                                        size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        ActionSet& s = actions_of(a);
        if (!conditions) throw std::invalid_argument("null conditions");
        if (width != 80 && width != 128) throw std::invalid_argument("action feature width must be 80 or 128");
        if (!rows || capacity < size_t(s.legal_count()) * width)
            throw std::invalid_argument("action feature capacity too small");
        std::vector<float> x;
        s.action_features(a->streams(), conditions, width, input_triangles, x);
        std::copy(x.begin(), x.end(), rows);
    });
}

blitz_status blitz_hip_policy_create(int32_t device, uint32_t architecture, uint32_t hidden_width,
                                     const float* weights, size_t count, blitz_hip_policy** out, char* error,
                                     size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null policy output");
        if (!blitz_hip_available(device)) throw Unavailable("no gfx950 HIP device at the requested index");
        auto p = std::make_unique<blitz_hip_policy>();
        p->impl = std::make_unique<Policy>(device, architecture, hidden_width, weights, count);
        *out = p.release();
    });
}

void blitz_hip_policy_destroy(blitz_hip_policy* p) { delete p; }

blitz_status blitz_hip_policy_predict(blitz_hip_policy* p, const float* input, size_t rows, float* output,
                                      size_t output_capacity, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!p || !p->impl) throw std::invalid_argument("null policy");
        if ((!input && rows) || (!output && rows)) throw std::invalid_argument("null policy rows");
        if (output_capacity < rows * p->impl->outputs()) throw std::invalid_argument("policy output capacity too small");
        const auto out = p->impl->predict(input, rows);
        std::copy(out.begin(), out.end(), output);
    });
}

blitz_status blitz_hip_placement_options_init(blitz_hip_placement_options* o, size_t size) {
    if (!o || size != sizeof(*o)) return BLITZ_INVALID_ARGUMENT;
    *o = blitz_hip_placement_options{};
    o->struct_size = sizeof(*o);
    return BLITZ_OK;
}

blitz_status blitz_hip_actions_placements(blitz_hip_actions* a, blitz_hip_policy* p, const float* conditions,
                                          uint64_t input_triangles, const blitz_hip_placement_options* o,
                                          float* out, size_t capacity, char* error, size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        ActionSet& s = actions_of(a);
        if (p && !p->impl) throw std::invalid_argument("null policy");
        if (p && !conditions) throw std::invalid_argument("null conditions");
        if (o && o->struct_size != sizeof(*o)) throw std::invalid_argument("placement options size mismatch");
        if (!out || capacity < size_t(s.legal_count()) * 9) throw std::invalid_argument("placement capacity too small");
        PlacementOptions options;
        if (o) {
            options.preserve_uv = o->preserve_uv != 0;
            options.quantized = o->quantized != 0;
            for (int k = 0; k < 3; ++k) {
                options.low[k] = o->low[k];
                options.extent[k] = o->extent[k];
            }
            options.exact = o->exact_bits;
            options.exact_words = o->exact_words;
        }
        const float zero[8] = {};
        std::vector<float> x;
        s.placements(p ? p->impl.get() : nullptr, a->streams(), conditions ? conditions : zero, input_triangles,
                     options, x);
        std::copy(x.begin(), x.end(), out);
    });
}

blitz_status blitz_hip_actions_mapping(blitz_hip_actions* a, uint32_t* out, size_t capacity, char* error,
                                       size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        ActionSet& s = actions_of(a);
        if (!out || capacity < size_t(s.legal_count()) * 5) throw std::invalid_argument("mapping capacity too small");
        std::vector<uint32_t> x;
        s.mapping(x);
        std::copy(x.begin(), x.end(), out);
    });
}

blitz_status blitz_hip_model_load(const char* path, blitz_hip_model** out, char* error, size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!path || !out) throw std::invalid_argument("null model path or destination");
        auto m = std::make_unique<blitz_hip_model>();
        m->data = load_model(path);
        *out = m.release();
    });
}

void blitz_hip_model_destroy(blitz_hip_model* m) { delete m; }

const char* blitz_hip_model_sha256(const blitz_hip_model* m) { return m ? m->data.sha256.c_str() : nullptr; }

blitz_status blitz_hip_model_info_of(const blitz_hip_model* m, blitz_hip_model_info* info) {
    if (!m || !info || info->struct_size != sizeof(blitz_hip_model_info)) return BLITZ_INVALID_ARGUMENT;
    info->architecture = m->data.architecture;
    info->hidden_width = m->data.hidden_width;
    info->use = m->data.use;
    info->weight_count = m->data.values.size();
    info->weights = m->data.values.data();
    info->provenance_size = m->data.provenance.size();
    info->provenance = m->data.provenance.data();
    return BLITZ_OK;
}

blitz_status blitz_hip_model_save(const char* path, uint32_t architecture, uint32_t hidden_width, uint32_t use,
                                  const float* weights, size_t count, const char* provenance,
                                  size_t provenance_size, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!path || (!weights && count) || (!provenance && provenance_size))
            throw std::invalid_argument("null model path, weights or provenance");
        ModelData m;
        m.architecture = architecture;
        m.hidden_width = hidden_width;
        m.use = use;
        m.values.assign(weights, weights + count);
        if (provenance_size) m.provenance.assign(provenance, provenance_size);
        save_model(path, m);
    });
}

blitz_status blitz_hip_policy_from_model(int32_t device, const blitz_hip_model* m, blitz_hip_policy** out,
                                         char* error, size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!m || !out) throw std::invalid_argument("null model or policy destination");
        if (!policy_inputs(m->data.architecture)) throw std::invalid_argument("model is not an action or placement policy");
        if (!blitz_hip_available(device)) throw Unavailable("no gfx950 HIP device at the requested index");
        auto p = std::make_unique<blitz_hip_policy>();
        p->impl = std::make_unique<Policy>(device, m->data.architecture, m->data.hidden_width, m->data.values.data(),
                                           m->data.values.size());
        *out = p.release();
    });
}

blitz_status blitz_hip_actions_rank_policy(blitz_hip_actions* a, blitz_hip_policy* p, const float* conditions,
                                           uint64_t input_triangles, int32_t preserve_uv, uint32_t* order,
                                           size_t capacity, char* error, size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        ActionSet& s = actions_of(a);
        if (!p || !p->impl) throw std::invalid_argument("null policy");
        if (!conditions) throw std::invalid_argument("null conditions");
        if (order && capacity < s.legal_count()) throw std::invalid_argument("rank order capacity too small");
        s.rank_policy(*p->impl, a->streams(), conditions, input_triangles, preserve_uv != 0, order);
    });
}

blitz_status blitz_hip_actions_independent(blitz_hip_actions* a, const uint32_t* order, size_t n, uint32_t maximum,
                                           uint64_t maximum_removed, uint32_t* selected, size_t* selected_count,
                                           char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if ((!order && n) || !selected || !selected_count) throw std::invalid_argument("null selection arrays");
        const auto picked = actions_of(a).independent(order, n, maximum, maximum_removed);
        std::copy(picked.begin(), picked.end(), selected);
        *selected_count = picked.size();
    });
}

blitz_status blitz_hip_actions_trial(blitz_hip_actions* a, const uint32_t* rows, size_t n, uint32_t* indices,
                                     size_t index_capacity, size_t* index_count, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!rows || !indices || !index_count) throw std::invalid_argument("null trial arrays");
        const auto out = actions_of(a).trial(rows, n);
        if (out.size() > index_capacity) throw std::invalid_argument("trial index capacity too small");
        std::memcpy(indices, out.data(), out.size() * 4);
        *index_count = out.size();
    });
}

blitz_status blitz_hip_actions_rank(blitz_hip_actions* a, uint8_t ranking, blitz_hip_scorer* scorer,
                                    blitz_hip_graph* graph, uint32_t* order, size_t capacity, char* error,
                                    size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        ActionSet& s = actions_of(a);
        if (!order || capacity < s.legal_count()) throw std::invalid_argument("order capacity too small");
        const Graph* g = nullptr;
        if (ranking == 0) {
            if (!scorer || !scorer->impl || !graph || !graph->graph || graph->owner != scorer)
                throw std::invalid_argument("learned ranking needs a scorer and its graph");
            if (scorer->impl->device() != s.device()) throw std::invalid_argument("scorer on another device");
            DeviceGuard guard(scorer->impl->device());
            hip_check(hipStreamSynchronize(scorer->impl->stream()), "scorer");
            g = graph->graph.get();
        }
        s.rank(ranking, s.diameter(), g, order);
    });
}

blitz_status blitz_hip_execute_actions(blitz_hip_audit* a, const blitz_mesh* input, const blitz_hip_bounds* b,
                                       const blitz_hip_eval_settings* gate, uint64_t target, uint32_t budget,
                                       uint8_t ranking, blitz_hip_scorer* scorer, uint8_t batch, uint32_t* indices,
                                       size_t index_capacity, size_t* index_count, blitz_hip_action_stats* stats,
                                       char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!indices || !index_count) throw std::invalid_argument("null output");
        if (stats && stats->struct_size != sizeof(*stats)) throw std::invalid_argument("stats size mismatch");
        Audit& audit = audit_of(a);
        const Settings settings = settings_of(gate);
        const Bounds bounds = bounds_from(b);
        validate_settings(bounds, settings);
        const HostMesh source = host_mesh(input);
        std::vector<float> uv;
        if (input->uv.count) {
            const size_t n = source.positions.size() / 3;
            uv.resize(2 * n);
            const auto* base = static_cast<const unsigned char*>(input->uv.data);
            for (size_t i = 0; i < n; ++i) std::memcpy(&uv[2 * i], base + i * input->uv.stride, 8);
        }
        auto reference = audit.upload(source);
        uint64_t audits = 0;
        // The gate reads only `passed`.
        Settings verdict = settings;
        verdict.verdict_only = true;

        // The gate: an exact audit of the candidate (the input's vertex
        // streams, already on the device, with the trial's triangles)
        // against the input.
        const auto derive = [&](const std::vector<uint32_t>& idx, const std::vector<uint16_t>& mats) {
            Digest d;
            d.a = reference->vertex_digest[0];
            d.b = reference->vertex_digest[1];
            digest_triangles(d, idx.data(), idx.size(), mats.data(), mats.size(), source.double_sided.data(),
                             source.double_sided.size());
            const uint64_t digest[2] = {d.a, d.b};
            return audit.derive(*reference, idx, mats, source.double_sided, digest);
        };
        const ActionGate accept = [&](const std::vector<uint32_t>& idx, const std::vector<uint16_t>& mats) {
            auto candidate = derive(idx, mats);
            ++audits;
            return audit.evaluate(*reference, *candidate, bounds, verdict).passed;
        };
        // The rest of a rejected batch's halving ladder, audited in one
        // batch (each candidate's measurement equals its single audit).
        const ActionLadderGate ladder = [&](const std::vector<std::vector<uint32_t>>& idx,
                                            const std::vector<std::vector<uint16_t>>& mats) {
            std::vector<std::shared_ptr<const DeviceMesh>> owned;
            std::vector<const DeviceMesh*> candidates;
            for (size_t i = 0; i < idx.size(); ++i) {
                owned.push_back(derive(idx[i], mats[i]));
                candidates.push_back(owned.back().get());
            }
            const auto m = audit.evaluate_batch(*reference, candidates, bounds, verdict);
            for (size_t i = 0; i < m.size(); ++i)
                if (m[i].passed) {
                    audits += i + 1;
                    return int(i);
                }
            audits += m.size();
            return -1;
        };
        ExecuteStats st;
        if (ranking == 0 && (!scorer || !scorer->impl)) throw std::invalid_argument("learned ranking needs a scorer");
        const auto out = execute_actions(audit.device(), source, uv.empty() ? nullptr : uv.data(), target, budget,
                                         ranking, batch, accept, st, nullptr, scorer ? scorer->impl.get() : nullptr,
                                         &bounds, nullptr, &ladder);
        if (out.size() > index_capacity) throw std::invalid_argument("index capacity too small");
        std::memcpy(indices, out.data(), out.size() * 4);
        *index_count = out.size();
        if (stats) {
            stats->ranked = st.ranked;
            stats->trials = st.trials;
            stats->accepted = st.accepted;
            stats->rejected = st.rejected;
            stats->states = st.states;
            stats->ms = st.ms;
            stats->build_ms = st.build_ms;
            stats->rank_ms = st.rank_ms;
            stats->select_ms = st.select_ms;
            stats->trial_ms = st.trial_ms;
            stats->gate_ms = st.gate_ms;
        }
    });
}

blitz_status blitz_hip_actions_patches(blitz_hip_actions* a, const uint32_t* rows, size_t n,
                                       uint32_t* removed_offsets, uint32_t* removed, size_t removed_capacity,
                                       uint32_t* added_offsets, uint32_t* added, size_t added_capacity,
                                       size_t* removed_count, size_t* added_count, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if ((!rows && n) || !removed_count || !added_count) throw std::invalid_argument("null patch arguments");
        std::vector<uint32_t> ro, rv, ao, av;
        actions_of(a).patches(rows, n, ro, rv, ao, av);
        *removed_count = rv.size();
        *added_count = av.size() / 3;
        if (!removed_offsets && !removed && !added_offsets && !added) return;
        if (!removed_offsets || !added_offsets || (!removed && !rv.empty()) || (!added && !av.empty()))
            throw std::invalid_argument("null patch output arrays");
        if (removed_capacity < rv.size() || added_capacity < av.size() / 3)
            throw std::invalid_argument("patch output capacity too small");
        std::memcpy(removed_offsets, ro.data(), ro.size() * 4);
        std::memcpy(added_offsets, ao.data(), ao.size() * 4);
        if (!rv.empty()) std::memcpy(removed, rv.data(), rv.size() * 4);
        if (!av.empty()) std::memcpy(added, av.data(), av.size() * 4);
    });
}

blitz_status blitz_hip_raster_coverage(blitz_hip_audit* a, const blitz_hip_device_mesh* m, const blitz_hip_bounds* b,
                                       const blitz_hip_camera* camera, double screen, uint8_t ss, uint8_t force_two,
                                       uint64_t* words, size_t word_capacity, uint32_t* size_out, uint8_t* clipped,
                                       char* error, size_t capacity) {
    bool limited = false;
    const blitz_status status = guarded(error, capacity, [&] {
        if (!camera || !words || !size_out || !clipped) throw std::invalid_argument("null raster output");
        Audit& audit = audit_of(a);
        Camera c{};
        std::memcpy(&c.right, camera->right, 12);
        std::memcpy(&c.up, camera->up, 12);
        std::memcpy(&c.forward, camera->forward, 12);
        c.distance = camera->distance;
        c.focal = camera->focal;
        c.scale = camera->scale;
        c.perspective = camera->perspective;
        std::vector<uint64_t> out;
        uint32_t size = 0;
        bool clip = false;
        if (!audit.raster_row_major(mesh_of(a, m), bounds_from(b), c, screen, ss, force_two != 0, out, size, clip)) {
            *size_out = size;
            limited = true;
            return;
        }
        if (out.size() > word_capacity) throw std::invalid_argument("word capacity too small");
        std::memcpy(words, out.data(), out.size() * 8);
        *size_out = size;
        *clipped = clip;
    });
    if (status == BLITZ_OK && limited) {
        message(error, capacity, "raster exceeds per-view memory budget");
        return BLITZ_BUDGET_LIMITED;
    }
    return status;
}

blitz_status blitz_hip_raster_pixels(blitz_hip_audit* a, const blitz_hip_device_mesh* m, const blitz_hip_bounds* b,
                                     const blitz_hip_camera* camera, double screen, uint8_t ss, uint8_t force_two,
                                     uint8_t* covered, uint8_t* visible, float* depth, float* normal, float* color,
                                     uint16_t* material, size_t sample_capacity, uint32_t* size_out, uint8_t* clipped,
                                     char* error, size_t capacity) {
    bool limited = false;
    const blitz_status status = guarded(error, capacity, [&] {
        if (!camera || !covered || !visible || !depth || !normal || !color || !material || !size_out || !clipped)
            throw std::invalid_argument("null raster output");
        Audit& audit = audit_of(a);
        Camera c{};
        std::memcpy(&c.right, camera->right, 12);
        std::memcpy(&c.up, camera->up, 12);
        std::memcpy(&c.forward, camera->forward, 12);
        c.distance = camera->distance;
        c.focal = camera->focal;
        c.scale = camera->scale;
        c.perspective = camera->perspective;
        Audit::PixelRaster r;
        if (!audit.raster_pixels(mesh_of(a, m), bounds_from(b), c, screen, ss, force_two != 0, r)) {
            *size_out = r.size;
            limited = true;
            return;
        }
        const size_t n = size_t(r.size) * r.size;
        if (n > sample_capacity) throw std::invalid_argument("sample capacity too small");
        std::memcpy(covered, r.covered.data(), n);
        std::memcpy(visible, r.visible.data(), n);
        std::memcpy(depth, r.depth.data(), n * 4);
        std::memcpy(normal, r.normal.data(), n * 12);
        std::memcpy(color, r.color.data(), n * 16);
        std::memcpy(material, r.material.data(), n * 2);
        *size_out = r.size;
        *clipped = r.clipped;
    });
    if (status == BLITZ_OK && limited) {
        message(error, capacity, "raster exceeds per-view memory budget");
        return BLITZ_BUDGET_LIMITED;
    }
    return status;
}

blitz_status blitz_hip_overlap(blitz_hip_audit* a, const blitz_hip_device_mesh* m, const blitz_hip_bounds* b,
                               double screen, uint16_t ortho, uint16_t persp, uint32_t seed, double* out, char* error,
                               size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null overlap output");
        *out = audit_of(a).overlap(mesh_of(a, m), bounds_from(b), screen, ortho, persp, seed);
    });
}

blitz_status blitz_hip_coverage_distance(blitz_hip_audit* a, const uint64_t* x, const uint64_t* y, uint32_t size,
                                         uint8_t ss, double* out, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!x || !y || !out) throw std::invalid_argument("null mask");
        if (!size || uint64_t(size) * size > 64000000ull) throw std::invalid_argument("invalid raster size");
        const size_t n = (size_t(size) * size + 63) / 64;
        std::vector<uint64_t> va(x, x + n), vb(y, y + n);
        *out = audit_of(a).coverage_distance(va, vb, size, ss);
    });
}

blitz_status blitz_hip_coverage_distance_noalloc(int32_t device, const uint64_t* a, const uint64_t* b,
                                                 uint32_t size, uint8_t supersample, double* out, void* stream,
                                                 char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        coverage_distance_noalloc(device, a, b, size, supersample, out, static_cast<hipStream_t>(stream));
    });
}

blitz_status blitz_hip_audit_last_timing(const blitz_hip_audit* a, blitz_hip_audit_timing* out) {
    if (!a || !a->impl || !out || out->struct_size != sizeof(*out)) return BLITZ_INVALID_ARGUMENT;
    const Timing& t = a->impl->last_timing();
    out->raster_tasks = t.raster_tasks;
    out->distance_views = t.distance_views;
    out->passes = t.passes;
    out->raster_ms = t.raster_ms;
    out->compare_ms = t.compare_ms;
    out->distance_ms = t.distance_ms;
    out->total_ms = t.total_ms;
    out->raster_triangles = t.raster_triangles;
    out->mask_bytes = t.mask_bytes;
    out->project_ms = t.project_ms;
    out->band_ms = t.band_ms;
    out->band_launches = t.band_launches;
    out->patch_fallbacks = t.patch_fallbacks;
    out->patch_ms = t.patch_ms;
    out->appearance_ms = t.appearance_ms;
    out->attributed_views = t.attributed_views;
    out->split_launches = t.split_launches;
    out->split_ms = t.split_ms;
    out->raster_chunks = t.raster_chunks;
    out->tuned_launches = t.tuned_launches;
    return BLITZ_OK;
}

blitz_status blitz_hip_reduce_reuse(const blitz_mesh* m, size_t target, uint32_t* out, size_t out_capacity,
                                    size_t* count, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out || !count) throw std::invalid_argument("null reduction output");
        const HostMesh h = host_mesh(m);
        const auto idx = reduce_reuse(h.positions.data(), h.positions.size() / 3, h.indices.data(),
                                      h.indices.size() / 3, target, nullptr);
        if (idx.size() > out_capacity) throw std::invalid_argument("index capacity too small");
        std::memcpy(out, idx.data(), idx.size() * 4);
        *count = idx.size();
    });
}
}

namespace {
Scorer& scorer_of(blitz_hip_scorer* s) {
    if (!s || !s->impl) throw std::invalid_argument("null scorer");
    return *s->impl;
}
Graph& graph_of(const blitz_hip_scorer* s, blitz_hip_graph* g) {
    if (!g || !g->graph) throw std::invalid_argument("null graph");
    if (g->owner != s) throw std::invalid_argument("graph belongs to another scorer");
    return *g->graph;
}
}  // namespace

extern "C" {
size_t blitz_hip_scorer_weight_count(void) { return kWeightCount; }

blitz_status blitz_hip_scorer_create(int32_t device, const float* weights, size_t count, blitz_hip_scorer** out,
                                     char* error, size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null scorer destination");
        if (!blitz_hip_available(device)) throw Unavailable("no gfx950 HIP device at the requested index");
        auto s = std::make_unique<blitz_hip_scorer>();
        s->impl = std::make_unique<Scorer>(device, weights, count);
        *out = s.release();
    });
}

void blitz_hip_scorer_destroy(blitz_hip_scorer* s) { delete s; }

blitz_status blitz_hip_scorer_set_weights(blitz_hip_scorer* s, const float* weights, size_t count, char* error,
                                          size_t capacity) {
    return guarded(error, capacity, [&] { scorer_of(s).set_weights(weights, count); });
}

blitz_status blitz_hip_graph_create(blitz_hip_scorer* s, const blitz_mesh* m, const blitz_hip_bounds* b,
                                    blitz_hip_graph** out, char* error, size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null graph destination");
        Scorer& scorer = scorer_of(s);
        const HostMesh h = host_mesh(m);
        auto g = std::make_unique<blitz_hip_graph>();
        g->owner = s;
        g->graph = scorer.graph(h, bounds_from(b));
        // The directed actions as the device topology holds them.
        const Graph& d = *g->graph;
        g->src.resize(d.actions);
        g->dst.resize(d.actions);
        if (d.actions) {
            hip_check(hipMemcpy(g->src.data(), d.src.p, size_t(d.actions) * 4, hipMemcpyDeviceToHost), "download actions");
            hip_check(hipMemcpy(g->dst.data(), d.neighbors.p, size_t(d.actions) * 4, hipMemcpyDeviceToHost),
                      "download actions");
        }
        *out = g.release();
    });
}

blitz_status blitz_hip_graph_concat(blitz_hip_scorer* s, blitz_hip_graph* const* parts, size_t count,
                                    blitz_hip_graph** out, char* error, size_t capacity) {
    if (out) *out = nullptr;
    return guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null graph destination");
        if (!parts || !count) throw std::invalid_argument("no graphs to join");
        Scorer& scorer = scorer_of(s);
        std::vector<const Graph*> graphs(count);
        for (size_t k = 0; k < count; ++k) graphs[k] = &graph_of(s, parts[k]);
        auto g = std::make_unique<blitz_hip_graph>();
        g->owner = s;
        g->graph = scorer.concat(graphs.data(), count);
        g->src.reserve(g->graph->actions);
        g->dst.reserve(g->graph->actions);
        uint32_t v0 = 0;
        for (size_t k = 0; k < count; ++k) {
            for (const uint32_t v : parts[k]->src) g->src.push_back(v + v0);
            for (const uint32_t v : parts[k]->dst) g->dst.push_back(v + v0);
            v0 += graphs[k]->vertices;
        }
        *out = g.release();
    });
}

void blitz_hip_graph_destroy(blitz_hip_graph* g) {
    if (!g) return;
    if (g->owner && g->owner->impl) {
        DeviceGuard guard(g->owner->impl->device());
        (void)hipStreamSynchronize(g->owner->impl->stream());
        g->graph.reset();
    }
    delete g;
}

blitz_status blitz_hip_graph_info(const blitz_hip_graph* g, uint32_t* vertices, uint32_t* actions) {
    if (!g || !g->graph || !vertices || !actions) return BLITZ_INVALID_ARGUMENT;
    *vertices = g->graph->vertices;
    *actions = g->graph->actions;
    return BLITZ_OK;
}

blitz_status blitz_hip_graph_features(blitz_hip_scorer* s, const blitz_hip_graph* g, float* out, size_t capacity) {
    return guarded(nullptr, 0, [&] {
        Graph& graph = graph_of(s, const_cast<blitz_hip_graph*>(g));
        if (!out || capacity < size_t(graph.vertices) * kFeatures) throw std::invalid_argument("feature capacity");
        scorer_of(s).download_features(graph, out);
    });
}

blitz_status blitz_hip_graph_actions(const blitz_hip_graph* g, uint32_t* src, uint32_t* dst, size_t capacity) {
    if (!g || !g->graph || !src || !dst || capacity < g->src.size()) return BLITZ_INVALID_ARGUMENT;
    std::memcpy(src, g->src.data(), g->src.size() * 4);
    std::memcpy(dst, g->dst.data(), g->dst.size() * 4);
    return BLITZ_OK;
}

blitz_status blitz_hip_scorer_forward(blitz_hip_scorer* s, blitz_hip_graph* g, uint8_t precision, float* outputs,
                                      size_t capacity, char* error, size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        Scorer& scorer = scorer_of(s);
        Graph& graph = graph_of(s, g);
        if (precision > 1) throw std::invalid_argument("unknown scorer precision");
        if (outputs && capacity < size_t(graph.actions) * kOutputs) throw std::invalid_argument("output capacity");
        scorer.forward(graph, Precision(precision));
        if (outputs) scorer.download_outputs(graph, outputs);
    });
}

blitz_status blitz_hip_scorer_rank(blitz_hip_scorer* s, blitz_hip_graph* g, uint8_t ranking, uint32_t* order,
                                   size_t capacity, char* error, size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        Graph& graph = graph_of(s, g);
        if (ranking > 4) throw std::invalid_argument("unknown ranking");
        if (!order || capacity < graph.actions) throw std::invalid_argument("order capacity");
        scorer_of(s).rank(graph, Ranking(ranking), order);
    });
}

blitz_status blitz_hip_scorer_decode(blitz_hip_scorer* s, blitz_hip_graph* g, const uint32_t* actions, size_t count,
                                     float* placements, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if ((!placements && count) || (!actions && count)) throw std::invalid_argument("null decode buffers");
        scorer_of(s).decode(graph_of(s, g), actions, count, placements);
    });
}

blitz_status blitz_hip_scorer_teacher(blitz_hip_scorer* s, blitz_hip_graph* g, const uint32_t* actions, size_t count,
                                      float* positions, float* targets, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (count && (!actions || !positions || !targets)) throw std::invalid_argument("null teacher buffers");
        scorer_of(s).teacher(graph_of(s, g), actions, count, positions, targets);
    });
}

blitz_status blitz_hip_scorer_last_forward_ms(const blitz_hip_scorer* s, double* ms) {
    if (!s || !s->impl || !ms) return BLITZ_INVALID_ARGUMENT;
    *ms = s->impl->last_forward_ms();
    return BLITZ_OK;
}

void blitz_hip_train_settings_init(blitz_hip_train_settings* s) {
    if (!s) return;
    const TrainSettings d;
    s->margin = d.margin;
    s->penalty = d.penalty;
    s->lr = d.lr;
    s->decay = d.decay;
    s->max_norm = d.max_norm;
    s->placement_weight = d.placement_weight;
    s->pool = d.pool;
    s->precision = d.precision;
}

blitz_status blitz_hip_scorer_train_step(blitz_hip_scorer* s, blitz_hip_graph* g, const uint32_t* actions,
                                         const uint8_t* labels, const float* targets, size_t count,
                                         const blitz_hip_train_settings* settings, blitz_hip_train_result* result,
                                         char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!settings || !result) throw std::invalid_argument("null training settings or result");
        TrainSettings t;
        t.margin = settings->margin;
        t.penalty = settings->penalty;
        t.lr = settings->lr;
        t.decay = settings->decay;
        t.max_norm = settings->max_norm;
        t.placement_weight = settings->placement_weight;
        t.pool = settings->pool;
        t.precision = settings->precision;
        Scorer& scorer = scorer_of(s);
        const TrainResult r = scorer.train_step(graph_of(s, g), actions, labels, targets, count, t);
        result->loss = r.loss;
        result->grad_norm = r.grad_norm;
        result->applied = r.applied;
        result->step = scorer.optimizer_steps();
    });
}

blitz_status blitz_hip_scorer_gradients(blitz_hip_scorer* s, float* out, size_t capacity) {
    return guarded(nullptr, 0, [&] {
        if (!out || capacity < kWeightCount) throw std::invalid_argument("gradient capacity");
        scorer_of(s).download_gradients(out);
    });
}

blitz_status blitz_hip_scorer_weights(blitz_hip_scorer* s, float* out, size_t capacity) {
    return guarded(nullptr, 0, [&] {
        if (!out || capacity < kWeightCount) throw std::invalid_argument("weight capacity");
        scorer_of(s).download_weights(out);
    });
}

blitz_status blitz_hip_scorer_reset_optimizer(blitz_hip_scorer* s) {
    return guarded(nullptr, 0, [&] { scorer_of(s).reset_optimizer(); });
}

blitz_status blitz_hip_scorer_last_step_ms(const blitz_hip_scorer* s, double* ms) {
    if (!s || !s->impl || !ms) return BLITZ_INVALID_ARGUMENT;
    *ms = s->impl->last_step_ms();
    return BLITZ_OK;
}

blitz_status blitz_hip_chain_options_init(blitz_hip_chain_options* o, size_t size) {
    if (!o || size != sizeof(*o)) return BLITZ_INVALID_ARGUMENT;
    *o = {};
    o->struct_size = sizeof(*o);
    o->version = BLITZ_HIP_API_VERSION;
    o->action_trials = 64;
    o->action_batch = 32;
    o->ranking = 4;
    o->origin = 0;
    o->output = 1;
    o->confirmation = 1;
    o->overdraw_tiebreak = 1;
    return BLITZ_OK;
}

blitz_status blitz_hip_generate(blitz_hip_audit* a, const blitz_mesh* source, const blitz_settings* settings,
                                const blitz_hip_chain_options* options, blitz_hip_scorer* scorer,
                                blitz_hip_chain** out, char* error, size_t capacity) {
    if (out) *out = nullptr;
    uint8_t status = 0;
    const blitz_status rc = guarded(error, capacity, [&] {
        if (!out) throw std::invalid_argument("null output");
        if (!options || options->struct_size != sizeof(*options) || options->version != BLITZ_HIP_API_VERSION)
            throw std::invalid_argument("null pointer, options size or version mismatch");
        if (options->origin > 2 || options->output > 1 || options->confirmation > 2)
            throw std::invalid_argument("invalid neural proposal origin, output or confirmation");
        if (options->origin != 0) throw Unavailable("the HIP chain proposes from the source only");
        if (options->output != 1) throw Unavailable("the HIP chain emits reuse (shared-vertex) LODs only");
        if (options->confirmation != 1) throw Unavailable("the HIP chain confirms on the GPU only");
        Audit& audit = audit_of(a);
        const ChainSettings s = chain_settings_of(settings);
        const HostMesh m = host_mesh(source);
        std::vector<float> uv;
        if (source->uv.count) {
            const size_t n = m.positions.size() / 3;
            uv.resize(2 * n);
            const auto* base = static_cast<const unsigned char*>(source->uv.data);
            for (size_t i = 0; i < n; ++i) std::memcpy(&uv[2 * i], base + i * source->uv.stride, 8);
        }
        ChainOptions o;
        o.action_trials = options->action_trials;
        o.action_batch = options->action_batch;
        o.ranking = options->ranking;
        o.overdraw_tiebreak = options->overdraw_tiebreak != 0;
        Scorer* sc = nullptr;
        if (scorer) {
            if (!scorer->impl) throw std::invalid_argument("null scorer");
            if (scorer->impl->device() != audit.device()) throw std::invalid_argument("scorer on another device");
            sc = scorer->impl.get();
        }
        auto chain = std::make_unique<blitz_hip_chain>();
        chain->result = generate_chain(audit, m, uv.empty() ? nullptr : uv.data(), s, o, sc);
        status = chain->result.status;
        *out = chain.release();
    });
    if (rc != BLITZ_OK) return rc;
    return status == 2 ? BLITZ_CANCELLED : status == 1 ? BLITZ_BUDGET_LIMITED : BLITZ_OK;
}

void blitz_hip_chain_destroy(blitz_hip_chain* c) { delete c; }

blitz_status blitz_hip_schedule(const blitz_settings* settings, const blitz_hip_bounds* b, double* out,
                                size_t capacity, char* error, size_t error_capacity) {
    return guarded(error, error_capacity, [&] {
        const auto steps = schedule(bounds_from(b), chain_settings_of(settings));
        if (!out || capacity < 3 * steps.size()) throw std::invalid_argument("schedule capacity too small");
        for (size_t i = 0; i < steps.size(); ++i) {
            out[3 * i] = steps[i].pixels;
            out[3 * i + 1] = steps[i].transition;
            out[3 * i + 2] = steps[i].source;
        }
    });
}

blitz_status blitz_hip_chain_info_of(const blitz_hip_chain* c, blitz_hip_chain_info* out) {
    return guarded(nullptr, 0, [&] {
        const ChainResult& r = chain_of(c);
        if (!out || out->struct_size != sizeof(*out)) throw std::invalid_argument("info size mismatch");
        *out = {};
        out->struct_size = sizeof(*out);
        out->levels = uint32_t(r.lods.size());
        out->status = r.status;
        out->fallback_levels = r.fallback_levels;
        out->candidate_evaluations = r.candidate_evaluations;
        for (int i = 0; i < 4; ++i) {
            out->audit_evaluations[i] = r.audit_evaluations[i];
            out->rejected_gates[i] = r.rejected_gates[i];
            out->area_rejected_gates[i] = r.area_rejected_gates[i];
        }
        out->decoded = r.decoded;
        out->legal_collapses = r.legal_collapses;
        out->rejected_collapses = r.rejected_collapses;
        out->action_ranked = r.action_ranked;
        out->action_trials = r.action_trials;
        out->reference_rejections = r.reference_rejections;
        out->bounded_audits = r.bounded_audits;
        out->finalists = uint32_t(r.candidates.size());
        out->selected = uint32_t(r.selected);
        out->proposals = uint32_t(r.proposals.size());
        out->gate_audits = r.gate_audits;
        out->confirm_audits = r.confirm_audits;
        out->ms = r.ms;
        out->executor_ms = r.executor_ms;
        out->gate_ms = r.gate_ms;
        out->build_ms = r.build_ms;
        out->rank_ms = r.rank_ms;
        out->select_ms = r.select_ms;
        out->trial_ms = r.trial_ms;
        out->speculative_ms = r.speculative_ms;
        out->speculative_build_ms = r.speculative_build_ms;
        out->offer_ms = r.offer_ms;
        out->confirm_ms = r.confirm_ms;
        out->tiebreak_ms = r.tiebreak_ms;
    });
}

blitz_status blitz_hip_chain_lod_of(const blitz_hip_chain* c, size_t level, blitz_hip_chain_lod* out) {
    return guarded(nullptr, 0, [&] {
        const ChainResult& r = chain_of(c);
        if (!out || out->struct_size != sizeof(*out)) throw std::invalid_argument("lod size mismatch");
        if (level >= r.lods.size()) throw std::invalid_argument("level out of range");
        const ChainLod& l = r.lods[level];
        *out = {};
        out->struct_size = sizeof(*out);
        out->indices = l.indices.data();
        out->index_count = l.indices.size();
        out->materials = l.materials.empty() ? nullptr : l.materials.data();
        out->material_count = l.materials.size();
        out->screen_pixels = l.schedule.pixels;
        out->transition_limit = l.schedule.transition;
        out->source_limit = l.schedule.source;
        to_c(l.adjacent, &out->adjacent);
        to_c(l.source_error, &out->source);
    });
}

blitz_status blitz_hip_chain_finalist(const blitz_hip_chain* c, size_t i, uint32_t* triangles, size_t capacity) {
    return guarded(nullptr, 0, [&] {
        const ChainResult& r = chain_of(c);
        if (i >= r.candidates.size()) throw std::invalid_argument("finalist out of range");
        if (!triangles || capacity < r.candidates[i].size()) throw std::invalid_argument("capacity too small");
        std::memcpy(triangles, r.candidates[i].data(), r.candidates[i].size() * 4);
    });
}

blitz_status blitz_hip_chain_proposal_of(const blitz_hip_chain* c, size_t i, blitz_hip_chain_proposal* out) {
    return guarded(nullptr, 0, [&] {
        const ChainResult& r = chain_of(c);
        if (!out || out->struct_size != sizeof(*out)) throw std::invalid_argument("proposal size mismatch");
        if (i >= r.proposals.size()) throw std::invalid_argument("proposal out of range");
        const ChainProposal& p = r.proposals[i];
        *out = {};
        out->struct_size = sizeof(*out);
        out->level = p.level;
        out->parent_triangles = p.parent_triangles;
        out->start_triangles = p.start_triangles;
        out->target_triangles = p.target_triangles;
        out->final_triangles = p.final_triangles;
        out->trials = p.trials;
        out->accepted_batches = p.accepted_batches;
        out->accepted = p.accepted;
        out->rejected = p.rejected;
        out->screen_pixels = p.screen_pixels;
        out->stop_reason = p.stop_reason;
        out->gate = p.gate;
    });
}
}

blitz_status blitz_hip_device_primitive(int32_t device, uint32_t kind, const void* keys, const uint32_t* values,
                                        uint32_t n, int32_t begin_bit, int32_t end_bit, void* keys_out,
                                        uint32_t* values_out, uint32_t* runs_out, char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (kind > 6) throw std::invalid_argument("unknown primitive");
        if (n && (!keys || !keys_out)) throw std::invalid_argument("null primitive arrays");
        if (kind == 5 && (!runs_out || (n && !values_out))) throw std::invalid_argument("run-length outputs required");
        if (kind <= 1 && (begin_bit < 0 || end_bit < begin_bit || end_bit > (kind ? 64 : 32)))
            throw std::invalid_argument("invalid key bits");
        if (!blitz_hip_available(device)) throw Unavailable("no gfx950 HIP device at the requested index");
        DeviceGuard guard(device);
        const size_t width = (kind == 1 || kind == 5 || kind == 6) ? 8 : 4;
        hipStream_t st{};
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        std::vector<void*> owned;
        auto alloc = [&](size_t bytes) {
            void* p = nullptr;
            hip_check(hipMalloc(&p, std::max<size_t>(bytes, 16)), "hipMalloc");
            owned.push_back(p);
            return p;
        };
        try {
            void* dk = alloc(n * width);
            void* dko = alloc(n * width);
            uint32_t* dv = values ? static_cast<uint32_t*>(alloc(n * 4)) : nullptr;
            uint32_t* dvo = values_out ? static_cast<uint32_t*>(alloc(n * 4)) : nullptr;
            uint32_t* druns = static_cast<uint32_t*>(alloc(4));
            if (n) hip_check(hipMemcpyAsync(dk, keys, n * width, hipMemcpyHostToDevice, st), "upload keys");
            if (n && dv) hip_check(hipMemcpyAsync(dv, values, n * 4, hipMemcpyHostToDevice, st), "upload values");
            const auto run = [&](auto call) {
                size_t bytes = 0;
                hip_check(call(nullptr, bytes), "workspace query");
                void* temp = alloc(bytes);
                hip_check(call(temp, bytes), "primitive");
            };
            switch (kind) {
            case 0:
                run([&](void* t, size_t& b) {
                    return radix_sort_pairs(t, b, static_cast<const uint32_t*>(dk), static_cast<uint32_t*>(dko), dv, dvo,
                                            n, begin_bit, end_bit, st);
                });
                break;
            case 1:
                run([&](void* t, size_t& b) {
                    return radix_sort_pairs(t, b, static_cast<const uint64_t*>(dk), static_cast<uint64_t*>(dko), dv, dvo,
                                            n, begin_bit, end_bit, st);
                });
                break;
            case 2:
                run([&](void* t, size_t& b) {
                    return radix_sort_pairs_descending(t, b, static_cast<const float*>(dk), static_cast<float*>(dko), dv,
                                                       dvo, n, st);
                });
                break;
            case 3:
                run([&](void* t, size_t& b) {
                    return exclusive_sum(t, b, static_cast<const uint32_t*>(dk), static_cast<uint32_t*>(dko), n, st);
                });
                break;
            case 4:
                run([&](void* t, size_t& b) {
                    return inclusive_max(t, b, static_cast<const uint32_t*>(dk), static_cast<uint32_t*>(dko), n, st);
                });
                break;
            case 5:
                run([&](void* t, size_t& b) {
                    return run_length_encode(t, b, static_cast<const uint64_t*>(dk), static_cast<uint64_t*>(dko), dvo,
                                             druns, n, st);
                });
                break;
            default:
                run([&](void* t, size_t& b) {
                    return exclusive_sum(t, b, static_cast<const uint64_t*>(dk), static_cast<uint64_t*>(dko), n, st);
                });
                break;
            }
            if (n) hip_check(hipMemcpyAsync(keys_out, dko, n * width, hipMemcpyDeviceToHost, st), "download keys");
            if (n && values_out) hip_check(hipMemcpyAsync(values_out, dvo, n * 4, hipMemcpyDeviceToHost, st), "download values");
            if (runs_out) hip_check(hipMemcpyAsync(runs_out, druns, 4, hipMemcpyDeviceToHost, st), "download runs");
            hip_check(hipStreamSynchronize(st), "primitive");
        } catch (...) {
            (void)hipStreamSynchronize(st);
            for (void* p : owned) (void)hipFree(p);
            (void)hipStreamDestroy(st);
            throw;
        }
        for (void* p : owned) (void)hipFree(p);
        (void)hipStreamDestroy(st);
    });
}

blitz_status blitz_hip_reduce_reuse_device(int32_t device, const blitz_mesh* m, size_t target, uint32_t* out,
                                           size_t out_capacity, size_t* count, blitz_hip_reduce_report* report,
                                           char* error, size_t capacity) {
    return guarded(error, capacity, [&] {
        if (!out || !count) throw std::invalid_argument("null reduction output");
        if (report && report->struct_size != sizeof(*report)) throw std::invalid_argument("report size mismatch");
        const HostMesh h = host_mesh(m);
        if (!blitz_hip_available(device)) throw Unavailable("no gfx950 HIP device at the requested index");
        ReduceReport r{};
        const auto idx = reduce_reuse_device(device, h, target, &r);
        if (idx.size() > out_capacity) throw std::invalid_argument("index capacity too small");
        std::memcpy(out, idx.data(), idx.size() * 4);
        *count = idx.size();
        if (report) {
            report->passes = r.passes;
            report->rounds = r.rounds;
            report->tail_collapses = r.tail_collapses;
            report->collapses = r.collapses;
        }
    });
}
