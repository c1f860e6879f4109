// Reference-side binding of the HIP backend: what a maintainer adds to the
// reference build (C++20, the reference's include/ on the include path, this
// directory too, linked with libblitz_hip.so) to route audits and neural
// generation to the HIP library when the neural options select
// raster_backend == BLITZ_NEURAL_RASTER_HIP.
//
// * AuditHip: the surface of AuditCuda the generator calls
//   (src/neural/internal.hpp:131-175): bind_reference, evaluate on host or
//   bound meshes, and certify for many candidates at once, forwarded to
//   blitz_hip_audit_evaluate_batch.
// * generate_neural_hip: blitz_generate_neural (include/blitz/blitz.h:165-167)
//   for the backend value 2, mapping blitz_neural_options_v2 onto the HIP
//   chain's options.
//
// Header-only; it throws the reference's exception types where the reference
// does (blitz::NeuralUnavailable, std::invalid_argument, std::bad_alloc,
// std::runtime_error), so src/c_api.cpp's status mapping applies unchanged.
#pragma once
#include <blitz/blitz.h>
#include <blitz/neural.hpp>

#include <functional>
#include <new>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "blitz_hip.h"

namespace blitz_hip_adapter {

inline blitz_stream stream_of(const std::byte* data, size_t count, size_t stride) {
    return {data, count, stride};
}

// blitz_mesh borrowing a MeshView's streams (the inverse of src/c_api.cpp's
// descriptor conversion).
inline blitz_mesh descriptor(const blitz::MeshView& v) {
    blitz_mesh m{};
    m.struct_size = sizeof(blitz_mesh);
    m.abi_version = BLITZ_ABI_VERSION;
    m.positions = stream_of(v.positions.data, v.positions.count, v.positions.stride);
    m.normals = stream_of(v.normals.data, v.normals.count, v.normals.stride);
    m.uv = stream_of(v.uv.data, v.uv.count, v.uv.stride);
    m.colors = stream_of(v.colors.data, v.colors.count, v.colors.stride);
    m.tangents = stream_of(v.tangents.data, v.tangents.count, v.tangents.stride);
    m.indices = v.indices.data();
    m.index_count = v.indices.size();
    m.materials = v.materials.data();
    m.material_count = v.materials.size();
    m.double_sided = v.double_sided.data();
    m.double_sided_count = v.double_sided.size();
    return m;
}

inline int poll_cancelled(void* user) {
    const auto* f = static_cast<const std::function<bool()>*>(user);
    return (*f)() ? 1 : 0;
}

// blitz::EvalSettings -> blitz_hip_eval_settings; the cancellation callback
// borrows `s.cancelled`, which must outlive the call.
inline blitz_hip_eval_settings to_hip(const blitz::EvalSettings& s) {
    blitz_hip_eval_settings e{};
    blitz_hip_eval_settings_init(&e, sizeof e);
    e.profile = uint8_t(s.profile);
    e.supersample = s.supersample;
    e.max_supersample = s.max_supersample;
    e.force_two_sided = s.force_two_sided;
    e.conservative_screen = s.conservative_screen;
    e.views_orthographic = s.views.orthographic;
    e.views_perspective = s.views.perspective;
    e.views_rotation_seed = s.views.rotation_seed;
    e.screen_size = s.screen_size;
    e.limit = s.limit;
    e.max_changed_area = s.max_changed_area;
    e.normal_weight = s.weights.normal;
    e.color_weight = s.weights.color;
    e.material_weight = s.weights.material;
    if (s.cancelled) {
        e.cancelled = &poll_cancelled;
        e.user_data = const_cast<std::function<bool()>*>(&s.cancelled);
    }
    return e;
}

inline blitz::Measurement from_hip(const blitz_hip_measurement& m) {
    blitz::Measurement r;
    r.error = m.error;
    r.coverage = m.coverage;
    r.coverage_upper = m.coverage_upper;
    r.changed_area = m.changed_area;
    r.normal_degrees = m.normal_degrees;
    r.worst_view = m.worst_view;
    r.changed_area_worst_view = m.changed_area_worst_view;
    r.views_evaluated = m.views_evaluated;
    r.supersample = m.supersample;
    r.complete = m.complete;
    r.passed = m.passed;
    r.resource_limited = m.resource_limited;
    r.cancelled = m.cancelled;
    return r;
}

inline blitz_hip_bounds to_hip(const blitz::Bounds& b) { return {{b.center.x, b.center.y, b.center.z}, b.radius}; }

// The reference's exception for each status (src/c_api.cpp:187-202 in reverse).
inline void check(blitz_status s, const char* message) {
    switch (s) {
        case BLITZ_OK:
            return;
        case BLITZ_UNAVAILABLE:
            throw blitz::NeuralUnavailable(message);
        case BLITZ_INVALID_ARGUMENT:
            throw std::invalid_argument(message);
        case BLITZ_OUT_OF_MEMORY:
            throw std::bad_alloc();
        default:
            throw std::runtime_error(message);
    }
}

class AuditHip {
  public:
    explicit AuditHip(const blitz::NeuralOptions& o) {
        blitz_hip_audit_options opt{};
        blitz_hip_audit_options_init(&opt, sizeof opt);
        opt.device = o.device;
        opt.memory_mib = o.memory_mib;
        opt.cache_rasters = o.cache_rasters;
        char err[256];
        check(blitz_hip_audit_create(&opt, &audit_, err, sizeof err), err);
    }
    ~AuditHip() {
        clear_reference();
        blitz_hip_audit_destroy(audit_);
    }
    AuditHip(const AuditHip&) = delete;
    AuditHip& operator=(const AuditHip&) = delete;

    // A validated immutable predecessor, uploaded once and kept until the
    // next binding (src/neural/internal.hpp:141-144).
    void bind_reference(blitz::MeshView reference) {
        clear_reference();
        const blitz_mesh m = descriptor(reference);
        char err[256];
        check(blitz_hip_mesh_upload(audit_, &m, &reference_, err, sizeof err), err);
    }
    void clear_reference() noexcept {
        blitz_hip_mesh_release(reference_);
        reference_ = nullptr;
    }

    // AuditCuda::evaluate(MeshView, MeshView, ...): host meshes.
    blitz::Measurement evaluate(blitz::MeshView reference, blitz::MeshView candidate, const blitz::Bounds& b,
                                const blitz::EvalSettings& s) {
        const blitz_mesh r = descriptor(reference), c = descriptor(candidate);
        const blitz_hip_eval_settings e = to_hip(s);
        const blitz_hip_bounds hb = to_hip(b);
        blitz_hip_measurement out{};
        char err[256];
        check(blitz_hip_audit_evaluate(audit_, &r, &c, &hb, &e, &out, err, sizeof err), err);
        return from_hip(out);
    }

    // certify_candidates (src/neural/evaluate.cu:2413-2578) against the bound
    // reference: every candidate's measurement in shared launches, without
    // the reference's 8-candidate limit; passed == the certificate's Pass.
    std::vector<blitz::Measurement> certify(std::span<const blitz::MeshView> candidates, const blitz::Bounds& b,
                                            const blitz::EvalSettings& s) {
        if (!reference_) throw std::invalid_argument("certify needs a bound reference");
        std::vector<blitz_hip_device_mesh*> uploaded;
        uploaded.reserve(candidates.size());
        char err[256];
        const auto release = [&] {
            for (auto* d : uploaded) blitz_hip_mesh_release(d);
        };
        for (const auto& c : candidates) {
            const blitz_mesh m = descriptor(c);
            blitz_hip_device_mesh* d = nullptr;
            const blitz_status st = blitz_hip_mesh_upload(audit_, &m, &d, err, sizeof err);
            if (st != BLITZ_OK) {
                release();
                check(st, err);
            }
            uploaded.push_back(d);
        }
        const blitz_hip_eval_settings e = to_hip(s);
        const blitz_hip_bounds hb = to_hip(b);
        std::vector<blitz_hip_measurement> out(candidates.size());
        const blitz_status st = blitz_hip_audit_evaluate_batch(audit_, reference_, uploaded.data(), uploaded.size(),
                                                               &hb, &e, out.data(), err, sizeof err);
        release();
        check(st, err);
        std::vector<blitz::Measurement> result;
        result.reserve(out.size());
        for (const auto& m : out) result.push_back(from_hip(m));
        return result;
    }

  private:
    blitz_hip_audit* audit_ = nullptr;
    blitz_hip_device_mesh* reference_ = nullptr;
};

// blitz_generate_neural for raster_backend == BLITZ_NEURAL_RASTER_HIP: the
// chain of `source` under the ABI-5 settings with the v2 options' trials,
// batch, ranking, origin and tie-break (explicit reuse output). `scorer` is
// the HIP scorer for learned ranking (null otherwise). The caller owns
// *out (blitz_hip_chain_destroy).
inline blitz_status generate_neural_hip(blitz_hip_audit* audit, const blitz_mesh* source,
                                        const blitz_settings* settings, const blitz_neural_options_v2* options,
                                        blitz_hip_scorer* scorer, blitz_hip_chain** out, char* error,
                                        size_t capacity) {
    if (!options || options->struct_size != sizeof(*options) || options->raster_backend != BLITZ_NEURAL_RASTER_HIP)
        return BLITZ_INVALID_ARGUMENT;
    blitz_hip_chain_options o{};
    blitz_hip_chain_options_init(&o, sizeof o);
    o.action_trials = options->action_trials;
    o.action_batch = options->action_batch;
    o.ranking = options->ranking;
    o.origin = options->origin;
    o.confirmation = options->confirmation;
    o.overdraw_tiebreak = options->overdraw_tiebreak;
    return blitz_hip_generate(audit, source, settings, &o, scorer, out, error, capacity);
}

}  // namespace blitz_hip_adapter
