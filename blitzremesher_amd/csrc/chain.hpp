// LOD chain generation with audited collapse proposals: the reference's
// chain search (detail::generate_with_hooks, src/chain.cpp:101-1188) driven
// by the neural generation hooks of src/neural/generate.cpp:118-577, with the
// device collapse executor as the proposer and the HIP audit as every gate.
//
// Scope: explicit reuse output (research.output = Reuse; every LOD shares
// the source's vertex streams), proposals from the source (NeuralOrigin
// Source, i.e. ChainMode::Direct), GPU confirmation, no packed vertex
// storage. Under those settings the reference path is: per level, re-offer
// each beam member, run up to candidate_budget proposals (the collapse
// executor from the source towards a logarithmic ladder of targets, each trial
// gated by four certifications: source-search, adjacent-search, source-audit,
// adjacent-audit), offer each proposal through the same four gates as full
// measurements, keep the beam, and finally confirm the selected chain level
// by level with full audits (src/neural/generate.cpp:456-562).
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "audit.hpp"

namespace blitz_hip {

class Scorer;

// blitz::Curve (include/blitz/remesher.hpp:12-15, src/chain.cpp:27-38).
struct CurvePoint {
    float x, y;
};
struct Curve {
    std::vector<CurvePoint> points{{0, 2}, {1, 3}};
    double at(double t) const;
};

// The blitz::Settings fields neural generation reads
// (include/blitz/remesher.hpp:32-54), defaults as there.
struct ChainSettings {
    uint8_t levels = 8;
    uint8_t profile = 1;
    uint8_t objective = 0;
    uint16_t triangle_overhead_bps = 0;
    double pixels_per_meter = 512, meters_per_unit = 1, last_pixels = 16;
    double base_pixels = 0, max_lod0_delta_px = 0;  // 0: not set
    Curve transition;
    Curve normal_importance{{{0, 1}, {1, 1}}}, attribute_importance{{{0, 1}, {1, 1}}};
    double normal_weight = 180.0 / (10 * 3.14159265358979323846), color_weight = 4, material_weight = 4;
    uint16_t search_orthographic = 42, search_perspective = 12;
    uint32_t search_seed = 0xB1172024u;
    uint16_t audit_orthographic = 642, audit_perspective = 64;
    uint32_t audit_seed = 0xB1172026u;
    uint8_t search_supersample = 4, audit_supersample = 8, max_supersample = 32;
    double max_changed_area = 1;
    uint16_t candidate_budget = 64;
    uint8_t beam_width = 8;
    std::function<bool()> cancelled;
};

// The blitz::NeuralOptions fields this path reads (include/blitz/neural.hpp:55-86).
struct ChainOptions {
    uint32_t action_trials = 64;
    uint8_t action_batch = 32;
    // 0 learned (BF16 logits of the scorer), 1 constant, 3 shortest edge,
    // 4 current plane (the plain classical order); 2 is unavailable.
    uint8_t ranking = 4;
    bool overdraw_tiebreak = true;
};

// blitz::ScheduleEntry (include/blitz/remesher.hpp:55-57).
struct ScheduleEntry {
    double pixels = 0, transition = 0, source = 0;
};

struct ChainLod {
    std::vector<uint32_t> indices;  // into the source's vertices
    std::vector<uint16_t> materials;
    ScheduleEntry schedule;
    Measurement adjacent, source_error;
};

// blitz::NeuralActionProposal (include/blitz/neural.hpp:31-38) plus the
// chain's offer gate for that proposal (ProposalTrace::gate: 0 accepted,
// 1..4 the four gates, 5 invalid, 6 growth).
struct ChainProposal {
    uint32_t level = 0, parent_triangles = 0, start_triangles = 0, target_triangles = 0, final_triangles = 0;
    uint32_t trials = 0, accepted_batches = 0;
    uint64_t accepted = 0, rejected = 0;
    double screen_pixels = 0;
    uint8_t stop_reason = 0, gate = 0;
};

struct ChainResult {
    std::vector<ChainLod> lods;
    uint8_t status = 0;  // blitz::Status: 0 complete, 1 budget limited, 2 cancelled
    // blitz::Result counters (include/blitz/remesher.hpp:119-151).
    uint64_t candidate_evaluations = 0;
    uint64_t audit_evaluations[4]{}, rejected_gates[4]{}, area_rejected_gates[4]{};
    Measurement worst_rejected[4];
    std::vector<std::vector<uint32_t>> candidates;  // triangles per level of each finalist
    size_t selected = 0;
    // blitz::NeuralStats (include/blitz/neural.hpp:100-110).
    uint64_t decoded = 0, legal_collapses = 0, rejected_collapses = 0, action_ranked = 0, action_trials = 0;
    uint64_t reference_rejections = 0, confirmation_resources = 0, confirmation_nonfinite = 0;
    uint32_t fallback_levels = 0, bounded_audits = 0;
    std::vector<ChainProposal> proposals;
    // Wall time of the whole generation and of its parts (ms).
    double ms = 0, executor_ms = 0, gate_ms = 0, offer_ms = 0, confirm_ms = 0, tiebreak_ms = 0;
    // Executor stages summed over proposals: state builds, rankings, batch
    // selection, trials, and the speculative next-state builds that overlap
    // the gates.
    double build_ms = 0, rank_ms = 0, select_ms = 0, trial_ms = 0, speculative_ms = 0, speculative_build_ms = 0;
    uint64_t gate_audits = 0, confirm_audits = 0;  // offers: the sum of audit_evaluations
};

// blitz::validate(const Settings&) (src/chain.cpp:40-100) for these fields;
// returns the reference's message or an empty string.
std::string validate_chain_settings(const ChainSettings& s);
// blitz::schedule (src/chain.cpp:101-118).
std::vector<ScheduleEntry> schedule(const Bounds& b, const ChainSettings& s);
// neural::bounded_refinement (src/neural/graph.cpp:5-18).
uint8_t bounded_refinement(const Settings& e);

// Generates the chain of `source` (optional per-vertex uv, 2 floats each,
// for the collapse rules). `scorer` is needed for ranking 0.
ChainResult generate_chain(Audit& audit, const HostMesh& source, const float* uv, const ChainSettings& s,
                           const ChainOptions& o, Scorer* scorer);

}  // namespace blitz_hip
