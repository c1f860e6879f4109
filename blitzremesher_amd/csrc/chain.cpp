// LOD chain generation with audited collapse proposals (see chain.hpp).
//
// The control flow restates the reference's chain search and neural hooks
// step by step, because the emitted chain (per-level triangle counts, index
// buffers and audit records) must equal the reference's for the same inputs:
// search_pass (src/chain.cpp:179-970) and confirm_search (:971-1066) under
// the hooks of src/neural/generate.cpp:215-224 (evaluate), 304-450
// (propose_guarded with the four-gate certification) and 456-562 (confirm).
// What runs is this build's own: every audit is the HIP audit, every
// proposal the device collapse executor, and candidates borrow the source's
// device vertex streams so projections are shared across all of them.
#include "chain.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

#include "actions.hpp"
#include "digest.hpp"
#include "scorer.hpp"

namespace blitz_hip {

double Curve::at(double t) const {
    if (points.empty()) throw std::invalid_argument("empty curve");
    if (t <= points.front().x) return points.front().y;
    for (size_t i = 1; i < points.size(); ++i)
        if (t <= points[i].x) {
            const double u = (t - points[i - 1].x) / (points[i].x - points[i - 1].x);
            return points[i - 1].y * (1 - u) + points[i].y * u;
        }
    return points.back().y;
}

std::string validate_chain_settings(const ChainSettings& s) {
    if (s.levels < 2 || s.levels > 32) return "level count must be 2..32";
    if (unsigned(s.objective) > 3 || unsigned(s.profile) > 2) return "unknown mode";
    if (s.triangle_overhead_bps > 10000) return "triangle overhead must be 0..10000 basis points";
    if (!std::isfinite(s.max_changed_area) || s.max_changed_area < 0 || s.max_changed_area > 1)
        return "maximum changed area must be in [0,1]";
    for (double v : {s.pixels_per_meter, s.meters_per_unit, s.last_pixels})
        if (!std::isfinite(v) || v <= 0) return "screen scale must be positive and finite";
    for (double v : {s.base_pixels, s.max_lod0_delta_px})
        if (v != 0 && (!std::isfinite(v) || v <= 0)) return "invalid optional pixel limit";
    for (const Curve* c : {&s.transition, &s.normal_importance, &s.attribute_importance}) {
        if (c->points.empty() || c->points.front().x != 0 || c->points.back().x != 1)
            return "curves must span [0,1]";
        double last = -1;
        for (auto p : c->points) {
            if (!std::isfinite(p.x) || !std::isfinite(p.y) || p.x <= last || p.x > 1 || p.y < 0 || p.y > 1e6)
                return "curve points must be ordered, finite and in 0..1e6";
            last = p.x;
        }
    }
    for (double v : {s.normal_weight, s.color_weight, s.material_weight})
        if (!std::isfinite(v) || v < 0 || v > 1e6) return "appearance weights must be in 0..1e6";
    if (!s.candidate_budget || !s.beam_width || s.beam_width > 32) return "invalid work budget or beam width";
    for (unsigned n : {s.search_supersample, s.audit_supersample, s.max_supersample})
        if (!n || n > 32 || (n & (n - 1))) return "supersampling must be a power of two in 1..32";
    if (s.max_supersample < std::max(s.search_supersample, s.audit_supersample))
        return "maximum supersampling is too small";
    if (!s.search_orthographic && !s.search_perspective) return "empty search cameras";
    if (!s.audit_orthographic && !s.audit_perspective) return "empty audit cameras";
    return {};
}

std::vector<ScheduleEntry> schedule(const Bounds& b, const ChainSettings& s) {
    if (auto e = validate_chain_settings(s); !e.empty()) throw std::invalid_argument(e);
    const double first = s.base_pixels != 0 ? s.base_pixels : b.radius * 2 * s.meters_per_unit * s.pixels_per_meter;
    if (!std::isfinite(first) || first < s.last_pixels || first > 16384)
        throw std::invalid_argument("base pixels must be >= last pixels and <=16384");
    std::vector<ScheduleEntry> out(s.levels);
    out[0].pixels = first;
    double cumulative = 0;
    for (unsigned i = 1; i < s.levels; ++i) {
        const double pixels = first * std::pow(s.last_pixels / first, double(i) / (s.levels - 1));
        const double delta = s.transition.at(s.levels == 2 ? 0 : double(i - 1) / (s.levels - 2));
        cumulative = cumulative * pixels / out[i - 1].pixels + delta;
        out[i] = {pixels, delta, s.max_lod0_delta_px != 0 ? std::min(cumulative, s.max_lod0_delta_px) : cumulative};
    }
    return out;
}

uint8_t bounded_refinement(const Settings& e) {
    if (!std::isfinite(e.screen_size) || e.screen_size <= 0 || e.screen_size > 16384 || !e.supersample ||
        e.max_supersample < e.supersample)
        throw std::invalid_argument("invalid bounded refinement settings");
    const uint64_t side = uint64_t(std::ceil(e.screen_size + 8));
    unsigned ss = e.supersample;
    while (ss < e.max_supersample) {
        const unsigned next = std::min<unsigned>(ss * 2, e.max_supersample);
        if (side * side * next * next > 64000000ull) break;
        ss = next;
    }
    return uint8_t(ss);
}

namespace {

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

// blitz::StorageStats (include/blitz/remesher.hpp:87-92).
struct Storage {
    uint64_t source_vertex_bytes = 0, added_vertex_bytes = 0, index_bytes = 0;
    uint64_t total() const { return source_vertex_bytes + added_vertex_bytes + index_bytes; }
};

struct Node {
    std::shared_ptr<Node> parent;
    ChainLod lod;
    uint64_t triangles = 0;
    Storage storage;
    std::shared_ptr<const DeviceMesh> device;  // the LOD on the device (source streams, own triangles)
};

// same_mesh_data of two views over the source's streams (src/mesh.cpp:57-61)
// and same_lod of two reuse LODs (src/chain.cpp:158-169): the triangle
// streams decide.
bool same_data(const ChainLod& a, const ChainLod& b) { return a.indices == b.indices && a.materials == b.materials; }

// blitz::EvalSettings of one level (eval_config, src/chain.cpp:179-199).
Settings eval_config(const ChainSettings& s, const ScheduleEntry& step, unsigned level, bool audit, double limit) {
    Settings e;
    e.profile = s.profile;
    e.screen_size = step.pixels;
    e.limit = limit;
    e.conservative_screen = false;  // research.conservative_screen is rejected in neural mode
    e.max_changed_area = audit ? s.max_changed_area : 1.0;
    const double t = s.levels == 2 ? 0 : double(level - 1) / (s.levels - 2);
    e.normal_weight = s.normal_weight * s.normal_importance.at(t);
    e.color_weight = s.color_weight * s.attribute_importance.at(t);
    e.material_weight = s.material_weight * s.attribute_importance.at(t);
    e.orthographic = audit ? s.audit_orthographic : s.search_orthographic;
    e.perspective = audit ? s.audit_perspective : s.search_perspective;
    e.rotation_seed = audit ? s.audit_seed : s.search_seed;
    e.supersample = audit ? s.audit_supersample : s.search_supersample;
    e.max_supersample = s.max_supersample;
    e.cancelled = s.cancelled;
    return e;
}

// Runs a fixed number of tasks at once: task 0 on the calling thread, task
// i > 0 on worker i - 1. The four gates of an executor trial are independent
// audits (each a pure function of its reference, candidate and settings), so
// they run side by side on four audit objects with their own HIP streams;
// their verdicts are then read in the reference's order.
class TaskPool {
  public:
    explicit TaskPool(size_t workers) {
        for (size_t w = 0; w < workers; ++w) threads_.emplace_back([this, w] { loop(w); });
    }
    ~TaskPool() {
        {
            std::lock_guard<std::mutex> lock(m_);
            stop_ = true;
        }
        wake_.notify_all();
        for (auto& t : threads_) t.join();
    }
    TaskPool(const TaskPool&) = delete;
    TaskPool& operator=(const TaskPool&) = delete;

    void run(std::vector<std::function<void()>>& tasks) {
        errors_.assign(tasks.size(), nullptr);
        {
            std::lock_guard<std::mutex> lock(m_);
            tasks_ = &tasks;
            pending_ = threads_.size();
            ++generation_;
        }
        wake_.notify_all();
        try {
            if (!tasks.empty()) tasks[0]();
        } catch (...) {
            errors_[0] = std::current_exception();
        }
        {
            std::unique_lock<std::mutex> lock(m_);
            done_.wait(lock, [&] { return pending_ == 0; });
            tasks_ = nullptr;
        }
        for (auto& e : errors_)
            if (e) std::rethrow_exception(e);
    }

  private:
    void loop(size_t w) {
        uint64_t seen = 0;
        for (;;) {
            std::function<void()>* task = nullptr;
            {
                std::unique_lock<std::mutex> lock(m_);
                wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                if (w + 1 < tasks_->size()) task = &(*tasks_)[w + 1];
            }
            if (task) try {
                    (*task)();
                } catch (...) {
                    errors_[w + 1] = std::current_exception();
                }
            {
                std::lock_guard<std::mutex> lock(m_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable wake_, done_;
    std::vector<std::function<void()>>* tasks_ = nullptr;
    std::vector<std::exception_ptr> errors_;
    uint64_t generation_ = 0;
    size_t pending_ = 0;
    bool stop_ = false;
};

// neural::audit_measurement_nonfinite (src/neural/audit_measurement.hpp:8-13).
bool nonfinite(const Measurement& m) {
    return std::isnan(m.error) || std::isnan(m.coverage) || std::isnan(m.coverage_upper) ||
           !std::isfinite(m.changed_area) || !std::isfinite(m.normal_degrees) ||
           (m.passed && (!std::isfinite(m.error) || !std::isfinite(m.coverage) || !std::isfinite(m.coverage_upper)));
}

class ChainSearch {
  public:
    ChainSearch(Audit& audit, const HostMesh& source, const float* uv, const ChainSettings& s, const ChainOptions& o,
                Scorer* scorer)
        : audit_(audit), source_(source), uv_(uv), s_(s), o_(o), scorer_(scorer), pool_(3) {
        gate_audit_[0] = &audit_;
        for (int g = 1; g < 4; ++g) {
            gate_owned_[g - 1] = std::make_unique<Audit>(audit.device(), audit.memory_bytes(), audit.cache_rasters());
            gate_audit_[g] = gate_owned_[g - 1].get();
        }
    }

    ChainResult run();

  private:
    bool cancelled() const { return r_.status == 2 || (s_.cancelled && s_.cancelled()); }
    std::shared_ptr<const DeviceMesh> derive(const ChainLod& lod);
    std::shared_ptr<const DeviceMesh> device_of(const std::shared_ptr<Node>& n);
    ChainLod propose(const std::shared_ptr<Node>& parent, size_t target, unsigned level);
    bool confirm();
    double tiebreak();
    void materialize(size_t selected);
    void incomplete(uint8_t status, const std::vector<ChainLod>& confirmed);

    Audit& audit_;
    const HostMesh& source_;
    const float* uv_;
    const ChainSettings& s_;
    const ChainOptions& o_;
    Scorer* scorer_;
    ChainResult r_;
    Bounds bounds_{};
    std::vector<ScheduleEntry> steps_;
    std::shared_ptr<const DeviceMesh> source_dev_;
    std::vector<std::shared_ptr<Node>> finalists_;
    std::vector<uint64_t> finalist_storage_;
    ExecutorCache executor_;
    // One audit object per executor gate (the first is the caller's).
    std::unique_ptr<Audit> gate_owned_[3];
    Audit* gate_audit_[4] = {};
    TaskPool pool_;
    // Per level: the four settings (source search, adjacent search, source
    // audit, adjacent audit).
    Settings gates_[4];
};

std::shared_ptr<const DeviceMesh> ChainSearch::derive(const ChainLod& lod) {
    if (lod.indices == source_.indices && lod.materials == source_.materials) return source_dev_;
    Digest d;
    d.a = source_.vertex_digest[0];
    d.b = source_.vertex_digest[1];
    digest_triangles(d, lod.indices.data(), lod.indices.size(), lod.materials.data(), lod.materials.size(),
                     source_.double_sided.data(), source_.double_sided.size());
    const uint64_t digest[2] = {d.a, d.b};
    return audit_.derive(*source_dev_, lod.indices, lod.materials, source_.double_sided, digest);
}

std::shared_ptr<const DeviceMesh> ChainSearch::device_of(const std::shared_ptr<Node>& n) {
    if (!n->device) n->device = derive(n->lod);
    return n->device;
}

// propose_guarded (src/neural/generate.cpp:304-450) for a direct reuse
// proposal: the collapse executor from the source towards `target`, every
// trial certified against the fixed source and the parent LOD under the
// search settings, then under the audit settings.
ChainLod ChainSearch::propose(const std::shared_ptr<Node>& parent, size_t target, unsigned level) {
    const auto previous = device_of(parent);
    // The four certifications of a trial in order: source search, adjacent
    // search, source audit, adjacent audit, each verdict-only with the
    // bounded refinement; cancellation is polled once per trial.
    Settings bounded[4];
    bool clamped[4];
    for (int g = 0; g < 4; ++g) {
        bounded[g] = gates_[g];
        bounded[g].max_supersample = bounded_refinement(gates_[g]);
        clamped[g] = bounded[g].max_supersample < gates_[g].max_supersample;
        bounded[g].verdict_only = true;
        bounded[g].cancelled = nullptr;
    }
    const DeviceMesh* refs[4] = {source_dev_.get(), previous.get(), source_dev_.get(), previous.get()};
    const auto gate = [&](const std::vector<uint32_t>& idx, const std::vector<uint16_t>& mats) {
        if (s_.cancelled && s_.cancelled()) return false;
        const auto t = Clock::now();
        ChainLod trial{idx, mats};
        const auto candidate = derive(trial);
        bool passed[4];
        std::vector<std::function<void()>> tasks;
        for (int g = 0; g < 4; ++g)
            tasks.push_back([&, g] {
                passed[g] = gate_audit_[g]->evaluate(*refs[g], *candidate, bounds_, bounded[g]).passed;
            });
        pool_.run(tasks);
        bool pass = true;
        for (int g = 0; g < 4 && pass; ++g) {
            ++r_.gate_audits;
            r_.bounded_audits += clamped[g];
            pass = passed[g];
        }
        r_.gate_ms += since(t);
        return pass;
    };
    // The rest of a rejected trial's halving ladder: each gate audits every
    // shorter prefix in one batch, and the gate audits are counted trial by
    // trial up to the first prefix all four accept, as the one-by-one gate
    // counts them.
    const ActionLadderGate ladder = [&](const std::vector<std::vector<uint32_t>>& idx,
                                        const std::vector<std::vector<uint16_t>>& mats) {
        if (s_.cancelled && s_.cancelled()) return -1;
        const auto t = Clock::now();
        std::vector<std::shared_ptr<const DeviceMesh>> owned;
        std::vector<const DeviceMesh*> candidates;
        for (size_t i = 0; i < idx.size(); ++i) {
            owned.push_back(derive(ChainLod{idx[i], mats[i]}));
            candidates.push_back(owned.back().get());
        }
        std::vector<Measurement> measured[4];
        std::vector<std::function<void()>> tasks;
        for (int g = 0; g < 4; ++g)
            tasks.push_back([&, g] {
                measured[g] = gate_audit_[g]->evaluate_batch(*refs[g], candidates, bounds_, bounded[g]);
            });
        pool_.run(tasks);
        int first = -1;
        for (size_t i = 0; i < candidates.size() && first < 0; ++i) {
            bool pass = true;
            for (int g = 0; g < 4 && pass; ++g) {
                ++r_.gate_audits;
                r_.bounded_audits += clamped[g];
                pass = measured[g][i].passed;
            }
            if (pass) first = int(i);
        }
        r_.gate_ms += since(t);
        return first;
    };
    ExecuteStats st;
    ChainLod out;
    const auto t = Clock::now();
    out.indices = execute_actions(audit_.device(), source_, uv_, target, o_.action_trials, o_.ranking, o_.action_batch,
                                  gate, st, &out.materials, scorer_, &bounds_, &executor_, &ladder);
    r_.executor_ms += since(t);
    r_.build_ms += st.build_ms;
    r_.rank_ms += st.rank_ms;
    r_.select_ms += st.select_ms;
    r_.trial_ms += st.trial_ms;
    r_.speculative_ms += st.speculative_ms;
    r_.speculative_build_ms += st.speculative_build_ms;
    ++r_.decoded;
    r_.legal_collapses += st.accepted;
    r_.rejected_collapses += st.rejected;
    r_.action_ranked += st.ranked;
    r_.action_trials += st.trials;
    ChainProposal p;
    p.level = level;
    p.parent_triangles = uint32_t(parent->lod.indices.size() / 3);
    p.start_triangles = uint32_t(source_.indices.size() / 3);
    p.target_triangles = uint32_t(target);
    p.final_triangles = uint32_t(out.indices.size() / 3);
    p.trials = uint32_t(st.trials);
    p.accepted_batches = st.accepted_batches;
    p.accepted = st.accepted;
    p.rejected = st.rejected;
    p.screen_pixels = gates_[2].screen_size;
    p.stop_reason = st.stop_reason;
    r_.proposals.push_back(p);
    return out;
}

// hooks.confirm (src/neural/generate.cpp:456-562) with GPU confirmation:
// full audits of every level against the source and its predecessor.
bool ChainSearch::confirm() {
    const auto t = Clock::now();
    bool good = true;
    std::vector<std::shared_ptr<const DeviceMesh>> devs(r_.lods.size());
    const auto dev = [&](size_t i) {
        if (!devs[i]) devs[i] = derive(r_.lods[i]);
        return devs[i];
    };
    const auto check = [&](const DeviceMesh& reference, const DeviceMesh& candidate, const Settings& e,
                           Measurement& output) {
        ++r_.confirm_audits;
        output = audit_.evaluate(reference, candidate, bounds_, e);
        const bool stopped = output.cancelled || (s_.cancelled && s_.cancelled());
        const bool bad = nonfinite(output);
        if (!stopped && !output.resource_limited && !bad && output.complete && output.passed) return true;
        if (stopped) {
        } else if (output.resource_limited)
            ++r_.confirmation_resources;
        else if (bad)
            ++r_.confirmation_nonfinite;
        else
            ++r_.reference_rejections;
        output.complete = false;
        output.passed = false;
        return false;
    };
    for (size_t i = 1; i < r_.lods.size() && good; ++i) {
        ChainLod& lod = r_.lods[i];
        Settings e = eval_config(s_, lod.schedule, unsigned(i), true, lod.schedule.source);
        e.screen_size = lod.schedule.pixels;
        good = check(*source_dev_, *dev(i), e, lod.source_error);
        if (good) {
            e.limit = lod.schedule.transition;
            good = check(*dev(i - 1), *dev(i), e, lod.adjacent);
        }
    }
    r_.confirm_ms += since(t);
    return good;
}

// hooks.tiebreak (src/neural/generate.cpp:563-570): the summed top-left
// overlap of levels 1.. under the search cameras.
double ChainSearch::tiebreak() {
    const auto t = Clock::now();
    double total = 0;
    for (size_t i = 1; i < r_.lods.size(); ++i) {
        const auto d = derive(r_.lods[i]);
        total += audit_.overlap(*d, bounds_, r_.lods[i].schedule.pixels, s_.search_orthographic,
                                s_.search_perspective, s_.search_seed);
    }
    r_.tiebreak_ms += since(t);
    return total;
}

void ChainSearch::materialize(size_t selected) {
    auto node = finalists_[selected];
    for (size_t i = r_.lods.size(); i-- > 0;) {
        r_.lods[i] = node->lod;
        node = node->parent;
    }
}

// incomplete_chain (src/chain.cpp:131-151) with generation hooks and no
// fallback proposal: the confirmed chain, else the source at every level.
void ChainSearch::incomplete(uint8_t status, const std::vector<ChainLod>& confirmed) {
    r_.status = status;
    if (!confirmed.empty())
        r_.lods = confirmed;
    else {
        r_.lods.clear();
        for (const auto& step : steps_) r_.lods.push_back({source_.indices, source_.materials, step, {}, {}});
    }
    std::vector<uint32_t> counts;
    for (const auto& lod : r_.lods) counts.push_back(uint32_t(lod.indices.size() / 3));
    r_.candidates = {counts};
    r_.selected = 0;
}

ChainResult ChainSearch::run() {
    const auto clock0 = Clock::now();
    bounds_ = bounds_of(source_.positions.data(), source_.positions.size() / 3);
    steps_ = schedule(bounds_, s_);
    const uint64_t source_triangles = source_.indices.size() / 3;
    source_dev_ = audit_.upload(source_);
    const uint64_t vertex_bytes = uint64_t(source_.positions.size() / 3) * 12 +
                                  uint64_t(source_.normals.size() / 3) * 12 + uint64_t(uv_ ? 8 : 0) *
                                  (source_.positions.size() / 3) + uint64_t(source_.colors.size() / 4) * 4;
    // An exact source chain is a valid incumbent (src/chain.cpp:215-219).
    for (const auto& step : steps_) r_.lods.push_back({source_.indices, source_.materials, step, {}, {}});
    auto root = std::make_shared<Node>();
    root->lod = {source_.indices, source_.materials, steps_[0], {}, {}};
    root->storage = {vertex_bytes, 0, uint64_t(source_.indices.size()) * 4};
    root->device = source_dev_;
    auto source_path = root;
    std::vector<std::shared_ptr<Node>> beam{root};
    const auto cancel = [&] {
        incomplete(2, {});
        r_.ms = since(clock0);
        return std::move(r_);
    };
    for (unsigned level = 1; level < s_.levels; ++level) {
        if (cancelled()) return cancel();
        std::vector<std::shared_ptr<Node>> next;
        const ScheduleEntry& step = steps_[level];
        gates_[0] = eval_config(s_, step, level, false, step.source);
        gates_[1] = eval_config(s_, step, level, false, step.transition);
        gates_[2] = eval_config(s_, step, level, true, step.source);
        gates_[3] = eval_config(s_, step, level, true, step.transition);
        const auto accepted = [&](const Measurement& m, size_t stage) {
            ++r_.audit_evaluations[stage];
            if (m.resource_limited) r_.status = 1;
            if (m.cancelled) {
                r_.status = 2;
                return false;
            }
            if (m.passed) return true;
            ++r_.rejected_gates[stage];
            if (stage >= 2 && m.error <= gates_[stage].limit && m.changed_area > s_.max_changed_area)
                ++r_.area_rejected_gates[stage];
            if (m.error >= r_.worst_rejected[stage].error) r_.worst_rejected[stage] = m;
            return false;
        };
        // offer (src/chain.cpp:383-470) without vertex budgets or the
        // duplicate cache (explicit output).
        const auto offer = [&](ChainLod candidate, const std::shared_ptr<Node>& parent) -> uint8_t {
            const auto t = Clock::now();
            const ChainLod& previous = parent->lod;
            if (candidate.indices.empty() || candidate.indices.size() % 3 ||
                (!candidate.materials.empty() && candidate.materials.size() != candidate.indices.size() / 3))
                return 5;
            if (candidate.indices.size() > previous.indices.size()) return 6;
            const bool changed = !same_data(candidate, previous);
            const auto prev_dev = device_of(parent);
            const auto cand_dev = changed ? derive(candidate) : prev_dev;
            // The four gates (hooks.evaluate, src/neural/generate.cpp:215-224,
            // with the refinement ceiling bounded by the sample cap as in
            // certify_device, :347-358) run at once, one per audit object;
            // their results are then taken in the reference's order (the
            // first failing gate ends the offer), so the counters are the
            // one-by-one ones.
            const DeviceMesh* refs[4] = {source_dev_.get(), prev_dev.get(), source_dev_.get(), prev_dev.get()};
            Measurement m[4];
            std::vector<std::function<void()>> tasks;
            for (int g = 0; g < 4; ++g)
                tasks.push_back([&, g] {
                    Settings bounded = gates_[g];
                    bounded.max_supersample = bounded_refinement(gates_[g]);
                    bounded.verdict_only = false;
                    m[g] = gate_audit_[g]->evaluate(*refs[g], *cand_dev, bounds_, bounded);
                });
            pool_.run(tasks);
            uint8_t gate = 0;
            for (int g = 0; g < 4 && !gate; ++g) {
                if (bounded_refinement(gates_[g]) < gates_[g].max_supersample) ++r_.bounded_audits;
                if (!accepted(m[g], size_t(g))) gate = uint8_t(g + 1);
            }
            candidate.source_error = m[2];
            candidate.adjacent = m[3];
            r_.offer_ms += since(t);
            if (gate) return gate;
            candidate.schedule = step;
            auto node = std::make_shared<Node>();
            node->parent = parent;
            node->triangles = parent->triangles + candidate.indices.size() / 3;
            node->storage = parent->storage;
            if (changed) node->storage.index_bytes += uint64_t(candidate.indices.size()) * 4;
            node->lod = std::move(candidate);
            node->device = cand_dev;
            next.push_back(std::move(node));
            return 0;
        };
        // Retain each incumbent (src/chain.cpp:471-474).
        for (auto& parent : beam) offer(parent->lod, parent);
        // One direct reuse slot per beam member, in beam order
        // (src/chain.cpp:488-535 with explicit reuse output and direct origin).
        const std::vector<std::shared_ptr<Node>> slots = beam;
        const unsigned rounds = std::max(1u, unsigned(s_.candidate_budget) / unsigned(slots.size()));
        unsigned proposals = 0;
        for (unsigned round = 0; proposals < s_.candidate_budget; ++round) {
            for (const auto& parent : slots) {
                if (proposals == s_.candidate_budget) break;
                if (cancelled()) return cancel();
                const size_t parent_triangles = parent->lod.indices.size() / 3;
                // Logarithmic target ladder (src/chain.cpp:624-633).
                const double fraction = round == rounds ? .1
                                        : rounds == 1   ? .5
                                                        : .02 * std::pow(45.0, double(round) / (rounds - 1));
                const size_t target = std::max<size_t>(1, size_t(parent_triangles * fraction));
                ChainLod candidate = propose(parent, target, level);
                ++proposals;
                ++r_.candidate_evaluations;
                r_.proposals.back().gate = offer(std::move(candidate), parent);
            }
        }
        if (cancelled()) return cancel();
        // Without a fallback hook the exact source history is kept as a node
        // of its own (src/chain.cpp:922-931).
        auto fallback = std::make_shared<Node>();
        fallback->parent = source_path;
        fallback->lod = {source_.indices, source_.materials, step, {}, {}};
        fallback->triangles = uint64_t(level) * source_triangles;
        fallback->storage = root->storage;
        fallback->device = source_dev_;
        source_path = fallback;
        next.push_back(fallback);
        std::stable_sort(next.begin(), next.end(), [](const auto& a, const auto& b) { return a->triangles < b->triangles; });
        std::vector<std::shared_ptr<Node>> distinct;
        for (auto& node : next)
            if (std::none_of(distinct.begin(), distinct.end(),
                             [&](const auto& n) { return same_data(node->lod, n->lod); }))
                distinct.push_back(node);
        next = std::move(distinct);
        for (auto& node : next)
            if (node->parent == source_path->parent && same_data(node->lod, source_path->lod)) {
                source_path = node;
                break;
            }
        if (level + 1 == s_.levels) finalists_ = next;
        if (next.size() > s_.beam_width) next.resize(s_.beam_width);
        if (std::find(next.begin(), next.end(), source_path) == next.end()) next.push_back(source_path);
        beam = std::move(next);
    }
    r_.candidates.clear();
    for (auto& node : finalists_) {
        std::vector<uint32_t> counts;
        for (auto p = node; p; p = p->parent) counts.push_back(uint32_t(p->lod.indices.size() / 3));
        std::reverse(counts.begin(), counts.end());
        r_.candidates.push_back(std::move(counts));
        finalist_storage_.push_back(node->storage.total());
    }
    // confirm_search (src/chain.cpp:971-1066): selection {0, 0} with
    // explicit output, confirmed and otherwise dropped in order.
    if (finalists_.empty()) {
        r_.ms = since(clock0);
        return std::move(r_);
    }
    for (;;) {
        materialize(0);
        if (confirm()) break;
        if (cancelled()) {
            incomplete(2, {});
            r_.ms = since(clock0);
            return std::move(r_);
        }
        finalists_.erase(finalists_.begin());
        r_.candidates.erase(r_.candidates.begin());
        finalist_storage_.erase(finalist_storage_.begin());
        if (finalists_.empty()) throw std::logic_error("exact source fallback failed reference audit");
    }
    size_t selected = 0;
    const auto tied = [&](size_t i) {
        return r_.candidates[i] == r_.candidates[selected] && finalist_storage_[i] == finalist_storage_[selected];
    };
    size_t ties = 0;
    for (size_t i = 0; i < r_.candidates.size(); ++i) ties += tied(i);
    if (o_.overdraw_tiebreak && ties > 1 && !cancelled()) {
        double best = tiebreak();
        auto confirmed = r_.lods;
        const size_t first = selected;
        for (size_t i = 0; i < finalists_.size(); ++i)
            if (i != first && r_.candidates[i] == r_.candidates[first] &&
                finalist_storage_[i] == finalist_storage_[first]) {
                if (cancelled()) break;
                materialize(i);
                if (!confirm()) continue;
                const double overlap = tiebreak();
                if (overlap < best) {
                    best = overlap;
                    selected = i;
                    confirmed = r_.lods;
                }
            }
        r_.lods = std::move(confirmed);
    }
    r_.selected = selected;
    if (cancelled()) r_.status = 2;
    for (size_t i = 1; i < r_.lods.size(); ++i) r_.fallback_levels += same_data(r_.lods[i], r_.lods[i - 1]);
    r_.ms = since(clock0);
    return std::move(r_);
}

}  // namespace

ChainResult generate_chain(Audit& audit, const HostMesh& source, const float* uv, const ChainSettings& s,
                           const ChainOptions& o, Scorer* scorer) {
    if (auto e = validate_chain_settings(s); !e.empty()) throw std::invalid_argument(e);
    if (!o.action_trials || o.action_trials > 65536) throw std::invalid_argument("invalid neural model options");
    if (!o.action_batch || o.action_batch > 64) throw std::invalid_argument("action batch outside 1..64");
    if (o.ranking == 2) throw Unavailable("shuffled ranking is not implemented by the HIP executor");
    if (o.ranking > 4) throw std::invalid_argument("invalid neural model options");
    if (o.ranking == 0 && !scorer) throw std::invalid_argument("learned ranking needs a scorer");
    // blitz::validate(MeshView) (src/mesh.cpp:24-38): finite positions and a
    // positive extent.
    for (float v : source.positions)
        if (!std::isfinite(v)) throw std::invalid_argument("nonfinite position or normal");
    const Bounds b = bounds_of(source.positions.data(), source.positions.size() / 3);
    if (!(b.radius > 0) || !std::isfinite(b.radius)) throw std::invalid_argument("zero or invalid source extent");
    if (source.indices.empty()) throw std::invalid_argument("mesh requires nonempty triangle indices");
    return ChainSearch(audit, source, uv, s, o, scorer).run();
}

}  // namespace blitz_hip
