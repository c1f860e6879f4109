// Host side of the bound-reference patch audit: binding a reference (per-view
// projection, mask, covered count, tile bins) and auditing many local patches
// of it in one launch, with the reference's view loop replayed on the host.
// Patches the device path cannot decide exactly (a clipped view, a view that
// would refine its supersampling, an oversized region) take the full audit.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "audit_kernels.hpp"
#include "patch_audit.hpp"
#include "digest.hpp"
#include "patch_kernels.hpp"
#include "patches.hpp"

namespace blitz_hip {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kMaxSamples = 64000000;  // src/evaluate.cpp:129
constexpr double kInf = std::numeric_limits<double>::infinity();

// Scratch slots of the patch path (the full path uses 0-10).
enum : size_t { kSlotCounts = 11, kSlotScan, kSlotPatchCsr, kSlotSummaries, kSlotPairs, kSlotViewSums };

struct Clock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

}  // namespace

std::unique_ptr<Binding> Audit::bind(const HostMesh& reference, const Bounds& b, const Settings& s) {
    validate_settings(b, s);
    if (s.profile != 0) throw Unavailable("the HIP audit backend implements the coverage profile only");
    auto bd = std::make_unique<Binding>();
    bd->owner = this;
    bd->host = reference;
    bd->bounds = b;
    bd->settings = s;
    bd->settings.cancelled = nullptr;
    bd->mesh = upload(reference);
    bd->cams = cameras(b, s.screen_size, s.orthographic, s.perspective, s.rotation_seed);
    bd->views = uint32_t(bd->cams.size());
    bd->ss = s.supersample;
    bd->size = raster_side(s.screen_size, s.supersample);
    bd->tiles = (bd->size + kTile - 1) / kTile;
    bd->supported = reference.materials.empty();
    bd->added_two_sided = reference.materials.empty() && !reference.double_sided.empty() && reference.double_sided[0];
    bd->clipped.assign(bd->views, 0);
    if (uint64_t(bd->size) * bd->size > kMaxSamples) {
        bd->limited = true;
        return bd;
    }
    if (!bd->views) return bd;
    DeviceGuard guard(device_);
    const uint32_t views = bd->views, vertices = bd->mesh->vertices, faces = bd->mesh->faces, tiles = bd->tiles;
    const size_t words = size_t(tiles) * tiles;
    bool perspective = false;
    for (const auto& c : bd->cams) perspective |= c.perspective != 0;
    bd->xy.resize(size_t(views) * std::max(vertices, 1u) * sizeof(double2));
    if (perspective) bd->z.resize(size_t(views) * std::max(vertices, 1u) * sizeof(double));
    std::vector<ProjectTask> proj(views);
    for (uint32_t v = 0; v < views; ++v) {
        ProjectTask& p = proj[v];
        p = {};
        p.positions = bd->mesh->positions;
        p.xy = bd->xy.as<double2>() + size_t(v) * vertices;
        p.z = bd->cams[v].perspective ? bd->z.as<double>() + size_t(v) * vertices : nullptr;
        p.camera = bd->cams[v];
        p.center = b.center;
        p.vertices = vertices;
        p.size = bd->size;
        p.ss = bd->ss;
    }
    ProjectTask* dp = scratch_as<ProjectTask>(8, views);
    hip_check(hipMemcpyAsync(dp, proj.data(), views * sizeof(ProjectTask), hipMemcpyHostToDevice, stream_),
              "upload projections");
    launch_project(dp, views, vertices, stream_);
    bd->masks.resize(size_t(views) * words * 8);
    launch_fill(bd->masks.as<uint64_t>(), size_t(views) * words, 0, stream_);
    uint32_t* status = scratch_as<uint32_t>(6, size_t(views) * 2);
    std::vector<uint32_t> st(size_t(views) * 2);
    for (uint32_t v = 0; v < views; ++v) {
        st[2 * v] = kNone;
        st[2 * v + 1] = 0;
    }
    hip_check(hipMemcpyAsync(status, st.data(), st.size() * 4, hipMemcpyHostToDevice, stream_), "upload status");
    std::vector<RasterTask> tasks(views);
    for (uint32_t v = 0; v < views; ++v) {
        RasterTask& t = tasks[v];
        t = {};
        t.xy = proj[v].xy;
        t.z = proj[v].z;
        t.indices = bd->mesh->indices;
        t.two_sided = bd->mesh->two_sided;
        t.mask = bd->masks.as<uint64_t>() + size_t(v) * words;
        t.status = status + 2 * v;
        t.faces = faces;
        t.size = bd->size;
        t.tiles = tiles;
        t.force_two = s.force_two_sided;
    }
    bd->tasks.resize(views * sizeof(RasterTask));
    hip_check(hipMemcpyAsync(bd->tasks.p, tasks.data(), views * sizeof(RasterTask), hipMemcpyHostToDevice, stream_),
              "upload raster tasks");
    const auto launch = [&](const RasterGrid& g) { launch_raster(bd->tasks.as<RasterTask>(), g, tiles, stream_); };
    launch(tuner_.grid(views, tiles, faces, stream_, launch));
    // Covered samples per view (compare of each mask with itself).
    std::vector<ViewPair> pairs(views);
    for (uint32_t v = 0; v < views; ++v) {
        pairs[v] = {};
        pairs[v].a = pairs[v].c = tasks[v].mask;
        pairs[v].size = bd->size;
        pairs[v].tiles = tiles;
    }
    ViewPair* dpairs = scratch_as<ViewPair>(kSlotPairs, views);
    ViewSummary* sums = scratch_as<ViewSummary>(kSlotViewSums, views);
    hip_check(hipMemcpyAsync(dpairs, pairs.data(), views * sizeof(ViewPair), hipMemcpyHostToDevice, stream_),
              "upload pairs");
    launch_fill(reinterpret_cast<uint64_t*>(sums), views * sizeof(ViewSummary) / 8, 0, stream_);
    launch_compare(dpairs, sums, views, tiles, stream_);
    // Tile bins of the drawable faces of every view.
    const size_t cells = size_t(views) * words + 1;
    auto* counts = scratch_as<unsigned long long>(kSlotCounts, cells);
    launch_fill(reinterpret_cast<uint64_t*>(counts), cells, 0, stream_);
    launch_bin_count(bd->tasks.as<RasterTask>(), views, faces, counts, stream_);
    bd->bin_offsets.resize(cells * 8);
    const size_t temp_bytes = scan_workspace(cells);
    void* temp = scratch(kSlotScan, std::max<size_t>(temp_bytes, 16));
    launch_exclusive_scan(temp, temp_bytes, counts, bd->bin_offsets.as<unsigned long long>(), cells, stream_);
    unsigned long long entries = 0;
    hip_check(hipMemcpyAsync(&entries, bd->bin_offsets.as<unsigned long long>() + cells - 1, 8, hipMemcpyDeviceToHost,
                             stream_),
              "download bin total");
    hip_check(hipMemcpyAsync(st.data(), status, st.size() * 4, hipMemcpyDeviceToHost, stream_), "download status");
    std::vector<ViewSummary> vs(views);
    hip_check(hipMemcpyAsync(vs.data(), sums, views * sizeof(ViewSummary), hipMemcpyDeviceToHost, stream_),
              "download counts");
    hip_check(hipStreamSynchronize(stream_), "bind");
    for (uint32_t v = 0; v < views; ++v) bd->clipped[v] = st[2 * v] != kNone || st[2 * v + 1] != 0;
    bd->count_a.resize(views);
    for (uint32_t v = 0; v < views; ++v) bd->count_a[v] = vs[v].count_a;
    bd->device_count_a.resize(views * 8);
    hip_check(hipMemcpyAsync(bd->device_count_a.p, bd->count_a.data(), views * 8, hipMemcpyHostToDevice, stream_),
              "upload counts");
    bd->bin_faces.resize(std::max<size_t>(1, entries) * 4);
    hip_check(hipMemcpyAsync(counts, bd->bin_offsets.p, cells * 8, hipMemcpyDeviceToDevice, stream_), "bin cursor");
    launch_bin_fill(bd->tasks.as<RasterTask>(), views, faces, counts, bd->bin_faces.as<uint32_t>(), stream_);
    hip_check(hipGetLastError(), "bind launch");
    hip_check(hipStreamSynchronize(stream_), "bind");
    return bd;
}

std::vector<Measurement> Audit::evaluate_patches(Binding& bd, const PatchSet& P, const Settings& s,
                                                 uint32_t* fallbacks_out) {
    if (bd.owner != this) throw std::invalid_argument("binding belongs to another audit");
    const Settings& k = bd.settings;
    if (s.profile != 0) throw Unavailable("the HIP audit backend implements the coverage profile only");
    if (s.screen_size != k.screen_size || s.supersample != k.supersample || s.orthographic != k.orthographic ||
        s.perspective != k.perspective || s.rotation_seed != k.rotation_seed ||
        s.force_two_sided != k.force_two_sided)
        throw std::invalid_argument("settings differ from the raster settings of the bound reference");
    validate_settings(bd.bounds, s);
    if (!bd.supported)
        throw Unavailable("patch audits need a reference without per-face materials (added faces have none)");
    const size_t n = P.count;
    const uint32_t faces = bd.mesh->faces, vertices = bd.mesh->vertices, views = bd.views;
    if (n && (!P.removed_offsets || !P.added_offsets)) throw std::invalid_argument("null patch offsets");
    if (n && (P.removed_offsets[0] != 0 || P.added_offsets[0] != 0))
        throw std::invalid_argument("patch offsets must start at 0");
    for (size_t i = 0; i < n; ++i) {
        const uint32_t r0 = P.removed_offsets[i], r1 = P.removed_offsets[i + 1];
        const uint32_t a0 = P.added_offsets[i], a1 = P.added_offsets[i + 1];
        if (r1 < r0 || a1 < a0) throw std::invalid_argument("patch offsets must not decrease");
        if ((r1 > r0 && !P.removed) || (a1 > a0 && !P.added)) throw std::invalid_argument("null patch arrays");
        for (uint32_t j = r0; j < r1; ++j)
            if (P.removed[j] >= faces || (j > r0 && P.removed[j] <= P.removed[j - 1]))
                throw std::invalid_argument("removed faces must be ascending face ids of the reference");
        for (size_t j = size_t(a0) * 3; j < size_t(a1) * 3; ++j)
            if (P.added[j] >= vertices) throw std::invalid_argument("added face outside the reference vertices");
    }
    Clock clock;
    timing_ = {};
    std::vector<Measurement> out(n);
    enum : uint8_t { kDevice = 0, kIdentical = 1, kFull = 2 };
    std::vector<uint8_t> route(n, kDevice);
    bool any_clipped = false;
    for (uint8_t c : bd.clipped) any_clipped |= c != 0;
    const uint32_t* ref_idx = bd.host.indices.data();
    for (size_t i = 0; i < n; ++i) {
        const uint32_t r0 = P.removed_offsets[i], nr = P.removed_offsets[i + 1] - r0;
        const uint32_t a0 = P.added_offsets[i], na = P.added_offsets[i + 1] - a0;
        // src/evaluate.cpp:77-82: identical index streams (the region is the
        // trailing faces, re-added unchanged) skip every view.
        bool same = nr == na;
        for (uint32_t j = 0; same && j < nr; ++j) {
            const uint32_t f = P.removed[r0 + j];
            same = f == faces - nr + j && std::memcmp(ref_idx + 3 * size_t(f), P.added + 3 * size_t(a0 + j), 12) == 0;
        }
        if (same)
            route[i] = kIdentical;
        else if (bd.limited || any_clipped || !views || nr > kPatchMaxFaces || na > kPatchMaxFaces)
            route[i] = kFull;
    }
    std::vector<size_t> device;
    for (size_t i = 0; i < n; ++i)
        if (route[i] == kDevice) device.push_back(i);
    double patch_ms = 0;
    std::vector<std::vector<Measurement>> per_view(n);
    if (!device.empty()) {
        DeviceGuard guard(device_);
        const size_t chunk = std::max<size_t>(1, std::min<size_t>(device.size(), (size_t(1) << 23) / views));
        for (size_t begin = 0; begin < device.size(); begin += chunk) {
            const size_t m = std::min(chunk, device.size() - begin);
            std::vector<uint32_t> ro{0}, rv, ao{0}, av;
            for (size_t j = 0; j < m; ++j) {
                const size_t i = device[begin + j];
                const uint32_t r0 = P.removed_offsets[i], r1 = P.removed_offsets[i + 1];
                const uint32_t a0 = P.added_offsets[i], a1 = P.added_offsets[i + 1];
                rv.insert(rv.end(), P.removed + r0, P.removed + r1);
                av.insert(av.end(), P.added + 3 * size_t(a0), P.added + 3 * size_t(a1));
                ro.push_back(uint32_t(rv.size()));
                ao.push_back(uint32_t(av.size() / 3));
            }
            auto* dcsr = scratch_as<uint32_t>(kSlotPatchCsr, ro.size() + ao.size() + rv.size() + av.size());
            uint32_t* d_ro = dcsr;
            uint32_t* d_ao = d_ro + ro.size();
            uint32_t* d_rv = d_ao + ao.size();
            uint32_t* d_av = d_rv + rv.size();
            hip_check(hipMemcpyAsync(d_ro, ro.data(), ro.size() * 4, hipMemcpyHostToDevice, stream_), "patches");
            hip_check(hipMemcpyAsync(d_ao, ao.data(), ao.size() * 4, hipMemcpyHostToDevice, stream_), "patches");
            if (!rv.empty())
                hip_check(hipMemcpyAsync(d_rv, rv.data(), rv.size() * 4, hipMemcpyHostToDevice, stream_), "patches");
            if (!av.empty())
                hip_check(hipMemcpyAsync(d_av, av.data(), av.size() * 4, hipMemcpyHostToDevice, stream_), "patches");
            auto* summaries = scratch_as<PatchSummary>(kSlotSummaries, size_t(views) * m);
            PatchInput in{};
            in.views = bd.tasks.as<RasterTask>();
            in.masks = bd.masks.as<uint64_t>();
            in.count_a = bd.device_count_a.as<unsigned long long>();
            in.bin_offsets = bd.bin_offsets.as<unsigned long long>();
            in.bin_faces = bd.bin_faces.as<uint32_t>();
            in.removed_offsets = d_ro;
            in.removed = d_rv;
            in.added_offsets = d_ao;
            in.added = d_av;
            in.out = summaries;
            in.view_count = views;
            in.patches = uint32_t(m);
            in.tiles = bd.tiles;
            in.added_two_sided = bd.added_two_sided;
            hip_check(hipEventRecord(events_[0], stream_), "event");
            launch_patches(in, stream_);
            hip_check(hipEventRecord(events_[1], stream_), "event");
            hip_check(hipGetLastError(), "patch launch");
            std::vector<PatchSummary> sum(size_t(views) * m);
            hip_check(hipMemcpyAsync(sum.data(), summaries, sum.size() * sizeof(PatchSummary), hipMemcpyDeviceToHost,
                                     stream_),
                      "download patch summaries");
            hip_check(hipStreamSynchronize(stream_), "patches");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, events_[0], events_[1]), "elapsed");
            patch_ms += ms;
            // Views in order up to the first failing one, as the reference
            // visits them; a flagged or refining view sends the patch to the
            // full audit.
            for (size_t j = 0; j < m; ++j) {
                const size_t i = device[begin + j];
                std::vector<Measurement>& per = per_view[i];
                for (uint32_t v = 0; v < views; ++v) {
                    const PatchSummary& r = sum[size_t(v) * m + j];
                    if (r.flags) {
                        route[i] = kFull;
                        break;
                    }
                    const unsigned long long ca = bd.count_a[v];
                    double coverage;
                    if (!r.changed)
                        coverage = 0;
                    else if ((ca != 0) != (r.count_c != 0))
                        coverage = kInf;
                    else
                        coverage = std::sqrt(double(float(std::max(r.max_ac, r.max_ca)))) / bd.ss;
                    const Measurement cur = finish_view(coverage, r.changed, r.total, false, s, bd.ss);
                    if (!settled(cur, s, bd.ss)) {
                        route[i] = kFull;
                        break;
                    }
                    per.push_back(cur);
                    if (!cur.passed) break;
                }
            }
        }
    }
    // Sequential assembly of the device-path patches, exactly as the
    // reference's view loop (src/evaluate.cpp:312-352).
    for (size_t i = 0; i < n; ++i) {
        Measurement& r = out[i];
        r.supersample = s.supersample;
        if (route[i] != kDevice) continue;
        const std::vector<Measurement>& per = per_view[i];
        for (uint32_t v = 0; v < per.size(); ++v) {
            if (s.cancelled && s.cancelled()) {
                r.complete = r.passed = false;
                r.cancelled = true;
                break;
            }
            const Measurement& cur = per[v];
            ++r.views_evaluated;
            if (cur.error > r.error) {
                r.worst_view = v;
                r.error = cur.error;
                r.supersample = cur.supersample;
            }
            r.coverage = std::max(r.coverage, cur.coverage);
            r.coverage_upper = std::max(r.coverage_upper, cur.coverage_upper);
            if (cur.changed_area > r.changed_area) {
                r.changed_area = cur.changed_area;
                r.changed_area_worst_view = v;
            }
            if (!cur.passed) {
                r.passed = r.complete = false;
                break;
            }
        }
    }
    // Full audit of the remaining patches, a few candidates per batch.
    std::vector<size_t> full;
    for (size_t i = 0; i < n; ++i)
        if (route[i] == kFull) full.push_back(i);
    Timing keep = timing_;
    for (size_t begin = 0; begin < full.size(); begin += 16) {
        const size_t m = std::min<size_t>(16, full.size() - begin);
        std::vector<std::unique_ptr<DeviceMesh>> meshes;
        std::vector<const DeviceMesh*> cands;
        for (size_t j = 0; j < m; ++j) {
            const size_t i = full[begin + j];
            const uint32_t r0 = P.removed_offsets[i], a0 = P.added_offsets[i];
            HostMesh c;
            c.positions = bd.host.positions;
            c.double_sided = bd.host.double_sided;
            c.indices = patch_candidate(ref_idx, faces, P.removed + r0, P.removed_offsets[i + 1] - r0,
                                        P.added ? P.added + 3 * size_t(a0) : nullptr, P.added_offsets[i + 1] - a0);
            c.normals = bd.host.normals;
            c.colors = bd.host.colors;
            c.pos_digest[0] = bd.host.pos_digest[0];
            c.pos_digest[1] = bd.host.pos_digest[1];
            c.vertex_digest[0] = bd.host.vertex_digest[0];
            c.vertex_digest[1] = bd.host.vertex_digest[1];
            // The identity digest of the candidate's own streams, as a fresh
            // upload of the same mesh would compute it.
            Digest d;
            d.a = c.vertex_digest[0];
            d.b = c.vertex_digest[1];
            digest_triangles(d, c.indices.data(), c.indices.size(), nullptr, 0, c.double_sided.data(),
                             c.double_sided.size());
            c.digest[0] = d.a;
            c.digest[1] = d.b;
            meshes.push_back(upload(c));
            cands.push_back(meshes.back().get());
        }
        const auto ms = evaluate_batch(*bd.mesh, cands, bd.bounds, s);
        for (size_t j = 0; j < m; ++j) out[full[begin + j]] = ms[j];
    }
    timing_ = keep;
    timing_.patch_ms = patch_ms;
    timing_.patch_fallbacks = uint32_t(full.size());
    timing_.total_ms = clock.ms();
    if (fallbacks_out) *fallbacks_out = uint32_t(full.size());
    return out;
}

}  // namespace blitz_hip
