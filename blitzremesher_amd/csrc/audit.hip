// Host controller of the HIP coverage audit.
//
// The GPU measures every (candidate, view) pair of one refinement round in a
// single set of launches; the host then replays the reference's sequential
// per-view control flow (src/evaluate.cpp:312-352) over those measurements:
// refinement while uncertain, stop at the first failing view, resource limit
// at the 64M-sample cap, cancellation polled once per view. Because each
// view's measurement depends only on that view, measuring views beyond the
// first failure (which the reference never visits) changes no output field.
#include "audit.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <new>
#include <tuple>

#include "appearance_kernels.hpp"
#include "audit_kernels.hpp"

namespace blitz_hip {
namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr uint64_t kMaxSamples = 64000000;  // src/evaluate.cpp:129, src/neural/internal.hpp:11
constexpr uint32_t kNone = 0xFFFFFFFFu;
// Meshes of at least this many faces get a triple table at upload and are
// split against their candidates when the raster is large (raster_launch).
constexpr uint32_t kSplitMinFaces = 4096;

double dot(F3 a, F3 b) { return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z; }
F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
F3 normalized(F3 a) {
    const double l = std::sqrt(dot(a, a));
    if (!(l > 0)) return {0, 0, 0};
    const double s = 1 / l;
    return {float(a.x * s), float(a.y * s), float(a.z * s)};
}
// Seeded rotation of the view directions (src/evaluate.cpp:16-20).
F3 rotate(F3 v, uint32_t seed) {
    const double a = (seed % 65521) * .000731, b = ((seed >> 8) % 32749) * .000913;
    const double x = v.x * std::cos(a) - v.y * std::sin(a), y = v.x * std::sin(a) + v.y * std::cos(a);
    return {float(x), float(y * std::cos(b) - v.z * std::sin(b)), float(y * std::sin(b) + v.z * std::cos(b))};
}

// Serials identify device meshes across every audit object of the process
// (caches and identity checks key on them), so one counter serves them all.
uint64_t next_serial() {
    static std::atomic<uint64_t> counter{1};
    return counter.fetch_add(1);
}

struct Clock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};
}  // namespace

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

DeviceGuard::DeviceGuard(int device) {
    hip_check(hipGetDevice(&previous), "hipGetDevice");
    if (previous != device) hip_check(hipSetDevice(device), "hipSetDevice");
}
DeviceGuard::~DeviceGuard() {
    if (previous >= 0) (void)hipSetDevice(previous);
}

// src/mesh.cpp:77-86
Bounds bounds_of(const float* p, size_t n) {
    if (!n) return {{0, 0, 0}, 0};
    F3 lo{p[0], p[1], p[2]}, hi = lo;
    for (size_t i = 1; i < n; ++i) {
        const F3 q{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
        lo = {std::min(lo.x, q.x), std::min(lo.y, q.y), std::min(lo.z, q.z)};
        hi = {std::max(hi.x, q.x), std::max(hi.y, q.y), std::max(hi.z, q.z)};
    }
    Bounds b{{float((double(lo.x) + hi.x) * .5), float((double(lo.y) + hi.y) * .5), float((double(lo.z) + hi.z) * .5)},
             0};
    for (size_t i = 0; i < n; ++i) {
        const double x = double(p[3 * i]) - b.center.x, y = double(p[3 * i + 1]) - b.center.y,
                     z = double(p[3 * i + 2]) - b.center.z;
        b.radius = std::max(b.radius, std::sqrt(x * x + y * y + z * z));
    }
    return b;
}

// Fibonacci-sphere cameras, host libm (src/evaluate.cpp:108-124).
std::vector<Camera> cameras(const Bounds& b, double screen, uint16_t ortho, uint16_t persp, uint32_t seed) {
    std::vector<Camera> out;
    out.reserve(size_t(ortho) + persp);
    for (int projection = 0; projection < 2; ++projection) {
        const unsigned count = projection ? persp : ortho;
        for (unsigned i = 0; i < count; ++i) {
            const double z = 1 - 2 * (i + .5) / count, r = std::sqrt(std::max(0.0, 1 - z * z));
            const double angle = i * kPi * (3 - std::sqrt(5.0));
            const F3 forward = normalized(
                rotate({float(r * std::cos(angle)), float(z), float(r * std::sin(angle))}, seed + uint32_t(projection) * 104729u));
            const F3 hint = std::abs(forward.y) > .99 ? F3{1, 0, 0} : F3{0, 1, 0};
            const F3 right = normalized(cross(hint, forward)), up = normalized(cross(forward, right));
            const double height = std::max(1024.0, 2 * screen), focal = height / (2 * std::tan(kPi / 6));
            const double distance = b.radius * std::sqrt(1 + std::pow(2 * focal / screen, 2));
            out.push_back({right, up, forward, distance, focal, screen / (b.radius * 2), uint8_t(projection)});
        }
    }
    return out;
}

// src/evaluation_settings.hpp:10-27
void validate_settings(const Bounds& b, const Settings& s) {
    if (!(s.screen_size > 0) || !std::isfinite(s.screen_size) || s.screen_size > 16384 || !(b.radius > 0) ||
        !std::isfinite(b.radius) || !std::isfinite(b.center.x) || !std::isfinite(b.center.y) ||
        !std::isfinite(b.center.z) || unsigned(s.profile) > 2 || !(s.limit >= 0) || !std::isfinite(s.limit) ||
        !s.supersample || s.max_supersample < s.supersample || s.max_supersample > 32 ||
        !std::isfinite(s.normal_weight) || !std::isfinite(s.color_weight) || !std::isfinite(s.material_weight) ||
        s.normal_weight < 0 || s.color_weight < 0 || s.material_weight < 0 || s.normal_weight > 1e12 ||
        s.color_weight > 1e12 || s.material_weight > 1e12 || !std::isfinite(s.max_changed_area) ||
        s.max_changed_area < 0 || s.max_changed_area > 1)
        throw std::invalid_argument("invalid evaluation settings");
    if (!s.orthographic && !s.perspective) throw std::invalid_argument("at least one camera required");
}

uint32_t raster_side(double screen, unsigned ss) { return uint32_t(std::ceil(screen + 8)) * ss; }

DeviceMesh::~DeviceMesh() {
    if (!borrowed_vertices) {
        (void)hipFree(positions);
        (void)hipFree(normals);
        (void)hipFree(colors);
    }
    (void)hipFree(indices);
    (void)hipFree(two_sided);
    (void)hipFree(materials);
    (void)hipFree(face_table);
}

Audit::Audit(int device, uint64_t memory_bytes, bool cache_rasters, uint8_t face_split, bool autotune)
    : device_(device), memory_bytes_(memory_bytes), cache_rasters_(cache_rasters), face_split_(face_split),
      tuner_(autotune) {
    DeviceGuard guard(device_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : events_) hip_check(hipEventCreate(&e), "hipEventCreate");
}

void Audit::cache_clear() {
    if (!cache_.masks.empty()) (void)hipStreamSynchronize(stream_);
    for (auto& kv : cache_.masks) (void)hipFree(kv.second.mask);
    cache_.masks.clear();
    cache_.bytes = 0;
}

void Audit::cache_context(const DeviceMesh& reference, const Bounds& b, const Settings& s) {
    RasterCache& c = cache_;
    const bool same = c.serial == reference.serial && c.bounds.center.x == b.center.x &&
                      c.bounds.center.y == b.center.y && c.bounds.center.z == b.center.z &&
                      c.bounds.radius == b.radius && c.screen == s.screen_size && c.ortho == s.orthographic &&
                      c.persp == s.perspective && c.seed == s.rotation_seed && c.two == s.force_two_sided;
    if (same) return;
    cache_clear();
    c.serial = reference.serial;
    c.bounds = b;
    c.screen = s.screen_size;
    c.ortho = s.orthographic;
    c.persp = s.perspective;
    c.seed = s.rotation_seed;
    c.two = s.force_two_sided;
}

Audit::~Audit() {
    DeviceGuard guard(device_);
    cache_clear();
    (void)hipStreamSynchronize(stream_);
    for (auto& s : slots_) (void)hipFree(s.first);
    for (auto& e : events_) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(stream_);
    (void)hipHostFree(pinned_status_);
}

void* Audit::scratch(size_t slot, size_t bytes) {
    if (slots_.size() <= slot) slots_.resize(slot + 1, {nullptr, 0});
    auto& s = slots_[slot];
    if (s.second < bytes) {
        hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        hip_check(hipFree(s.first), "hipFree");
        const size_t grown = std::max(bytes, s.second + s.second / 2);
        s = {nullptr, 0};
        hip_check(hipMalloc(&s.first, grown), "hipMalloc");
        s.second = grown;
    }
    return s.first;
}

std::unique_ptr<DeviceMesh> Audit::upload(const HostMesh& m) {
    DeviceGuard guard(device_);
    auto d = std::make_unique<DeviceMesh>();
    d->owner = this;
    d->vertices = uint32_t(m.positions.size() / 3);
    d->faces = uint32_t(m.indices.size() / 3);
    d->digest[0] = m.digest[0];
    d->digest[1] = m.digest[1];
    d->pos_digest[0] = m.pos_digest[0];
    d->pos_digest[1] = m.pos_digest[1];
    d->vertex_digest[0] = m.vertex_digest[0];
    d->vertex_digest[1] = m.vertex_digest[1];
    d->serial = next_serial();
    d->positions_serial = d->serial;
    if (!m.normals.empty()) d->raw_normals = std::make_shared<const std::vector<float>>(m.normals);
    if (!m.colors.empty()) d->raw_colors = std::make_shared<const std::vector<uint8_t>>(m.colors);
    d->raw_double_sided = m.double_sided;
    std::vector<float4> pos(d->vertices);
    for (uint32_t i = 0; i < d->vertices; ++i)
        pos[i] = make_float4(m.positions[3 * i], m.positions[3 * i + 1], m.positions[3 * i + 2], 0.f);
    hip_check(hipMalloc(&d->positions, std::max<size_t>(1, pos.size()) * sizeof(float4)), "hipMalloc");
    hip_check(hipMalloc(&d->indices, std::max<size_t>(1, m.indices.size()) * sizeof(uint32_t)), "hipMalloc");
    if (!pos.empty())
        hip_check(hipMemcpyAsync(d->positions, pos.data(), pos.size() * sizeof(float4), hipMemcpyHostToDevice, stream_),
                  "upload positions");
    if (!m.indices.empty())
        hip_check(hipMemcpyAsync(d->indices, m.indices.data(), m.indices.size() * 4, hipMemcpyHostToDevice, stream_),
                  "upload indices");
    // Per-face sidedness (include/blitz/mesh.hpp:103-104).
    std::vector<uint8_t> two(d->faces, 0);
    bool any = false;
    if (!m.double_sided.empty())
        for (uint32_t f = 0; f < d->faces; ++f) {
            const uint16_t mat = m.materials.empty() ? 0 : m.materials[f];
            two[f] = mat < m.double_sided.size() && m.double_sided[mat];
            any |= two[f] != 0;
        }
    if (any) {
        hip_check(hipMalloc(&d->two_sided, two.size()), "hipMalloc");
        hip_check(hipMemcpyAsync(d->two_sided, two.data(), two.size(), hipMemcpyHostToDevice, stream_),
                  "upload sidedness");
    }
    // Appearance streams as the attribute raster reads them
    // (src/evaluate.cpp:141-143): unit normals and colours / 255 per vertex.
    std::vector<float4> nrm, col;
    if (!m.normals.empty() && d->vertices) {
        nrm.resize(d->vertices);
        for (uint32_t i = 0; i < d->vertices; ++i) {
            const F3 n = normalized({m.normals[3 * i], m.normals[3 * i + 1], m.normals[3 * i + 2]});
            nrm[i] = make_float4(n.x, n.y, n.z, 0.f);
        }
        hip_check(hipMalloc(&d->normals, nrm.size() * sizeof(float4)), "hipMalloc");
        hip_check(hipMemcpyAsync(d->normals, nrm.data(), nrm.size() * sizeof(float4), hipMemcpyHostToDevice, stream_),
                  "upload normals");
    }
    if (!m.colors.empty() && d->vertices) {
        col.resize(d->vertices);
        for (uint32_t i = 0; i < d->vertices; ++i)
            col[i] = make_float4(m.colors[4 * i] / 255.f, m.colors[4 * i + 1] / 255.f, m.colors[4 * i + 2] / 255.f,
                                 m.colors[4 * i + 3] / 255.f);
        hip_check(hipMalloc(&d->colors, col.size() * sizeof(float4)), "hipMalloc");
        hip_check(hipMemcpyAsync(d->colors, col.data(), col.size() * sizeof(float4), hipMemcpyHostToDevice, stream_),
                  "upload colours");
    }
    if (!m.materials.empty() && d->faces) {
        hip_check(hipMalloc(&d->materials, size_t(d->faces) * 2), "hipMalloc");
        hip_check(hipMemcpyAsync(d->materials, m.materials.data(), size_t(d->faces) * 2, hipMemcpyHostToDevice,
                                 stream_),
                  "upload materials");
    }
    // The triple table of a mesh that may be the reference of a face split
    // (raster_launch): built once per upload.
    if (face_split_ == 2 || (face_split_ == 0 && d->faces >= kSplitMinFaces)) {
        hip_check(hipMalloc(&d->face_table, size_t(face_table_slots(d->faces)) * sizeof(uint4)), "hipMalloc");
        launch_face_table(d->indices, d->two_sided, d->faces, d->face_table, stream_);
        hip_check(hipGetLastError(), "face table launch");
    }
    hip_check(hipStreamSynchronize(stream_), "upload");
    return d;
}

std::unique_ptr<DeviceMesh> Audit::derive(const DeviceMesh& base, const std::vector<uint32_t>& indices,
                                          const std::vector<uint16_t>& materials,
                                          const std::vector<uint8_t>& double_sided, const uint64_t digest[2]) {
    if (base.owner != this) throw std::invalid_argument("base mesh belongs to another audit");
    if (indices.size() % 3 || indices.size() / 3 >= UINT32_MAX)
        throw std::invalid_argument("mesh requires 32-bit representable counts and whole triangles");
    if (!materials.empty() && materials.size() != indices.size() / 3)
        throw std::invalid_argument("material count differs from triangle count");
    for (uint32_t v : indices)
        if (v >= base.vertices) throw std::invalid_argument("index outside vertex stream");
    DeviceGuard guard(device_);
    auto d = std::make_unique<DeviceMesh>();
    d->owner = this;
    d->borrowed_vertices = true;
    d->positions = base.positions;
    d->normals = base.normals;
    d->colors = base.colors;
    d->vertices = base.vertices;
    d->faces = uint32_t(indices.size() / 3);
    d->digest[0] = digest[0];
    d->digest[1] = digest[1];
    d->pos_digest[0] = base.pos_digest[0];
    d->pos_digest[1] = base.pos_digest[1];
    d->vertex_digest[0] = base.vertex_digest[0];
    d->vertex_digest[1] = base.vertex_digest[1];
    d->serial = next_serial();
    d->positions_serial = base.positions_serial;
    d->raw_normals = base.raw_normals;
    d->raw_colors = base.raw_colors;
    d->raw_double_sided = double_sided;
    hip_check(hipMalloc(&d->indices, std::max<size_t>(1, indices.size()) * sizeof(uint32_t)), "hipMalloc");
    if (!indices.empty())
        hip_check(hipMemcpyAsync(d->indices, indices.data(), indices.size() * 4, hipMemcpyHostToDevice, stream_),
                  "upload indices");
    std::vector<uint8_t> two(d->faces, 0);
    bool any = false;
    if (!double_sided.empty())
        for (uint32_t f = 0; f < d->faces; ++f) {
            const uint16_t mat = materials.empty() ? 0 : materials[f];
            two[f] = mat < double_sided.size() && double_sided[mat];
            any |= two[f] != 0;
        }
    if (any) {
        hip_check(hipMalloc(&d->two_sided, two.size()), "hipMalloc");
        hip_check(hipMemcpyAsync(d->two_sided, two.data(), two.size(), hipMemcpyHostToDevice, stream_),
                  "upload sidedness");
    }
    if (!materials.empty() && d->faces) {
        hip_check(hipMalloc(&d->materials, size_t(d->faces) * 2), "hipMalloc");
        hip_check(hipMemcpyAsync(d->materials, materials.data(), size_t(d->faces) * 2, hipMemcpyHostToDevice,
                                 stream_),
                  "upload materials");
    }
    hip_check(hipStreamSynchronize(stream_), "derive");
    return d;
}

void Audit::release(DeviceMesh*) {}

bool Audit::device_equal(const void* a, const void* b, size_t bytes) {
    if (a == b || !bytes) return true;
    DeviceGuard guard(device_);
    uint64_t* flag = scratch_as<uint64_t>(48, 1);  // a slot no other pass uses
    launch_fill(flag, 1, 0, stream_);
    launch_differ(a, b, bytes, flag, stream_);
    uint64_t differs = 1;
    hip_check(hipMemcpyAsync(&differs, flag, 8, hipMemcpyDeviceToHost, stream_), "download compare flag");
    hip_check(hipStreamSynchronize(stream_), "compare streams");
    return differs == 0;
}

bool Audit::same_positions(const DeviceMesh& a, const DeviceMesh& b) {
    if (a.positions == b.positions || a.positions_serial == b.positions_serial) return a.vertices == b.vertices;
    if (a.vertices != b.vertices || a.pos_digest[0] != b.pos_digest[0] || a.pos_digest[1] != b.pos_digest[1])
        return false;
    const auto key = std::minmax(a.positions_serial, b.positions_serial);
    auto it = position_pairs_.find(key);
    if (it == position_pairs_.end())
        it = position_pairs_.emplace(key, device_equal(a.positions, b.positions, size_t(a.vertices) * sizeof(float4)))
                 .first;
    return it->second;
}

bool Audit::identical(const DeviceMesh& a, const DeviceMesh& b) {
    if (&a == &b) return true;
    if (a.digest[0] != b.digest[0] || a.digest[1] != b.digest[1] || a.vertices != b.vertices || a.faces != b.faces)
        return false;
    const auto same_host = [](const auto& x, const auto& y) { return x == y || (x && y && *x == *y); };
    if (a.raw_double_sided != b.raw_double_sided || !same_host(a.raw_normals, b.raw_normals) ||
        !same_host(a.raw_colors, b.raw_colors) || !a.materials != !b.materials)
        return false;
    return same_positions(a, b) && device_equal(a.indices, b.indices, size_t(a.faces) * 12) &&
           (!a.materials || device_equal(a.materials, b.materials, size_t(a.faces) * 2));
}

// Coverage-profile view finish (src/evaluate.cpp:271-298).
Measurement finish_view(double coverage, uint64_t changed, uint64_t total, bool clipped, const Settings& s, uint8_t ss) {
    Measurement cur;
    cur.supersample = ss;
    cur.coverage = coverage;
    cur.coverage_upper = cur.coverage + 2 * std::sqrt(2.0) / ss + 1e-6;
    if (clipped) cur.coverage_upper = kInf;
    cur.error = cur.coverage_upper;
    cur.changed_area = total ? double(changed) / double(total) : 0;
    cur.passed = cur.error <= s.limit && cur.changed_area <= s.max_changed_area;
    if (s.conservative_screen) {
        cur.error = clipped ? 0 : std::max(0.0, cur.coverage - 2 * std::sqrt(2.0) / ss - 1e-6);
        cur.passed = cur.error <= s.limit;
    }
    return cur;
}
// Normals and attributes profiles (src/evaluate.cpp:282-298).
Measurement finish_view_attributed(double coverage, uint64_t changed, uint64_t total, bool clipped,
                                   double normal_acos, double attributed, double previous_degrees,
                                   const Settings& s, uint8_t ss) {
    Measurement cur = finish_view(coverage, changed, total, clipped, s, ss);
    cur.normal_degrees = std::max(previous_degrees, normal_acos * 180 / kPi);
    cur.error = cur.coverage_upper > s.limit ? cur.coverage_upper : std::max(cur.coverage_upper, attributed);
    cur.passed = cur.error <= s.limit && cur.changed_area <= s.max_changed_area;
    return cur;
}

bool attributed_weights(const Settings& s) {
    return s.profile != 0 && (s.normal_weight != 0 || (s.profile == 2 && (s.color_weight != 0 || s.material_weight != 0)));
}

// Refinement stop rule (src/evaluate.cpp:336-337).
// Verdict-only audits: the smallest squared sample distance whose coverage
// (sqrt(float(d)) / ss, as the host rounds it) settles a coverage view as
// failing at this supersampling, by the rules of finish_view and settled():
// below the largest supersampling, coverage - 2*sqrt(2)/ss > limit (no
// refinement can pass it); at the largest, coverage_upper > limit. A
// conservative screen never refines and fails a view only on
// max(0, coverage - 2*sqrt(2)/ss - 1e-6) > limit (a clipped view passes
// whatever its distance), at every supersampling.
static uint32_t failing_distance(const Settings& s, unsigned ss) {
    const auto fails = [&](uint32_t d) {
        const double c = std::sqrt(double(float(d))) / ss;
        if (s.conservative_screen) return std::max(0.0, c - 2 * std::sqrt(2.0) / ss - 1e-6) > s.limit;
        return ss >= s.max_supersample ? c + 2 * std::sqrt(2.0) / ss + 1e-6 > s.limit
                                       : c - 2 * std::sqrt(2.0) / ss > s.limit;
    };
    uint32_t lo = 0, hi = 0xFFFFFFFEu;  // fails(hi) holds for any finite limit
    if (!fails(hi)) return 0xFFFFFFFFu;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (fails(mid))
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

bool settled(const Measurement& cur, const Settings& s, unsigned ss) {
    return s.conservative_screen || cur.passed || ss >= s.max_supersample ||
           cur.coverage - 2 * std::sqrt(2.0) / ss > s.limit || !std::isfinite(cur.error) ||
           cur.coverage_upper <= s.limit;
}

Measurement Audit::evaluate(const DeviceMesh& reference, const DeviceMesh& candidate, const Bounds& b,
                            const Settings& s) {
    return evaluate_batch(reference, {&candidate}, b, s)[0];
}

std::vector<Measurement> Audit::evaluate_batch(const DeviceMesh& reference,
                                               const std::vector<const DeviceMesh*>& candidates, const Bounds& b,
                                               const Settings& s) {
    validate_settings(b, s);
    // Conservative screens use the coverage rule under every profile
    // (src/evaluate.cpp:324-326).
    const bool attributed = s.profile != 0 && !s.conservative_screen;
    Clock clock;
    timing_ = {};
    const auto cams = cameras(b, s.screen_size, s.orthographic, s.perspective, s.rotation_seed);
    const uint32_t views = uint32_t(cams.size());
    const size_t n = candidates.size();
    struct State {
        uint8_t ss = 0;
        bool final = false, limited = false;
        Measurement m;
    };
    std::vector<std::vector<State>> state(n, std::vector<State>(views));
    std::vector<bool> identical(n, false);
    for (size_t c = 0; c < n; ++c) {
        const DeviceMesh& m = *candidates[c];
        // src/evaluate.cpp:77-82,317: byte-identical streams skip every view.
        identical[c] = this->identical(m, reference);
        for (auto& st : state[c]) st.ss = s.supersample;
    }
    DeviceGuard guard(device_);
    // Views per candidate and round. The reference walks its views in order
    // and stops at the first failing one (src/evaluate.cpp:319-350). Under
    // the appearance profiles (attribute rasters and windowed searches) the
    // first round takes the first pending view and each later round twice
    // as many, so work after an early failure is skipped while a fully
    // passing audit needs only about log2(views) more rounds. The coverage
    // profile takes every pending view at once. Results do not depend on the
    // window.
    uint32_t window = attributed ? 1u : 0u;
    for (;; window = window ? std::min(window * 2, 1u << 16) : 0u) {
        // Per candidate, the first `window` pending views in view order (all
        // of them for window 0), stopping at a settled failing view.
        std::vector<Item> items;
        for (size_t c = 0; c < n; ++c) {
            if (identical[c]) continue;
            uint32_t taken = 0;
            for (uint32_t v = 0; v < views; ++v) {
                State& st = state[c][v];
                if (st.final) {
                    if (st.limited || !st.m.passed) break;
                    continue;
                }
                const uint64_t side = raster_side(s.screen_size, st.ss);
                if (side * side > kMaxSamples) {
                    st.final = st.limited = true;
                    break;
                }
                items.push_back({uint32_t(c), v, st.ss});
                if (window && ++taken == window) break;
            }
        }
        if (items.empty()) break;
        // One pass per supersampling among the round's items (with a window,
        // views refined earlier run beside views seen for the first time).
        std::map<uint8_t, std::vector<Item>> by_ss;
        for (const Item& it : items) by_ss[it.ss].push_back(it);
        for (const auto& group : by_ss) {
            const std::vector<Item>& pass = group.second;
            std::vector<ViewResult> results;
            run_items(reference, candidates, cams, b, s, pass, results);
            ++timing_.passes;
            for (size_t i = 0; i < pass.size(); ++i) {
                const Item& it = pass[i];
                State& st = state[it.candidate][it.view];
                const ViewResult& res = results[i];
                st.m = attributed ? finish_view_attributed(res.coverage, res.changed, res.total, res.clipped,
                                                           res.normal_acos, res.attributed, st.m.normal_degrees, s,
                                                           it.ss)
                                  : finish_view(res.coverage, res.changed, res.total, res.clipped, s, it.ss);
                if (settled(st.m, s, it.ss))
                    st.final = true;
                else
                    st.ss = uint8_t(std::min<unsigned>(s.max_supersample, unsigned(it.ss) * 2));
            }
        }
    }
    // Sequential assembly, exactly as the reference's view loop.
    std::vector<Measurement> out(n);
    for (size_t c = 0; c < n; ++c) {
        Measurement& r = out[c];
        r.supersample = s.supersample;
        if (identical[c]) continue;
        for (uint32_t v = 0; v < views; ++v) {
            if (s.cancelled && s.cancelled()) {
                r.complete = r.passed = false;
                r.cancelled = true;
                break;
            }
            const State& st = state[c][v];
            if (st.limited) {
                r.complete = r.passed = false;
                r.resource_limited = true;
                r.error = kInf;
                break;
            }
            const Measurement& cur = st.m;
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
            r.normal_degrees = std::max(r.normal_degrees, cur.normal_degrees);
            if (!cur.passed) {
                r.passed = r.complete = false;
                break;
            }
        }
    }
    timing_.total_ms = clock.ms();
    return out;
}

void Audit::launch_tasks(const std::vector<RasterTask>& list, size_t slot, uint32_t tiles) {
    RasterTask* d = scratch_as<RasterTask>(slot, std::max<size_t>(1, list.size()));
    hip_check(hipMemcpyAsync(d, list.data(), list.size() * sizeof(RasterTask), hipMemcpyHostToDevice, stream_),
              "upload raster tasks");
    uint32_t faces = 0;
    for (const auto& t : list) faces = std::max(faces, t.faces);
    const auto launch = [&](const RasterGrid& g) { launch_raster(d, g, tiles, stream_); };
    const RasterGrid g = tuner_.grid(uint32_t(list.size()), tiles, faces, stream_, launch);
    launch(g);
    timing_.raster_chunks = g.chunks;
    timing_.tuned_launches += tuner_.last_tuned();
}

FaceSplit Audit::split_faces(const DeviceMesh& a, const DeviceMesh& b) {
    const size_t fa = std::max<uint32_t>(1, a.faces), fb = std::max<uint32_t>(1, b.faces);
    uint32_t* blocks = scratch_as<uint32_t>(40, 3 * size_t(std::max(1u, face_split_blocks(a.faces, b.faces))));
    uint8_t* matched = scratch_as<uint8_t>(41, (fa + 7) / 8 * 8);
    uint32_t* match = scratch_as<uint32_t>(42, fb);
    FaceSplit sp{};
    sp.shared_tri = scratch_as<uint32_t>(43, 3 * fb);
    sp.shared_ids_a = scratch_as<uint32_t>(44, 2 * fb);
    sp.shared_ids_b = sp.shared_ids_a + fb;
    sp.a_tri = scratch_as<uint32_t>(45, 4 * fa);
    sp.a_ids = sp.a_tri + 3 * fa;
    sp.b_tri = scratch_as<uint32_t>(46, 4 * fb);
    sp.b_ids = sp.b_tri + 3 * fb;
    sp.counts = scratch_as<uint32_t>(47, 4);
    launch_face_split(a.indices, a.two_sided, a.faces, a.face_table, b.indices, b.two_sided, b.faces, match, matched,
                      blocks, sp, stream_);
    return sp;
}

void Audit::raster_launch(const std::vector<RasterJob>& jobs, const std::vector<Camera>& cams, const Bounds& b,
                          uint8_t ss, uint32_t size, bool force_two, bool want_depth, RasterLaunch& out,
                          const std::vector<std::pair<uint32_t, uint32_t>>* pairs) {
    const uint32_t tiles = (size + kTile - 1) / kTile;
    const size_t words = size_t(tiles) * tiles;
    // One projection per (vertex stream contents, view).
    struct Key {
        uint64_t d0, d1;
        uint32_t vertices, view;
        bool operator<(const Key& o) const {
            return std::tie(d0, d1, vertices, view) < std::tie(o.d0, o.d1, o.vertices, o.view);
        }
    };
    // Equal digests only propose sharing; same_positions() confirms it.
    std::map<Key, std::vector<uint32_t>> slot;
    std::vector<std::pair<const DeviceMesh*, uint32_t>> sets;
    std::vector<uint32_t> job_slot(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
        const Key k{jobs[j].mesh->pos_digest[0], jobs[j].mesh->pos_digest[1], jobs[j].mesh->vertices, jobs[j].view};
        auto& list = slot[k];
        uint32_t found = kNone;
        for (uint32_t si : list)
            if (same_positions(*sets[si].first, *jobs[j].mesh)) {
                found = si;
                break;
            }
        if (found == kNone) {
            found = uint32_t(sets.size());
            list.push_back(found);
            sets.push_back({jobs[j].mesh, jobs[j].view});
        }
        job_slot[j] = found;
    }
    std::vector<size_t> offset(sets.size() + 1, 0);
    uint32_t max_vertices = 0;
    bool perspective = false;
    for (size_t i = 0; i < sets.size(); ++i) {
        offset[i + 1] = offset[i] + sets[i].first->vertices;
        max_vertices = std::max(max_vertices, sets[i].first->vertices);
        perspective |= cams[sets[i].second].perspective != 0;
    }
    double2* xy = scratch_as<double2>(9, std::max<size_t>(1, offset.back()));
    // View depths: perspective views divide by them; attribute rasters
    // interpolate them in every view.
    double* zs = perspective || want_depth ? scratch_as<double>(10, std::max<size_t>(1, offset.back())) : nullptr;
    std::vector<ProjectTask> projections(sets.size());
    for (size_t i = 0; i < sets.size(); ++i) {
        ProjectTask& p = projections[i];
        p = {};
        p.positions = sets[i].first->positions;
        p.xy = xy + offset[i];
        p.z = cams[sets[i].second].perspective || want_depth ? zs + offset[i] : nullptr;
        p.camera = cams[sets[i].second];
        p.center = b.center;
        p.vertices = sets[i].first->vertices;
        p.size = size;
        p.ss = ss;
    }
    std::vector<RasterTask> tasks(jobs.size());
    uint32_t max_faces = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
        const DeviceMesh& m = *jobs[j].mesh;
        RasterTask& t = tasks[j];
        t = {};
        t.xy = xy + offset[job_slot[j]];
        t.z = cams[jobs[j].view].perspective ? zs + offset[job_slot[j]] : nullptr;
        t.indices = m.indices;
        t.two_sided = m.two_sided;
        t.mask = jobs[j].mask;
        t.faces = m.faces;
        t.size = size;
        t.tiles = tiles;
        t.force_two = force_two;
        max_faces = std::max(max_faces, m.faces);
    }
    uint32_t* status = scratch_as<uint32_t>(6, std::max<size_t>(1, jobs.size() * 2));
    for (size_t j = 0; j < jobs.size(); ++j) tasks[j].status = status + 2 * j;
    std::vector<uint32_t> init(jobs.size() * 2);
    for (size_t j = 0; j < jobs.size(); ++j) {
        init[2 * j] = kNone;
        init[2 * j + 1] = 0;
    }
    hip_check(hipMemcpyAsync(status, init.data(), init.size() * 4, hipMemcpyHostToDevice, stream_), "upload status");
    ProjectTask* dp = scratch_as<ProjectTask>(8, std::max<size_t>(1, projections.size()));
    hip_check(hipMemcpyAsync(dp, projections.data(), projections.size() * sizeof(ProjectTask), hipMemcpyHostToDevice,
                             stream_),
              "upload projections");
    // Reference/candidate pairs over one vertex set: their faces split into
    // shared, reference-only and candidate-only lists (one split for every
    // pair: all pairs hold the same two meshes), each list one raster task.
    // The split costs a hash build and probe over both meshes' faces, so
    // only rasters with at least 4M face-views are split.
    std::vector<RasterTask> list;
    bool split = false;
    if (pairs && !pairs->empty()) {
        const DeviceMesh& ra = *jobs[pairs->front().first].mesh;
        const DeviceMesh& cb = *jobs[pairs->front().second].mesh;
        split = ra.face_table != nullptr &&
                (face_split_ == 2 || (face_split_ == 0 && cb.faces >= kSplitMinFaces &&
                                      uint64_t(ra.faces + cb.faces) * pairs->size() >= (uint64_t(1) << 22)));
        for (const auto& p : *pairs)
            split = split && jobs[p.first].mesh == &ra && jobs[p.second].mesh == &cb &&
                    job_slot[p.first] == job_slot[p.second];
    }
    if (split) {
        hip_check(hipEventRecord(events_[4], stream_), "event");
        const DeviceMesh& ra = *jobs[pairs->front().first].mesh;
        const DeviceMesh& cb = *jobs[pairs->front().second].mesh;
        const FaceSplit sp = split_faces(ra, cb);
        hip_check(hipEventRecord(events_[10], stream_), "event");
        std::vector<int32_t> pair_of(jobs.size(), -1);
        for (size_t k = 0; k < pairs->size(); ++k) {
            pair_of[(*pairs)[k].first] = int32_t(k);
            pair_of[(*pairs)[k].second] = int32_t(k);
        }
        for (size_t j = 0; j < jobs.size(); ++j) {
            if (pair_of[j] < 0) {
                list.push_back(tasks[j]);
                continue;
            }
            if ((*pairs)[size_t(pair_of[j])].first != j) continue;  // emitted with its reference job
            const RasterTask& r = tasks[j];
            const RasterTask& c = tasks[(*pairs)[size_t(pair_of[j])].second];
            RasterTask t = r;
            t.two_sided = nullptr;
            t.or_flush = 1;
            // Shared faces: both masks, both meshes' status words.
            t.indices = sp.shared_tri;
            t.ids = sp.shared_ids_a;
            t.mask2 = c.mask;
            t.status2 = c.status;
            t.ids2 = sp.shared_ids_b;
            t.faces = cb.faces;
            t.count = sp.counts;
            list.push_back(t);
            t.mask2 = nullptr;
            t.status2 = nullptr;
            t.ids2 = nullptr;
            // Reference only.
            t.indices = sp.a_tri;
            t.ids = sp.a_ids;
            t.faces = ra.faces;
            t.count = sp.counts + 1;
            list.push_back(t);
            // Candidate only.
            t.indices = sp.b_tri;
            t.ids = sp.b_ids;
            t.mask = c.mask;
            t.status = c.status;
            t.faces = cb.faces;
            t.count = sp.counts + 2;
            list.push_back(t);
        }
    }
    hip_check(hipEventRecord(events_[5], stream_), "event");
    launch_project(dp, uint32_t(projections.size()), max_vertices, stream_);
    hip_check(hipEventRecord(events_[6], stream_), "event");
    launch_tasks(split ? list : tasks, 4, tiles);
    hip_check(hipEventRecord(events_[7], stream_), "event");
    hip_check(hipGetLastError(), "raster launch");
    // Status words go to pinned memory without a wait: the caller may queue
    // more work on the masks before raster_finish() reads them.
    if (pinned_status_words_ < init.size()) {
        hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        (void)hipHostFree(pinned_status_);
        pinned_status_ = nullptr;
        pinned_status_words_ = 0;
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&pinned_status_), std::max<size_t>(64, init.size()) * 4,
                                hipHostMallocDefault),
                  "hipHostMalloc");
        pinned_status_words_ = std::max<size_t>(64, init.size());
    }
    hip_check(hipMemcpyAsync(pinned_status_, status, init.size() * 4, hipMemcpyDeviceToHost, stream_),
              "download status");
    out.tasks = std::move(tasks);
    out.split = split;
    out.words = words;
    out.tiles = tiles;
    out.depth.assign(jobs.size(), nullptr);
    if (zs)
        for (size_t j = 0; j < jobs.size(); ++j) out.depth[j] = zs + offset[job_slot[j]];
}

bool Audit::raster_finish(RasterLaunch& l, std::vector<uint32_t>& status_out, std::vector<RasterTask>* drawn,
                          std::vector<const double*>* depth) {
    const size_t n = l.tasks.size();
    status_out.assign(pinned_status_, pinned_status_ + 2 * n);
    {
        float t_p = 0, t_b = 0;
        hip_check(hipEventElapsedTime(&t_p, events_[5], events_[6]), "elapsed");
        hip_check(hipEventElapsedTime(&t_b, events_[6], events_[7]), "elapsed");
        timing_.project_ms += t_p;
        timing_.band_ms += t_b;
        ++timing_.band_launches;
        if (l.split) {
            float t_s = 0;
            hip_check(hipEventElapsedTime(&t_s, events_[4], events_[10]), "elapsed");
            timing_.split_ms += t_s;
            ++timing_.split_launches;
        }
    }
    timing_.raster_tasks += uint32_t(n);
    for (const auto& t : l.tasks) timing_.raster_triangles += t.faces;
    timing_.mask_bytes += uint64_t(n) * l.words * 8;
    // The reference stops rasterizing at the first out-of-frame face
    // (src/evaluate.cpp:159-160): redraw such masks from the faces before it.
    std::vector<RasterTask> redo;
    for (size_t j = 0; j < n; ++j)
        if (status_out[2 * j] != kNone) {
            RasterTask t = l.tasks[j];
            t.faces = status_out[2 * j];
            redo.push_back(t);
            launch_fill(t.mask, l.words, 0, stream_);
        }
    if (!redo.empty()) {
        launch_tasks(redo, 7, l.tiles);
        hip_check(hipGetLastError(), "raster redo launch");
        hip_check(hipStreamSynchronize(stream_), "raster redo");
    }
    if (drawn) {
        *drawn = l.tasks;
        for (size_t j = 0; j < n; ++j)
            if (status_out[2 * j] != kNone) (*drawn)[j].faces = status_out[2 * j];
    }
    if (depth) *depth = l.depth;
    return !redo.empty();
}

void Audit::rasterize(const std::vector<RasterJob>& jobs, const std::vector<Camera>& cams, const Bounds& b,
                      uint8_t ss, uint32_t size, bool force_two, std::vector<uint32_t>& status_out,
                      std::vector<RasterTask>* drawn, std::vector<const double*>* depth) {
    RasterLaunch l;
    raster_launch(jobs, cams, b, ss, size, force_two, drawn != nullptr || depth != nullptr, l);
    hip_check(hipStreamSynchronize(stream_), "raster");
    raster_finish(l, status_out, drawn, depth);
}

void Audit::run_items(const DeviceMesh& reference, const std::vector<const DeviceMesh*>& candidates,
                      const std::vector<Camera>& cams, const Bounds& b, const Settings& s,
                      const std::vector<Item>& all, std::vector<ViewResult>& out) {
    out.assign(all.size(), {});
    const uint8_t ss = all.front().ss;
    const uint32_t size = raster_side(s.screen_size, ss), tiles = (size + kTile - 1) / kTile;
    const size_t words = size_t(tiles) * tiles;
    const size_t samples = size_t(size) * size;
    // Per item: candidate mask, two row-distance planes, projection, summary.
    // Per view: one reference mask shared by every candidate of the chunk.
    const bool attributed = s.profile != 0 && !s.conservative_screen;
    // Appearance profiles add two attribute rasters (normal, colour: 32 bytes
    // per sample) per item and per reference view, and the items' face bins.
    const size_t per_item = words * 8 + samples * 4 + sizeof(ViewSummary) + sizeof(RasterTask) + sizeof(ViewPair) +
                            24 * size_t(std::max(reference.vertices, 1u)) +
                            (attributed ? words * 64 * 64 + words * 16 : 0);
    const uint64_t budget = memory_bytes_ ? memory_bytes_ : (uint64_t(8) << 30);
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(all.size(), budget / (per_item + words * 8)));
    for (size_t begin = 0; begin < all.size(); begin += chunk) {
        const size_t count = std::min(chunk, all.size() - begin);
        std::map<uint32_t, uint32_t> ref_slot;
        for (size_t i = 0; i < count; ++i) ref_slot.emplace(all[begin + i].view, 0u);
        uint32_t next = 0;
        for (auto& kv : ref_slot) kv.second = next++;
        const uint32_t nref = next, ncand = uint32_t(count);
        uint64_t* ref_masks = scratch_as<uint64_t>(0, size_t(nref) * words);
        uint64_t* cand_masks = scratch_as<uint64_t>(1, size_t(ncand) * words);
        uint16_t* dist = scratch_as<uint16_t>(2, size_t(ncand) * samples * 2);
        ViewSummary* summaries = scratch_as<ViewSummary>(3, ncand);
        ViewPair* dpairs = scratch_as<ViewPair>(5, ncand);
        // Reference masks kept from earlier audits of the same reference
        // (coverage rule only) are not drawn again.
        const bool cached = cache_rasters_ && !attributed;
        if (cached) cache_context(reference, b, s);
        std::vector<const CachedMask*> hit(nref, nullptr);
        if (cached)
            for (auto& kv : ref_slot) {
                auto it = cache_.masks.find({ss, kv.first});
                if (it != cache_.masks.end()) hit[kv.second] = &it->second;
            }
        // Jobs in view order, each view's reference first: consecutive raster
        // tasks then share one projection (and one XCD's L2).
        std::vector<RasterJob> jobs;
        std::vector<uint32_t> ref_job(nref, kNone), cand_job(ncand);
        {
            std::vector<std::vector<uint32_t>> by_view(nref);
            for (uint32_t i = 0; i < ncand; ++i) by_view[ref_slot[all[begin + i].view]].push_back(i);
            for (auto& kv : ref_slot) {
                if (!hit[kv.second]) {
                    ref_job[kv.second] = uint32_t(jobs.size());
                    jobs.push_back({&reference, kv.first, ref_masks + size_t(kv.second) * words});
                }
                for (uint32_t i : by_view[kv.second]) {
                    cand_job[i] = uint32_t(jobs.size());
                    jobs.push_back({candidates[all[begin + i].candidate], kv.first, cand_masks + size_t(i) * words});
                }
            }
        }
        const uint32_t fail_sq = failing_distance(s, ss);
        std::vector<ViewPair> pairs(ncand);
        for (uint32_t i = 0; i < ncand; ++i) {
            const Item& it = all[begin + i];
            ViewPair& p = pairs[i];
            const uint32_t slot = ref_slot[it.view];
            p.a = hit[slot] ? hit[slot]->mask : ref_masks + size_t(slot) * words;
            p.c = cand_masks + size_t(i) * words;
            p.dist_a = dist + size_t(i) * samples * 2;
            p.dist_c = p.dist_a + samples;
            p.size = size;
            p.tiles = tiles;
            p.fail_sq = s.verdict_only && !attributed ? fail_sq : 0xFFFFFFFFu;
        }
        hip_check(hipEventRecord(events_[0], stream_), "event");
        launch_fill(ref_masks, size_t(nref) * words, 0, stream_);
        launch_fill(cand_masks, size_t(ncand) * words, 0, stream_);
        std::vector<uint32_t> st;
        std::vector<RasterTask> drawn;
        std::vector<const double*> zdepth;
        // Coverage audits queue the summaries and distances (and the copies
        // of new reference masks into the cache) right behind the raster and
        // read its status words with them (one wait); masks cut at an
        // out-of-frame face are then redrawn, copied and measured again.
        // Appearance rasters need the final masks first.
        const bool deferred = !attributed;
        // Each drawn reference view with its first candidate, when every
        // such pair holds the same candidate mesh.
        std::vector<std::pair<uint32_t, uint32_t>> split_pairs;
        if (deferred) {
            const DeviceMesh* only = nullptr;
            bool one = true;
            for (uint32_t i = 0; i < ncand; ++i) {
                const DeviceMesh* c = candidates[all[begin + i].candidate];
                one = one && (!only || only == c);
                only = c;
            }
            std::vector<uint8_t> paired(nref, 0);
            for (uint32_t i = 0; one && i < ncand; ++i) {
                const uint32_t r = ref_slot[all[begin + i].view];
                if (ref_job[r] == kNone || paired[r]) continue;
                paired[r] = 1;
                split_pairs.push_back({ref_job[r], cand_job[i]});
            }
            if (!one) split_pairs.clear();
        }
        RasterLaunch rl;
        if (deferred)
            raster_launch(jobs, cams, b, ss, size, s.force_two_sided, false, rl, &split_pairs);
        else
            rasterize(jobs, cams, b, ss, size, s.force_two_sided, st, attributed ? &drawn : nullptr,
                      attributed ? &zdepth : nullptr);
        // Reference clip flags, and new reference masks into the cache while
        // it stays within a quarter of the memory budget.
        std::vector<uint8_t> ref_clipped(nref, 0);
        const auto clip_flags = [&] {
            for (uint32_t r = 0; r < nref; ++r)
                ref_clipped[r] = hit[r] ? hit[r]->clipped
                                        : (st[2 * ref_job[r]] != kNone || st[2 * ref_job[r] + 1] != 0);
        };
        if (!deferred) clip_flags();
        // Cache entries made by this pass (reference slot, entry); their clip
        // flags are set once the status words are read.
        std::vector<std::pair<uint32_t, CachedMask*>> fresh;
        const auto copy_fresh = [&] {
            for (const auto& e : fresh)
                hip_check(hipMemcpyAsync(e.second->mask, ref_masks + size_t(e.first) * words, words * 8,
                                         hipMemcpyDeviceToDevice, stream_),
                          "cache reference mask");
        };
        if (cached) {
            for (auto& kv : ref_slot) {
                const uint32_t r = kv.second;
                if (hit[r] || cache_.bytes + words * 8 > budget / 4) continue;
                uint64_t* keep = nullptr;
                hip_check(hipMalloc(&keep, words * 8), "hipMalloc");
                CachedMask& entry = cache_.masks[{ss, kv.first}];
                entry = {keep, false};
                fresh.push_back({r, &entry});
                cache_.bytes += words * 8;
            }
            copy_fresh();
        }
        hip_check(hipEventRecord(events_[1], stream_), "event");
        hip_check(hipMemcpyAsync(dpairs, pairs.data(), pairs.size() * sizeof(ViewPair), hipMemcpyHostToDevice, stream_),
                  "upload pairs");
        const auto measure = [&] {
            launch_fill(reinterpret_cast<uint64_t*>(summaries), ncand * sizeof(ViewSummary) / 8, 0, stream_);
            launch_compare(dpairs, summaries, ncand, tiles, stream_);
            hip_check(hipEventRecord(events_[2], stream_), "event");
            launch_distance(dpairs, summaries, ncand, size, tiles, stream_);
            hip_check(hipEventRecord(events_[3], stream_), "event");
            hip_check(hipGetLastError(), "audit launch");
        };
        measure();
        // Appearance profiles: attribute rasters of every job, then per item
        // the normal angle (src/evaluate.cpp:282-290).
        const size_t per = words * 64;
        std::vector<AttrPair> apairs;
        unsigned long long* aout = nullptr;
        std::vector<unsigned long long> angles;
        if (attributed) {
            hip_check(hipEventRecord(events_[8], stream_), "event");
            float4* anorm = scratch_as<float4>(30, jobs.size() * per);
            // Colours and materials matter to the attributes profile only.
            float4* acol = s.profile == 2 ? scratch_as<float4>(31, jobs.size() * per) : nullptr;
            attribute_rasters(jobs, drawn, zdepth, cams, b, anorm, acol, nullptr);
            aout = scratch_as<unsigned long long>(32, 2 * size_t(ncand));
            launch_fill(reinterpret_cast<uint64_t*>(aout), 2 * size_t(ncand), 0, stream_);
            apairs.resize(ncand);
            for (uint32_t i = 0; i < ncand; ++i) {
                const size_t ra = ref_job[ref_slot[all[begin + i].view]], rc = cand_job[i];
                apairs[i] = {anorm + ra * per, acol ? acol + ra * per : nullptr, anorm + rc * per,
                             acol ? acol + rc * per : nullptr, aout + 2 * size_t(i), size, tiles};
            }
            AttrPair* dap = scratch_as<AttrPair>(33, ncand);
            hip_check(hipMemcpyAsync(dap, apairs.data(), ncand * sizeof(AttrPair), hipMemcpyHostToDevice, stream_),
                      "upload attribute pairs");
            for (uint32_t i0 = 0; i0 < ncand; i0 += 65535)
                launch_normal_angle(dap + i0, std::min<uint32_t>(65535, ncand - i0), tiles, stream_);
            hip_check(hipGetLastError(), "normal angle launch");
            angles.resize(2 * size_t(ncand));
            hip_check(hipMemcpyAsync(angles.data(), aout, angles.size() * 8, hipMemcpyDeviceToHost, stream_),
                      "download angles");
        }
        std::vector<ViewSummary> sum(ncand);
        hip_check(hipMemcpyAsync(sum.data(), summaries, ncand * sizeof(ViewSummary), hipMemcpyDeviceToHost, stream_),
                  "download summaries");
        hip_check(hipStreamSynchronize(stream_), "audit");
        if (deferred) {
            if (raster_finish(rl, st)) {
                copy_fresh();
                measure();
                hip_check(hipMemcpyAsync(sum.data(), summaries, ncand * sizeof(ViewSummary), hipMemcpyDeviceToHost,
                                         stream_),
                          "download summaries");
                hip_check(hipStreamSynchronize(stream_), "audit");
            }
            clip_flags();
            for (const auto& e : fresh) e.second->clipped = ref_clipped[e.first] != 0;
        }
        float t_r = 0, t_c = 0, t_d = 0;
        hip_check(hipEventElapsedTime(&t_r, events_[0], events_[1]), "elapsed");
        hip_check(hipEventElapsedTime(&t_c, events_[1], events_[2]), "elapsed");
        hip_check(hipEventElapsedTime(&t_d, events_[2], events_[3]), "elapsed");
        timing_.raster_ms += t_r;
        timing_.compare_ms += t_c;
        timing_.distance_ms += t_d;
        timing_.distance_views += ncand;
        for (uint32_t i = 0; i < ncand; ++i) {
            const Item& it = all[begin + i];
            const uint32_t rc = cand_job[i];
            const bool clipped = ref_clipped[ref_slot[it.view]] || st[2 * rc] != kNone || st[2 * rc + 1];
            const ViewSummary& v = sum[i];
            ViewResult& r = out[begin + i];
            r.changed = v.changed;
            r.total = v.total;
            r.clipped = clipped;
            // src/evaluate.cpp:202,208: identical masks give 0, exactly one
            // empty mask gives infinity, otherwise the directed maxima of the
            // float distance field, rounded once here (rounding is monotone).
            if (!v.changed)
                r.coverage = 0;
            else if ((v.count_a != 0) != (v.count_c != 0))
                r.coverage = kInf;
            else
                r.coverage = std::sqrt(double(float(std::max(v.max_ac, v.max_ca)))) / ss;
            r.normal_acos = 0;
            r.attributed = 0;
            if (attributed) {
                double a;
                std::memcpy(&a, &angles[2 * size_t(i)], 8);
                r.normal_acos = a;
            }
        }
        if (!attributed) continue;
        // The attributed distance, only where the coverage bound is within
        // the limit (src/evaluate.cpp:292-293) and the weights can make it
        // non-zero (src/evaluate.cpp:218-219).
        std::vector<uint32_t> need;
        std::vector<AttrPair> list;
        if (attributed_weights(s))
            for (uint32_t i = 0; i < ncand; ++i) {
                const ViewResult& r = out[begin + i];
                if (finish_view(r.coverage, r.changed, r.total, r.clipped, s, ss).coverage_upper <= s.limit)
                    need.push_back(i);
            }
        if (!need.empty()) {
            list.resize(need.size());
            for (size_t k = 0; k < need.size(); ++k) list[k] = apairs[need[k]];
            AttrPair* dl = scratch_as<AttrPair>(34, list.size());
            hip_check(hipMemcpyAsync(dl, list.data(), list.size() * sizeof(AttrPair), hipMemcpyHostToDevice, stream_),
                      "upload attributed pairs");
            CostSettings cs{s.normal_weight, s.color_weight, s.material_weight, s.limit, s.profile, ss};
            for (size_t k0 = 0; k0 < list.size(); k0 += 65535)
                launch_attributed(dl + k0, uint32_t(std::min<size_t>(65535, list.size() - k0)), tiles, cs, stream_);
            hip_check(hipGetLastError(), "attributed launch");
            hip_check(hipMemcpyAsync(angles.data(), aout, angles.size() * 8, hipMemcpyDeviceToHost, stream_),
                      "download attributed");
        }
        hip_check(hipEventRecord(events_[9], stream_), "event");
        hip_check(hipStreamSynchronize(stream_), "attributed");
        {
            float t_a = 0;
            hip_check(hipEventElapsedTime(&t_a, events_[8], events_[9]), "elapsed");
            timing_.appearance_ms += t_a;
            timing_.attributed_views += uint32_t(need.size());
        }
        const double limit2 = s.limit * s.limit;
        for (uint32_t i : need) {
            double m;
            std::memcpy(&m, &angles[2 * size_t(i) + 1], 8);
            // src/evaluate.cpp:262-266: any sample beyond the limit makes the
            // distance infinite; otherwise the root of the largest cost.
            out[begin + i].attributed = m > limit2 ? kInf : std::sqrt(m);
        }
    }
}

bool Audit::raster_row_major(const DeviceMesh& m, const Bounds& b, const Camera& c, double screen, uint8_t ss,
                             bool force_two, std::vector<uint64_t>& words_out, uint32_t& size_out, bool& clipped) {
    if (!ss || !std::isfinite(screen) || screen <= 0 || screen > 16384)
        throw std::invalid_argument("invalid raster extent");
    const uint32_t size = raster_side(screen, ss);
    size_out = size;
    if (uint64_t(size) * size > kMaxSamples) return false;
    DeviceGuard guard(device_);
    const uint32_t tiles = (size + kTile - 1) / kTile;
    const size_t words = size_t(tiles) * tiles;
    uint64_t* mask = scratch_as<uint64_t>(0, words);
    launch_fill(mask, words, 0, stream_);
    std::vector<uint32_t> st;
    rasterize({{&m, 0, mask}}, {c}, b, ss, size, force_two, st);
    clipped = st[0] != kNone || st[1];
    std::vector<uint64_t> tiled(words);
    hip_check(hipMemcpy(tiled.data(), mask, words * 8, hipMemcpyDeviceToHost), "download mask");
    words_out.assign((size_t(size) * size + 63) / 64, 0);
    for (uint32_t ty = 0; ty < tiles; ++ty)
        for (uint32_t tx = 0; tx < tiles; ++tx) {
            uint64_t w = tiled[size_t(ty) * tiles + tx];
            while (w) {
                const int bit = __builtin_ctzll(w);
                w &= w - 1;
                const size_t x = tx * 8 + (bit & 7), y = ty * 8 + (bit >> 3);
                const size_t index = y * size + x;
                words_out[index / 64] |= uint64_t(1) << (index % 64);
            }
        }
    return true;
}

double Audit::coverage_distance(const std::vector<uint64_t>& a, const std::vector<uint64_t>& b, uint32_t size,
                                uint8_t ss) {
    if (!ss) throw std::invalid_argument("raster shapes differ");
    const size_t need = (size_t(size) * size + 63) / 64;
    if (a.size() < need || b.size() < need) throw std::invalid_argument("raster shapes differ");
    DeviceGuard guard(device_);
    const uint32_t tiles = (size + kTile - 1) / kTile;
    const size_t words = size_t(tiles) * tiles, samples = size_t(size) * size;
    std::vector<uint64_t> ta(words, 0), tb(words, 0);
    for (size_t i = 0; i < samples; ++i) {
        const size_t x = i % size, y = i / size, w = (y / 8) * tiles + x / 8;
        const uint64_t bit = uint64_t(1) << ((y & 7) * 8 + (x & 7));
        if ((a[i / 64] >> (i % 64)) & 1) ta[w] |= bit;
        if ((b[i / 64] >> (i % 64)) & 1) tb[w] |= bit;
    }
    uint64_t* masks = scratch_as<uint64_t>(0, words * 2);
    uint16_t* dist = scratch_as<uint16_t>(2, samples * 2);
    ViewSummary* summary = scratch_as<ViewSummary>(3, 1);
    ViewPair* pair = scratch_as<ViewPair>(5, 1);
    ViewPair p{masks, masks + words, dist, dist + samples, size, tiles};
    hip_check(hipMemcpyAsync(masks, ta.data(), words * 8, hipMemcpyHostToDevice, stream_), "upload a");
    hip_check(hipMemcpyAsync(masks + words, tb.data(), words * 8, hipMemcpyHostToDevice, stream_), "upload b");
    hip_check(hipMemcpyAsync(pair, &p, sizeof p, hipMemcpyHostToDevice, stream_), "upload pair");
    launch_fill(reinterpret_cast<uint64_t*>(summary), sizeof(ViewSummary) / 8, 0, stream_);
    launch_compare(pair, summary, 1, tiles, stream_);
    launch_distance(pair, summary, 1, size, tiles, stream_);
    hip_check(hipGetLastError(), "distance launch");
    ViewSummary v{};
    hip_check(hipMemcpyAsync(&v, summary, sizeof v, hipMemcpyDeviceToHost, stream_), "download summary");
    hip_check(hipStreamSynchronize(stream_), "distance");
    if (!v.changed) return 0;
    if ((v.count_a != 0) != (v.count_c != 0)) return kInf;
    return std::sqrt(double(float(std::max(v.max_ac, v.max_ca)))) / ss;
}

}  // namespace blitz_hip
