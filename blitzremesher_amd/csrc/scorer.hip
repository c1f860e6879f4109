// Host controller of the edge-graph proposal scorer (see scorer.hpp).
#include "scorer.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "graph_topology.hpp"
#include "scorer_kernels.hpp"

namespace blitz_hip {

DeviceBuffer::~DeviceBuffer() { (void)hipFree(p); }

void DeviceBuffer::resize(size_t n) {
    if (n <= bytes && p) return;
    hip_check(hipFree(p), "hipFree");
    p = nullptr;
    bytes = 0;
    hip_check(hipMalloc(&p, std::max<size_t>(n, 16)), "hipMalloc");
    bytes = std::max<size_t>(n, 16);
}

GraphView Graph::view() const {
    return {offsets.as<uint32_t>(), neighbors.as<uint32_t>(), src.as<uint32_t>(), reverse.as<uint32_t>(), vertices,
            actions};
}

namespace {
template <class T> void upload(DeviceBuffer& d, const std::vector<T>& v, hipStream_t s) {
    d.resize(v.size() * sizeof(T));
    if (!v.empty()) hip_check(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}
}  // namespace

Scorer::Scorer(int device, const float* weights, size_t count) : device_(device) {
    DeviceGuard guard(device_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&begin_), "hipEventCreate");
    hip_check(hipEventCreate(&end_), "hipEventCreate");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device_), "hipGetDeviceProperties");
    blocks_ = uint32_t(std::max(1, prop.multiProcessorCount)) * 2;
    set_weights(weights, count);
}

Scorer::~Scorer() {
    DeviceGuard guard(device_);
    (void)hipStreamSynchronize(stream_);
    (void)hipEventDestroy(begin_);
    (void)hipEventDestroy(end_);
    (void)hipStreamDestroy(stream_);
}

void Scorer::set_weights(const float* w, size_t count) {
    if (!w || count != kWeightCount) throw std::invalid_argument("scorer weight count mismatch");
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(w[i])) throw std::invalid_argument("nonfinite scorer weight");
    DeviceGuard guard(device_);
    weights_.resize(count * 4);
    hip_check(hipMemcpyAsync(weights_.p, w, count * 4, hipMemcpyHostToDevice, stream_), "upload weights");
    refresh_bf16();
    reset_optimizer();
    hip_check(hipStreamSynchronize(stream_), "weights");
}

// BF16 copy of the four MFMA layer matrices from the FP32 master weights
// (biases and the 4-wide head stay FP32).
void Scorer::refresh_bf16() {
    weights_bf16_.resize(size_t(kBf16WeightCount) * 2);
    for (int l = 0; l < 4; ++l)
        launch_convert_bf16(weights_.as<float>() + layer_offset(l), weights_bf16_.as<__bf16>() + bf16_offset(l),
                            size_t(kHidden) * kLayerIn[l], stream_);
}

// Sizes the per-vertex and per-action buffers of the forward pass.
static void size_forward(Graph& g) {
    const size_t nv = g.vertices;
    g.h_a.resize(nv * kHidden * 4);
    g.h_b.resize(nv * kHidden * 4);
    g.z.resize(std::max(nv * 2 * kHidden, size_t(g.actions) * kHidden) * 4);
    g.hq_a.resize(nv * kHidden * 2);
    g.hq_b.resize(nv * kHidden * 2);
    g.outputs.resize(size_t(g.actions) * kOutputs * 4);
}

std::unique_ptr<Graph> Scorer::graph(const HostMesh& m, const Bounds& b) {
    auto g = std::make_unique<Graph>();
    rebuild(*g, m, b);
    return g;
}

void Scorer::rebuild(Graph& graph, const HostMesh& m, const Bounds& b, bool same_positions) {
    if (!(b.radius > 0) || !std::isfinite(b.radius)) throw std::invalid_argument("invalid bounds");
    DeviceGuard guard(device_);
    Graph* g = &graph;
    g->has_outputs = false;
    const size_t nv = m.positions.size() / 3, nf = m.indices.size() / 3;
    if (nv > UINT32_MAX - 1 || nf * 3 > UINT32_MAX) throw std::invalid_argument("graph too large");
    same_positions = same_positions && g->vertices == nv && g->positions.p;
    g->vertices = uint32_t(nv);
    g->faces = uint32_t(nf);
    g->diameter = b.radius * 2;
    if (!same_positions) {
        std::vector<float> pos(nv * 4, 0.f);
        for (size_t i = 0; i < nv; ++i) std::memcpy(&pos[4 * i], &m.positions[3 * i], 12);
        upload(g->positions, pos, stream_);
    }
    upload(g->indices, m.indices, stream_);
    build_graph_topology(*g, stream_);
    g->features.resize(nv * kFeatures * 4);
    FeatureInput in{};
    in.positions = g->positions.as<float4>();
    in.indices = g->indices.as<uint32_t>();
    in.face_offsets = g->face_offsets.as<uint32_t>();
    in.faces = g->faces_of.as<uint32_t>();
    in.offsets = g->offsets.as<uint32_t>();
    in.neighbors = g->neighbors.as<uint32_t>();
    in.boundary = g->boundary.as<uint32_t>();
    in.cx = b.center.x;
    in.cy = b.center.y;
    in.cz = b.center.z;
    in.diameter = g->diameter;
    in.vertices = g->vertices;
    in.out = g->features.as<float>();
    launch_features(in, stream_);
    g->features_bf16.resize(nv * kFeatures * 2);
    launch_convert_bf16(g->features.as<float>(), g->features_bf16.as<__bf16>(), nv * kFeatures, stream_);
    size_forward(*g);
    hip_check(hipGetLastError(), "features launch");
    hip_check(hipStreamSynchronize(stream_), "graph");
}

std::unique_ptr<Graph> Scorer::concat(const Graph* const* parts, size_t count) {
    if (!parts || !count) throw std::invalid_argument("no graphs to join");
    uint64_t nv = 0, na = 0, nf = 0, nc = 0;
    for (size_t k = 0; k < count; ++k) {
        if (!parts[k]) throw std::invalid_argument("null graph");
        nv += parts[k]->vertices;
        na += parts[k]->actions;
        nf += parts[k]->faces;
        nc += parts[k]->corners;
    }
    if (std::max({nv + 1, na, 3 * nf, nc}) > UINT32_MAX) throw std::invalid_argument("joined graph exceeds 32-bit ids");
    DeviceGuard guard(device_);
    auto g = std::make_unique<Graph>();
    g->vertices = uint32_t(nv);
    g->actions = uint32_t(na);
    g->faces = uint32_t(nf);
    g->corners = uint32_t(nc);
    g->positions.resize(nv * 16);
    g->indices.resize(nf * 12);
    g->offsets.resize((nv + 1) * 4);
    g->neighbors.resize(na * 4);
    g->src.resize(na * 4);
    g->reverse.resize(na * 4);
    g->face_offsets.resize((nv + 1) * 4);
    g->faces_of.resize(nc * 4);
    g->boundary.resize(nv * 4);
    g->features.resize(nv * kFeatures * 4);
    g->features_bf16.resize(nv * kFeatures * 2);
    uint32_t v0 = 0, a0 = 0, f0 = 0, c0 = 0;
    const auto copy = [&](DeviceBuffer& to, size_t at, const DeviceBuffer& from, size_t bytes) {
        if (bytes)
            hip_check(hipMemcpyAsync(static_cast<char*>(to.p) + at, from.p, bytes, hipMemcpyDeviceToDevice, stream_),
                      "join graphs");
    };
    for (size_t k = 0; k < count; ++k) {
        const Graph& p = *parts[k];
        copy(g->positions, size_t(v0) * 16, p.positions, size_t(p.vertices) * 16);
        copy(g->boundary, size_t(v0) * 4, p.boundary, size_t(p.vertices) * 4);
        copy(g->features, size_t(v0) * kFeatures * 4, p.features, size_t(p.vertices) * kFeatures * 4);
        copy(g->features_bf16, size_t(v0) * kFeatures * 2, p.features_bf16, size_t(p.vertices) * kFeatures * 2);
        launch_offset_copy(p.indices.as<uint32_t>(), g->indices.as<uint32_t>() + 3 * size_t(f0), 3 * size_t(p.faces),
                           v0, stream_);
        launch_offset_copy(p.offsets.as<uint32_t>(), g->offsets.as<uint32_t>() + v0, p.vertices, a0, stream_);
        launch_offset_copy(p.neighbors.as<uint32_t>(), g->neighbors.as<uint32_t>() + a0, p.actions, v0, stream_);
        launch_offset_copy(p.src.as<uint32_t>(), g->src.as<uint32_t>() + a0, p.actions, v0, stream_);
        launch_offset_copy(p.reverse.as<uint32_t>(), g->reverse.as<uint32_t>() + a0, p.actions, a0, stream_);
        launch_offset_copy(p.face_offsets.as<uint32_t>(), g->face_offsets.as<uint32_t>() + v0, p.vertices, c0, stream_);
        launch_offset_copy(p.faces_of.as<uint32_t>(), g->faces_of.as<uint32_t>() + c0, p.corners, f0, stream_);
        v0 += p.vertices;
        a0 += p.actions;
        f0 += p.faces;
        c0 += p.corners;
    }
    const uint32_t ends[2] = {a0, c0};
    hip_check(hipMemcpyAsync(g->offsets.as<uint32_t>() + v0, &ends[0], 4, hipMemcpyHostToDevice, stream_), "join");
    hip_check(hipMemcpyAsync(g->face_offsets.as<uint32_t>() + v0, &ends[1], 4, hipMemcpyHostToDevice, stream_), "join");
    size_forward(*g);
    hip_check(hipGetLastError(), "join launch");
    hip_check(hipStreamSynchronize(stream_), "join graphs");
    return g;
}

void Scorer::forward(Graph& g, Precision p) {
    DeviceGuard guard(device_);
    hip_check(hipEventRecord(begin_, stream_), "event");
    if (p == Precision::Fp32)
        launch_forward_fp32(g.features.as<float>(), g.view(), weights_.as<float>(), g.h_a.as<float>(), g.h_b.as<float>(),
                            g.z.as<float>(), g.outputs.as<float>(), stream_);
    else {
        g.zq.resize(size_t(g.vertices) * 2 * kHidden * 2);
        g.pq.resize(size_t(g.vertices) * 2 * kHidden * 2);
        launch_forward_bf16(g.features_bf16.as<__bf16>(), g.view(), weights_bf16_.as<__bf16>(), weights_.as<float>(),
                            g.hq_a.as<__bf16>(), g.hq_b.as<__bf16>(), g.zq.as<__bf16>(), g.pq.as<__bf16>(),
                            g.outputs.as<float>(), stream_);
    }
    hip_check(hipEventRecord(end_, stream_), "event");
    hip_check(hipGetLastError(), "forward launch");
    hip_check(hipStreamSynchronize(stream_), "forward");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, begin_, end_), "elapsed");
    forward_ms_ = ms;
    g.has_outputs = true;
}

void Scorer::download_outputs(const Graph& g, float* out) {
    if (!g.has_outputs) throw std::invalid_argument("forward() has not run on this graph");
    DeviceGuard guard(device_);
    hip_check(hipMemcpy(out, g.outputs.p, size_t(g.actions) * kOutputs * 4, hipMemcpyDeviceToHost), "download");
}

void Scorer::download_features(const Graph& g, float* out) {
    DeviceGuard guard(device_);
    hip_check(hipMemcpy(out, g.features.p, size_t(g.vertices) * kFeatures * 4, hipMemcpyDeviceToHost), "download");
}

void Scorer::rank(Graph& g, Ranking r, uint32_t* order) {
    if (r == Ranking::Shuffled) throw Unavailable("shuffled ranking is not implemented by the HIP scorer");
    if (r == Ranking::Learned && !g.has_outputs) throw std::invalid_argument("forward() has not run on this graph");
    DeviceGuard guard(device_);
    const uint32_t n = g.actions;
    g.keys.resize(size_t(n) * 4);
    g.order.resize(size_t(n) * 4);
    g.keys_out.resize(size_t(n) * 4);
    g.order_out.resize(size_t(n) * 4);
    ScoreInput in{};
    in.outputs = g.outputs.as<float>();
    in.positions = g.positions.as<float4>();
    in.indices = g.indices.as<uint32_t>();
    in.face_offsets = g.face_offsets.as<uint32_t>();
    in.faces = g.faces_of.as<uint32_t>();
    in.src = g.src.as<uint32_t>();
    in.dst = g.neighbors.as<uint32_t>();
    in.scale = g.diameter;
    in.actions = n;
    in.mode = r == Ranking::Learned ? 0u : r == Ranking::ShortestEdge ? 1u : r == Ranking::CurrentPlane ? 2u : 3u;
    in.keys = g.keys.as<float>();
    in.order = g.order.as<uint32_t>();
    launch_scores(in, stream_);
    const size_t temp = sort_workspace(n);
    g.sort_temp.resize(temp);
    hip_check(launch_sort(g.sort_temp.p, temp, g.keys.as<float>(), g.keys_out.as<float>(), g.order.as<uint32_t>(),
                          g.order_out.as<uint32_t>(), n, stream_),
              "radix sort");
    hip_check(hipGetLastError(), "rank launch");
    hip_check(hipMemcpyAsync(order, g.order_out.p, size_t(n) * 4, hipMemcpyDeviceToHost, stream_), "download order");
    hip_check(hipStreamSynchronize(stream_), "rank");
}

void Scorer::decode(Graph& g, const uint32_t* selection, size_t count, float* out) {
    if (!g.has_outputs) throw std::invalid_argument("forward() has not run on this graph");
    for (size_t i = 0; i < count; ++i)
        if (selection && selection[i] >= g.actions) throw std::invalid_argument("action outside graph");
    if (!selection && count > g.actions) throw std::invalid_argument("action count outside graph");
    if (!count) return;
    DeviceGuard guard(device_);
    DeviceBuffer sel, res, err;
    if (selection) {
        sel.resize(count * 4);
        hip_check(hipMemcpyAsync(sel.p, selection, count * 4, hipMemcpyHostToDevice, stream_), "upload selection");
    }
    res.resize(count * 3 * 4);
    err.resize(4);
    hip_check(hipMemsetAsync(err.p, 0, 4, stream_), "clear error");
    DecodeInput in{};
    in.outputs = g.outputs.as<float>();
    in.positions = g.positions.as<float4>();
    in.src = g.src.as<uint32_t>();
    in.dst = g.neighbors.as<uint32_t>();
    in.selection = selection ? sel.as<uint32_t>() : nullptr;
    in.count = uint32_t(count);
    in.out = res.as<float>();
    in.error = err.as<uint32_t>();
    launch_decode(in, stream_);
    hip_check(hipGetLastError(), "decode launch");
    uint32_t error = 0;
    hip_check(hipMemcpyAsync(&error, err.p, 4, hipMemcpyDeviceToHost, stream_), "download error");
    hip_check(hipMemcpyAsync(out, res.p, count * 3 * 4, hipMemcpyDeviceToHost, stream_), "download placements");
    hip_check(hipStreamSynchronize(stream_), "decode");
    if (error) throw std::runtime_error("non-finite placement prediction");
}

void Scorer::teacher(Graph& g, const uint32_t* selection, size_t count, float* positions, float* targets) {
    if (!selection && count) throw std::invalid_argument("null action list");
    for (size_t i = 0; i < count; ++i)
        if (selection[i] >= g.actions) throw std::invalid_argument("action outside graph");
    if (!count) return;
    DeviceGuard guard(device_);
    DeviceBuffer sel, pos, tgt;
    sel.resize(count * 4);
    hip_check(hipMemcpyAsync(sel.p, selection, count * 4, hipMemcpyHostToDevice, stream_), "upload selection");
    const size_t values = count * kTeacherCandidates * 3;
    pos.resize(values * 4);
    tgt.resize(values * 4);
    TeacherInput in{};
    in.positions = g.positions.as<float4>();
    in.indices = g.indices.as<uint32_t>();
    in.face_offsets = g.face_offsets.as<uint32_t>();
    in.faces = g.faces_of.as<uint32_t>();
    in.src = g.src.as<uint32_t>();
    in.dst = g.neighbors.as<uint32_t>();
    in.selection = sel.as<uint32_t>();
    in.count = uint32_t(count);
    in.positions_out = pos.as<float>();
    in.targets_out = tgt.as<float>();
    launch_teacher(in, stream_);
    hip_check(hipGetLastError(), "teacher launch");
    hip_check(hipMemcpyAsync(positions, pos.p, values * 4, hipMemcpyDeviceToHost, stream_), "download positions");
    hip_check(hipMemcpyAsync(targets, tgt.p, values * 4, hipMemcpyDeviceToHost, stream_), "download targets");
    hip_check(hipStreamSynchronize(stream_), "teacher");
}

}  // namespace blitz_hip
