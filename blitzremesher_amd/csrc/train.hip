// One training step of the edge-graph scorer: forward with every activation
// kept (FP32 compensated or BF16 MFMA), loss over labelled actions, FP32
// backward, global-norm clip and AdamW on FP32 master weights, then a fresh
// BF16 copy of the weights. The counterpart of the reference's update step
// (training/fused.cu:95-120 forward/backward, training/update_cuda.cu:159-393
// loss and optimizer), for the build's network.
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "scorer.hpp"
#include "scorer_kernels.hpp"
#include "train_kernels.hpp"
#include <algorithm>

namespace blitz_hip {
namespace {
template <class T> T* grow(DeviceBuffer& b, size_t count) {
    b.resize(std::max<size_t>(count, 1) * sizeof(T));
    return b.as<T>();
}
}  // namespace

void Scorer::reset_optimizer() {
    DeviceGuard guard(device_);
    grow<float>(grads_, kWeightCount);
    grow<float>(mean_, kWeightCount);
    grow<float>(variance_, kWeightCount);
    hip_check(hipMemsetAsync(grads_.p, 0, size_t(kWeightCount) * 4, stream_), "clear");
    hip_check(hipMemsetAsync(mean_.p, 0, size_t(kWeightCount) * 4, stream_), "clear");
    hip_check(hipMemsetAsync(variance_.p, 0, size_t(kWeightCount) * 4, stream_), "clear");
    hip_check(hipStreamSynchronize(stream_), "reset optimizer");
    step_ = 0;
}

void Scorer::download_gradients(float* out) {
    DeviceGuard guard(device_);
    hip_check(hipMemcpy(out, grads_.p, size_t(kWeightCount) * 4, hipMemcpyDeviceToHost), "download gradients");
}

void Scorer::download_weights(float* out) {
    DeviceGuard guard(device_);
    hip_check(hipMemcpy(out, weights_.p, size_t(kWeightCount) * 4, hipMemcpyDeviceToHost), "download weights");
}

TrainResult Scorer::train_step(Graph& g, const uint32_t* actions, const uint8_t* labels, const float* targets,
                               size_t count, const TrainSettings& s) {
    if (!actions || !labels || !count || count > g.actions) throw std::invalid_argument("invalid training batch");
    if (!s.pool || s.pool > 256) throw std::invalid_argument("ranking pool must hold 1..256 entries");
    if (s.precision > 1 || !std::isfinite(s.margin) || s.margin < 0 || !std::isfinite(s.penalty) || s.penalty < 0 ||
        !std::isfinite(s.lr) || s.lr <= 0 || s.lr > 1 || !std::isfinite(s.decay) || s.decay < 0 ||
        !std::isfinite(s.max_norm) || s.max_norm <= 0 || !std::isfinite(s.placement_weight) ||
        s.placement_weight < 0)
        throw std::invalid_argument("invalid training settings");
    // Loss normalisers from the labels (host).
    LossInput li{};
    li.count = uint32_t(count);
    li.pool = s.pool;
    li.margin = s.margin;
    li.penalty = s.penalty;
    li.placement_weight = s.placement_weight;
    for (size_t i = 0; i < count; ++i)
        if (actions[i] >= g.actions) throw std::invalid_argument("action outside graph");
    for (size_t p0 = 0; p0 < count; p0 += s.pool) {
        uint32_t valid = 0, preferred = 0;
        for (size_t j = p0; j < std::min(count, p0 + s.pool); ++j) {
            valid += (labels[j] & 1) != 0;
            preferred += (labels[j] & 3) == 3;
            li.placed += (labels[j] & 4) != 0;
        }
        li.pairs += preferred * (valid - preferred);
        li.valid += valid;
    }
    if (li.placed && !targets) throw std::invalid_argument("placement labels without targets");
    DeviceGuard guard(device_);
    const uint32_t nv = g.vertices, nb = uint32_t(count);
    const hipStream_t st = stream_;
    const GraphView gv = g.view();
    uint32_t* d_actions = grow<uint32_t>(g.batch_actions, nb);
    uint32_t* d_src = grow<uint32_t>(g.batch_src, nb);
    uint32_t* d_dst = grow<uint32_t>(g.batch_dst, nb);
    uint8_t* d_labels = grow<uint8_t>(g.labels, nb);
    float* d_targets = grow<float>(g.targets, size_t(nb) * 3);
    hip_check(hipMemcpyAsync(d_actions, actions, size_t(nb) * 4, hipMemcpyHostToDevice, st), "upload batch");
    hip_check(hipMemcpyAsync(d_labels, labels, nb, hipMemcpyHostToDevice, st), "upload labels");
    if (targets) hip_check(hipMemcpyAsync(d_targets, targets, size_t(nb) * 12, hipMemcpyHostToDevice, st), "targets");
    launch_batch_endpoints(d_actions, gv, d_src, d_dst, nb, st);
    GraphView batch = gv;
    batch.src = d_src;
    batch.neighbors = d_dst;
    batch.actions = nb;
    const float* w = weights_.as<float>();
    float* h[4];
    h[0] = g.features.as<float>();
    for (int l = 1; l < 4; ++l) h[l] = grow<float>(g.t_h[l], size_t(nv) * kHidden);
    float* z[3];
    z[0] = grow<float>(g.t_z[0], size_t(nv) * 2 * kFeatures);
    z[1] = grow<float>(g.t_z[1], size_t(nv) * 2 * kHidden);
    z[2] = grow<float>(g.t_z[2], size_t(nv) * 2 * kHidden);
    float* hidden = grow<float>(g.hidden, size_t(nb) * kHidden);
    float* out = grow<float>(g.batch_out, size_t(nb) * kOutputs);
    hip_check(hipEventRecord(begin_, st), "event");
    // ---- forward, every activation kept
    if (s.precision == 0) {
        uint32_t width = kFeatures;
        for (int l = 0; l < 3; ++l) {
            launch_aggregate_fp32(h[l], gv, width, z[l], st);
            launch_compensated(z[l], nullptr, nullptr, w + layer_offset(l), w + layer_offset(l) + kHidden * kLayerIn[l],
                               h[l + 1], nv, kLayerIn[l], kHidden, true, st);
            width = kHidden;
        }
        launch_compensated(h[3], d_src, d_dst, w + layer_offset(3), w + layer_offset(3) + kHidden * kEdgeIn, hidden,
                           nb, kEdgeIn, kHidden, true, st);
        launch_compensated(hidden, nullptr, nullptr, w + layer_offset(4), w + layer_offset(4) + kOutputs * kHidden,
                           out, nb, kHidden, kOutputs, false, st);
    } else {
        __bf16* hq[4];
        hq[0] = g.features_bf16.as<__bf16>();
        for (int l = 1; l < 4; ++l) hq[l] = grow<__bf16>(g.t_hq[l], size_t(nv) * kHidden);
        __bf16* zq = grow<__bf16>(g.zq, size_t(nv) * 2 * kHidden);
        __bf16* pq = grow<__bf16>(g.pq, size_t(nv) * 2 * kHidden);
        for (int l = 0; l < 3; ++l) launch_mp_bf16(l, hq[l], gv, weights_bf16_.as<__bf16>(), w, hq[l + 1], zq, st);
        launch_edge_bf16(hq[3], batch, weights_bf16_.as<__bf16>(), w, out, hidden, pq, st);
        // FP32 views of the BF16 activations for the backward pass.
        for (int l = 1; l < 4; ++l) launch_bf16_to_fp32(hq[l], h[l], size_t(nv) * kHidden, st);
        uint32_t width = kFeatures;
        for (int l = 0; l < 3; ++l) {
            launch_aggregate_fp32(h[l], gv, width, z[l], st);
            width = kHidden;
        }
    }
    // ---- loss
    float* accum = grow<float>(g.accum, 2);  // [0] loss, [1] squared gradient norm
    // Reductions without float atomics: per-wave loss and norm partials,
    // per-slice weight-gradient tiles, per-vertex gathers of the edge
    // gradients, each added in a fixed order (one workspace, sized here
    // before any launch).
    size_t gemm_work = 0;
    for (int l = 0; l < 3; ++l) gemm_work = std::max(gemm_work, gemm_at_b_workspace(nv, kHidden, kLayerIn[l]));
    gemm_work = std::max({gemm_work, gemm_at_b_workspace(nb, kOutputs, kHidden),
                          gemm_at_b_workspace(nb, kHidden, kEdgeIn)});
    const size_t parts = std::max<size_t>(loss_parts(li), 4096);
    const size_t scan_bytes = gather_edge_scan_bytes(nv);
    float* work = grow<float>(g.reduce_work, gemm_work + parts + (scan_bytes + 3) / 4);
    float* part = work + gemm_work;
    void* scan_temp = part + parts;
    uint32_t* buckets = grow<uint32_t>(g.edge_buckets, 3 * size_t(nv) + 2 + 2 * size_t(nb));
    hip_check(hipMemsetAsync(grads_.p, 0, size_t(kWeightCount) * 4, st), "clear gradients");
    float* dy = grow<float>(g.dy, size_t(nb) * kOutputs);
    launch_loss(out, d_labels, d_targets, dy, part, accum, li, st);
    // ---- backward: head and edge layer
    float* gr = grads_.as<float>();
    float* t_e = grow<float>(g.t_e, size_t(nb) * kHidden);
    float* z_e = grow<float>(g.z_e, size_t(nb) * kEdgeIn);
    float* dz_e = grow<float>(g.dz_e, size_t(nb) * kEdgeIn);
    launch_gemm_at_b(dy, kOutputs, nullptr, hidden, kHidden, gr + layer_offset(4),
                     gr + layer_offset(4) + kOutputs * kHidden, work, nb, kOutputs, kHidden, st);
    launch_gemm_a_w(dy, nullptr, w + layer_offset(4), t_e, nb, kOutputs, kHidden, st);
    // z_e = [h3(src) ; h3(dst)] for the edge weight gradient.
    launch_gather_edge(h[3], d_src, d_dst, z_e, nb, st);
    launch_gemm_at_b(t_e, kHidden, hidden, z_e, kEdgeIn, gr + layer_offset(3), gr + layer_offset(3) + kHidden * kEdgeIn,
                     work, nb, kHidden, kEdgeIn, st);
    launch_gemm_a_w(t_e, hidden, w + layer_offset(3), dz_e, nb, kHidden, kEdgeIn, st);
    float* dh = grow<float>(g.dh_a, size_t(nv) * kHidden);
    float* dh_next = grow<float>(g.dh_b, size_t(nv) * kHidden);
    launch_gather_edge_backward(dz_e, d_src, d_dst, nb, nv, buckets, buckets + nv + 1, buckets + 2 * size_t(nv) + 2,
                                buckets + 3 * size_t(nv) + 2, scan_temp, scan_bytes, dh, st);
    // ---- backward: message-passing layers
    float* t_v = grow<float>(g.t_v, size_t(nv) * 2 * kHidden);
    for (int l = 2; l >= 0; --l) {
        const uint32_t in = kLayerIn[l];
        launch_gemm_at_b(dh, kHidden, h[l + 1], z[l], in, gr + layer_offset(l), gr + layer_offset(l) + kHidden * in, work,
                         nv, kHidden, in, st);
        if (l > 0) {
            launch_gemm_a_w(dh, h[l + 1], w + layer_offset(l), t_v, nv, kHidden, in, st);
            launch_aggregate_backward(t_v, gv, in / 2, dh_next, st);
            std::swap(dh, dh_next);
        }
    }
    // ---- clip and AdamW (training/update_cuda.cu:358-393)
    launch_norm(gr, kWeightCount, part, accum + 1, st);
    float host[2] = {0, 0};
    hip_check(hipMemcpyAsync(host, accum, 8, hipMemcpyDeviceToHost, st), "download loss");
    hip_check(hipStreamSynchronize(st), "train step");
    TrainResult r;
    r.loss = host[0];
    r.grad_norm = std::sqrt(host[1]);
    if (std::isfinite(r.grad_norm) && std::isfinite(r.loss)) {
        AdamStep a{};
        a.clip = std::fmin(1.f, s.max_norm / (r.grad_norm + 1e-6f));
        a.lr = s.lr;
        a.decay = s.decay;
        a.correction1 = float(double(s.lr) / (1 - std::pow(.9, double(step_) + 1)));
        a.correction2 = float(std::sqrt(1 - std::pow(.999, double(step_) + 1)));
        launch_adamw(weights_.as<float>(), gr, mean_.as<float>(), variance_.as<float>(), kWeightCount, a, st);
        refresh_bf16();
        ++step_;
        r.applied = true;
    }
    hip_check(hipEventRecord(end_, st), "event");
    hip_check(hipGetLastError(), "train launch");
    hip_check(hipStreamSynchronize(st), "train step");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, begin_, end_), "elapsed");
    step_ms_ = ms;
    return r;
}

}  // namespace blitz_hip
