// Policy inference: ActionCuda::predict / predict_device
// (src/neural/inference.cu:236-290) on the compensated linear layer of the
// scorer (the reference's compensated_linear_kernel arithmetic), rows taken
// in tiles so the activations stay bounded.
#include "policy.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "audit.hpp"
#include "scorer_kernels.hpp"

namespace blitz_hip {

namespace {
constexpr uint32_t kActionSchema = 2, kPlacementSchema = 3, kConditionedPlacementSchema = 4;
constexpr uint32_t kTileRows = 65536;  // rows per layer pass (ActionCuda's batch bound)
}  // namespace

uint32_t policy_inputs(uint32_t architecture) {
    return architecture == kActionSchema ? 80u
           : (architecture == kPlacementSchema || architecture == kConditionedPlacementSchema) ? 128u
                                                                                               : 0u;
}

uint32_t policy_outputs(uint32_t architecture) {
    return architecture == kActionSchema ? 3u
           : (architecture == kPlacementSchema || architecture == kConditionedPlacementSchema) ? 12u
                                                                                               : 0u;
}

size_t policy_weights(uint32_t architecture, uint32_t hidden) {
    const size_t in = policy_inputs(architecture), out = policy_outputs(architecture);
    if (!in || (hidden != 64 && hidden != 128 && hidden != 256)) return 0;
    return size_t(hidden) * (in + 1) + size_t(hidden) * (hidden + 1) + out * (hidden + 1);
}

Policy::Policy(int device, uint32_t architecture, uint32_t hidden, const float* weights, size_t count)
    : device_(device), architecture_(architecture), hidden_(hidden), inputs_(policy_inputs(architecture)),
      outputs_(policy_outputs(architecture)) {
    const size_t need = policy_weights(architecture, hidden);
    if (!need || count != need || !weights) throw std::invalid_argument("invalid action model or batch dimensions");
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(weights[i])) throw std::invalid_argument("nonfinite action weights");
    DeviceGuard guard(device_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "policy stream");
    weights_.resize(count * 4);
    hip_check(hipMemcpy(weights_.p, weights, count * 4, hipMemcpyHostToDevice), "upload policy weights");
}

Policy::~Policy() {
    DeviceGuard guard(device_);
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
}

void Policy::predict_device(const float* input, float* output, uint32_t rows, hipStream_t stream) {
    if (!rows) return;
    const uint32_t tile = std::min(rows, kTileRows);
    a_.resize(size_t(tile) * hidden_ * 4);
    b_.resize(size_t(tile) * hidden_ * 4);
    const float* w = weights_.as<float>();
    for (uint32_t row = 0; row < rows; row += kTileRows) {
        const uint32_t n = std::min(kTileRows, rows - row);
        const float* in = input + size_t(row) * inputs_;
        size_t at = 0;
        for (uint32_t layer = 0; layer < 3; ++layer) {
            const uint32_t columns = layer ? hidden_ : inputs_, channels = layer == 2 ? outputs_ : hidden_;
            float* target = layer == 0 ? a_.as<float>() : layer == 1 ? b_.as<float>() : output + size_t(row) * outputs_;
            launch_compensated(in, nullptr, nullptr, w + at, w + at + size_t(columns) * channels, target, n, columns,
                               channels, layer < 2, stream);
            in = target;
            at += size_t(channels) * (columns + 1);
        }
    }
    hip_check(hipGetLastError(), "policy launch");
}

std::vector<float> Policy::predict(const float* input, size_t rows) {
    for (size_t i = 0; i < rows * inputs_; ++i)
        if (!std::isfinite(input[i])) throw std::invalid_argument("nonfinite action input");
    if (rows > UINT32_MAX) throw std::invalid_argument("too many policy rows");
    std::vector<float> out(rows * outputs_);
    if (!rows) return out;
    DeviceGuard guard(device_);
    input_.resize(rows * inputs_ * 4);
    output_.resize(rows * outputs_ * 4);
    hip_check(hipMemcpyAsync(input_.p, input, rows * inputs_ * 4, hipMemcpyHostToDevice, stream_), "upload inputs");
    predict_device(input_.as<float>(), output_.as<float>(), uint32_t(rows), stream_);
    hip_check(hipMemcpyAsync(out.data(), output_.p, out.size() * 4, hipMemcpyDeviceToHost, stream_), "download outputs");
    hip_check(hipStreamSynchronize(stream_), "policy");
    for (float v : out)
        if (!std::isfinite(v)) throw std::runtime_error("nonfinite action prediction");
    return out;
}

}  // namespace blitz_hip
