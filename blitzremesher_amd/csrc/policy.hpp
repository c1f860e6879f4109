// The action and placement policies of the reference's GPU inference
// (ActionCuda, src/neural/action.hpp:70-87, src/neural/inference.cu:183-290;
// shapes src/neural/placement.hpp:8-31): three compensated FP32 layers with
// ReLU after the first two, weights [out, in] then bias per layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "device_buffer.hpp"

namespace blitz_hip {

// Architecture 2: 80 action features -> 3 outputs; 3 and 4: 128 placement
// features -> 12 outputs (4 carries preserve_uv in feature 79); 0 otherwise.
uint32_t policy_inputs(uint32_t architecture);
uint32_t policy_outputs(uint32_t architecture);
// Weight count for hidden width 64, 128 or 256; 0 for other shapes.
size_t policy_weights(uint32_t architecture, uint32_t hidden);

class Policy {
  public:
    Policy(int device, uint32_t architecture, uint32_t hidden, const float* weights, size_t count);
    ~Policy();
    Policy(const Policy&) = delete;
    Policy& operator=(const Policy&) = delete;

    uint32_t architecture() const { return architecture_; }
    uint32_t inputs() const { return inputs_; }
    uint32_t outputs() const { return outputs_; }
    int device() const { return device_; }
    // ActionCuda::predict_device: `rows` rows of inputs() floats on the
    // device to rows of outputs() floats, queued on `stream`. The layers'
    // activations live in this object: calls on one policy are ordered on
    // one stream.
    void predict_device(const float* input, float* output, uint32_t rows, hipStream_t stream);
    // ActionCuda::predict: host rows; inputs and outputs must be finite.
    std::vector<float> predict(const float* input, size_t rows);

  private:
    int device_;
    uint32_t architecture_, hidden_, inputs_, outputs_;
    DeviceBuffer weights_, a_, b_, input_, output_;
    hipStream_t stream_{};
};

}  // namespace blitz_hip
