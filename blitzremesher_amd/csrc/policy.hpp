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
#include "policy_shapes.hpp"

namespace blitz_hip {

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
