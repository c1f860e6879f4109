// Shapes of the policy architectures (src/neural/placement.hpp:8-31).
#pragma once
#include <cstddef>
#include <cstdint>

namespace blitz_hip {

// Architecture 2: 80 action features -> 3 outputs; 3 and 4: 128 placement
// features -> 12 outputs (4 carries preserve_uv in feature 79); 0 otherwise.
uint32_t policy_inputs(uint32_t architecture);
uint32_t policy_outputs(uint32_t architecture);
// Weight count for hidden width 64, 128 or 256; 0 for other shapes.
size_t policy_weights(uint32_t architecture, uint32_t hidden);

}  // namespace blitz_hip
