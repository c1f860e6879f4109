// Neural model files (BLZNET01/02/03): the reference's load_weights /
// save_weights (src/neural/model_io.cpp:61-153) for every architecture the
// reference writes: 1 (the graph encoder), 2 (action policy), 3 and 4
// (placement policies).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace blitz_hip {

struct ModelData {
    uint32_t architecture = 1, hidden_width = 64;
    uint32_t use = 0;  // 0 unrestricted, 1 endpoint-reuse only (architecture 4)
    std::vector<float> values;
    std::string provenance;
    std::string sha256;  // of the whole file (load only)
};

// Weight count of an architecture and hidden width; 0 when the pair is not
// one the format admits.
size_t model_weight_count(uint32_t architecture, uint32_t hidden_width);
ModelData load_model(const std::string& path);
void save_model(const std::string& path, const ModelData& model);

}  // namespace blitz_hip
