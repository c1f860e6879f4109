// BLZNET model files (src/neural/model_io.cpp:61-153). Layout, all integers
// little-endian u32:
//   "BLZNET01" | architecture | weight count | provenance bytes
//   "BLZNET02" adds the hidden width, "BLZNET03" the hidden width and the
//   use scope (endpoint-reuse only, architecture 4 alone);
//   provenance text; float32 weights; 64 lower-case hex digits of the
//   SHA-256 of everything before them.
// A file is written as BLZNET01 unless its width is not 64 or its scope is
// restricted, so re-saving a loaded model gives the same bytes.
#include "model_io.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <stdexcept>

#include "policy_shapes.hpp"
#include "sha256.hpp"

namespace blitz_hip {
namespace {

constexpr uint32_t kEncoderSchema = 1, kConditionedPlacementSchema = 4, kEndpointReuseOnly = 1;
constexpr size_t kMaxProvenance = 65536;

// The graph encoder's layers (src/neural/internal.hpp:59-68): inputs
// 48, 128, 128, 72, 64; outputs 64, 64, 64, 64, 4.
size_t encoder_weight_count() {
    const size_t in[5] = {48, 128, 128, 72, 64}, out[5] = {64, 64, 64, 64, 4};
    size_t n = 0;
    for (int i = 0; i < 5; ++i) n += out[i] * (in[i] + 1);
    return n;
}

uint32_t get_u32(const std::vector<unsigned char>& a, size_t at) {
    return uint32_t(a[at]) | uint32_t(a[at + 1]) << 8 | uint32_t(a[at + 2]) << 16 | uint32_t(a[at + 3]) << 24;
}

void put_u32(std::vector<unsigned char>& out, uint32_t x) {
    for (int i = 0; i < 4; ++i) out.push_back(static_cast<unsigned char>(x >> (8 * i)));
}

}  // namespace

size_t model_weight_count(uint32_t architecture, uint32_t hidden_width) {
    if (architecture == kEncoderSchema) return hidden_width == 64 ? encoder_weight_count() : 0;
    return policy_weights(architecture, hidden_width);
}

ModelData load_model(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::invalid_argument("cannot open neural model: " + path);
    const std::streamoff size = f.tellg();
    const std::streamoff largest =
        std::streamoff(92 + model_weight_count(kConditionedPlacementSchema, 256) * 4 + kMaxProvenance);
    if (size < 84 || size > largest) throw std::invalid_argument("neural model size invalid");
    std::vector<unsigned char> data(static_cast<size_t>(size));
    f.seekg(0);
    if (!f.read(reinterpret_cast<char*>(data.data()), size)) throw std::invalid_argument("truncated neural model");
    const bool scoped = !std::memcmp(data.data(), "BLZNET03", 8);
    const bool shaped = scoped || !std::memcmp(data.data(), "BLZNET02", 8);
    const bool plain = !std::memcmp(data.data(), "BLZNET01", 8);
    const uint32_t width = shaped ? get_u32(data, 20) : 64;
    const size_t header = scoped ? 28 : shaped ? 24 : 20;
    const uint32_t architecture = get_u32(data, 8);
    const uint32_t use = scoped ? get_u32(data, 24) : 0;
    if (scoped && (use != kEndpointReuseOnly || architecture != kConditionedPlacementSchema))
        throw std::invalid_argument("incompatible neural model use scope");
    const size_t count = model_weight_count(architecture, width);
    if ((!shaped && !plain) || !count || get_u32(data, 12) != count)
        throw std::invalid_argument("incompatible neural model schema/architecture");
    const size_t provenance = get_u32(data, 16);
    if (provenance > kMaxProvenance || data.size() != header + provenance + count * 4 + 64)
        throw std::invalid_argument("invalid neural model payload");
    const std::string checksum = sha256_hex(data.data(), data.size() - 64);
    if (std::memcmp(checksum.data(), data.data() + data.size() - 64, 64))
        throw std::invalid_argument("neural model checksum mismatch");
    ModelData m;
    m.architecture = architecture;
    m.hidden_width = width;
    m.use = use;
    m.provenance.assign(reinterpret_cast<const char*>(data.data() + header), provenance);
    m.values.resize(count);
    for (size_t i = 0; i < count; ++i) {
        const uint32_t bits = get_u32(data, header + provenance + 4 * i);
        std::memcpy(&m.values[i], &bits, 4);
        if (!std::isfinite(m.values[i])) throw std::invalid_argument("nonfinite model weights");
    }
    m.sha256 = sha256_hex(data.data(), data.size());
    return m;
}

void save_model(const std::string& path, const ModelData& m) {
    const size_t count = model_weight_count(m.architecture, m.hidden_width);
    if (!count || m.values.size() != count || m.provenance.size() > kMaxProvenance)
        throw std::invalid_argument("invalid model weight dimensions or provenance");
    if (m.use > kEndpointReuseOnly || (m.use == kEndpointReuseOnly && m.architecture != kConditionedPlacementSchema))
        throw std::invalid_argument("incompatible neural model use scope");
    std::vector<unsigned char> data(reinterpret_cast<const unsigned char*>("BLZNET01"),
                                    reinterpret_cast<const unsigned char*>("BLZNET01") + 8);
    put_u32(data, m.architecture);
    put_u32(data, uint32_t(count));
    put_u32(data, uint32_t(m.provenance.size()));
    if (m.use || m.hidden_width != 64) {
        data[7] = m.use ? '3' : '2';
        put_u32(data, m.hidden_width);
        if (m.use) put_u32(data, m.use);
    }
    data.insert(data.end(), m.provenance.begin(), m.provenance.end());
    for (float v : m.values) {
        if (!std::isfinite(v)) throw std::invalid_argument("nonfinite export");
        uint32_t bits;
        std::memcpy(&bits, &v, 4);
        put_u32(data, bits);
    }
    const std::string sum = sha256_hex(data.data(), data.size());
    data.insert(data.end(), sum.begin(), sum.end());
    // Written beside the target and renamed over it, so a reader never sees
    // a partial file.
    static uint64_t serial = 0;
    const std::string temp = path + ".tmp." +
                             std::to_string(std::chrono::steady_clock::now().time_since_epoch().count()) + "." +
                             std::to_string(__atomic_fetch_add(&serial, 1, __ATOMIC_RELAXED));
    try {
        std::ofstream out(temp, std::ios::binary | std::ios::trunc);
        out.write(reinterpret_cast<const char*>(data.data()), std::streamsize(data.size()));
        out.close();
        if (!out) throw std::runtime_error("model write failed");
        std::filesystem::rename(temp, path);
    } catch (...) {
        std::error_code ec;
        std::filesystem::remove(temp, ec);
        throw;
    }
}

}  // namespace blitz_hip
