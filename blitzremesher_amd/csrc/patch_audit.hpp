// Bound reference of the patch audit (patch_audit.hip).
#pragma once
#include <memory>
#include <vector>

#include "audit.hpp"
#include "device_buffer.hpp"

namespace blitz_hip {

struct Binding {
    Audit* owner = nullptr;
    HostMesh host;                      // reference mesh (host copy)
    std::unique_ptr<DeviceMesh> mesh;   // reference mesh in HBM
    Bounds bounds{};
    Settings settings;                  // raster-defining fields of the bind
    std::vector<Camera> cams;
    uint8_t ss = 0;
    uint32_t size = 0, tiles = 0, views = 0;
    bool limited = false;               // raster over the sample cap: patches take the full path
    bool supported = true;              // false with per-face materials (patches unavailable)
    bool added_two_sided = false;       // sidedness of added faces (material 0)
    std::vector<uint8_t> clipped;       // per view: reference clipped or behind the camera
    std::vector<unsigned long long> count_a;  // per view: samples covered by the reference
    DeviceBuffer xy, z, masks, device_count_a, tasks, bin_offsets, bin_faces;
};

}  // namespace blitz_hip
