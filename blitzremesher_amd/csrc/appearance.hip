// Host side of the normals and attributes audit profiles: per-tile face bins
// of every drawn raster job, the attribute raster launch, the row-major
// attribute raster of one view (src/evaluate.cpp:125-189) and the
// top-left overlap measure (src/neural/evaluate.cu:2665-2685).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "appearance_kernels.hpp"
#include "audit.hpp"
#include "audit_kernels.hpp"
#include "patch_kernels.hpp"

namespace blitz_hip {
namespace {
enum : size_t {
    kSlotDrawn = 20,
    kSlotBinCounts,
    kSlotBinOffsets,
    kSlotBinScan,
    kSlotBinFaces,
    kSlotAttrTasks,
    kSlotPixelNormals,
    kSlotPixelColors,
    kSlotPixelDepth,
    kSlotPixelMask,
};
constexpr uint32_t kMaxGridY = 65535;
constexpr uint64_t kMaxSamples = 64000000;  // src/evaluate.cpp:129
constexpr uint32_t kNone = 0xFFFFFFFFu;
}  // namespace

void Audit::build_bins(const std::vector<RasterTask>& drawn, const unsigned long long*& offsets_out,
                       const uint32_t*& faces_out) {
    const size_t n = drawn.size();
    const uint32_t tiles = drawn[0].tiles;
    const size_t words = size_t(tiles) * tiles;
    uint32_t max_faces = 0;
    for (const auto& t : drawn) max_faces = std::max(max_faces, t.faces);
    // Per (job, tile) face bins: count, exclusive scan, fill.
    RasterTask* dtasks = scratch_as<RasterTask>(kSlotDrawn, n);
    hip_check(hipMemcpyAsync(dtasks, drawn.data(), n * sizeof(RasterTask), hipMemcpyHostToDevice, stream_),
              "upload drawn tasks");
    const size_t cells = n * words + 1;
    auto* counts = scratch_as<unsigned long long>(kSlotBinCounts, cells);
    auto* offsets = scratch_as<unsigned long long>(kSlotBinOffsets, cells);
    launch_fill(reinterpret_cast<uint64_t*>(counts), cells, 0, stream_);
    for (size_t j0 = 0; j0 < n; j0 += kMaxGridY)
        launch_bin_count(dtasks + j0, uint32_t(std::min<size_t>(kMaxGridY, n - j0)), max_faces, counts + j0 * words,
                         stream_);
    const size_t temp_bytes = scan_workspace(cells);
    void* temp = scratch(kSlotBinScan, std::max<size_t>(temp_bytes, 16));
    launch_exclusive_scan(temp, temp_bytes, counts, offsets, cells, stream_);
    unsigned long long entries = 0;
    hip_check(hipMemcpyAsync(&entries, offsets + cells - 1, 8, hipMemcpyDeviceToHost, stream_), "download bin total");
    hip_check(hipStreamSynchronize(stream_), "bin count");
    uint32_t* faces = scratch_as<uint32_t>(kSlotBinFaces, std::max<size_t>(1, entries));
    hip_check(hipMemcpyAsync(counts, offsets, cells * 8, hipMemcpyDeviceToDevice, stream_), "bin cursor");
    for (size_t j0 = 0; j0 < n; j0 += kMaxGridY)
        launch_bin_fill(dtasks + j0, uint32_t(std::min<size_t>(kMaxGridY, n - j0)), max_faces, counts + j0 * words,
                        faces, stream_);
    offsets_out = offsets;
    faces_out = faces;
}

void Audit::attribute_rasters(const std::vector<RasterJob>& jobs, const std::vector<RasterTask>& drawn,
                              const std::vector<const double*>& depth, const std::vector<Camera>& cams,
                              const Bounds& b, float4* normal_out, float4* color_out, float* depth_out) {
    const size_t n = jobs.size();
    if (!n) return;
    const uint32_t tiles = drawn[0].tiles;
    const size_t words = size_t(tiles) * tiles, per = words * 64;
    const unsigned long long* offsets = nullptr;
    const uint32_t* faces = nullptr;
    build_bins(drawn, offsets, faces);
    std::vector<AttrTask> tasks(n);
    for (size_t j = 0; j < n; ++j) {
        const DeviceMesh& m = *jobs[j].mesh;
        AttrTask& t = tasks[j];
        t = {};
        t.raster = drawn[j];
        t.depth_z = depth[j];
        t.positions = m.positions;
        t.normals = m.normals;
        t.colors = m.colors;
        t.materials = m.materials;
        t.bin_offsets = offsets + j * words;
        t.bin_faces = faces;
        t.normal_out = normal_out + j * per;
        t.color_out = color_out ? color_out + j * per : nullptr;
        t.depth_out = depth_out ? depth_out + j * per : nullptr;
        t.diameter = b.radius * 2;  // src/evaluate.cpp:135
        t.perspective = cams[jobs[j].view].perspective;
    }
    AttrTask* dattr = scratch_as<AttrTask>(kSlotAttrTasks, n);
    hip_check(hipMemcpyAsync(dattr, tasks.data(), n * sizeof(AttrTask), hipMemcpyHostToDevice, stream_),
              "upload attribute tasks");
    for (size_t j0 = 0; j0 < n; j0 += kMaxGridY)
        launch_attribute_raster(dattr + j0, uint32_t(std::min<size_t>(kMaxGridY, n - j0)), tiles, stream_);
    hip_check(hipGetLastError(), "attribute raster launch");
}

// overlap_cuda (src/neural/evaluate.cu:2665-2685): every view is drawn at one
// sample per pixel on an even raster side (src/neural/evaluate.cu:599); a view
// with a face behind the camera or out of frame makes the result infinite.
double Audit::overlap(const DeviceMesh& m, const Bounds& b, double screen, uint16_t ortho, uint16_t persp,
                      uint32_t seed) {
    if (!std::isfinite(screen) || screen <= 0 || screen > 16384) throw std::invalid_argument("invalid raster extent");
    const auto cams = cameras(b, screen, ortho, persp, seed);
    if (cams.empty()) throw std::invalid_argument("overlap requires views");
    const uint32_t size = (uint32_t(std::ceil(screen + 8)) + 1) & ~1u;
    if (uint64_t(size) * size > kMaxSamples) throw std::invalid_argument("raster exceeds per-view sample cap");
    DeviceGuard guard(device_);
    const uint32_t tiles = (size + kTile - 1) / kTile;
    const size_t words = size_t(tiles) * tiles;
    const size_t n = cams.size();
    uint64_t* masks = scratch_as<uint64_t>(kSlotPixelMask, n * words);
    launch_fill(masks, n * words, 0, stream_);
    std::vector<RasterJob> jobs(n);
    for (size_t v = 0; v < n; ++v) jobs[v] = {&m, uint32_t(v), masks + v * words};
    std::vector<uint32_t> st;
    std::vector<RasterTask> drawn;
    std::vector<const double*> zs;
    rasterize(jobs, cams, b, 1, size, false, st, &drawn, &zs);
    for (size_t v = 0; v < n; ++v)
        if (st[2 * v] != kNone || st[2 * v + 1]) return std::numeric_limits<double>::infinity();
    // Per-tile face bins of every view, then the per-sample counts.
    std::vector<AttrTask> tasks;
    const unsigned long long* offsets = nullptr;
    const uint32_t* faces = nullptr;
    build_bins(drawn, offsets, faces);
    tasks.resize(n);
    for (size_t v = 0; v < n; ++v) {
        AttrTask& t = tasks[v];
        t = {};
        t.raster = drawn[v];
        t.bin_offsets = offsets + v * words;
        t.bin_faces = faces;
    }
    AttrTask* dattr = scratch_as<AttrTask>(kSlotAttrTasks, n);
    hip_check(hipMemcpyAsync(dattr, tasks.data(), n * sizeof(AttrTask), hipMemcpyHostToDevice, stream_),
              "upload overlap tasks");
    auto* sums = scratch_as<unsigned long long>(kSlotPixelDepth, 2 * n);
    launch_fill(reinterpret_cast<uint64_t*>(sums), 2 * n, 0, stream_);
    for (size_t j0 = 0; j0 < n; j0 += kMaxGridY)
        launch_overlap(dattr + j0, uint32_t(std::min<size_t>(kMaxGridY, n - j0)), tiles, sums + 2 * j0, stream_);
    hip_check(hipGetLastError(), "overlap launch");
    std::vector<unsigned long long> host(2 * n);
    hip_check(hipMemcpyAsync(host.data(), sums, host.size() * 8, hipMemcpyDeviceToHost, stream_), "download overlap");
    hip_check(hipStreamSynchronize(stream_), "overlap");
    double total = 0;
    for (size_t v = 0; v < n; ++v) total += host[2 * v + 1] ? double(host[2 * v]) / host[2 * v + 1] : 0;
    return total / n;
}

bool Audit::raster_pixels(const DeviceMesh& m, const Bounds& b, const Camera& c, double screen, uint8_t ss,
                          bool force_two, PixelRaster& out) {
    if (!ss || !std::isfinite(screen) || screen <= 0 || screen > 16384)
        throw std::invalid_argument("invalid raster extent");
    const uint32_t size = raster_side(screen, ss);
    out.size = size;
    if (uint64_t(size) * size > kMaxSamples) return false;
    DeviceGuard guard(device_);
    const uint32_t tiles = (size + kTile - 1) / kTile;
    const size_t words = size_t(tiles) * tiles, per = words * 64;
    uint64_t* mask = scratch_as<uint64_t>(kSlotPixelMask, words);
    float4* nrm = scratch_as<float4>(kSlotPixelNormals, per);
    float4* col = scratch_as<float4>(kSlotPixelColors, per);
    float* dep = scratch_as<float>(kSlotPixelDepth, per);
    launch_fill(mask, words, 0, stream_);
    std::vector<RasterJob> jobs{{&m, 0, mask}};
    const std::vector<Camera> cams{c};
    std::vector<uint32_t> st;
    std::vector<RasterTask> drawn;
    std::vector<const double*> zs;
    rasterize(jobs, cams, b, ss, size, force_two, st, &drawn, &zs);
    out.clipped = st[0] != kNone || st[1];
    attribute_rasters(jobs, drawn, zs, cams, b, nrm, col, dep);
    std::vector<uint64_t> tm(words);
    std::vector<float4> hn(per), hc(per);
    std::vector<float> hd(per);
    hip_check(hipMemcpyAsync(tm.data(), mask, words * 8, hipMemcpyDeviceToHost, stream_), "download mask");
    hip_check(hipMemcpyAsync(hn.data(), nrm, per * sizeof(float4), hipMemcpyDeviceToHost, stream_), "download normals");
    hip_check(hipMemcpyAsync(hc.data(), col, per * sizeof(float4), hipMemcpyDeviceToHost, stream_), "download colours");
    hip_check(hipMemcpyAsync(hd.data(), dep, per * 4, hipMemcpyDeviceToHost, stream_), "download depth");
    hip_check(hipStreamSynchronize(stream_), "attribute raster");
    const size_t samples = size_t(size) * size;
    out.covered.assign(samples, 0);
    out.visible.assign(samples, 0);
    out.depth.assign(samples, 0);
    out.normal.assign(samples * 3, 0);
    out.color.assign(samples * 4, 0);
    out.material.assign(samples, 0);
    for (uint32_t y = 0; y < size; ++y)
        for (uint32_t x = 0; x < size; ++x) {
            const size_t i = size_t(y) * size + x;
            const size_t w = size_t(y >> 3) * tiles + (x >> 3);
            const uint32_t bit = (y & 7) * 8 + (x & 7);
            const size_t k = w * 64 + bit;
            out.covered[i] = uint8_t((tm[w] >> bit) & 1);
            out.depth[i] = hd[k];
            // Pixel defaults where nothing is visible (include/blitz/evaluate.hpp:40-44).
            out.color[4 * i + 3] = 1.f;
            if (hn[k].w == 0) continue;
            out.visible[i] = 1;
            out.normal[3 * i] = hn[k].x;
            out.normal[3 * i + 1] = hn[k].y;
            out.normal[3 * i + 2] = hn[k].z;
            out.color[4 * i] = hc[k].x;
            out.color[4 * i + 1] = hc[k].y;
            out.color[4 * i + 2] = hc[k].z;
            uint32_t material;
            std::memcpy(&material, &hc[k].w, 4);
            out.material[i] = uint16_t(material);
        }
    return true;
}

}  // namespace blitz_hip
