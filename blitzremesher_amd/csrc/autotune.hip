// Launch shapes of the band raster picked by measurement.
//
// A raster launch's shape is its face chunks per task (each workgroup walks
// one chunk of one task's faces with the task's band in LDS and flushes the
// band once). Few large chunks pay fewer band clears and flushes and fewer
// workgroup start-ups; many small ones fill the chip and balance better at
// the end of the launch. Which wins depends on the task count, the faces
// per task and the raster side, so the tuner times the candidates on the
// first launch of each shape class and keeps the fastest for the process.
//
// Timing on live data is safe because a raster launch is idempotent: every
// candidate ORs the same bits into the same masks (zeroed by the caller
// before the first) and min-reduces the same status words, so the masks
// after the tuning runs are exactly those of one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "autotune.hpp"

namespace blitz_hip {
namespace {

struct Key {
    int device;
    uint32_t tiles, tasks_log2, faces_log2;
    bool operator<(const Key& o) const {
        return std::tie(device, tiles, tasks_log2, faces_log2) < std::tie(o.device, o.tiles, o.tasks_log2, o.faces_log2);
    }
};

uint32_t ceil_log2(uint32_t v) {
    uint32_t l = 0;
    while ((1ull << l) < v) ++l;
    return l;
}

std::mutex mutex;
std::map<Key, uint32_t> chosen;  // shape class -> chunks per task

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Chunk counts worth timing around the heuristic's choice: an eighth, a
// quarter, half, the same and twice as many, multiples of 8 above 8 (one
// chunk set per XCD), at most one batch of faces per chunk.
std::vector<uint32_t> candidates(const RasterGrid& base, uint32_t max_faces) {
    const uint32_t most = std::max(1u, (max_faces + kRasterThreads - 1) / kRasterThreads);
    std::vector<uint32_t> out;
    for (uint32_t c : {base.chunks / 8, base.chunks / 4, base.chunks / 2, base.chunks, base.chunks * 2}) {
        c = std::min(std::max(c, 1u), most);
        if (c > 8) c = (c + 7) / 8 * 8;
        if (std::find(out.begin(), out.end(), c) == out.end()) out.push_back(c);
    }
    return out;
}

}  // namespace

RasterGrid RasterTuner::grid(uint32_t tasks, uint32_t tiles, uint32_t max_faces, hipStream_t stream,
                             const std::function<void(const RasterGrid&)>& launch) {
    RasterGrid g = make_raster_grid(tasks, tiles, max_faces);
    last_tuned_ = false;
    if (!enabled_ || !tasks || !max_faces) return g;
    int device = 0;
    check(hipGetDevice(&device), "hipGetDevice");
    const Key key{device, tiles, ceil_log2(tasks), ceil_log2(max_faces)};
    {
        std::lock_guard<std::mutex> lock(mutex);
        auto it = chosen.find(key);
        if (it != chosen.end()) {
            g.chunks = it->second;
            return g;
        }
    }
    const std::vector<uint32_t> list = candidates(g, max_faces);
    if (list.size() < 2) return g;
    if (!events_[0]) {
        for (auto& e : events_) check(hipEventCreate(&e), "hipEventCreate");
    }
    // One untimed run (first-touch effects), then each candidate once.
    launch(g);
    float best_ms = 0;
    uint32_t best = g.chunks;
    for (size_t i = 0; i < list.size(); ++i) {
        RasterGrid c = g;
        c.chunks = list[i];
        check(hipEventRecord(events_[0], stream), "hipEventRecord");
        launch(c);
        check(hipEventRecord(events_[1], stream), "hipEventRecord");
        check(hipEventSynchronize(events_[1]), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, events_[0], events_[1]), "hipEventElapsedTime");
        if (i == 0 || ms < best_ms) {
            best_ms = ms;
            best = list[i];
        }
    }
    {
        std::lock_guard<std::mutex> lock(mutex);
        chosen[key] = best;
    }
    last_tuned_ = true;
    g.chunks = best;
    return g;
}

RasterTuner::~RasterTuner() {
    for (auto& e : events_)
        if (e) (void)hipEventDestroy(e);
}

}  // namespace blitz_hip
