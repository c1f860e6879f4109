// Measured choice of the band raster's launch shape (autotune.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

#include "audit_kernels.hpp"

namespace blitz_hip {

class RasterTuner {
  public:
    explicit RasterTuner(bool enabled = true) : enabled_(enabled) {}
    ~RasterTuner();
    RasterTuner(const RasterTuner&) = delete;
    RasterTuner& operator=(const RasterTuner&) = delete;
    // The grid for `tasks` raster tasks of up to `max_faces` faces on a
    // `tiles` x `tiles` word raster. The first time a shape class (device,
    // raster side, task count and face count rounded up to powers of two) is
    // seen, every candidate chunk count is launched through `launch` on
    // `stream` and timed, and the fastest is kept for the process; the masks
    // and status words are left as one launch leaves them. Disabled: the
    // heuristic grid (make_raster_grid).
    RasterGrid grid(uint32_t tasks, uint32_t tiles, uint32_t max_faces, hipStream_t stream,
                    const std::function<void(const RasterGrid&)>& launch);
    bool last_tuned() const { return last_tuned_; }

  private:
    bool enabled_;
    bool last_tuned_ = false;
    hipEvent_t events_[2]{};
};

}  // namespace blitz_hip
