// Collapse actions on the device: the counterpart of the reference's GPU
// action state (src/neural/action_gpu.cu:178-343 topology, 487-643 legality,
// 757-808 trial edits, 865-899 batch selection) with the CPU semantics of
// blitz::neural::ActionState (src/neural/action.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "audit.hpp"
#include "device_buffer.hpp"

namespace blitz_hip {

// One legal directed half-edge collapse `from` -> `to` (geometric vertices:
// the smallest index among bit-equal positions) and its wedge mapping: in
// every face holding both, index wedge_from[i] becomes wedge_to[i].
struct LegalAction {
    uint32_t from, to;
    uint32_t wedge_from[2], wedge_to[2];
    uint32_t wedges;   // 1 or 2
    uint32_t removed;  // faces the collapse removes (1 boundary, 2 interior)
};

class ActionSet {
  public:
    // Builds the topology of `m` (optional per-vertex UVs, 2 floats each) and
    // enumerates its legal actions, edges ascending, a->b then b->a
    // (src/neural/action.cpp:371-383).
    ActionSet(int device, const HostMesh& m, const float* uv);
    ~ActionSet();
    ActionSet(const ActionSet&) = delete;
    ActionSet& operator=(const ActionSet&) = delete;

    uint32_t legal_count() const { return legal_count_; }
    uint32_t edge_count() const { return edges_; }
    void download_legal(uint32_t* pairs) const;  // from, to per legal action
    // ActionState::independent (src/neural/action.cpp:280-304): greedy batch
    // of non-overlapping footprints in `order` (indices of legal actions).
    std::vector<uint32_t> independent(const uint32_t* order, size_t n, uint32_t maximum, uint64_t removable);
    // ActionState::trial (src/neural/action.cpp:308-344): the index buffer of
    // the mesh with the legal actions `rows` applied.
    std::vector<uint32_t> trial(const uint32_t* rows, size_t n);
    // Each legal action `rows[i]` as a patch of the source: the faces around
    // `from` (ascending) removed, their remapped survivors added.
    void patches(const uint32_t* rows, size_t n, std::vector<uint32_t>& removed_offsets,
                 std::vector<uint32_t>& removed, std::vector<uint32_t>& added_offsets,
                 std::vector<uint32_t>& added);
    double build_ms() const { return build_ms_; }
    int device() const { return device_; }

  private:
    void claim(const uint32_t* rows, size_t n, uint32_t maximum, uint64_t removable, bool strict,
               std::vector<uint32_t>& picked);

    int device_;
    hipStream_t stream_{};
    uint32_t vertices_ = 0, faces_ = 0, edges_ = 0, legal_count_ = 0;
    double build_ms_ = 0;
    DeviceBuffer positions_, uv_, indices_, materials_;
    DeviceBuffer geometry_, invalid_, boundary_;
    DeviceBuffer offsets_, neighbors_, face_offsets_, face_lists_;
    DeviceBuffer edge_keys_, edge_counts_;
    DeviceBuffer legal_;  // LegalAction per legal action
    DeviceBuffer used_, selection_, scratch_a_, scratch_b_, scratch_c_, temp_;
};

}  // namespace blitz_hip
