// Owning device allocation that grows on demand (contents not kept).
#pragma once
#include <cstddef>

namespace blitz_hip {

struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer();
    void resize(size_t n);
    template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace blitz_hip
