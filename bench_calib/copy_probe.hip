// Plain device-to-device copy of a known byte count, used only to calibrate
// the HBM byte counters (FETCH_SIZE / WRITE_SIZE) that the profiler reports
// for the library's kernels: the copy moves exactly `bytes` in and out, so
// counter / bytes is the counter's scale on this device.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}
__global__ void fill_kernel(uint4* p, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = make_uint4(uint32_t(i), 1u, 2u, 3u);
}
}  // namespace

extern "C" int blitz_copy_probe(int device, size_t bytes, int repeats, double* ms_per_copy) {
    if (bytes % 16 || !bytes || repeats < 1 || !ms_per_copy) return 1;
    if (hipSetDevice(device) != hipSuccess) return 2;
    uint4 *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess) return 3;
    if (hipMalloc(&dst, bytes) != hipSuccess) {
        (void)hipFree(src);
        return 3;
    }
    const size_t n = bytes / 16;
    const unsigned blocks = 4096;
    fill_kernel<<<blocks, 256>>>(src, n);
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, nullptr);
    for (int r = 0; r < repeats; ++r) copy_kernel<<<blocks, 256>>>(src, dst, n);
    (void)hipEventRecord(b, nullptr);
    int rc = hipEventSynchronize(b) == hipSuccess && hipGetLastError() == hipSuccess ? 0 : 4;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    *ms_per_copy = double(ms) / repeats;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(src);
    (void)hipFree(dst);
    return rc;
}
