// blitz_hip_coverage_distance_noalloc: the directed-Hausdorff coverage
// distance of two device-resident masks with every temporary in one
// fixed-size array inside the library, for callers that cannot allocate
// device memory. It has no counterpart in the reference.
//
// The array (`scratch`) is a file-scope __device__ array: each GPU the
// library runs on holds its own copy until the process ends. Host code never
// passes the array's name to a kernel (that is a host-side stand-in, not a
// device address): the first call on a device, with that device current,
// looks up the device address with hipGetSymbolAddress and keeps it per
// device, and every kernel argument uses that address. Calls that overlap on
// one GPU share the array and race; calls on one stream are ordered.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <stdexcept>

#include "audit.hpp"
#include "audit_kernels.hpp"
#include "noalloc.hpp"

namespace blitz_hip {
namespace {

constexpr uint32_t kMaxTiles = kNoAllocMaxSide / kTile;
constexpr size_t kMaxWords = size_t(kMaxTiles) * kMaxTiles;
// Layout: the two tile-major masks, then one ViewPair and one ViewSummary
// (each rounded up to whole 8-byte words).
constexpr size_t kPairWords = (sizeof(ViewPair) + 7) / 8;
constexpr size_t kSummaryWords = (sizeof(ViewSummary) + 7) / 8;
constexpr size_t kScratchWords = 2 * kMaxWords + kPairWords + kSummaryWords;

__device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
uint64_t* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), looked up once.
uint64_t* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        g_scratch[device] = static_cast<uint64_t*>(p);
    }
    return g_scratch[device];
}

// Row-major packed mask (bit index y*size + x) -> tile-major words (word
// ty*tiles + tx, bit (y&7)*8 + (x&7)); padding bits zero. One thread per
// tile word.
__global__ __launch_bounds__(256) void tile_kernel(const uint64_t* __restrict__ src, uint32_t size, uint32_t tiles,
                                                   uint64_t* __restrict__ dst) {
    const size_t w = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (w >= size_t(tiles) * tiles) return;
    const uint32_t tx = uint32_t(w % tiles), ty = uint32_t(w / tiles);
    const uint32_t x0 = tx * kTile;
    const uint32_t columns = min(kTile, size - x0);
    const uint32_t keep = (1u << columns) - 1;
    uint64_t word = 0;
    for (uint32_t r = 0; r < kTile; ++r) {
        const uint32_t y = ty * kTile + r;
        if (y >= size) break;
        const uint64_t first = uint64_t(y) * size + x0, last = first + columns - 1;
        const uint64_t lo = src[first >> 6] >> (first & 63);
        const uint64_t hi = (last >> 6) != (first >> 6) ? src[last >> 6] << (64 - (first & 63)) : 0;
        word |= uint64_t(uint32_t(lo | hi) & keep) << (r * kTile);
    }
    dst[w] = word;
}

__global__ void pair_kernel(uint64_t* masks, uint32_t size, uint32_t tiles, ViewPair* pair) {
    const size_t words = size_t(tiles) * tiles;
    ViewPair p;
    p.a = masks;
    p.c = masks + words;
    p.dist_a = nullptr;  // no row-distance planes: the Hausdorff pass searches the masks
    p.dist_c = nullptr;
    p.size = size;
    p.tiles = tiles;
    p.fail_sq = 0xFFFFFFFFu;
    *pair = p;
}

// The host rule of the coverage distance (src/evaluate.cpp:202-208):
// identical masks 0, exactly one empty +inf, else sqrt(float(d^2)) / ss with
// correctly rounded square root and division.
__global__ void finish_kernel(const ViewSummary* s, uint32_t supersample, double* out) {
    double v;
    if (!s->changed)
        v = 0;
    else if ((s->count_a != 0) != (s->count_c != 0))
        v = __builtin_huge_val();
    else
        v = __ddiv_rn(__dsqrt_rn(double(float(max(s->max_ac, s->max_ca)))), double(supersample));
    *out = v;
}

}  // namespace

void coverage_distance_noalloc(int device, const uint64_t* a, const uint64_t* b, uint32_t size, uint8_t supersample,
                               double* out, hipStream_t stream) {
    if (!a || !b || !out) throw std::invalid_argument("null mask or output");
    if (!size || !supersample) throw std::invalid_argument("invalid raster size or supersampling");
    if (size > kNoAllocMaxSide)
        throw std::invalid_argument("raster side exceeds the no-allocation maximum of 8192 samples");
    DeviceGuard guard(device);
    uint64_t* base = scratch_address(device);
    const uint32_t tiles = (size + kTile - 1) / kTile;
    const size_t words = size_t(tiles) * tiles;
    uint64_t* masks = base;
    auto* pair = reinterpret_cast<ViewPair*>(base + 2 * kMaxWords);
    auto* summary = reinterpret_cast<ViewSummary*>(base + 2 * kMaxWords + kPairWords);
    const unsigned blocks = unsigned((words + 255) / 256);
    tile_kernel<<<blocks, 256, 0, stream>>>(a, size, tiles, masks);
    tile_kernel<<<blocks, 256, 0, stream>>>(b, size, tiles, masks + words);
    pair_kernel<<<1, 1, 0, stream>>>(masks, size, tiles, pair);
    // The summary starts at zero on every call (the array keeps the last
    // call's contents).
    launch_fill(reinterpret_cast<uint64_t*>(summary), kSummaryWords, 0, stream);
    launch_compare(pair, summary, 1, tiles, stream);
    launch_hausdorff(pair, summary, 1, tiles, stream);
    finish_kernel<<<1, 1, 0, stream>>>(summary, supersample, out);
    hip_check(hipGetLastError(), "no-allocation distance launch");
}

}  // namespace blitz_hip
