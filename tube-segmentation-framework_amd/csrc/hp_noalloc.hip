// Histogram-pyramid compaction without device allocation.
//
// The cube counts and their scan live in one static device array per GPU
// (`hp_scratch`), sized for volumes whose largest extent is at most
// HP_NOALLOC_MAX_EXTENT: (1024/16)^3 = 2^18 cubes, 1 MiB of counts and
// 2 MiB + 8 B of offsets, held by every GPU the library runs on until the
// process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress and kept; the array's host-side name is never handed
// to a kernel or a copy.  Every entry the kernels read is written earlier in
// the same call, so nothing is cleared between calls.  Calls that overlap on
// one GPU (two streams, or two host threads) share the array and race.
#include <mutex>

#include "hp.hpp"

namespace tsf {

constexpr int HP_NOALLOC_MAX_EXTENT = 1024;
constexpr int HP_NOALLOC_LEVELS = 6;  // 16 * 2^6 = 1024
constexpr size_t HP_NOALLOC_CUBES = (size_t)1 << (3 * HP_NOALLOC_LEVELS);
constexpr size_t HP_NOALLOC_BYTES =
    ((HP_NOALLOC_CUBES * sizeof(int) + 255) & ~(size_t)255) + (HP_NOALLOC_CUBES + 1) * sizeof(long long);
static __device__ __align__(256) unsigned char hp_scratch[HP_NOALLOC_BYTES];

constexpr int HP_MAX_DEVICES = 64;
static std::mutex g_hp_scratch_mu;
static void* g_hp_scratch_ptr[HP_MAX_DEVICES];

// Device address of the current device's hp_scratch, looked up on the first
// call on that device.
static tsf_status hp_scratch_ptr(void** out, const char* W) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: hipGetDevice: %s", W, hipGetErrorString(e));
    if (dev < 0 || dev >= HP_MAX_DEVICES) return set_error(TSF_ERR_HIP, "%s: device index %d out of range", W, dev);
    std::lock_guard<std::mutex> lock(g_hp_scratch_mu);
    if (!g_hp_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(hp_scratch));
        if (e != hipSuccess || !p)
            return set_error(TSF_ERR_HIP, "%s: hipGetSymbolAddress(hp_scratch): %s", W, hipGetErrorString(e));
        g_hp_scratch_ptr[dev] = p;
    }
    *out = g_hp_scratch_ptr[dev];
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_hp_compact_workspace_size(tsf_dims dims);

tsf_status tsf_hp_compact_noalloc(const uint8_t* mask, tsf_dims dims, int32_t* positions, int64_t capacity,
                                  int64_t* count, hipStream_t stream) {
    const char* W = "tsf_hp_compact_noalloc";
    TSF_TRY(hp_check(mask, dims, positions, capacity, count, W));
    if (dims.x > HP_NOALLOC_MAX_EXTENT || dims.y > HP_NOALLOC_MAX_EXTENT || dims.z > HP_NOALLOC_MAX_EXTENT)
        return set_error(TSF_ERR_DIMENSIONS, "%s: %dx%dx%d exceeds the largest extent %d of the static scratch", W,
                         dims.x, dims.y, dims.z, HP_NOALLOC_MAX_EXTENT);
    const size_t need = tsf_hp_compact_workspace_size(dims);
    if (need > HP_NOALLOC_BYTES)
        return set_error(TSF_ERR_DIMENSIONS, "%s: scratch of %zu bytes needed, %zu available", W, need,
                         HP_NOALLOC_BYTES);
    void* ws = nullptr;
    TSF_TRY(hp_scratch_ptr(&ws, W));
    Dims d = D(dims);
    return hp_run(mask, d, hp_plan(d), ws, positions, capacity, count, stream, W);
}

}  // extern "C"
