// Host-side runtime pieces of the C ABI: error reporting, device memory,
// streams and events, scratch allocation.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "tsf_common.hpp"

namespace tsf {

static thread_local char g_err[512] = "";

tsf_status set_error(tsf_status st, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return st;
}

tsf_status check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return TSF_OK;
}

tsf_status check_dims(tsf_dims d, int min_extent, const char* what) {
    if (d.x < min_extent || d.y < min_extent || d.z < min_extent)
        return set_error(TSF_ERR_DIMENSIONS, "%s: dimensions %dx%dx%d below the minimum extent %d", what, d.x, d.y,
                         d.z, min_extent);
    if ((long long)d.x * d.y * d.z > (1LL << 34))
        return set_error(TSF_ERR_DIMENSIONS, "%s: volume of %dx%dx%d voxels is too large", what, d.x, d.y, d.z);
    // the stencil kernels put y/4 and z on grid dimensions limited to 65535
    if (d.x > (1 << 21) || d.y > 65535 || d.z > 65535)
        return set_error(TSF_ERR_DIMENSIONS, "%s: extent above the supported maximum", what);
    return TSF_OK;
}

tsf_status check_field(const tsf_field* f, bool need_w, const char* what) {
    if (!f) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: field descriptor is NULL", what);
    if (!f->x || !f->y || !f->z) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: field plane pointer is NULL", what);
    if (need_w && !f->w) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: field plane w is required", what);
    if (f->storage < TSF_STORAGE_F32 || f->storage > TSF_STORAGE_SN16)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: unknown storage type %d", what, f->storage);
    if (!(f->scale > 0.0f) || f->scale > 1e30f)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: field scale must be positive", what);
    if ((f->storage == TSF_STORAGE_F32 || f->storage == TSF_STORAGE_SN16) && f->scale != 1.0f)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: F32 and SN16 fields must have scale 1", what);
    int ex;
    if (frexpf(f->scale, &ex) != 0.5f)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: field scale must be a power of two", what);
    return TSF_OK;
}

size_t storage_bytes(int storage) { return storage == TSF_STORAGE_F32 ? 4 : 2; }

Scratch::~Scratch() {
    if (owned && ptr) (void)hipFree(ptr);
}

tsf_status Scratch::get(void* ws, size_t ws_bytes, size_t need, const char* what) {
    if (need == 0) return TSF_OK;
    if (ws) {
        if (ws_bytes < need)
            return set_error(TSF_ERR_OUT_OF_MEMORY, "%s: workspace of %zu bytes, %zu needed", what, ws_bytes, need);
        if ((uintptr_t)ws % 256 != 0)
            return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: workspace must be 256-byte aligned", what);
        ptr = ws;
        return TSF_OK;
    }
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess) {
        ptr = nullptr;
        return set_error(TSF_ERR_OUT_OF_MEMORY, "%s: hipMalloc(%zu) failed: %s", what, need, hipGetErrorString(e));
    }
    owned = true;
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

#define HIPCALL(call, what)                                                                       \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", what, hipGetErrorString(_e)); \
    } while (0)

extern "C" {

const char* tsf_last_error(void) { return g_err; }
const char* tsf_version(void) { return "tsf-mi355x 0.1 (gfx950)"; }

tsf_status tsf_malloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_malloc: ptr is NULL");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return set_error(TSF_ERR_OUT_OF_MEMORY, "tsf_malloc(%zu): %s", bytes, hipGetErrorString(e));
    return TSF_OK;
}
tsf_status tsf_free(void* ptr) {
    HIPCALL(hipFree(ptr), "tsf_free");
    return TSF_OK;
}
tsf_status tsf_memcpy_htod(void* dst, const void* src, size_t bytes, hipStream_t s) {
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "tsf_memcpy_htod");
    HIPCALL(hipStreamSynchronize(s), "tsf_memcpy_htod");
    return TSF_OK;
}
tsf_status tsf_memcpy_dtoh(void* dst, const void* src, size_t bytes, hipStream_t s) {
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "tsf_memcpy_dtoh");
    HIPCALL(hipStreamSynchronize(s), "tsf_memcpy_dtoh");
    return TSF_OK;
}
tsf_status tsf_memcpy_dtod(void* dst, const void* src, size_t bytes, hipStream_t s) {
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s), "tsf_memcpy_dtod");
    return TSF_OK;
}
tsf_status tsf_memset(void* dst, int value, size_t bytes, hipStream_t s) {
    HIPCALL(hipMemsetAsync(dst, value, bytes, s), "tsf_memset");
    return TSF_OK;
}
tsf_status tsf_stream_create(hipStream_t* s) {
    if (!s) return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_stream_create: NULL");
    HIPCALL(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "tsf_stream_create");
    return TSF_OK;
}
tsf_status tsf_stream_destroy(hipStream_t s) {
    HIPCALL(hipStreamDestroy(s), "tsf_stream_destroy");
    return TSF_OK;
}
tsf_status tsf_stream_synchronize(hipStream_t s) {
    HIPCALL(hipStreamSynchronize(s), "tsf_stream_synchronize");
    return TSF_OK;
}
tsf_status tsf_event_create(hipEvent_t* ev) {
    if (!ev) return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_event_create: NULL");
    HIPCALL(hipEventCreate(ev), "tsf_event_create");
    return TSF_OK;
}
tsf_status tsf_event_destroy(hipEvent_t ev) {
    HIPCALL(hipEventDestroy(ev), "tsf_event_destroy");
    return TSF_OK;
}
tsf_status tsf_event_record(hipEvent_t ev, hipStream_t s) {
    HIPCALL(hipEventRecord(ev, s), "tsf_event_record");
    return TSF_OK;
}
tsf_status tsf_event_elapsed_ms(hipEvent_t a, hipEvent_t b, float* ms) {
    if (!ms) return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_event_elapsed_ms: NULL");
    HIPCALL(hipEventSynchronize(b), "tsf_event_elapsed_ms");
    HIPCALL(hipEventElapsedTime(ms, a, b), "tsf_event_elapsed_ms");
    return TSF_OK;
}
tsf_status tsf_stream_wait_event(hipStream_t s, hipEvent_t ev) {
    HIPCALL(hipStreamWaitEvent(s, ev, 0), "tsf_stream_wait_event");
    return TSF_OK;
}
tsf_status tsf_device_count(int* count) {
    if (!count) return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_device_count: NULL");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return set_error(TSF_ERR_HIP, "tsf_device_count: %s", hipGetErrorString(e));
    }
    return TSF_OK;
}
tsf_status tsf_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_mem_info: NULL");
    HIPCALL(hipMemGetInfo(free_bytes, total_bytes), "tsf_mem_info");
    return TSF_OK;
}

}  // extern "C"
