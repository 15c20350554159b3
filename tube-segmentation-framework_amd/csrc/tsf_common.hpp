// Shared device/host helpers of the MI355X (gfx950) volume-filter library.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/tsf.h"

namespace tsf {

typedef long long i64;

struct Dims {
    int X, Y, Z;
    __host__ __device__ i64 n() const { return (i64)X * Y * Z; }
    __host__ __device__ i64 at(int x, int y, int z) const {
        return (i64)x + (i64)X * ((i64)y + (i64)Y * z);
    }
};

__host__ __device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Mirror remap of the GVF/FMG kernels: 0 -> 2, >= N-1 -> N-3.
__host__ __device__ inline int remap(int p, int N) { return p == 0 ? 2 : (p >= N - 1 ? N - 3 : p); }

// ---------------------------------------------------------------- storage
// Vector-field planes are stored as float, half or bfloat16.  Reduced
// precision planes hold value * scale (scale is a power of two, so the
// scaling itself is exact); arithmetic is always float.
struct F32 {
    typedef float T;
    static __device__ __forceinline__ float ld(const T* p, i64 i, float) { return p[i]; }
    static __device__ __forceinline__ void st(T* p, i64 i, float v, float) { p[i] = v; }
};
struct F16 {
    typedef __half T;
    static __device__ __forceinline__ float ld(const T* p, i64 i, float inv) { return __half2float(p[i]) * inv; }
    static __device__ __forceinline__ void st(T* p, i64 i, float v, float s) { p[i] = __float2half_rn(v * s); }
};
struct BF16 {
    typedef __hip_bfloat16 T;
    static __device__ __forceinline__ float ld(const T* p, i64 i, float inv) {
        return __bfloat162float(p[i]) * inv;
    }
    static __device__ __forceinline__ void st(T* p, i64 i, float v, float s) {
        p[i] = __float2bfloat16(v * s);
    }
};

// Host-side error plumbing.  Every entry point returns a tsf_status and leaves
// a message retrievable with tsf_last_error().
tsf_status set_error(tsf_status st, const char* fmt, ...);
tsf_status check_launch(const char* what);
tsf_status check_dims(tsf_dims d, int min_extent, const char* what);
tsf_status check_field(const tsf_field* f, bool need_w, const char* what);
size_t storage_bytes(int storage);

inline Dims D(tsf_dims d) { return Dims{d.x, d.y, d.z}; }

// Scratch allocation for entry points called with workspace == NULL.
struct Scratch {
    void* ptr = nullptr;
    bool owned = false;
    ~Scratch();
    tsf_status get(void* ws, size_t ws_bytes, size_t need, const char* what);
};

}  // namespace tsf

#define TSF_TRY(expr)                          \
    do {                                       \
        tsf_status _s = (expr);                \
        if (_s != TSF_OK) return _s;           \
    } while (0)
