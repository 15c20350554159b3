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

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round-robin
// by linear workgroup id.  xcd_tile() renumbers them so that XCD k works
// through a contiguous range of tiles: tiles that share halos then meet in
// the same L2.  A bijection on [0, gridDim.x * gridDim.y * gridDim.z).
constexpr int XCDS = 8;
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz) {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned total = gx * gy * gridDim.z;
    const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned q = total / XCDS, r = total % XCDS, k = id % XCDS, j = id / XCDS;
    const unsigned t = k * q + (k < r ? k : r) + j;
    bx = (int)(t % gx);
    by = (int)((t / gx) % gy);
    bz = (int)(t / (gx * gy));
}

// Word 3 of a raw buffer resource on gfx9-family chips (gfx950 included):
// 32-bit data format, no swizzle.  With stride 0, num_records counts bytes
// and a load at or past it returns 0.
constexpr int BUFFER_WORD3 = 0x00020000;

// Mirror remap of the GVF/FMG kernels: 0 -> 2, >= N-1 -> N-3.
__host__ __device__ inline int remap(int p, int N) { return p == 0 ? 2 : (p >= N - 1 ? N - 3 : p); }

// ---------------------------------------------------------------- storage
// Vector-field planes are stored as float, half or bfloat16.  Reduced
// precision planes hold value * scale (scale is a power of two, so the
// scaling itself is exact); arithmetic is always float.
// ld2/st2 move two consecutive elements with one access (i must be even and
// the plane base suitably aligned: 8 bytes for F32, 4 bytes otherwise).
struct F32 {
    typedef float T;
    static __device__ __forceinline__ float ld(const T* p, i64 i, float) { return p[i]; }
    static __device__ __forceinline__ void st(T* p, i64 i, float v, float) { p[i] = v; }
    static __device__ __forceinline__ void ld2(const T* p, i64 i, float, float& a, float& b) {
        const float2 v = *reinterpret_cast<const float2*>(p + i);
        a = v.x;
        b = v.y;
    }
    static __device__ __forceinline__ void st2(T* p, i64 i, float a, float b, float) {
        *reinterpret_cast<float2*>(p + i) = make_float2(a, b);
    }
};
struct F16 {
    typedef __half T;
    static __device__ __forceinline__ float ld(const T* p, i64 i, float inv) { return __half2float(p[i]) * inv; }
    static __device__ __forceinline__ void st(T* p, i64 i, float v, float s) { p[i] = __float2half_rn(v * s); }
    static __device__ __forceinline__ void ld2(const T* p, i64 i, float inv, float& a, float& b) {
        const __half2 h = *reinterpret_cast<const __half2*>(p + i);
        a = __low2float(h) * inv;
        b = __high2float(h) * inv;
    }
    static __device__ __forceinline__ void st2(T* p, i64 i, float a, float b, float s) {
        *reinterpret_cast<__half2*>(p + i) = __halves2half2(__float2half_rn(a * s), __float2half_rn(b * s));
    }
};
struct BF16 {
    typedef __hip_bfloat16 T;
    static __device__ __forceinline__ float ld(const T* p, i64 i, float inv) {
        return __bfloat162float(p[i]) * inv;
    }
    static __device__ __forceinline__ void st(T* p, i64 i, float v, float s) {
        p[i] = __float2bfloat16(v * s);
    }
    static __device__ __forceinline__ void ld2(const T* p, i64 i, float inv, float& a, float& b) {
        const unsigned u = *reinterpret_cast<const unsigned*>(p + i);
        a = __uint_as_float(u << 16) * inv;
        b = __uint_as_float(u & 0xffff0000u) * inv;
    }
    static __device__ __forceinline__ void st2(T* p, i64 i, float a, float b, float s) {
        const __hip_bfloat16 x = __float2bfloat16(a * s), y = __float2bfloat16(b * s);
        const unsigned u = (unsigned)__bfloat16_as_ushort(x) | ((unsigned)__bfloat16_as_ushort(y) << 16);
        *reinterpret_cast<unsigned*>(p + i) = u;
    }
};

// The reference's 16-bit vector format (kernels_no_3d_write.cl:9-21):
// stored as convert_short_sat_rte(v * 32767), read as max(-1, s / 32767).
// The quotient is computed as q0 = s * r, e = fma(-q0, 32767, s), q0 +
// e * r (one fma) with r = 1/32767 rounded: for every one of the 65536
// codes this equals the correctly rounded s / 32767.
__device__ __forceinline__ float sn16_to_f(int s) {
    const float sf = (float)s, r = 1.0f / 32767.0f;
    const float q0 = sf * r;
    const float e = __builtin_fmaf(-q0, 32767.0f, sf);
    return fmaxf(-1.0f, __builtin_fmaf(e, r, q0));
}
__device__ __forceinline__ short f_to_sn16(float v) {
    const float x = v * 32767.0f;
    // round to nearest even, saturate; NaN converts to 0
    const float r = rintf(fminf(fmaxf(x, -32768.0f), 32767.0f));
    return (short)(x == x ? (int)r : 0);
}
// The same conversions on values in units of 2^-15 (u = value * 32768), for
// kernels that compute in storage units: fma(s, 1/32767, s) is the correctly
// rounded 32768 * s / 32767 for every one of the 65536 codes, and
// u * (32767/32768) is exactly the product value * 32767.
__device__ __forceinline__ float sn16u_to_f(int s) {
    const float sf = (float)s;
    return fmaxf(-32768.0f, __builtin_fmaf(sf, 1.0f / 32767.0f, sf));
}
__device__ __forceinline__ short f_to_sn16u(float u) {
    const float x = u * (32767.0f / 32768.0f);
    const float r = rintf(fminf(fmaxf(x, -32768.0f), 32767.0f));
    return (short)(x == x ? (int)r : 0);
}
struct SN16 {
    typedef short T;
    static __device__ __forceinline__ float ld(const T* p, i64 i, float) { return sn16_to_f(p[i]); }
    static __device__ __forceinline__ void st(T* p, i64 i, float v, float) { p[i] = f_to_sn16(v); }
    static __device__ __forceinline__ void ld2(const T* p, i64 i, float, float& a, float& b) {
        const unsigned u = *reinterpret_cast<const unsigned*>(p + i);
        a = sn16_to_f((short)(u & 0xffffu));
        b = sn16_to_f((short)(u >> 16));
    }
    static __device__ __forceinline__ void st2(T* p, i64 i, float a, float b, float) {
        const unsigned u = (unsigned)(unsigned short)f_to_sn16(a) | ((unsigned)(unsigned short)f_to_sn16(b) << 16);
        *reinterpret_cast<unsigned*>(p + i) = u;
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
