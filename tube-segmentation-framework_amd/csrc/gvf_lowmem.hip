// Gradient vector flow one component at a time (runLowMemoryGVF,
// gradientVectorFlow.cpp:1173-1445; kernels GVF3DInit_one_component,
// GVF3DIteration_one_component, GVF3DFinish_one_component,
// kernels_no_3d_write.cl:1669-1744).
//
// The components of the GVF update couple only through |f|^2, so each one
// can iterate alone given f's component and |f|^2: the device then holds
// f, one |f|^2 plane and a single ping-pong plane instead of a whole second
// field.  As in the reference, |f|^2 is stored in the field's format: for
// F32 the path is the fast path's arithmetic exactly; in the 16-bit mode
// |f|^2 is rounded to SNORM16 once, so the result differs from the fast
// path's there as the reference's two paths differ.
//
// One step per launch: a workgroup of 64 x 4 threads owns a 64-column,
// 4-row tile and marches a chunk of planes in z, reading its stencil
// through the caches (the path is for memory, not speed).
#include <algorithm>

#include "tsf_common.hpp"

namespace tsf {

constexpr int LMX = 64, LMY = 4, LMZ = 32;

// Values of one plane element: F32 as is, SNORM16 as max(-1, s / 32767).
template <class S>
__device__ __forceinline__ float lm_ld(const typename S::T* p, i64 i) {
    return S::ld(p, i, 1.0f);
}

// |f|^2 = x x + y y + z z of f's values, stored in the field's format.
template <class S>
__global__ void k_lm_init(const typename S::T* __restrict__ fx, const typename S::T* __restrict__ fy,
                          const typename S::T* __restrict__ fz, typename S::T* __restrict__ m2, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const float x = lm_ld<S>(fx, i), y = lm_ld<S>(fy, i), z = lm_ld<S>(fz, i);
        S::st(m2, i, x * x + y * y + z * z, 1.0f);
    }
}

// One Euler step of one component: v' = v + (mu lap(v) - (v - f) m2) at the
// mirror voxel of each voxel (0 -> 2, N-1 -> N-3), lap in the reference's
// order (-6 v, then x+1, x-1, y+1, y-1, z+1, z-1).
template <class S>
__global__ void __launch_bounds__(LMX* LMY)
    k_lm_step(const typename S::T* __restrict__ f, const typename S::T* __restrict__ m2,
              const typename S::T* __restrict__ v, typename S::T* __restrict__ out, Dims d, float mu) {
    const int x = blockIdx.x * LMX + threadIdx.x, y = blockIdx.y * LMY + threadIdx.y;
    if (x >= d.X || y >= d.Y) return;
    const i64 sy = d.X, sz = (i64)d.X * d.Y;
    const int qx = remap(x, d.X), qy = remap(y, d.Y);
    const int z0 = blockIdx.z * LMZ, z1 = min(z0 + LMZ, d.Z);
    for (int z = z0; z < z1; z++) {
        const i64 q = d.at(qx, qy, remap(z, d.Z));
        const float vc = lm_ld<S>(v, q);
        float lap = -6.0f * vc;
        lap = lap + lm_ld<S>(v, q + 1);
        lap = lap + lm_ld<S>(v, q - 1);
        lap = lap + lm_ld<S>(v, q + sy);
        lap = lap + lm_ld<S>(v, q - sy);
        lap = lap + lm_ld<S>(v, q + sz);
        lap = lap + lm_ld<S>(v, q - sz);
        const float r = vc + (mu * lap - (vc - lm_ld<S>(f, q)) * lm_ld<S>(m2, q));
        S::st(out, d.at(x, y, z), r, 1.0f);
    }
}

// f's component as the output (no iterations), through the field's values.
template <class S>
__global__ void k_lm_copy(const typename S::T* __restrict__ f, typename S::T* __restrict__ o, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        S::st(o, i, lm_ld<S>(f, i), 1.0f);
}

// w = |v| if > 0, else 1 (GVF3DFinish), in the field's format.
template <class S>
__global__ void k_lm_finish(const typename S::T* __restrict__ x, const typename S::T* __restrict__ y,
                            const typename S::T* __restrict__ z, typename S::T* __restrict__ w, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const float a = lm_ld<S>(x, i), b = lm_ld<S>(y, i), c = lm_ld<S>(z, i);
        const float l = sqrtf(a * a + b * b + c * c);
        S::st(w, i, l > 0.0f ? l : 1.0f, 1.0f);
    }
}

template <class S>
static tsf_status run_lowmem(const tsf_field* f, const tsf_field* out, Dims d, float mu, int iters, void* ws,
                             hipStream_t stream) {
    typedef typename S::T T;
    const i64 n = d.n();
    T* M2 = (T*)ws;
    T* P = M2 + n;  // the ping-pong plane
    const unsigned G = (unsigned)std::min<i64>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_lm_init<S>, G, 256, 0, stream, (const T*)f->x, (const T*)f->y, (const T*)f->z, M2, n);
    TSF_TRY(check_launch("tsf_gvf_euler_low_memory"));
    const T* F[3] = {(const T*)f->x, (const T*)f->y, (const T*)f->z};
    T* O[3] = {(T*)out->x, (T*)out->y, (T*)out->z};
    const dim3 grid((d.X + LMX - 1) / LMX, (d.Y + LMY - 1) / LMY, (d.Z + LMZ - 1) / LMZ), block(LMX, LMY);
    for (int c = 0; c < 3; c++) {
        if (iters == 0) {
            hipLaunchKernelGGL(k_lm_copy<S>, G, 256, 0, stream, F[c], O[c], n);
            TSF_TRY(check_launch("tsf_gvf_euler_low_memory"));
            continue;
        }
        // ping-pong so that the last step lands in the output plane
        const T* src = F[c];
        for (int k = 0; k < iters; k++) {
            T* dst = ((iters - 1 - k) % 2 == 0) ? O[c] : P;
            hipLaunchKernelGGL(k_lm_step<S>, grid, block, 0, stream, F[c], (const T*)M2, src, dst, d, mu);
            TSF_TRY(check_launch("tsf_gvf_euler_low_memory"));
            src = dst;
        }
    }
    hipLaunchKernelGGL(k_lm_finish<S>, G, 256, 0, stream, (const T*)O[0], (const T*)O[1], (const T*)O[2],
                       (T*)out->w, n);
    return check_launch("tsf_gvf_euler_low_memory");
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_gvf_euler_low_memory_workspace_size(tsf_dims dims, int32_t storage) {
    // the |f|^2 plane and one ping-pong plane
    return 2 * (size_t)D(dims).n() * (storage == TSF_STORAGE_F32 ? 4 : 2);
}

tsf_status tsf_gvf_euler_low_memory(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu,
                                    int32_t iterations, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_gvf_euler_low_memory";
    TSF_TRY(check_field(f, false, W));
    TSF_TRY(check_field(out, true, W));
    TSF_TRY(check_dims(dims, 4, W));
    if (f->storage != out->storage)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f and out must share a storage type", W);
    if (f->storage != TSF_STORAGE_F32 && f->storage != TSF_STORAGE_SN16)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: F32 or SNORM16 fields (the reference's formats)", W);
    if (iterations < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative iteration count", W);
    if (!(mu >= 0.0f) || !__builtin_isfinite(mu)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: mu must be >= 0", W);
    for (const void* p : {out->x, out->y, out->z})
        if (p == f->x || p == f->y || p == f->z) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: out aliases f", W);
    const Dims d = D(dims);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_gvf_euler_low_memory_workspace_size(dims, f->storage), W));
    tsf_status st = f->storage == TSF_STORAGE_F32 ? run_lowmem<F32>(f, out, d, mu, iterations, sc.ptr, stream)
                                                  : run_lowmem<SN16>(f, out, d, mu, iterations, sc.ptr, stream);
    if (st == TSF_OK && sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return st;
}

}  // extern "C"
