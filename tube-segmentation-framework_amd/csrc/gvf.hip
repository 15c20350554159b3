// Gradient vector flow, explicit Euler (Jacobi) iterations.
//
// Layout: f (the initial field) and v (the iterate) are SoA planes.  One
// launch advances the iterate by one step; every thread owns one (x, y)
// column and marches a run of ZRUN planes in z, keeping the three z-planes
// of its column in registers, so each v value is fetched from memory once per
// step by its owner plus the x/y neighbour reads that hit L1/L2.
#include "tsf_common.hpp"

namespace tsf {

constexpr int GX = 64, GY = 4, ZRUN = 8;

template <class S>
struct Planes3 {
    const typename S::T* x;
    const typename S::T* y;
    const typename S::T* z;
};
template <class S>
struct Out4 {
    typename S::T* x;
    typename S::T* y;
    typename S::T* z;
    typename S::T* w;
};

// One Jacobi step of GVF3DIteration (kernels_no_3d_write.cl:1746-1787).  The
// arithmetic order follows the reference:
//   lap = -6 v + v(x+1) + v(x-1) + v(y+1) + v(y-1) + v(z+1) + v(z-1)
//   v'  = v + (mu lap - (v - f) |f|^2)
// With FINISH the step also writes w = |v'| (or 1), GVF3DFinish 1798-1807.
template <class S, bool FINISH>
__global__ void __launch_bounds__(GX* GY) k_gvf_step(Planes3<S> f, Planes3<S> vin, Out4<S> vout, Dims d, float mu,
                                                     float inv_f, float inv_v, float sc_v) {
    const int x = blockIdx.x * GX + threadIdx.x;
    const int y = blockIdx.y * GY + threadIdx.y;
    if (x >= d.X || y >= d.Y) return;
    const int qx = remap(x, d.X), qy = remap(y, d.Y);
    const i64 sx = 1, sy = d.X, sz = (i64)d.X * d.Y;
    const int z0 = blockIdx.z * ZRUN;
    const int z1 = min(z0 + ZRUN, d.Z);
    for (int z = z0; z < z1; z++) {
        const int qz = remap(z, d.Z);
        const i64 q = d.at(qx, qy, qz);
        const float ix = S::ld(f.x, q, inv_f), iy = S::ld(f.y, q, inv_f), iz = S::ld(f.z, q, inv_f);
        const float m2 = ix * ix + iy * iy + iz * iz;
        float r[3];
        const typename S::T* P[3] = {vin.x, vin.y, vin.z};
        const float F[3] = {ix, iy, iz};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const typename S::T* p = P[c];
            const float vc = S::ld(p, q, inv_v);
            float lap = -6.0f * vc;
            lap = lap + S::ld(p, q + sx, inv_v);
            lap = lap + S::ld(p, q - sx, inv_v);
            lap = lap + S::ld(p, q + sy, inv_v);
            lap = lap + S::ld(p, q - sy, inv_v);
            lap = lap + S::ld(p, q + sz, inv_v);
            lap = lap + S::ld(p, q - sz, inv_v);
            r[c] = vc + (mu * lap - (vc - F[c]) * m2);
        }
        const i64 w = d.at(x, y, z);
        S::st(vout.x, w, r[0], sc_v);
        S::st(vout.y, w, r[1], sc_v);
        S::st(vout.z, w, r[2], sc_v);
        if (FINISH) {
            // the magnitude is taken of the stored (rounded) components
            float a = S::ld(vout.x, w, inv_v), b = S::ld(vout.y, w, inv_v), c = S::ld(vout.z, w, inv_v);
            float l = sqrtf(a * a + b * b + c * c);
            S::st(vout.w, w, l > 0.0f ? l : 1.0f, sc_v);
        }
    }
}

// iterations == 0: v = f, w = |f| or 1.
template <class S>
__global__ void k_gvf_copy_finish(Planes3<S> f, Out4<S> vout, i64 n, float inv_f, float inv_v, float sc_v) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        S::st(vout.x, i, S::ld(f.x, i, inv_f), sc_v);
        S::st(vout.y, i, S::ld(f.y, i, inv_f), sc_v);
        S::st(vout.z, i, S::ld(f.z, i, inv_f), sc_v);
        float a = S::ld(vout.x, i, inv_v), b = S::ld(vout.y, i, inv_v), c = S::ld(vout.z, i, inv_v);
        float l = sqrtf(a * a + b * b + c * c);
        S::st(vout.w, i, l > 0.0f ? l : 1.0f, sc_v);
    }
}

template <class S>
static tsf_status run_euler(const tsf_field* f, const tsf_field* out, Dims d, float mu, int iters, void* ws,
                            hipStream_t stream) {
    typedef typename S::T T;
    const i64 N = d.n();
    const float inv_f = 1.0f / f->scale, inv_v = 1.0f / out->scale, sc_v = out->scale;
    Planes3<S> F{(const T*)f->x, (const T*)f->y, (const T*)f->z};
    Out4<S> O{(T*)out->x, (T*)out->y, (T*)out->z, (T*)out->w};
    if (iters == 0) {
        hipLaunchKernelGGL(k_gvf_copy_finish<S>, 4096, 256, 0, stream, F, O, N, inv_f, inv_v, sc_v);
        return check_launch("tsf_gvf_euler");
    }
    T* A = (T*)ws;
    Out4<S> Wk{A, A + N, A + 2 * N, nullptr};
    Planes3<S> Wr{A, A + N, A + 2 * N};
    Planes3<S> Or{O.x, O.y, O.z};
    dim3 grid((d.X + GX - 1) / GX, (d.Y + GY - 1) / GY, (d.Z + ZRUN - 1) / ZRUN);
    dim3 block(GX, GY);
    for (int i = 1; i <= iters; i++) {
        const bool to_out = ((iters - i) % 2) == 0;
        Planes3<S> src = (i == 1) ? F : (to_out ? Wr : Or);
        const float inv_src = (i == 1) ? inv_f : inv_v;
        if (to_out) {
            if (i == iters)
                hipLaunchKernelGGL((k_gvf_step<S, true>), grid, block, 0, stream, F, src, O, d, mu, inv_f, inv_src, sc_v);
            else
                hipLaunchKernelGGL((k_gvf_step<S, false>), grid, block, 0, stream, F, src, O, d, mu, inv_f, inv_src, sc_v);
        } else {
            hipLaunchKernelGGL((k_gvf_step<S, false>), grid, block, 0, stream, F, src, Wk, d, mu, inv_f, inv_src, sc_v);
        }
        TSF_TRY(check_launch("tsf_gvf_euler"));
    }
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_gvf_euler_workspace_size(tsf_dims dims, int32_t storage) {
    if (storage < 0 || storage > 2) return 0;
    return (size_t)3 * (size_t)D(dims).n() * storage_bytes(storage);
}

tsf_status tsf_gvf_euler(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu, int32_t iterations,
                         int32_t fuse, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_gvf_euler";
    TSF_TRY(check_field(f, false, W));
    TSF_TRY(check_field(out, true, W));
    TSF_TRY(check_dims(dims, 4, W));
    if (f->storage != out->storage)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f and out must use the same storage type", W);
    if (iterations < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative iteration count", W);
    if (fuse < 1) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: fuse must be >= 1", W);
    if (!(mu >= 0.0f) || !__builtin_isfinite(mu)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: mu must be >= 0", W);
    if (f->x == out->x || f->y == out->y || f->z == out->z)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f and out planes must not alias", W);
    Dims d = D(dims);
    Scratch sc;
    if (iterations > 1) TSF_TRY(sc.get(workspace, workspace_bytes, tsf_gvf_euler_workspace_size(dims, out->storage), W));
    tsf_status st;
    switch (out->storage) {
        case TSF_STORAGE_F32: st = run_euler<F32>(f, out, d, mu, iterations, sc.ptr, stream); break;
        case TSF_STORAGE_F16: st = run_euler<F16>(f, out, d, mu, iterations, sc.ptr, stream); break;
        default: st = run_euler<BF16>(f, out, d, mu, iterations, sc.ptr, stream); break;
    }
    if (st == TSF_OK && sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return st;
}

}  // extern "C"
