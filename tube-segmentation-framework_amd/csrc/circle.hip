// Circle-fitting tube detection filter (TDF).
//
// circleFittingTDF (kernels_no_3d_write.cl:1503-1572) samples the field with
// trilinear interpolation at 32 points per radius: 8 gathered texels per
// sample, each needing x, y, z (and w for the normalised Hessian).  The field
// arrives as SoA planes, which would make every texel three or four separate
// gathers, so a first pass packs it into an AoS "texel" array in workspace
// (float4 for F32 fields, 4 x 16 bit for F16/BF16, values unchanged) and the
// filter then fetches one 16-byte (or 8-byte) texel per tap.  Workgroups
// cover compact 16 x 4 x 4 voxel bricks so neighbouring threads sample
// neighbouring circles out of the same L1/L2 lines.
#include "eigen3.hpp"
#include "field.hpp"

namespace tsf {

template <class S>
struct Texel;
template <>
struct Texel<F32> {
    typedef float4 P;
    static __device__ __forceinline__ float4 unpack(P p, float) { return p; }
};
template <>
struct Texel<F16> {
    typedef uint2 P;
    static __device__ __forceinline__ float h(unsigned b) { return __half2float(__ushort_as_half((unsigned short)b)); }
    static __device__ __forceinline__ float4 unpack(P p, float inv) {
        return make_float4(h(p.x & 0xffffu) * inv, h(p.x >> 16) * inv, h(p.y & 0xffffu) * inv, h(p.y >> 16) * inv);
    }
};
template <>
struct Texel<BF16> {
    typedef uint2 P;
    static __device__ __forceinline__ float b(unsigned v) { return __uint_as_float(v << 16); }
    static __device__ __forceinline__ float4 unpack(P p, float inv) {
        return make_float4(b(p.x & 0xffffu) * inv, b(p.x >> 16) * inv, b(p.y & 0xffffu) * inv, b(p.y >> 16) * inv);
    }
};

// Pack SoA planes (raw stored values) into texels.  w == NULL packs 1.
__global__ void k_pack_f32(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                           const float* __restrict__ w, float4* __restrict__ out, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        out[i] = make_float4(x[i], y[i], z[i], w ? w[i] : 1.0f);
}
__global__ void k_pack_16(const uint16_t* __restrict__ x, const uint16_t* __restrict__ y,
                          const uint16_t* __restrict__ z, const uint16_t* __restrict__ w, uint16_t one,
                          uint2* __restrict__ out, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const unsigned ww = w ? w[i] : one;
        out[i] = make_uint2((unsigned)x[i] | ((unsigned)y[i] << 16), (unsigned)z[i] | (ww << 16));
    }
}

template <class S>
struct Tex {
    const typename Texel<S>::P* __restrict__ p;
    float inv;
    __device__ __forceinline__ float4 at(i64 i) const { return Texel<S>::unpack(p[i], inv); }
};

template <class S, bool NORM>
__device__ __forceinline__ void hessian_tex(const Tex<S>& T, const Dims& d, int x, int y, int z, float H[3][3]) {
    const int xm = max(x - 1, 0), xp = min(x + 1, d.X - 1);
    const int ym = max(y - 1, 0), yp = min(y + 1, d.Y - 1);
    const int zm = max(z - 1, 0), zp = min(z + 1, d.Z - 1);
    const float4 XP = T.at(d.at(xp, y, z)), XM = T.at(d.at(xm, y, z));
    const float4 YP = T.at(d.at(x, yp, z)), YM = T.at(d.at(x, ym, z));
    const float4 ZP = T.at(d.at(x, y, zp)), ZM = T.at(d.at(x, y, zm));
    float a0 = XP.x, a1 = XM.x;
    float b0 = XP.y, b1 = XM.y, b2 = YP.y, b3 = YM.y;
    float c0 = XP.z, c1 = XM.z, c2 = YP.z, c3 = YM.z, c4 = ZP.z, c5 = ZM.z;
    if (NORM) {
        a0 = a0 / XP.w; a1 = a1 / XM.w;
        b0 = b0 / XP.w; b1 = b1 / XM.w; b2 = b2 / YP.w; b3 = b3 / YM.w;
        c0 = c0 / XP.w; c1 = c1 / XM.w; c2 = c2 / YP.w; c3 = c3 / YM.w; c4 = c4 / ZP.w; c5 = c5 / ZM.w;
    }
    const float fxx = 0.5f * (a0 - a1);
    const float fyx = 0.5f * (b0 - b1), fyy = 0.5f * (b2 - b3);
    const float fzx = 0.5f * (c0 - c1), fzy = 0.5f * (c2 - c3), fzz = 0.5f * (c4 - c5);
    H[0][0] = fxx; H[0][1] = fyx; H[0][2] = fzx;
    H[1][0] = fyx; H[1][1] = fyy; H[1][2] = fzy;
    H[2][0] = fzx; H[2][1] = fzy; H[2][2] = fzz;
}

// OpenCL CLK_FILTER_LINEAR, unnormalised coordinates, clamp-to-edge: sample
// at u - 0.5, the eight terms in the OpenCL specification's order.
template <class S>
__device__ __forceinline__ void trilinear_tex(const Tex<S>& T, const Dims& d, float u, float v, float w,
                                              float out[3]) {
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    const int i1 = clampi(i0 + 1, 0, d.X - 1), j1 = clampi(j0 + 1, 0, d.Y - 1), k1 = clampi(k0 + 1, 0, d.Z - 1);
    i0 = clampi(i0, 0, d.X - 1);
    j0 = clampi(j0, 0, d.Y - 1);
    k0 = clampi(k0, 0, d.Z - 1);
    const i64 r00 = (i64)d.X * ((i64)j0 + (i64)d.Y * k0), r10 = (i64)d.X * ((i64)j1 + (i64)d.Y * k0);
    const i64 r01 = (i64)d.X * ((i64)j0 + (i64)d.Y * k1), r11 = (i64)d.X * ((i64)j1 + (i64)d.Y * k1);
    const float4 t0 = T.at(r00 + i0), t1 = T.at(r00 + i1), t2 = T.at(r10 + i0), t3 = T.at(r10 + i1);
    const float4 t4 = T.at(r01 + i0), t5 = T.at(r01 + i1), t6 = T.at(r11 + i0), t7 = T.at(r11 + i1);
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    float sx = wt0 * t0.x, sy = wt0 * t0.y, sz = wt0 * t0.z;
    sx = sx + wt1 * t1.x; sy = sy + wt1 * t1.y; sz = sz + wt1 * t1.z;
    sx = sx + wt2 * t2.x; sy = sy + wt2 * t2.y; sz = sz + wt2 * t2.z;
    sx = sx + wt3 * t3.x; sy = sy + wt3 * t3.y; sz = sz + wt3 * t3.z;
    sx = sx + wt4 * t4.x; sy = sy + wt4 * t4.y; sz = sz + wt4 * t4.z;
    sx = sx + wt5 * t5.x; sy = sy + wt5 * t5.y; sz = sz + wt5 * t5.z;
    sx = sx + wt6 * t6.x; sy = sy + wt6 * t6.y; sz = sz + wt6 * t6.z;
    sx = sx + wt7 * t7.x; sy = sy + wt7 * t7.y; sz = sz + wt7 * t7.z;
    out[0] = sx;
    out[1] = sy;
    out[2] = sz;
}

constexpr int CBX = 16, CBY = 4, CBZ = 4;

// circleFittingTDF, kernels_no_3d_write.cl:1503-1572.
template <class S, bool NORM>
__global__ void __launch_bounds__(CBX* CBY* CBZ) k_tdf(Tex<S> T, Dims d, float rmin, float rmax, float rstep,
                                                       CircleTables tab, float* __restrict__ tdf,
                                                       float* __restrict__ rad) {
    const int x = blockIdx.x * CBX + threadIdx.x, y = blockIdx.y * CBY + threadIdx.y;
    const int z = blockIdx.z * CBZ + threadIdx.z;
    if (x >= d.X || y >= d.Y || z >= d.Z) return;
    float H[3][3], V[3][3], ev[3];
    hessian_tex<S, NORM>(T, d, x, y, z, H);
    eigen3(H, V, ev);
    const float e2x = V[0][1], e2y = V[1][1], e2z = V[2][1];
    const float e3x = V[0][2], e3y = V[1][2], e3z = V[2][2];
    float maxSum = 0.0f, maxRadius = 0.0f;
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    for (float radius = rmin; radius <= rmax; radius += rstep) {
        float sum = 0.0f;
        for (int j = 0; j < 32; j++) {
            const float cj = tab.c[j], sj = tab.s[j];
            const float vx = cj * e3x + sj * e2x, vy = cj * e3y + sj * e2y, vz = cj * e3z + sj * e2z;
            float s[3];
            trilinear_tex<S>(T, d, fx + radius * vx, fy + radius * vy, fz + radius * vz, s);
            sum += (-s[0]) * vx + (-s[1]) * vy + (-s[2]) * vz;
        }
        sum /= 32.0f;
        if (sum > maxSum) {
            maxSum = sum;
            maxRadius = radius;
        } else {
            break;
        }
    }
    const i64 i = d.at(x, y, z);
    tdf[i] = maxSum;
    rad[i] = maxRadius;
}

static unsigned grid1d(i64 n) {
    i64 g = (n + 255) / 256;
    return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

template <class S, bool NORM>
static void launch_tdf(const void* tex, float inv, Dims d, float rmin, float rmax, float rstep, float* tdf,
                       float* radius, hipStream_t stream) {
    static const CircleTables tab = make_circle_tables();
    Tex<S> T{(const typename Texel<S>::P*)tex, inv};
    dim3 grid((d.X + CBX - 1) / CBX, (d.Y + CBY - 1) / CBY, (d.Z + CBZ - 1) / CBZ), block(CBX, CBY, CBZ);
    hipLaunchKernelGGL((k_tdf<S, NORM>), grid, block, 0, stream, T, d, rmin, rmax, rstep, tab, tdf, radius);
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_tdf_circle_workspace_size(tsf_dims dims, int32_t storage) {
    if (storage < 0 || storage > 2) return 0;
    return (size_t)D(dims).n() * (storage == TSF_STORAGE_F32 ? 16 : 8);
}

tsf_status tsf_tdf_circle(const tsf_field* field, tsf_dims dims, float rmin, float rmax, float rstep, float* tdf,
                          float* radius, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_tdf_circle";
    const bool norm = !(rmax < 4.0f);
    TSF_TRY(check_field(field, norm, W));
    TSF_TRY(check_dims(dims, 1, W));
    if (!tdf || !radius) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL output", W);
    if (!(rstep > 0.0f) || !__builtin_isfinite(rmin) || !__builtin_isfinite(rmax))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: radius range must be finite with rstep > 0", W);
    if ((double)(rmax - rmin) / rstep > 100000.0)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: more than 1e5 radii", W);
    Dims d = D(dims);
    const i64 n = d.n();
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_tdf_circle_workspace_size(dims, field->storage), W));
    const float inv = 1.0f / field->scale;
    switch (field->storage) {
        case TSF_STORAGE_F32:
            hipLaunchKernelGGL(k_pack_f32, grid1d(n), 256, 0, stream, (const float*)field->x,
                               (const float*)field->y, (const float*)field->z, (const float*)field->w,
                               (float4*)sc.ptr, n);
            if (norm) launch_tdf<F32, true>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            else launch_tdf<F32, false>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            break;
        case TSF_STORAGE_F16: {
            const uint16_t one = 0x3c00;  // only read when w is absent and the Hessian is not normalised
            hipLaunchKernelGGL(k_pack_16, grid1d(n), 256, 0, stream, (const uint16_t*)field->x,
                               (const uint16_t*)field->y, (const uint16_t*)field->z, (const uint16_t*)field->w, one,
                               (uint2*)sc.ptr, n);
            if (norm) launch_tdf<F16, true>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            else launch_tdf<F16, false>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            break;
        }
        default: {
            const uint16_t one = 0x3f80;
            hipLaunchKernelGGL(k_pack_16, grid1d(n), 256, 0, stream, (const uint16_t*)field->x,
                               (const uint16_t*)field->y, (const uint16_t*)field->z, (const uint16_t*)field->w, one,
                               (uint2*)sc.ptr, n);
            if (norm) launch_tdf<BF16, true>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            else launch_tdf<BF16, false>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
        }
    }
    TSF_TRY(check_launch(W));
    if (sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

}  // extern "C"
