// Volume preparation: Gaussian blur, toFloat, vector-field creation, combine,
// candidate threshold.  All are single-pass HBM-bound kernels.
#include <math.h>
#include <string.h>

#include "tsf_common.hpp"

namespace tsf {

// ---------------------------------------------------------------- blur
// A workgroup of 32 x 8 threads owns a 32 x 8 x BZV voxel brick: each thread
// one (x, y) column of BZV consecutive voxels in z.  The brick's input with
// an m-voxel apron (clamp-to-edge reads) is staged in LDS; every load is
// issued before the first LDS store, so one memory latency covers them all.
// For every (a, b) offset a thread reads the BZV + 2m values of its shifted
// column once and feeds them to its BZV sums; each voxel's sum still takes
// its (2m+1)^3 products in the reference's order (x offset outermost, z
// offset innermost), so the float result is the reference's.  m is a
// template parameter (1..5): the loops unroll and the mask weights are
// uniform scalars, passed in the kernel arguments (m <= 4, up to 729
// weights) or in device memory (m = 5).
constexpr int BX = 32, BY = 8;
// voxels per thread in z: small bricks for the 27-point mask keep the
// registers (and so the waves per SIMD that hide the loads) in bounds
__host__ __device__ constexpr int blur_zv(int m) { return m == 1 ? 4 : 8; }

template <int M>
struct BlurMask {
    float w[M * M * M];
};

template <int m, bool ARG, bool BUF>
__global__ void __launch_bounds__(BX* BY) k_blur(const float* __restrict__ in, float* __restrict__ out, Dims d,
                                                 BlurMask<ARG ? 2 * m + 1 : 1> W, const float* __restrict__ gw) {
    constexpr int BZV = blur_zv(m), M = 2 * m + 1, TX = BX + 2 * m, TY = BY + 2 * m, TZ = BZV + 2 * m;
    constexpr int NY = (TY + BY - 1) / BY;  // tile rows per thread
    __shared__ float tile[TX * TY * TZ];
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);  // bricks sharing aprons meet in one L2
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = tbx * BX - m, y0 = tby * BY - m, zb = tbz * BZV, z0 = zb - m;
    // Thread row ly = ty (+ BY, ...) takes every plane of its row: the
    // lane's x and the row's y are clamped once, each plane's offset is
    // uniform.  Lanes tx < 2m also fetch the row's apron columns.
    const int gx = min(tbx * BX + tx, d.X - 1);
    const int alx = tx < m ? tx : BX + tx;  // apron column of lanes tx < 2m
    const int ax = clampi(x0 + alx, 0, d.X - 1);
    const bool apron = tx < 2 * m;
    const i64 plane = (i64)d.X * d.Y;
    // BUF: buffer loads from a base at the brick's first clamped plane, with
    // a per-lane 32-bit row offset and a uniform plane offset (the host
    // picks this form when TZ planes span less than 2 GiB)
    const int zlo = clampi(z0, 0, d.Z - 1);
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(in + plane * zlo), 0, (int)(BUF ? TZ * plane * 4 : 0), BUFFER_WORD3);
    auto rowok = [&](int j) { return j == 0 || ty + BY * j < TY; };
    auto ld = [&](int ry, int lz, int col) -> float {
        const int gz = clampi(z0 + lz, 0, d.Z - 1);
        if constexpr (BUF)
            return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rin, (int)(4u * (unsigned)(ry * d.X + col)),
                                                                        (int)(4 * plane * (gz - zlo)), 0));
        else
            return in[plane * gz + (i64)ry * d.X + col];
    };
    float v[NY][TZ], va[NY][TZ];
#pragma unroll
    for (int j = 0; j < NY; j++)
        if (rowok(j)) {
            const int ry = clampi(y0 + ty + BY * j, 0, d.Y - 1);
#pragma unroll
            for (int lz = 0; lz < TZ; lz++) v[j][lz] = ld(ry, lz, gx);
        }
    if (apron)
#pragma unroll
        for (int j = 0; j < NY; j++)
            if (rowok(j)) {
                const int ry = clampi(y0 + ty + BY * j, 0, d.Y - 1);
#pragma unroll
                for (int lz = 0; lz < TZ; lz++) va[j][lz] = ld(ry, lz, ax);
            }
#pragma unroll
    for (int j = 0; j < NY; j++)
        if (rowok(j))
#pragma unroll
            for (int lz = 0; lz < TZ; lz++) tile[TX * (ty + BY * j + TY * lz) + m + tx] = v[j][lz];
    if (apron)
#pragma unroll
        for (int j = 0; j < NY; j++)
            if (rowok(j))
#pragma unroll
                for (int lz = 0; lz < TZ; lz++) tile[TX * (ty + BY * j + TY * lz) + alx] = va[j][lz];
    __syncthreads();

    const int x = tbx * BX + tx, y = tby * BY + ty;
    if (x >= d.X || y >= d.Y) return;
    float sum[BZV];
#pragma unroll
    for (int q = 0; q < BZV; q++) sum[q] = 0.0f;
    // m = 1 unrolls completely (the weights stay in scalar registers); for
    // larger masks the x offset loop stays rolled so that only the weights
    // of one x offset are live at once
    constexpr int UA = m == 1 ? M : 1;
#pragma unroll UA
    for (int a = 0; a < M; a++)
#pragma unroll
        for (int b = 0; b < M; b++) {
            const float* t = tile + (tx + a) + TX * (ty + b);
            float col[TZ];
#pragma unroll
            for (int j = 0; j < TZ; j++) col[j] = t[TX * TY * j];
#pragma unroll
            for (int c = 0; c < M; c++) {
                // The mask is a Gaussian of the squared offset, so z offsets
                // c and 2m - c carry the same weight: reading it from one
                // index lets w * col[j] be computed once for both (the same
                // product, so the same float)
                const int i = a + b * M + (c <= m ? c : 2 * m - c) * M * M;
                const float w = ARG ? W.w[i] : gw[i];
#pragma unroll
                for (int q = 0; q < BZV; q++) sum[q] += w * col[q + c];
            }
        }
    if constexpr (BUF) {
        const __amdgpu_buffer_rsrc_t ro =
            __builtin_amdgcn_make_buffer_rsrc((void*)(out + plane * zb), 0, (int)(BZV * plane * 4), BUFFER_WORD3);
        const unsigned vo = 4u * (unsigned)(y * d.X + x);
#pragma unroll
        for (int q = 0; q < BZV; q++)
            if (zb + q < d.Z) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(sum[q]), ro, (int)vo, (int)(4 * plane * q), 0);
    } else {
#pragma unroll
        for (int q = 0; q < BZV; q++)
            if (zb + q < d.Z) out[d.at(x, y, zb + q)] = sum[q];
    }
}

template <int m, bool ARG, bool BUF, class W>
static void launch_blur(const float* in, float* out, Dims d, const W& w, const float* dmask, hipStream_t stream) {
    constexpr int BZV = blur_zv(m);
    const dim3 grid((d.X + BX - 1) / BX, (d.Y + BY - 1) / BY, (d.Z + BZV - 1) / BZV), block(BX, BY);
    hipLaunchKernelGGL((k_blur<m, ARG, BUF>), grid, block, 0, stream, in, out, d, w, dmask);
}

template <int m>
static void launch_blur(const float* in, float* out, Dims d, const float* hmask, const float* dmask,
                        hipStream_t stream) {
    constexpr int M = 2 * m + 1, TZ = blur_zv(m) + 2 * m;
    const bool buf = (i64)TZ * d.X * d.Y * 4 < ((i64)1 << 31);
    if constexpr (m == 5) {
        BlurMask<1> none{};
        if (buf) launch_blur<m, false, true>(in, out, d, none, dmask, stream);
        else launch_blur<m, false, false>(in, out, d, none, dmask, stream);
    } else {
        BlurMask<M> w;
        for (int i = 0; i < M * M * M; i++) w.w[i] = hmask[i];
        if (buf) launch_blur<m, true, true>(in, out, d, w, (const float*)nullptr, stream);
        else launch_blur<m, true, false>(in, out, d, w, (const float*)nullptr, stream);
    }
}

// ---------------------------------------------------------------- toFloat
template <typename TIn>
__global__ void k_to_float(const TIn* __restrict__ in, float* __restrict__ out, i64 n, float mn, float mx) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        float v = (float)in[i];
        v = v > mx ? mx : v;
        v = v < mn ? mn : v;
        out[i] = (v - mn) / (mx - mn);
    }
}

// ---------------------------------------------------------------- vector field
template <class S>
__global__ void __launch_bounds__(256) k_vector_field(const float* __restrict__ I, typename S::T* __restrict__ fx,
                                                      typename S::T* __restrict__ fy, typename S::T* __restrict__ fz,
                                                      typename S::T* __restrict__ fw, Dims d, float fmax, float sign,
                                                      float scale) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= d.X || y >= d.Y) return;
    const int xm = max(x - 1, 0), xp = min(x + 1, d.X - 1);
    const int ym = max(y - 1, 0), yp = min(y + 1, d.Y - 1);
    const int zm = max(z - 1, 0), zp = min(z + 1, d.Z - 1);
    float gx = 0.5f * (I[d.at(xp, y, z)] - I[d.at(xm, y, z)]);
    float gy = 0.5f * (I[d.at(x, yp, z)] - I[d.at(x, ym, z)]);
    float gz = 0.5f * (I[d.at(x, y, zp)] - I[d.at(x, y, zm)]);
    gx = sign * gx;
    gy = sign * gy;
    gz = sign * gz;
    const float l = sqrtf(gx * gx + gy * gy + gz * gz);
    const float s = l < fmax ? fmax : l;
    const i64 i = d.at(x, y, z);
    S::st(fx, i, gx / s, scale);
    S::st(fy, i, gy / s, scale);
    S::st(fz, i, gz / s, scale);
    S::st(fw, i, 1.0f, scale);
}

// ---------------------------------------------------------------- combine / candidates
__global__ void k_combine(const float* __restrict__ ts, const float* __restrict__ rs, float* __restrict__ tl,
                          float* __restrict__ rl, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        float a = ts[i];
        if (tl[i] < a) {
            tl[i] = a;
            rl[i] = rs[i];
        }
    }
}

// The reference's 16-bit TDF format (kernels_no_3d_write.cl:19-20, 1570):
// stored as convert_ushort_sat_rte(t * 65535), read as c / 65535.  The
// quotient uses one fma correction of c * (1/65535), which equals the
// correctly rounded c / 65535 for every code.
__device__ __forceinline__ float unorm16_round(float t) {
    const float x = t * 65535.0f;
    const float c = x == x ? rintf(fminf(fmaxf(x, 0.0f), 65535.0f)) : 0.0f;
    const float r = 1.0f / 65535.0f;
    const float q0 = c * r;
    return __builtin_fmaf(__builtin_fmaf(-q0, 65535.0f, c), r, q0);
}

__global__ void k_unorm16_round(float* __restrict__ t, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        t[i] = unorm16_round(t[i]);
}

__global__ void k_candidates(const float* __restrict__ tdf, uint8_t* __restrict__ mask, i64 n, float thr) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        mask[i] = tdf[i] < thr ? 0 : 1;
}

static unsigned grid1d(i64 n, int block) {
    i64 g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    return (unsigned)(g < 1 ? 1 : g);
}

// Host mask, as createBlurMask computes it (float, expf).
static void blur_mask(float sigma, float* mask, int* maskSize) {
    int m = (int)ceilf(sigma / 0.5f);
    if (m < 1) m = 1;
    if (m > 5) m = 5;
    const int M = 2 * m + 1;
    float sum = 0.0f;
    for (int a = -m; a <= m; a++)
        for (int b = -m; b <= m; b++)
            for (int c = -m; c <= m; c++) {
                float w = expf(-((float)(a * a + b * b + c * c) / (2 * sigma * sigma)));
                sum += w;
                mask[(a + m) + (b + m) * M + (c + m) * M * M] = w;
            }
    for (int i = 0; i < M * M * M; i++) mask[i] = mask[i] / sum;
    *maskSize = m;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

tsf_status tsf_create_blur_mask(float sigma, float* mask, int32_t* mask_size) {
    if (!mask || !mask_size) return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_create_blur_mask: NULL argument");
    if (!(sigma > 0.0f) || !__builtin_isfinite(sigma))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "tsf_create_blur_mask: sigma must be positive");
    int m;
    blur_mask(sigma, mask, &m);
    *mask_size = m;
    return TSF_OK;
}

tsf_status tsf_blur_gaussian(const float* in, float* out, tsf_dims dims, float sigma, hipStream_t stream) {
    const char* W = "tsf_blur_gaussian";
    if (!in || !out) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL volume", W);
    if (in == out) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: in and out must differ", W);
    TSF_TRY(check_dims(dims, 1, W));
    float hmask[11 * 11 * 11];
    int m;
    TSF_TRY(tsf_create_blur_mask(sigma, hmask, &m));
    const Dims d = D(dims);
    switch (m) {
        case 1: launch_blur<1>(in, out, d, hmask, nullptr, stream); break;
        case 2: launch_blur<2>(in, out, d, hmask, nullptr, stream); break;
        case 3: launch_blur<3>(in, out, d, hmask, nullptr, stream); break;
        case 4: launch_blur<4>(in, out, d, hmask, nullptr, stream); break;
        default: {
            // 1331 weights: staged in device memory with a blocking copy (the
            // scratch is freed on return)
            const size_t bytes = sizeof(float) * 11 * 11 * 11;
            Scratch sc;
            TSF_TRY(sc.get(nullptr, 0, bytes, W));
            hipError_t e = hipMemcpyAsync(sc.ptr, hmask, bytes, hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
            launch_blur<5>(in, out, d, nullptr, (const float*)sc.ptr, stream);
            TSF_TRY(check_launch(W));
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
            return TSF_OK;
        }
    }
    return check_launch(W);
}

tsf_status tsf_to_float(const void* in, int32_t in_type, float* out, tsf_dims dims, float minimum, float maximum,
                        hipStream_t stream) {
    const char* W = "tsf_to_float";
    if (!in || !out) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL volume", W);
    TSF_TRY(check_dims(dims, 1, W));
    if (!(maximum != minimum)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: maximum == minimum", W);
    i64 n = D(dims).n();
    unsigned g = grid1d(n, 256);
    switch (in_type) {
        case 0: hipLaunchKernelGGL(k_to_float<uint8_t>, g, 256, 0, stream, (const uint8_t*)in, out, n, minimum, maximum); break;
        case 1: hipLaunchKernelGGL(k_to_float<int8_t>, g, 256, 0, stream, (const int8_t*)in, out, n, minimum, maximum); break;
        case 2: hipLaunchKernelGGL(k_to_float<uint16_t>, g, 256, 0, stream, (const uint16_t*)in, out, n, minimum, maximum); break;
        case 3: hipLaunchKernelGGL(k_to_float<int16_t>, g, 256, 0, stream, (const int16_t*)in, out, n, minimum, maximum); break;
        case 4: hipLaunchKernelGGL(k_to_float<float>, g, 256, 0, stream, (const float*)in, out, n, minimum, maximum); break;
        default: return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: unknown input type %d", W, in_type);
    }
    return check_launch(W);
}

tsf_status tsf_create_vector_field(const float* volume, const tsf_field* out, tsf_dims dims, float fmax, int32_t sign,
                                   hipStream_t stream) {
    const char* W = "tsf_create_vector_field";
    if (!volume) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL volume", W);
    TSF_TRY(check_field(out, true, W));
    TSF_TRY(check_dims(dims, 1, W));
    if (!(fmax > 0.0f)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: fmax must be positive", W);
    if (sign != 1 && sign != -1) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: sign must be +1 or -1", W);
    Dims d = D(dims);
    dim3 grid((d.X + 63) / 64, (d.Y + 3) / 4, d.Z);
    float sg = (float)sign;
    switch (out->storage) {
        case TSF_STORAGE_F32:
            hipLaunchKernelGGL(k_vector_field<F32>, grid, dim3(64, 4), 0, stream, volume, (float*)out->x,
                               (float*)out->y, (float*)out->z, (float*)out->w, d, fmax, sg, out->scale);
            break;
        case TSF_STORAGE_F16:
            hipLaunchKernelGGL(k_vector_field<F16>, grid, dim3(64, 4), 0, stream, volume, (__half*)out->x,
                               (__half*)out->y, (__half*)out->z, (__half*)out->w, d, fmax, sg, out->scale);
            break;
        case TSF_STORAGE_SN16:
            hipLaunchKernelGGL(k_vector_field<SN16>, grid, dim3(64, 4), 0, stream, volume, (short*)out->x,
                               (short*)out->y, (short*)out->z, (short*)out->w, d, fmax, sg, out->scale);
            break;
        default:
            hipLaunchKernelGGL(k_vector_field<BF16>, grid, dim3(64, 4), 0, stream, volume,
                               (__hip_bfloat16*)out->x, (__hip_bfloat16*)out->y, (__hip_bfloat16*)out->z,
                               (__hip_bfloat16*)out->w, d, fmax, sg, out->scale);
    }
    return check_launch(W);
}

tsf_status tsf_combine(const float* ts, const float* rs, float* tl, float* rl, int64_t count, hipStream_t stream) {
    const char* W = "tsf_combine";
    if (!ts || !rs || !tl || !rl) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", W);
    if (count < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative count", W);
    if (count == 0) return TSF_OK;
    hipLaunchKernelGGL(k_combine, grid1d(count, 256), 256, 0, stream, ts, rs, tl, rl, (i64)count);
    return check_launch(W);
}

tsf_status tsf_tdf_round_unorm16(float* tdf, int64_t count, hipStream_t stream) {
    const char* W = "tsf_tdf_round_unorm16";
    if (!tdf) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", W);
    if (count < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative count", W);
    if (count == 0) return TSF_OK;
    hipLaunchKernelGGL(k_unorm16_round, grid1d(count, 256), 256, 0, stream, tdf, (i64)count);
    return check_launch(W);
}

tsf_status tsf_candidates(const float* tdf, uint8_t* mask, int64_t count, float threshold, hipStream_t stream) {
    const char* W = "tsf_candidates";
    if (!tdf || !mask) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", W);
    if (count < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative count", W);
    if (count == 0) return TSF_OK;
    hipLaunchKernelGGL(k_candidates, grid1d(count, 256), 256, 0, stream, tdf, mask, (i64)count, threshold);
    return check_launch(W);
}

}  // extern "C"
