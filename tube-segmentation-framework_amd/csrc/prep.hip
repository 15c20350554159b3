// Volume preparation: Gaussian blur, toFloat, vector-field creation, combine,
// candidate threshold.  All are single-pass HBM-bound kernels.
#include <math.h>
#include <string.h>

#include "tsf_common.hpp"

namespace tsf {

// ---------------------------------------------------------------- blur
// Output tile per workgroup: BX x BY x BZ voxels; the input tile with an
// m-voxel apron is staged in LDS once (clamp-to-edge reads), then every
// thread sums the (2m+1)^3 products in the reference's order (x offset
// outermost, z offset innermost) so the float result is the reference's.
constexpr int BX = 16, BY = 8, BZ = 4;
constexpr int MMAX = 5;
constexpr int TILE_MAX = (BX + 2 * MMAX) * (BY + 2 * MMAX) * (BZ + 2 * MMAX);

__global__ void __launch_bounds__(BX* BY* BZ) k_blur(const float* __restrict__ in, float* __restrict__ out, Dims d,
                                                     int m, const float* __restrict__ mask) {
    __shared__ float tile[TILE_MAX];
    const int TX = BX + 2 * m, TY = BY + 2 * m, TZ = BZ + 2 * m;
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);  // tiles sharing aprons meet in one L2
    const int x0 = tbx * BX - m, y0 = tby * BY - m, z0 = tbz * BZ - m;
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    const int total = TX * TY * TZ;
    for (int i = tid; i < total; i += BX * BY * BZ) {
        int tx = i % TX, ty = (i / TX) % TY, tz = i / (TX * TY);
        int gx = clampi(x0 + tx, 0, d.X - 1), gy = clampi(y0 + ty, 0, d.Y - 1), gz = clampi(z0 + tz, 0, d.Z - 1);
        tile[i] = in[d.at(gx, gy, gz)];
    }
    __syncthreads();
    const int x = tbx * BX + threadIdx.x, y = tby * BY + threadIdx.y, z = tbz * BZ + threadIdx.z;
    if (x >= d.X || y >= d.Y || z >= d.Z) return;
    const int M = 2 * m + 1;
    float sum = 0.0f;
    for (int a = 0; a < M; a++)
        for (int b = 0; b < M; b++)
            for (int c = 0; c < M; c++) {
                float w = mask[a + b * M + c * M * M];
                float v = tile[(threadIdx.x + a) + TX * ((threadIdx.y + b) + TY * (threadIdx.z + c))];
                sum += w * v;
            }
    out[d.at(x, y, z)] = sum;
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
    const int M = 2 * m + 1;
    Scratch sc;
    // the mask is small; stage it with a blocking copy into scratch
    TSF_TRY(sc.get(nullptr, 0, sizeof(float) * M * M * M, W));
    hipError_t e = hipMemcpyAsync(sc.ptr, hmask, sizeof(float) * M * M * M, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    Dims d = D(dims);
    dim3 grid((d.X + BX - 1) / BX, (d.Y + BY - 1) / BY, (d.Z + BZ - 1) / BZ);
    hipLaunchKernelGGL(k_blur, grid, dim3(BX, BY, BZ), 0, stream, in, out, d, m, (const float*)sc.ptr);
    TSF_TRY(check_launch(W));
    e = hipStreamSynchronize(stream);  // scratch is freed on return
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    return TSF_OK;
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
