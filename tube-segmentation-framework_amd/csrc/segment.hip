// Segmentation from the centerline (segmentation.cpp:5-221, 223-307;
// kernels initGrowing, grow, dilate, erode, sphereSegmentation,
// kernels_no_3d_write.cl:1100-1234, 1370-1412).
//
// Inverse-gradient region growing: every centerline voxel seeds a cube of
// candidates (half size round(radius / 2) clamped to 1..4); a candidate X is
// accepted when some neighbour Y, not yet accepted, with a larger field
// magnitude than X's, has its normalised field pointing (among Y's 26
// neighbour directions) most nearly at X -- and Y becomes a candidate.
// Candidates that find no such Y are dropped.  The grown mask is closed
// (3x3x3 dilation, then erosion).
//
// The reference writes the next state in place from many threads: a
// candidate's own verdict and its neighbours' nominations of it land on the
// same byte in an unspecified order.  Here one pass records the verdicts
// (accepted) and the nominations apart and a second pass merges them with a
// fixed precedence, accepted > nominated > dropped: accepted voxels stay
// accepted and every pass that grows accepts a new voxel, so the iteration
// ends (as in the reference: at least two passes, then until a pass grows
// nothing).  The dilation starts from an empty volume (the image path
// initialises it; the buffer path leaves it uninitialised).
#include "field.hpp"

namespace tsf {

__device__ __forceinline__ bool inside(const Dims& d, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < d.X && y < d.Y && z < d.Z;
}
// clamp-to-edge read (the reference's image sampler)
__device__ __forceinline__ i64 clamped(const Dims& d, int x, int y, int z) {
    return d.at(clampi(x, 0, d.X - 1), clampi(y, 0, d.Y - 1), clampi(z, 0, d.Z - 1));
}

#define FOR_VOXEL(d)                                                              \
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y; \
    const int z = blockIdx.z;                                                     \
    if (x >= d.X || y >= d.Y) return;                                             \
    const i64 i = d.at(x, y, z);

// initGrowing: `out` holds a copy of the centerline; the seeds become 2.
__global__ void k_seg_seed(const uint8_t* __restrict__ cl, const float* __restrict__ radius, uint8_t* out, Dims d) {
    FOR_VOXEL(d)
    if (cl[i] != 1) return;
    const int N = min(max(1, (int)roundf(radius[i] / 2.0f)), 4);
    for (int a = -N; a <= N; a++)
        for (int b = -N; b <= N; b++)
            for (int c = -N; c <= N; c++) {
                const int nx = x + a, ny = y + b, nz = z + c;
                if (cl[clamped(d, nx, ny, nz)] == 0 && inside(d, nx, ny, nz)) out[d.at(nx, ny, nz)] = 2;
            }
}

// grow, first half: the verdicts of the candidates and their nominations
template <class S>
__device__ __forceinline__ void decide_voxel(const uint8_t* __restrict__ cur, const Field<S>& F, const Dims& d, int x,
                                             int y, int z, i64 i, uint8_t* __restrict__ acc, uint8_t* nom,
                                             int32_t* grew) {
    const float wx = F.mag(i);
    bool cont = false;
    for (int a = -1; a < 2; a++)
        for (int b = -1; b < 2; b++)
            for (int c = -1; c < 2; c++) {
                if (a == 0 && b == 0 && c == 0) continue;
                const int yx = x + a, yy = y + b, yz = z + c;
                const i64 j = clamped(d, yx, yy, yz);
                if (cur[j] == 1) continue;
                const float wy = F.mag(j);
                const float fx = F.c(0, j) / wy, fy = F.c(1, j) / wy, fz = F.c(2, j) / wy;
                if (!(wy > wx)) continue;
                // the neighbour direction of Y its field points at most
                float best = -2.0f;
                int za = 0, zb = 0, zc = 0;
                for (int a2 = -1; a2 < 2; a2++)
                    for (int b2 = -1; b2 < 2; b2++)
                        for (int c2 = -1; c2 < 2; c2++) {
                            if (a2 == 0 && b2 == 0 && c2 == 0) continue;
                            const float l = sqrtf((float)(a2 * a2 + b2 * b2 + c2 * c2));
                            const float v = fx * ((float)a2 / l) + fy * ((float)b2 / l) + fz * ((float)c2 / l);
                            if (v > best) {
                                best = v;
                                za = a2;
                                zb = b2;
                                zc = c2;
                            }
                        }
                if (a + za == 0 && b + zb == 0 && c + zc == 0 && inside(d, yx, yy, yz)) {
                    nom[d.at(yx, yy, yz)] = 1;
                    cont = true;
                }
            }
    if (cont) {
        acc[i] = 1;
        *grew = 1;
    }
}

template <class S>
__global__ void k_seg_decide(const uint8_t* __restrict__ cur, Field<S> F, Dims d, uint8_t* __restrict__ acc,
                             uint8_t* nom, int32_t* grew) {
    FOR_VOXEL(d)
    if (cur[i] == 2) decide_voxel<S>(cur, F, d, x, y, z, i, acc, nom, grew);
}

// grow, second half: accepted > nominated > dropped; clears the marks.
// Elementwise over the 256-byte-aligned workspace arrays (their lengths are
// multiples of 256), 16 voxels per thread; the padding past the last voxel
// is merged too and never read as a voxel.
__device__ __forceinline__ unsigned merge_word(unsigned c, unsigned a, unsigned m) {
    unsigned o = 0;
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const unsigned cb = (c >> k) & 0xffu, ab = (a >> k) & 0xffu, mb = (m >> k) & 0xffu;
        o |= (cb == 1u || ab ? 1u : (mb ? 2u : 0u)) << k;
    }
    return o;
}
__global__ void __launch_bounds__(256) k_seg_merge(const uint4* __restrict__ cur, uint4* __restrict__ acc,
                                                   uint4* __restrict__ nom, uint4* __restrict__ next, i64 n16) {
    const i64 t = (i64)blockIdx.x * 256 + threadIdx.x;
    if (t >= n16) return;
    const uint4 c = cur[t], a = acc[t], m = nom[t];
    next[t] = make_uint4(merge_word(c.x, a.x, m.x), merge_word(c.y, a.y, m.y), merge_word(c.z, a.z, m.z),
                         merge_word(c.w, a.w, m.w));
    acc[t] = make_uint4(0, 0, 0, 0);
    nom[t] = make_uint4(0, 0, 0, 0);
}

// The 3x3x3 closing, one axis per pass (a box is the product of its three
// axis ranges, and both the inside-the-volume test of the dilation and the
// clamp-to-edge of the erosion act per axis):
// dilate: 1 where a voxel of the axis neighbourhood inside the volume is 1;
template <int AXIS>
__global__ void k_seg_dilate_axis(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, Dims d) {
    FOR_VOXEL(d)
    const int p = AXIS == 0 ? x : (AXIS == 1 ? y : z), e = AXIS == 0 ? d.X : (AXIS == 1 ? d.Y : d.Z);
    const i64 s = AXIS == 0 ? 1 : (AXIS == 1 ? (i64)d.X : (i64)d.X * d.Y);
    const bool v = in[i] == 1 || (p > 0 && in[i - s] == 1) || (p + 1 < e && in[i + s] == 1);
    out[i] = v ? 1 : 0;
}
// erode: 1 where the axis neighbourhood (clamp-to-edge) is all 1.
template <int AXIS>
__global__ void k_seg_erode_axis(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, Dims d) {
    FOR_VOXEL(d)
    const int p = AXIS == 0 ? x : (AXIS == 1 ? y : z), e = AXIS == 0 ? d.X : (AXIS == 1 ? d.Y : d.Z);
    const i64 s = AXIS == 0 ? 1 : (AXIS == 1 ? (i64)d.X : (i64)d.X * d.Y);
    const bool v = in[i] == 1 && in[p > 0 ? i - s : i] == 1 && in[p + 1 < e ? i + s : i] == 1;
    out[i] = v ? 1 : 0;
}

// The same closing passes on 32-bit words (4 voxels along x) when X is a
// multiple of 4, so that no word crosses a row.  Each byte of `is1` is 1
// where the input byte is 1; along x the neighbour bytes come from the
// adjacent words (little-endian: byte k is x0 + k), zero (dilation) or the
// edge byte itself (erosion) past the row ends.
__device__ __forceinline__ unsigned is1(unsigned w) {
    unsigned o = 0;
#pragma unroll
    for (int k = 0; k < 32; k += 8) o |= (((w >> k) & 0xffu) == 1u ? 1u : 0u) << k;
    return o;
}
constexpr int CLOSE_ZC = 8;  // planes per workgroup of the word passes
template <int AXIS, bool DILATE>
__global__ void __launch_bounds__(256) k_seg_close_words(const unsigned* __restrict__ in, unsigned* __restrict__ out,
                                                         Dims d) {
    const int XW = d.X >> 2;
    const int xw = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (xw >= XW || y >= d.Y) return;
    const int z1 = min(d.Z, (int)(blockIdx.z + 1) * CLOSE_ZC);
    for (int z = blockIdx.z * CLOSE_ZC; z < z1; z++) {
    const i64 i = (i64)xw + (i64)XW * ((i64)y + (i64)d.Y * z);
    const unsigned c = is1(in[i]);
    unsigned r;
    if (AXIS == 0) {
        unsigned l = xw > 0 ? is1(in[i - 1]) >> 24 : (DILATE ? 0u : c & 0xffu);
        unsigned h = xw + 1 < XW ? (is1(in[i + 1]) & 0xffu) << 24 : (DILATE ? 0u : c & 0xff000000u);
        const unsigned left = (c << 8) | l, right = (c >> 8) | h;
        r = DILATE ? (c | left | right) : (c & left & right);
    } else {
        const int p = AXIS == 1 ? y : z, e = AXIS == 1 ? d.Y : d.Z;
        const i64 st = AXIS == 1 ? (i64)XW : (i64)XW * d.Y;
        const unsigned lo = p > 0 ? is1(in[i - st]) : (DILATE ? 0u : c);
        const unsigned hi = p + 1 < e ? is1(in[i + st]) : (DILATE ? 0u : c);
        r = DILATE ? (c | lo | hi) : (c & lo & hi);
    }
    out[i] = r;
    }
}

// sphereSegmentation: a ball of the voxel's radius (|offset| < r) around
// every centerline voxel
__global__ void k_seg_spheres(const uint8_t* __restrict__ cl, const float* __restrict__ radius, uint8_t* out, Dims d) {
    FOR_VOXEL(d)
    if (cl[i] == 0) return;
    const float r = radius[i];
    const int N = (int)ceilf(r);
    for (int a = -N; a <= N; a++)
        for (int b = -N; b <= N; b++)
            for (int c = -N; c <= N; c++)
                if (sqrtf((float)a * (float)a + (float)b * (float)b + (float)c * (float)c) < r &&
                    inside(d, x + a, y + b, z + c))
                    out[d.at(x + a, y + b, z + c)] = 1;
}

template <class S>
static void launch_decide(const tsf_field* f, const uint8_t* cur, Dims d, uint8_t* acc, uint8_t* nom, int32_t* grew,
                          hipStream_t stream) {
    typedef typename S::T T;
    Field<S> F{(const T*)f->x, (const T*)f->y, (const T*)f->z, (const T*)f->w, 1.0f / f->scale};
    dim3 grid((d.X + 63) / 64, (d.Y + 3) / 4, d.Z), block(64, 4);
    hipLaunchKernelGGL(k_seg_decide<S>, grid, block, 0, stream, cur, F, d, acc, nom, grew);
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_segment_inverse_gradient_workspace_size(tsf_dims dims) {
    const size_t n = (size_t)D(dims).n();
    return 4 * ((n + 255) & ~(size_t)255) + 256;
}

tsf_status tsf_segment_inverse_gradient(const uint8_t* centerlines, const float* radius, const tsf_field* field,
                                        tsf_dims dims, uint8_t* segmentation, int32_t* iterations, void* workspace,
                                        size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_segment_inverse_gradient";
    if (!centerlines || !radius || !segmentation) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    TSF_TRY(check_field(field, true, W));
    TSF_TRY(check_dims(dims, 1, W));
    Dims d = D(dims);
    const i64 n = d.n();
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_segment_inverse_gradient_workspace_size(dims), W));
    const size_t nb = ((size_t)n + 255) & ~(size_t)255;
    uint8_t* A = (uint8_t*)sc.ptr;
    uint8_t* B = A + nb;
    uint8_t* acc = B + nb;
    uint8_t* nom = acc + nb;
    int32_t* grew = (int32_t*)(nom + nb);
    dim3 grid((d.X + 63) / 64, (d.Y + 3) / 4, d.Z), block(64, 4);
    hipError_t e = hipMemcpyAsync(A, centerlines, n, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipMemsetAsync(acc, 0, 2 * nb, stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    hipLaunchKernelGGL(k_seg_seed, grid, block, 0, stream, centerlines, radius, A, d);
    TSF_TRY(check_launch(W));
    // at least two passes, then until a pass grows nothing; a pass that
    // grows accepts a new voxel, so there are at most n + 2 of them
    i64 pass = 0;
    for (;; pass++) {
        int32_t g = 0;
        e = hipMemsetAsync(grew, 0, sizeof(int32_t), stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
        switch (field->storage) {
            case TSF_STORAGE_F32: launch_decide<F32>(field, A, d, acc, nom, grew, stream); break;
            case TSF_STORAGE_F16: launch_decide<F16>(field, A, d, acc, nom, grew, stream); break;
            case TSF_STORAGE_SN16: launch_decide<SN16>(field, A, d, acc, nom, grew, stream); break;
            default: launch_decide<BF16>(field, A, d, acc, nom, grew, stream);
        }
        hipLaunchKernelGGL(k_seg_merge, (unsigned)((nb / 16 + 255) / 256), 256, 0, stream, (const uint4*)A,
                           (uint4*)acc, (uint4*)nom, (uint4*)B, (i64)(nb / 16));
        TSF_TRY(check_launch(W));
        std::swap(A, B);
        if (pass == 0) continue;
        e = hipMemcpyAsync(&g, grew, sizeof g, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
        if (g == 0 || pass > n + 2) break;
    }
    if (iterations) *iterations = (int32_t)(pass + 1);
    // closing: A -> acc -> nom -> B (dilation), B -> acc -> nom -> segmentation (erosion)
    if (d.X % 4 == 0 && ((uintptr_t)segmentation & 3u) == 0) {
        typedef unsigned* W;
        const dim3 gw((d.X / 4 + 63) / 64, (d.Y + 3) / 4, (d.Z + CLOSE_ZC - 1) / CLOSE_ZC);
        hipLaunchKernelGGL((k_seg_close_words<0, true>), gw, block, 0, stream, (const unsigned*)A, (W)acc, d);
        hipLaunchKernelGGL((k_seg_close_words<1, true>), gw, block, 0, stream, (const unsigned*)acc, (W)nom, d);
        hipLaunchKernelGGL((k_seg_close_words<2, true>), gw, block, 0, stream, (const unsigned*)nom, (W)B, d);
        hipLaunchKernelGGL((k_seg_close_words<0, false>), gw, block, 0, stream, (const unsigned*)B, (W)acc, d);
        hipLaunchKernelGGL((k_seg_close_words<1, false>), gw, block, 0, stream, (const unsigned*)acc, (W)nom, d);
        hipLaunchKernelGGL((k_seg_close_words<2, false>), gw, block, 0, stream, (const unsigned*)nom,
                           (W)segmentation, d);
    } else {
        hipLaunchKernelGGL(k_seg_dilate_axis<0>, grid, block, 0, stream, A, acc, d);
        hipLaunchKernelGGL(k_seg_dilate_axis<1>, grid, block, 0, stream, acc, nom, d);
        hipLaunchKernelGGL(k_seg_dilate_axis<2>, grid, block, 0, stream, nom, B, d);
        hipLaunchKernelGGL(k_seg_erode_axis<0>, grid, block, 0, stream, B, acc, d);
        hipLaunchKernelGGL(k_seg_erode_axis<1>, grid, block, 0, stream, acc, nom, d);
        hipLaunchKernelGGL(k_seg_erode_axis<2>, grid, block, 0, stream, nom, segmentation, d);
    }
    TSF_TRY(check_launch(W));
    if (sc.owned) {
        e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

tsf_status tsf_segment_spheres(const uint8_t* centerlines, const float* radius, tsf_dims dims, uint8_t* segmentation,
                               hipStream_t stream) {
    const char* W = "tsf_segment_spheres";
    if (!centerlines || !radius || !segmentation) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    TSF_TRY(check_dims(dims, 1, W));
    Dims d = D(dims);
    hipError_t e = hipMemsetAsync(segmentation, 0, (size_t)d.n(), stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    dim3 grid((d.X + 63) / 64, (d.Y + 3) / 4, d.Z), block(64, 4);
    hipLaunchKernelGGL(k_seg_spheres, grid, block, 0, stream, centerlines, radius, segmentation, d);
    return check_launch(W);
}

}  // extern "C"
