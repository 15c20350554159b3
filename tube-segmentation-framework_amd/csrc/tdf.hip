// Hessian eigenanalysis, candidate non-maximum suppression and cube maximum.
// Gather/VALU-bound per-voxel kernels; the eigensolver runs in registers
// (eigen3.hpp).  The circle-fitting TDF lives in circle.hip.
#include "eigen3.hpp"
#include "field.hpp"

namespace tsf {

template <class S, bool NORM>
__global__ void __launch_bounds__(256) k_hessian_eigen(Field<S> F, Dims d, float* __restrict__ lam,
                                                       float* __restrict__ vec) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= d.X || y >= d.Y) return;
    float H[3][3], V[3][3], ev[3];
    hessian<NORM>(F, d, x, y, z, H);
    eigen3(H, V, ev);
    const i64 i = d.at(x, y, z), N = d.n();
    lam[i] = ev[0];
    lam[N + i] = ev[1];
    lam[2 * N + i] = ev[2];
    if (vec) {
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) vec[(3 * j + k) * N + i] = V[k][j];
    }
}

// doEigen (eigenanalysisOfHessian.cpp:264-389): samples divided by the
// host-style magnitude sqrt(Fx^2+Fy^2+Fz^2) of their voxel; no clamping.
__global__ void k_do_eigen_points(const float* __restrict__ Fx, const float* __restrict__ Fy,
                                  const float* __restrict__ Fz, Dims d, const int32_t* __restrict__ pos, i64 count,
                                  float* __restrict__ out) {
    i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int x = pos[3 * t], y = pos[3 * t + 1], z = pos[3 * t + 2];
    float* o = out + 12 * t;
    if (x < 1 || y < 1 || z < 1 || x > d.X - 2 || y > d.Y - 2 || z > d.Z - 2) {
        // the reference reads outside the arrays here; report NaN instead
#pragma unroll
        for (int k = 0; k < 12; k++) o[k] = __builtin_nanf("");
        return;
    }
    auto nv = [&](const float* c, int xx, int yy, int zz) -> float {
        i64 i = d.at(xx, yy, zz);
        return c[i] / sqrtf(Fx[i] * Fx[i] + Fy[i] * Fy[i] + Fz[i] * Fz[i]);
    };
    const float fxx = 0.5f * (nv(Fx, x + 1, y, z) - nv(Fx, x - 1, y, z));
    const float fyx = 0.5f * (nv(Fy, x + 1, y, z) - nv(Fy, x - 1, y, z));
    const float fyy = 0.5f * (nv(Fy, x, y + 1, z) - nv(Fy, x, y - 1, z));
    const float fzx = 0.5f * (nv(Fz, x + 1, y, z) - nv(Fz, x - 1, y, z));
    const float fzy = 0.5f * (nv(Fz, x, y + 1, z) - nv(Fz, x, y - 1, z));
    const float fzz = 0.5f * (nv(Fz, x, y, z + 1) - nv(Fz, x, y, z - 1));
    float H[3][3] = {{fxx, fyx, fzx}, {fyx, fyy, fzy}, {fzx, fzy, fzz}};
    float V[3][3], ev[3];
    eigen3(H, V, ev);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        o[k] = ev[k];
        o[3 + k] = V[k][0];
        o[6 + k] = V[k][1];
        o[9 + k] = V[k][2];
    }
}

// findCandidateCenterpoints2, kernels_no_3d_write.cl:1019-1084.
template <class S>
__global__ void __launch_bounds__(256) k_candidate_nms(const float* __restrict__ radius, Field<S> F, Dims d,
                                                       const int32_t* __restrict__ pos, i64 count, int cap,
                                                       uint8_t* __restrict__ out) {
    i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int x = pos[3 * t], y = pos[3 * t + 1], z = pos[3 * t + 2];
    if (x < 0 || y < 0 || z < 0 || x >= d.X || y >= d.Y || z >= d.Z) return;  // not a voxel: ignored
    const float rr = radius[d.at(x, y, z)];
    const int maxD = (int)fmaxf(fminf(roundf(rr), (float)cap), 1.0f);
    float H[3][3], V[3][3], ev[3];
    hessian<true>(F, d, x, y, z, H);
    eigen3(H, V, ev);
    const float e1x = V[0][0], e1y = V[1][0], e1z = V[2][0];
    const float wp = F.mag(d.at(x, y, z));
    bool invalid = false;
    for (int a = -maxD; a <= maxD; a++)
        for (int b = -maxD; b <= maxD; b++)
            for (int c = -maxD; c <= maxD; c++) {
                if (a == 0 && b == 0 && c == 0) continue;
                const float rx = (float)a, ry = (float)b, rz = (float)c;
                const float dp = e1x * rx + e1y * ry + e1z * rz;
                const float px = rx - e1x * dp, py = ry - e1y * dp, pz = rz - e1z * dp;
                const float lr = sqrtf(rx * rx + ry * ry + rz * rz);
                const float lp = sqrtf(px * px + py * py + pz * pz);
                const float cs = (rx / lr) * (px / lp) + (ry / lr) * (py / lp) + (rz / lr) * (pz / lp);
                const float theta = acosf(cs);
                if (theta < 0.5f && lr < (float)maxD) {
                    const int nx = clampi(x + a, 0, d.X - 1), ny = clampi(y + b, 0, d.Y - 1),
                              nz = clampi(z + c, 0, d.Z - 1);
                    if (F.mag(d.at(nx, ny, nz)) < wp) invalid = true;
                }
            }
    out[d.at(x, y, z)] = invalid ? 0 : 1;
}

// dd, kernels_no_3d_write.cl:970-1004 (in-range voxels only).
__global__ void k_cube_max(const float* __restrict__ tdf, const uint8_t* __restrict__ cand, Dims d, int cube, int cx,
                           int cy, int cz, uint8_t* __restrict__ out) {
    i64 cell = blockIdx.x * (i64)blockDim.x + threadIdx.x;
    if (cell >= (i64)cx * cy * cz) return;
    const int gx = (int)(cell % cx), gy = (int)((cell / cx) % cy), gz = (int)(cell / ((i64)cx * cy));
    float best = 0.0f;
    i64 bi = -1;
    for (int a = 0; a < cube; a++) {
        const int x = gx * cube + a;
        if (x >= d.X) break;
        for (int b = 0; b < cube; b++) {
            const int y = gy * cube + b;
            if (y >= d.Y) break;
            for (int c = 0; c < cube; c++) {
                const int z = gz * cube + c;
                if (z >= d.Z) break;
                const i64 i = d.at(x, y, z);
                if (cand[i] == 1) {
                    const float t = tdf[i];
                    if (t > best) {
                        best = t;
                        bi = i;
                    }
                }
            }
        }
    }
    if (bi >= 0) out[bi] = 1;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

tsf_status tsf_hessian_eigen(const tsf_field* field, tsf_dims dims, int32_t normalize, float* lambda, float* vectors,
                             hipStream_t stream) {
    const char* W = "tsf_hessian_eigen";
    TSF_TRY(check_field(field, normalize != 0, W));
    TSF_TRY(check_dims(dims, 1, W));
    if (!lambda) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL lambda", W);
    Dims d = D(dims);
    dim3 grid((d.X + 63) / 64, (d.Y + 3) / 4, d.Z), block(64, 4);
    const float inv = 1.0f / field->scale;
    switch (field->storage) {
        case TSF_STORAGE_F32: {
            Field<F32> F{(const float*)field->x, (const float*)field->y, (const float*)field->z, (const float*)field->w, inv};
            if (normalize) hipLaunchKernelGGL((k_hessian_eigen<F32, true>), grid, block, 0, stream, F, d, lambda, vectors);
            else hipLaunchKernelGGL((k_hessian_eigen<F32, false>), grid, block, 0, stream, F, d, lambda, vectors);
            break;
        }
        case TSF_STORAGE_F16: {
            typedef __half T;
            Field<F16> F{(const T*)field->x, (const T*)field->y, (const T*)field->z, (const T*)field->w, inv};
            if (normalize) hipLaunchKernelGGL((k_hessian_eigen<F16, true>), grid, block, 0, stream, F, d, lambda, vectors);
            else hipLaunchKernelGGL((k_hessian_eigen<F16, false>), grid, block, 0, stream, F, d, lambda, vectors);
            break;
        }
        case TSF_STORAGE_SN16: {
            typedef short T;
            Field<SN16> F{(const T*)field->x, (const T*)field->y, (const T*)field->z, (const T*)field->w, inv};
            if (normalize) hipLaunchKernelGGL((k_hessian_eigen<SN16, true>), grid, block, 0, stream, F, d, lambda, vectors);
            else hipLaunchKernelGGL((k_hessian_eigen<SN16, false>), grid, block, 0, stream, F, d, lambda, vectors);
            break;
        }
        default: {
            typedef __hip_bfloat16 T;
            Field<BF16> F{(const T*)field->x, (const T*)field->y, (const T*)field->z, (const T*)field->w, inv};
            if (normalize) hipLaunchKernelGGL((k_hessian_eigen<BF16, true>), grid, block, 0, stream, F, d, lambda, vectors);
            else hipLaunchKernelGGL((k_hessian_eigen<BF16, false>), grid, block, 0, stream, F, d, lambda, vectors);
        }
    }
    return check_launch(W);
}

tsf_status tsf_do_eigen_points(const float* fx, const float* fy, const float* fz, tsf_dims dims,
                               const int32_t* positions, int64_t count, float* out, hipStream_t stream) {
    const char* W = "tsf_do_eigen_points";
    if (!fx || !fy || !fz || !out || (count > 0 && !positions))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    TSF_TRY(check_dims(dims, 3, W));
    if (count < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative count", W);
    if (count == 0) return TSF_OK;
    hipLaunchKernelGGL(k_do_eigen_points, (unsigned)((count + 255) / 256), 256, 0, stream, fx, fy, fz, D(dims),
                       positions, (i64)count, out);
    return check_launch(W);
}

tsf_status tsf_candidate_nms(const float* tdf, const float* radius, const tsf_field* field, tsf_dims dims,
                             const int32_t* positions, int64_t count, int32_t maxd_cap, uint8_t* out,
                             hipStream_t stream) {
    const char* W = "tsf_candidate_nms";
    (void)tdf;
    TSF_TRY(check_field(field, true, W));
    TSF_TRY(check_dims(dims, 1, W));
    if (!radius || !out || (count > 0 && !positions))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (maxd_cap < 1 || maxd_cap > 64) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: maxd_cap out of [1, 64]", W);
    if (count < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative count", W);
    if (count == 0) return TSF_OK;
    Dims d = D(dims);
    const float inv = 1.0f / field->scale;
    unsigned g = (unsigned)((count + 255) / 256);
    switch (field->storage) {
        case TSF_STORAGE_F32: {
            Field<F32> F{(const float*)field->x, (const float*)field->y, (const float*)field->z, (const float*)field->w, inv};
            hipLaunchKernelGGL(k_candidate_nms<F32>, g, 256, 0, stream, radius, F, d, positions, (i64)count, maxd_cap, out);
            break;
        }
        case TSF_STORAGE_F16: {
            typedef __half T;
            Field<F16> F{(const T*)field->x, (const T*)field->y, (const T*)field->z, (const T*)field->w, inv};
            hipLaunchKernelGGL(k_candidate_nms<F16>, g, 256, 0, stream, radius, F, d, positions, (i64)count, maxd_cap, out);
            break;
        }
        case TSF_STORAGE_SN16: {
            typedef short T;
            Field<SN16> F{(const T*)field->x, (const T*)field->y, (const T*)field->z, (const T*)field->w, inv};
            hipLaunchKernelGGL(k_candidate_nms<SN16>, g, 256, 0, stream, radius, F, d, positions, (i64)count, maxd_cap, out);
            break;
        }
        default: {
            typedef __hip_bfloat16 T;
            Field<BF16> F{(const T*)field->x, (const T*)field->y, (const T*)field->z, (const T*)field->w, inv};
            hipLaunchKernelGGL(k_candidate_nms<BF16>, g, 256, 0, stream, radius, F, d, positions, (i64)count, maxd_cap, out);
        }
    }
    return check_launch(W);
}

tsf_status tsf_cube_max(const float* tdf, const uint8_t* candidates, tsf_dims dims, int32_t cube, uint8_t* out,
                        hipStream_t stream) {
    const char* W = "tsf_cube_max";
    if (!tdf || !candidates || !out) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    TSF_TRY(check_dims(dims, 1, W));
    if (cube < 1 || cube > 64) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: cube size out of [1, 64]", W);
    Dims d = D(dims);
    const int cx = (d.X + cube - 1) / cube, cy = (d.Y + cube - 1) / cube, cz = (d.Z + cube - 1) / cube;
    const i64 cells = (i64)cx * cy * cz;
    hipLaunchKernelGGL(k_cube_max, (unsigned)((cells + 255) / 256), 256, 0, stream, tdf, candidates, d, cube, cx, cy, cz,
                       out);
    return check_launch(W);
}

}  // extern "C"
