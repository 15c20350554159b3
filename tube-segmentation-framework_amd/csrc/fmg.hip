// Gradient vector flow by the full-multigrid (FMG) solver.
//
// Per component c, `iterations` defect-correction steps are run on the
// linear system  mu*lap(v) - |F|^2 v = -|F|^2 F_c  (mirror boundary):
//   r = fmgResidual(f);  e = FMG(r);  f += e
// (runFMGGVF, gradientVectorFlow.cpp:729-1005).  FMG recurses over
// levels+1 cubic grids (calculateNewSize: next power of two of the largest
// extent, halved), each smoothed by red-black Gauss-Seidel sweeps.  All grid
// buffers are allocated once per call (the reference allocates a fresh image
// per helper call) and the restricted |F|^2 pyramid is built once, since it
// does not change between cycles.  Arithmetic is float in the reference's
// operation order; the smoother divides with an IEEE division where the
// reference uses native_divide.
#include <math.h>
#include <vector>

#include "tsf_common.hpp"

namespace tsf {

#define STENCIL_GRID(d) dim3(((d).X + 63) / 64, ((d).Y + 3) / 4, (d).Z), dim3(64, 4)

__device__ __forceinline__ bool in_dims(const Dims& d, int x, int y) { return x < d.X && y < d.Y; }

__global__ void k_sqrmag(const float* __restrict__ fx, const float* __restrict__ fy, const float* __restrict__ fz,
                         float* __restrict__ s, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        s[i] = fx[i] * fx[i] + fy[i] * fy[i] + fz[i] * fz[i];
}

// fmgResidual, kernels_no_3d_write.cl:2303-2348 (finest grid, h = 1).
__global__ void k_fmg_residual(const float* __restrict__ Fx, const float* __restrict__ Fy,
                               const float* __restrict__ Fz, const float* __restrict__ f, Dims d, float mu, int comp,
                               float* __restrict__ r) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (!in_dims(d, x, y)) return;
    const i64 q = d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z));
    const i64 sy = d.X, sz = (i64)d.X * d.Y;
    const float fx = Fx[q], fy = Fy[q], fz = Fz[q];
    const float v0 = comp == 0 ? fx : (comp == 1 ? fy : fz);
    const float sq = fx * fx + fy * fy + fz * fz;
    const float c = f[q];
    float nb = f[q + 1];
    nb = nb + f[q - 1];
    nb = nb + f[q + sy];
    nb = nb + f[q - sy];
    nb = nb + f[q + sz];
    nb = nb + f[q - sz];
    const float residue = (mu * (nb - 6.0f * c)) / (1.0f * 1.0f);
    r[d.at(x, y, z)] = -sq * v0 - (residue - sq * c);
}

__device__ __forceinline__ float gs_value(const float* __restrict__ v, const float* __restrict__ r,
                                          const float* __restrict__ s, i64 q, i64 sy, i64 sz, float mu, float h) {
    float nb = v[q + 1];
    nb = nb + v[q - 1];
    nb = nb + v[q + sy];
    nb = nb + v[q - sy];
    nb = nb + v[q + sz];
    nb = nb + v[q - sz];
    const float num = 2.0f * mu * nb - 2.0f * h * h * r[q];
    const float den = 12.0f * mu + h * h * s[q];
    return num / den;
}

// GVFgaussSeidel (red half-sweep), kernels_no_3d_write.cl:2071-2107.
__global__ void k_gs_red(const float* __restrict__ v, const float* __restrict__ r, const float* __restrict__ s,
                         Dims d, float mu, float h, float* __restrict__ t) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (!in_dims(d, x, y)) return;
    const int px = remap(x, d.X), py = remap(y, d.Y), pz = remap(z, d.Z);
    if (((px + py + pz) & 1) != 0) return;
    t[d.at(x, y, z)] = gs_value(v, r, s, d.at(px, py, pz), d.X, (i64)d.X * d.Y, mu, h);
}

// GVFgaussSeidel2 (black half-sweep; red cells copy the remapped value),
// kernels_no_3d_write.cl:2109-2150.
__global__ void k_gs_black(const float* __restrict__ t, const float* __restrict__ r, const float* __restrict__ s,
                           Dims d, float mu, float h, float* __restrict__ v) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (!in_dims(d, x, y)) return;
    const int px = remap(x, d.X), py = remap(y, d.Y), pz = remap(z, d.Z);
    const i64 q = d.at(px, py, pz);
    v[d.at(x, y, z)] = (((px + py + pz) & 1) == 0) ? t[q] : gs_value(t, r, s, q, d.X, (i64)d.X * d.Y, mu, h);
}

// residual, kernels_no_3d_write.cl:2268-2301.
__global__ void k_residual(const float* __restrict__ r, const float* __restrict__ v, const float* __restrict__ s,
                           Dims d, float mu, float h, float* __restrict__ p) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (!in_dims(d, x, y)) return;
    const i64 q = d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z));
    const i64 sy = d.X, sz = (i64)d.X * d.Y;
    const float c = v[q];
    float nb = v[q + 1];
    nb = nb + v[q - 1];
    nb = nb + v[q + sy];
    nb = nb + v[q - sy];
    nb = nb + v[q + sz];
    nb = nb + v[q - sz];
    const float lap = (mu * (nb - 6.0f * c)) / (h * h);
    p[d.at(x, y, z)] = r[q] - (lap - s[q] * c);
}

__device__ __forceinline__ float rz(const float* __restrict__ p, const Dims& d, int x, int y, int z) {
    return (x < d.X && y < d.Y && z < d.Z) ? p[d.at(x, y, z)] : 0.0f;
}

// restrictVolume, kernels_no_3d_write.cl:2218-2246: 2x2x2 mean, zero outside.
__global__ void k_restrict(const float* __restrict__ fine, Dims fd, Dims cd, float* __restrict__ coarse) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (!in_dims(cd, x, y)) return;
    const int X = 2 * x, Y = 2 * y, Z = 2 * z;
    float s = rz(fine, fd, X, Y, Z);
    s = s + rz(fine, fd, X + 1, Y, Z);
    s = s + rz(fine, fd, X, Y + 1, Z);
    s = s + rz(fine, fd, X, Y, Z + 1);
    s = s + rz(fine, fd, X + 1, Y + 1, Z);
    s = s + rz(fine, fd, X, Y + 1, Z + 1);
    s = s + rz(fine, fd, X + 1, Y + 1, Z + 1);
    s = s + rz(fine, fd, X + 1, Y, Z + 1);
    coarse[cd.at(x, y, z)] = 0.125f * s;
}

// prolongate (ADD) / prolongate2, kernels_no_3d_write.cl:2248-2266: injection.
template <bool ADD>
__global__ void k_prolong(const float* __restrict__ coarse, Dims cd, Dims fd, float* v) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (!in_dims(fd, x, y)) return;
    const float c = rz(coarse, cd, x / 2, y / 2, z / 2);
    const i64 i = fd.at(x, y, z);
    v[i] = ADD ? v[i] + c : c;
}

__global__ void k_add_inplace(float* __restrict__ f, const float* __restrict__ e, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) f[i] = f[i] + e[i];
}

__global__ void k_fill(float* __restrict__ p, i64 n, float v) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) p[i] = v;
}

// MGGVFFinish, kernels_no_3d_write.cl:2202-2216: w = |v|, no zero guard.
__global__ void k_fmg_finish(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                             float* __restrict__ w, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        w[i] = sqrtf(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
}

// Sum of squares in double: per-block partials, then one block folds them.
__global__ void __launch_bounds__(256) k_sumsq_partial(const float* __restrict__ a, i64 n, double* __restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double v = (double)a[i];
        s += v * v;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}
constexpr int NPART = 1024;

static unsigned g1(i64 n) {
    i64 g = (n + 255) / 256;
    return (unsigned)(g > 16384 ? 16384 : (g < 1 ? 1 : g));
}

static Dims coarser(Dims d) {
    int m = d.X > d.Y ? d.X : d.Y;
    m = m > d.Z ? m : d.Z;
    int p = 1;
    while (p < m) p *= 2;
    int s = p / 2 > 1 ? p / 2 : 1;
    return Dims{s, s, s};
}

struct Level {
    Dims d;
    float *r, *v, *s, *t, *p;
};

struct Solver {
    std::vector<Level> L;
    float mu;
    int v0, v1, v2;
    hipStream_t st;
    tsf_status err = TSF_OK;

    void chk() {
        if (err == TSF_OK) err = check_launch("tsf_gvf_fmg");
    }
    void zero(float* p, Dims d) {
        hipLaunchKernelGGL(k_fill, g1(d.n()), 256, 0, st, p, d.n(), 0.0f);
        chk();
    }
    void smooth(int l, int iters) {
        Level& a = L[l];
        const float h = (float)(1 << l);
        for (int i = 0; i < iters; i++) {
            hipLaunchKernelGGL(k_gs_red, STENCIL_GRID(a.d), 0, st, a.v, a.r, a.s, a.d, mu, h, a.t);
            hipLaunchKernelGGL(k_gs_black, STENCIL_GRID(a.d), 0, st, a.t, a.r, a.s, a.d, mu, h, a.v);
            chk();
        }
    }
    void vcycle(int l) {
        Level& a = L[l];
        smooth(l, v1);
        if (l + 1 < (int)L.size()) {
            Level& b = L[l + 1];
            const float h = (float)(1 << l);
            hipLaunchKernelGGL(k_residual, STENCIL_GRID(a.d), 0, st, a.r, a.v, a.s, a.d, mu, h, a.p);
            hipLaunchKernelGGL(k_restrict, STENCIL_GRID(b.d), 0, st, a.p, a.d, b.d, b.r);
            chk();
            zero(b.v, b.d);
            vcycle(l + 1);
            hipLaunchKernelGGL(k_prolong<true>, STENCIL_GRID(a.d), 0, st, b.v, b.d, a.d, a.v);
            chk();
        }
        smooth(l, v2);
    }
    void fmg(int l) {
        Level& a = L[l];
        if (l + 1 < (int)L.size()) {
            Level& b = L[l + 1];
            hipLaunchKernelGGL(k_restrict, STENCIL_GRID(b.d), 0, st, a.r, a.d, b.d, b.r);
            chk();
            fmg(l + 1);
            hipLaunchKernelGGL(k_prolong<false>, STENCIL_GRID(a.d), 0, st, b.v, b.d, a.d, a.v);
            chk();
        } else {
            zero(a.v, a.d);
        }
        for (int i = 0; i < v0; i++) vcycle(l);
    }
};

static std::vector<Dims> level_dims(Dims d, int levels) {
    std::vector<Dims> v{d};
    for (int l = 0; l < levels; l++) v.push_back(coarser(v.back()));
    return v;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_gvf_fmg_workspace_size(tsf_dims dims, int32_t levels) {
    if (levels < 0 || levels > 12) return 0;
    size_t b = 0;
    for (Dims d : level_dims(D(dims), levels)) b += 5 * (((size_t)d.n() * 4 + 255) & ~(size_t)255);
    return b + NPART * sizeof(double) + 256;
}

tsf_status tsf_gvf_fmg(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu, int32_t iterations,
                       int32_t levels, int32_t v0, int32_t v1, int32_t v2, double* residual_norms, void* workspace,
                       size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_gvf_fmg";
    TSF_TRY(check_field(f, false, W));
    TSF_TRY(check_field(out, true, W));
    if (f->storage != TSF_STORAGE_F32 || out->storage != TSF_STORAGE_F32)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: FMG supports F32 storage only", W);
    TSF_TRY(check_dims(dims, 4, W));
    if (iterations < 0 || levels < 0 || levels > 12 || v0 < 0 || v1 < 0 || v2 < 0)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative iteration, level or sweep count", W);
    if (!(mu > 0.0f) || !__builtin_isfinite(mu)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: mu must be > 0", W);
    std::vector<Dims> ld = level_dims(D(dims), levels);
    if (ld.back().X < 4) return set_error(TSF_ERR_DIMENSIONS, "%s: coarsest grid below 4^3", W);
    if (f->x == out->x || f->y == out->y || f->z == out->z)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f and out planes must not alias", W);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_gvf_fmg_workspace_size(dims, levels), W));
    Solver S;
    S.mu = mu;
    S.v0 = v0;
    S.v1 = v1;
    S.v2 = v2;
    S.st = stream;
    char* p = (char*)sc.ptr;
    for (Dims d : ld) {
        size_t b = ((size_t)d.n() * 4 + 255) & ~(size_t)255;
        Level lv;
        lv.d = d;
        float** bufs[5] = {&lv.r, &lv.v, &lv.s, &lv.t, &lv.p};
        for (float** q : bufs) {
            *q = (float*)p;
            p += b;
        }
        S.L.push_back(lv);
    }
    double* part = (double*)p;
    double* norm_out = part + NPART;
    const float *Fx = (const float*)f->x, *Fy = (const float*)f->y, *Fz = (const float*)f->z;
    float* O[3] = {(float*)out->x, (float*)out->y, (float*)out->z};
    const Dims d0 = ld[0];
    const i64 n0 = d0.n();
    // |F|^2 at every level (restricted once)
    hipLaunchKernelGGL(k_sqrmag, g1(n0), 256, 0, stream, Fx, Fy, Fz, S.L[0].s, n0);
    S.chk();
    for (size_t l = 1; l < S.L.size(); l++)
        hipLaunchKernelGGL(k_restrict, STENCIL_GRID(S.L[l].d), 0, stream, S.L[l - 1].s, S.L[l - 1].d, S.L[l].d,
                           S.L[l].s);
    S.chk();
    for (int c = 0; c < 3 && S.err == TSF_OK; c++) {
        S.zero(O[c], d0);
        for (int it = 0; it < iterations && S.err == TSF_OK; it++) {
            hipLaunchKernelGGL(k_fmg_residual, STENCIL_GRID(d0), 0, stream, Fx, Fy, Fz, O[c], d0, mu, c, S.L[0].r);
            S.chk();
            if (residual_norms) {
                hipLaunchKernelGGL(k_sumsq_partial, NPART, 256, 0, stream, S.L[0].r, n0, part);
                hipLaunchKernelGGL(k_sum_partials, 1, 256, 0, stream, part, NPART, norm_out);
                S.chk();
                double ss = 0.0;
                hipError_t e = hipMemcpyAsync(&ss, norm_out, sizeof ss, hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipStreamSynchronize(stream);
                if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
                residual_norms[c * iterations + it] = sqrt(ss);
            }
            S.fmg(0);
            hipLaunchKernelGGL(k_add_inplace, g1(n0), 256, 0, stream, O[c], S.L[0].v, n0);
            S.chk();
        }
    }
    if (S.err != TSF_OK) return S.err;
    hipLaunchKernelGGL(k_fmg_finish, g1(n0), 256, 0, stream, O[0], O[1], O[2], (float*)out->w, n0);
    TSF_TRY(check_launch(W));
    if (sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

}  // extern "C"
