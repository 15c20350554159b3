// CPU oracle for the volume-filter hot path of the Tube Segmentation Framework.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
// compared against; it is never linked into, loaded by, or called from the
// product path (tube-segmentation-framework_amd/).  Only tests/, the smoke
// check in __graft_entry__.py and the cpu_baseline leg of bench.py use it.
//
// Every function restates, in plain scalar C++, the arithmetic of one function
// of the reference (file:line citations are relative to the reference tree).
// Each function is instantiated twice:
//   *_f32  -- float arithmetic in the reference's operation order, compiled
//             with -ffp-contract=off (no FMA) so it is a deterministic IEEE
//             restatement of the reference's FP32 buffer path;
//   *_f64  -- the same algorithm in double, used as the error reference for
//             the reduced-precision storage forms.
// Parameters that are part of the algorithm's definition (radius sequence,
// the 32 circle directions, mu, Fmax, blur mask) stay float in both modes.
//
// Where the reference's behaviour is undefined or can hang, the oracle fixes a
// rule and the HIP path follows the same rule:
//   * tql2: the small-subdiagonal search stops at index n-1 (the reference
//     relies on e[n-1] == 0; with NaN input it would read past the array),
//     and the QL iteration is capped at TSF_QL_MAX_ITER sweeps (the reference
//     has no cap, kernels_no_3d_write.cl:1973).
//   * cube max (dd): only in-range voxels are scanned (the reference's last
//     partial cells read clamped coordinates, kernels_no_3d_write.cl:983-1003;
//     strict '>' makes the clamped duplicates lose, so the result is equal).
//   * GS smoother: native_divide is replaced by an IEEE division.
//
// Pinning status of each function is listed in the header of tests/test_oracle.py.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>
#include <algorithm>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TSF_QL_MAX_ITER 100

namespace {

typedef long long i64;

struct Dims {
    int X, Y, Z;
    i64 n() const { return (i64)X * Y * Z; }
    i64 at(int x, int y, int z) const { return (i64)x + (i64)X * ((i64)y + (i64)Y * z); }
};

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Nearest read with CLK_ADDRESS_CLAMP_TO_EDGE (kernels_no_3d_write.cl:1).
template <typename T>
static inline T rd_edge(const T* p, const Dims& d, int x, int y, int z) {
    return p[d.at(clampi(x, 0, d.X - 1), clampi(y, 0, d.Y - 1), clampi(z, 0, d.Z - 1))];
}
// Nearest read with CLK_ADDRESS_CLAMP: zero outside (kernels_no_3d_write.cl:4).
template <typename T>
static inline T rd_zero(const T* p, const Dims& d, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= d.X || y >= d.Y || z >= d.Z) return T(0);
    return p[d.at(x, y, z)];
}
// Mirror remap used by every GVF / FMG kernel: 0 -> 2, >= N-1 -> N-3
// (kernels_no_3d_write.cl:1759-1760, 2086-2087).
static inline int remap(int p, int N) {
    if (p == 0) return 2;
    if (p >= N - 1) return N - 3;
    return p;
}

// ---------------------------------------------------------------- blur (A1)
// createBlurMask, tube-segmentation.cpp:188-211 (float arithmetic, expf).
static void blur_mask(float sigma, float* mask, int* maskSize) {
    int m = (int)std::ceil(sigma / 0.5f);
    if (m < 1) m = 1;
    if (m > 5) m = 5;
    const int M = 2 * m + 1;
    float sum = 0.0f;
    for (int a = -m; a <= m; a++)
        for (int b = -m; b <= m; b++)
            for (int c = -m; c <= m; c++) {
                float w = std::exp(-((float)(a * a + b * b + c * c) / (2 * sigma * sigma)));
                sum += w;
                mask[(a + m) + (b + m) * M + (c + m) * M * M] = w;
            }
    for (int i = 0; i < M * M * M; i++) mask[i] = mask[i] / sum;
    *maskSize = m;
}

// blurVolumeWithGaussian, kernels_no_3d_write.cl:1442-1465.  Loop a (x) outer,
// b (y), c (z) inner; clamp-to-edge reads.
template <typename T>
static void blur(const T* in, T* out, Dims d, int m, const float* mask) {
    const int M = 2 * m + 1;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                T sum = 0;
                for (int a = -m; a <= m; a++)
                    for (int b = -m; b <= m; b++)
                        for (int c = -m; c <= m; c++)
                            sum += (T)mask[(a + m) + (b + m) * M + (c + m) * M * M] *
                                   rd_edge(in, d, x + a, y + b, z + c);
                out[d.at(x, y, z)] = sum;
            }
}

// ---------------------------------------------------------------- toFloat (A2)
// toFloat, kernels_no_3d_write.cl:1414-1440.
template <typename T>
static void to_float(const double* in, T* out, i64 n, float mn, float mx) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; i++) {
        T v = (T)(float)in[i];
        v = v > (T)mx ? (T)mx : v;
        v = v < (T)mn ? (T)mn : v;
        out[i] = (v - (T)mn) / ((T)mx - (T)mn);
    }
}

// ---------------------------------------------------------------- vector field (A3)
// createVectorField + gradient, kernels_no_3d_write.cl:1470-1498 and 146-200.
// field is SoA: plane 0..2 = F.xyz, plane 3 = w (= 1).
template <typename T>
static void vector_field(const T* I, T* field, Dims d, float Fmax, int sign) {
    const i64 N = d.n();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                T gx = (T)0.5 * (rd_edge(I, d, x + 1, y, z) - rd_edge(I, d, x - 1, y, z));
                T gy = (T)0.5 * (rd_edge(I, d, x, y + 1, z) - rd_edge(I, d, x, y - 1, z));
                T gz = (T)0.5 * (rd_edge(I, d, x, y, z + 1) - rd_edge(I, d, x, y, z - 1));
                gx = (T)sign * gx; gy = (T)sign * gy; gz = (T)sign * gz;
                T l = std::sqrt(gx * gx + gy * gy + gz * gz);
                T s = l < (T)Fmax ? (T)Fmax : l;
                i64 i = d.at(x, y, z);
                field[i] = gx / s;
                field[N + i] = gy / s;
                field[2 * N + i] = gz / s;
                field[3 * N + i] = (T)1;
            }
}

// ---------------------------------------------------------------- GVF Euler (A4)
// One Jacobi step of GVF3DIteration, kernels_no_3d_write.cl:1746-1787.
// f and v are 3-plane SoA.  The data-term magnitude |f(q)|^2 is recomputed.
template <typename T>
static void gvf_step(const T* f, const T* v, T* out, Dims d, T mu) {
    const i64 N = d.n();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                const int qx = remap(x, d.X), qy = remap(y, d.Y), qz = remap(z, d.Z);
                const i64 q = d.at(qx, qy, qz);
                const T ix = f[q], iy = f[N + q], iz = f[2 * N + q];
                const T m2 = ix * ix + iy * iy + iz * iz;
                const i64 w = d.at(x, y, z);
                const i64 sx = 1, sy = d.X, sz = (i64)d.X * d.Y;
                for (int c = 0; c < 3; c++) {
                    const T* p = v + c * N;
                    const T vc = p[q];
                    T lap = (T)-6 * vc;
                    lap = lap + p[q + sx];
                    lap = lap + p[q - sx];
                    lap = lap + p[q + sy];
                    lap = lap + p[q - sy];
                    lap = lap + p[q + sz];
                    lap = lap + p[q - sz];
                    const T fc = c == 0 ? ix : (c == 1 ? iy : iz);
                    out[c * N + w] = vc + (mu * lap - (vc - fc) * m2);
                }
            }
}

// GVF3DFinish, kernels_no_3d_write.cl:1798-1807: w = |v| if > 0 else 1.
template <typename T>
static void gvf_finish(T* v, Dims d) {
    const i64 N = d.n();
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < N; i++) {
        T l = std::sqrt(v[i] * v[i] + v[N + i] * v[N + i] + v[2 * N + i] * v[2 * N + i]);
        v[3 * N + i] = l > (T)0 ? l : (T)1;
    }
}

// runFastGVF, gradientVectorFlow.cpp:1007-1110: v0 = f, `iters` Jacobi steps.
template <typename T>
static void gvf_euler(const T* f4, T* v4, Dims d, float mu, int iters) {
    const i64 N = d.n();
    std::vector<T> a(f4, f4 + 3 * N), b(3 * N);
    for (int it = 0; it < iters; it++) {
        gvf_step(f4, a.data(), b.data(), d, (T)mu);
        a.swap(b);
    }
    std::memcpy(v4, a.data(), sizeof(T) * 3 * N);
    gvf_finish(v4, d);
}

// ---------------------------------------------------------------- 16-bit mode
// The reference's 16-bit vector format (kernels_no_3d_write.cl:9-21, the
// VECTORS_16BIT branch): every store is convert_short_sat_rte(v * 32767.0f)
// (round half to even, saturate, NaN -> 0), every load max(-1.0f,
// (float)s / 32767.0f).
static inline int16_t sn16_store(float v) {
    const float x = v * 32767.0f;
    if (x != x) return 0;
    const float r = std::nearbyint(x);  // default rounding mode: to nearest even
    return (int16_t)(r > 32767.0f ? 32767 : (r < -32768.0f ? -32768 : (int)r));
}
static inline float sn16_load(int16_t s) { return std::max(-1.0f, (float)s / 32767.0f); }
static inline float sn16_round(float v) { return sn16_load(sn16_store(v)); }

// runFastGVF in the 16-bit mode: f arrives already in the format (its
// values are SNORM16 values); every iterate is stored in the format
// (GVF3DIteration, kernels_no_3d_write.cl:1746-1787) and the finish stores
// w = |v| (1 where 0) in the format too (GVF3DFinish 1798-1807), so w
// saturates at 1.
static void gvf_euler_sn16(const float* f4, float* v4, Dims d, float mu, int iters) {
    const i64 N = d.n();
    std::vector<float> a(f4, f4 + 3 * N), b(3 * N);
    for (int it = 0; it < iters; it++) {
        gvf_step<float>(f4, a.data(), b.data(), d, mu);
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < 3 * N; i++) b[i] = sn16_round(b[i]);
        a.swap(b);
    }
    std::memcpy(v4, a.data(), sizeof(float) * 3 * N);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < N; i++) {
        const float l = std::sqrt(v4[i] * v4[i] + v4[N + i] * v4[N + i] + v4[2 * N + i] * v4[2 * N + i]);
        v4[3 * N + i] = sn16_round(l > 0.0f ? l : 1.0f);
    }
}

// runLowMemoryGVF, gradientVectorFlow.cpp:1173-1445 (GVF3DInit_one_component,
// GVF3DIteration_one_component, GVF3DFinish_one_component,
// kernels_no_3d_write.cl:1669-1744): each component iterates alone with the
// pair (f_c, |f|^2) of the init kernel, |f|^2 = x x + y y + z z stored in the
// field's format (rounded to SNORM16 in the 16-bit mode), the update
// v + (mu lap - (v - f_c) |f|^2) at the mirror voxel, every iterate stored
// in the field's format; then w = |v| (1 where 0).  f4 holds the field's
// values (SNORM16 values when sn16).
static void gvf_euler_low_memory(const float* f4, float* v4, Dims d, float mu, int iters, bool sn16) {
    const i64 N = d.n(), sy = d.X, sz = (i64)d.X * d.Y;
    auto st = [sn16](float v) { return sn16 ? sn16_round(v) : v; };
    std::vector<float> m2(N), a(N), b(N);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < N; i++) {
        const float x = f4[i], y = f4[N + i], z = f4[2 * N + i];
        m2[i] = st(x * x + y * y + z * z);
    }
    for (int c = 0; c < 3; c++) {
        const float* f = f4 + c * N;
        for (i64 i = 0; i < N; i++) a[i] = st(f[i]);
        for (int it = 0; it < iters; it++) {
#pragma omp parallel for schedule(static)
            for (int z = 0; z < d.Z; z++)
                for (int y = 0; y < d.Y; y++)
                    for (int x = 0; x < d.X; x++) {
                        const i64 q = d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z));
                        const float vc = a[q];
                        float lap = -6.0f * vc;
                        lap = lap + a[q + 1];
                        lap = lap + a[q - 1];
                        lap = lap + a[q + sy];
                        lap = lap + a[q - sy];
                        lap = lap + a[q + sz];
                        lap = lap + a[q - sz];
                        b[d.at(x, y, z)] = st(vc + (mu * lap - (vc - f[q]) * m2[q]));
                    }
            a.swap(b);
        }
        std::memcpy(v4 + c * N, a.data(), sizeof(float) * N);
    }
    for (i64 i = 0; i < N; i++) {
        const float l = std::sqrt(v4[i] * v4[i] + v4[N + i] * v4[N + i] + v4[2 * N + i] * v4[2 * N + i]);
        v4[3 * N + i] = st(l > 0.0f ? l : 1.0f);
    }
}

// ---------------------------------------------------------------- FMG GVF (A7)
// Multigrid solver of gradientVectorFlow.cpp:5-441, 631-1005 with the kernels
// of kernels_no_3d_write.cl:2071-2354 (FP32 path: no SNORM16 quantisation).
template <typename T>
struct FMG {
    T mu;
    int v0, v1, v2, lmax;

    // calculateNewSize, gradientVectorFlow.cpp:371-395: cubic, next power of
    // two of the largest dimension, halved.
    static Dims coarser(Dims d) {
        int mx = std::max(d.X, std::max(d.Y, d.Z));
        int p = 1;
        while (p < mx) p *= 2;
        int s = std::max(1, p / 2);
        return Dims{s, s, s};
    }

    // GVFgaussSeidel + GVFgaussSeidel2 (red then black), cl:2071-2150.
    void smooth(std::vector<T>& v, const std::vector<T>& r, const std::vector<T>& s, int iters,
                Dims d, T h) {
        const i64 N = d.n();
        std::vector<T> v2(N, (T)0), out(N);
        const i64 sx = 1, sy = d.X, sz = (i64)d.X * d.Y;
        for (int it = 0; it < iters; it++) {
            // red
#pragma omp parallel for schedule(static)
            for (int z = 0; z < d.Z; z++)
                for (int y = 0; y < d.Y; y++)
                    for (int x = 0; x < d.X; x++) {
                        int px = remap(x, d.X), py = remap(y, d.Y), pz = remap(z, d.Z);
                        if ((px + py + pz) % 2 != 0) continue;
                        i64 q = d.at(px, py, pz);
                        T nb = v[q + sx];
                        nb = nb + v[q - sx];
                        nb = nb + v[q + sy];
                        nb = nb + v[q - sy];
                        nb = nb + v[q + sz];
                        nb = nb + v[q - sz];
                        T num = (T)2 * mu * nb - (T)2 * h * h * r[q];
                        T den = (T)12 * mu + h * h * s[q];
                        v2[d.at(x, y, z)] = num / den;
                    }
            // black (red cells copy the value at the remapped position)
#pragma omp parallel for schedule(static)
            for (int z = 0; z < d.Z; z++)
                for (int y = 0; y < d.Y; y++)
                    for (int x = 0; x < d.X; x++) {
                        int px = remap(x, d.X), py = remap(y, d.Y), pz = remap(z, d.Z);
                        i64 q = d.at(px, py, pz);
                        T val;
                        if ((px + py + pz) % 2 == 0) {
                            val = v2[q];
                        } else {
                            T nb = v2[q + sx];
                            nb = nb + v2[q - sx];
                            nb = nb + v2[q + sy];
                            nb = nb + v2[q - sy];
                            nb = nb + v2[q + sz];
                            nb = nb + v2[q - sz];
                            T num = (T)2 * mu * nb - (T)2 * h * h * r[q];
                            T den = (T)12 * mu + h * h * s[q];
                            val = num / den;
                        }
                        out[d.at(x, y, z)] = val;
                    }
            v.swap(out);
        }
    }

    // residual, cl:2268-2301.
    std::vector<T> residual(const std::vector<T>& r, const std::vector<T>& v, const std::vector<T>& s,
                            Dims d, T h) {
        std::vector<T> p(d.n());
#pragma omp parallel for schedule(static)
        for (int z = 0; z < d.Z; z++)
            for (int y = 0; y < d.Y; y++)
                for (int x = 0; x < d.X; x++) {
                    int px = remap(x, d.X), py = remap(y, d.Y), pz = remap(z, d.Z);
                    T c = rd_zero(v.data(), d, px, py, pz);
                    T nb = rd_zero(v.data(), d, px + 1, py, pz);
                    nb = nb + rd_zero(v.data(), d, px - 1, py, pz);
                    nb = nb + rd_zero(v.data(), d, px, py + 1, pz);
                    nb = nb + rd_zero(v.data(), d, px, py - 1, pz);
                    nb = nb + rd_zero(v.data(), d, px, py, pz + 1);
                    nb = nb + rd_zero(v.data(), d, px, py, pz - 1);
                    T lap = (mu * (nb - (T)6 * c)) / (h * h);
                    p[d.at(x, y, z)] = rd_zero(r.data(), d, px, py, pz) -
                                       (lap - rd_zero(s.data(), d, px, py, pz) * c);
                }
        return p;
    }

    // restrictVolume, cl:2218-2246: 2x2x2 mean, zero outside the fine grid.
    static std::vector<T> restrict_(const std::vector<T>& v, Dims fine, Dims coarse) {
        std::vector<T> o(coarse.n());
#pragma omp parallel for schedule(static)
        for (int z = 0; z < coarse.Z; z++)
            for (int y = 0; y < coarse.Y; y++)
                for (int x = 0; x < coarse.X; x++) {
                    int X = 2 * x, Y = 2 * y, Zz = 2 * z;
                    const T* p = v.data();
                    T s = rd_zero(p, fine, X, Y, Zz);
                    s = s + rd_zero(p, fine, X + 1, Y, Zz);
                    s = s + rd_zero(p, fine, X, Y + 1, Zz);
                    s = s + rd_zero(p, fine, X, Y, Zz + 1);
                    s = s + rd_zero(p, fine, X + 1, Y + 1, Zz);
                    s = s + rd_zero(p, fine, X, Y + 1, Zz + 1);
                    s = s + rd_zero(p, fine, X + 1, Y + 1, Zz + 1);
                    s = s + rd_zero(p, fine, X + 1, Y, Zz + 1);
                    o[coarse.at(x, y, z)] = (T)0.125 * s;
                }
        return o;
    }

    // prolongate (add = true) and prolongate2 (add = false), cl:2248-2266:
    // injection from floor(p/2), zero outside the coarse grid.
    static std::vector<T> prolong(const std::vector<T>* fineAdd, const std::vector<T>& c, Dims fine,
                                  Dims coarse) {
        std::vector<T> o(fine.n());
#pragma omp parallel for schedule(static)
        for (int z = 0; z < fine.Z; z++)
            for (int y = 0; y < fine.Y; y++)
                for (int x = 0; x < fine.X; x++) {
                    T cv = rd_zero(c.data(), coarse, x / 2, y / 2, z / 2);
                    i64 i = fine.at(x, y, z);
                    o[i] = fineAdd ? (*fineAdd)[i] + cv : cv;
                }
        return o;
    }

    // multigridVcycle, gradientVectorFlow.cpp:397-441.
    void vcycle(const std::vector<T>& r, std::vector<T>& v, const std::vector<T>& s, int l, Dims d, T h) {
        smooth(v, r, s, v1, d, h);
        if (l < lmax) {
            Dims c = coarser(d);
            std::vector<T> p = residual(r, v, s, d, h);
            std::vector<T> r1 = restrict_(p, d, c);
            std::vector<T> s1 = restrict_(s, d, c);
            std::vector<T> vc(c.n(), (T)0);
            vcycle(r1, vc, s1, l + 1, c, h * 2);
            v = prolong(&v, vc, d, c);
        }
        smooth(v, r, s, v2, d, h);
    }

    // fullMultigrid, gradientVectorFlow.cpp:694-727.
    std::vector<T> fmg(const std::vector<T>& r, const std::vector<T>& s, int l, Dims d, T h) {
        std::vector<T> v;
        if (l < lmax) {
            Dims c = coarser(d);
            std::vector<T> r1 = restrict_(r, d, c);
            std::vector<T> s1 = restrict_(s, d, c);
            std::vector<T> vc = fmg(r1, s1, l + 1, c, h * 2);
            v = prolong(nullptr, vc, d, c);
        } else {
            v.assign(d.n(), (T)0);
        }
        for (int i = 0; i < v0; i++) vcycle(r, v, s, l, d, h);
        return v;
    }

    // fmgResidual, cl:2303-2348 (h = 1 at the finest level).
    static std::vector<T> fmg_residual(const T* F, const std::vector<T>& f, Dims d, T mu, int comp) {
        const i64 N = d.n();
        std::vector<T> r(N);
#pragma omp parallel for schedule(static)
        for (int z = 0; z < d.Z; z++)
            for (int y = 0; y < d.Y; y++)
                for (int x = 0; x < d.X; x++) {
                    int px = remap(x, d.X), py = remap(y, d.Y), pz = remap(z, d.Z);
                    i64 q = d.at(px, py, pz);
                    T fx = F[q], fy = F[N + q], fz = F[2 * N + q];
                    T v0c = comp == 0 ? fx : (comp == 1 ? fy : fz);
                    T sq = fx * fx + fy * fy + fz * fz;
                    T c = rd_zero(f.data(), d, px, py, pz);
                    T nb = rd_zero(f.data(), d, px + 1, py, pz);
                    nb = nb + rd_zero(f.data(), d, px - 1, py, pz);
                    nb = nb + rd_zero(f.data(), d, px, py + 1, pz);
                    nb = nb + rd_zero(f.data(), d, px, py - 1, pz);
                    nb = nb + rd_zero(f.data(), d, px, py, pz + 1);
                    nb = nb + rd_zero(f.data(), d, px, py, pz - 1);
                    T residue = (mu * (nb - (T)6 * c)) / ((T)1 * (T)1);
                    r[d.at(x, y, z)] = -sq * v0c - (residue - sq * c);
                }
        return r;
    }
};

// runFMGGVF, gradientVectorFlow.cpp:729-1005.
template <typename T>
static void gvf_fmg(const T* F, T* out, Dims d, float mu, int iters, int v0, int v1, int v2, int lmax) {
    const i64 N = d.n();
    FMG<T> M{(T)mu, v0, v1, v2, lmax};
    std::vector<T> s(N);
    for (i64 i = 0; i < N; i++) s[i] = F[i] * F[i] + F[N + i] * F[N + i] + F[2 * N + i] * F[2 * N + i];
    for (int c = 0; c < 3; c++) {
        std::vector<T> f(N, (T)0);
        for (int it = 0; it < iters; it++) {
            std::vector<T> r = FMG<T>::fmg_residual(F, f, d, (T)mu, c);
            std::vector<T> e = M.fmg(r, s, 0, d, (T)1);
            for (i64 i = 0; i < N; i++) f[i] = f[i] + e[i];
        }
        std::memcpy(out + c * N, f.data(), sizeof(T) * N);
    }
    // MGGVFFinish, cl:2202-2216: w = |v| with no zero guard.
    for (i64 i = 0; i < N; i++)
        out[3 * N + i] = std::sqrt(out[i] * out[i] + out[N + i] * out[N + i] + out[2 * N + i] * out[2 * N + i]);
}

// ---------------------------------------------------------------- eigen (A9)
// Householder tridiagonalisation followed by the implicit QL method with
// eigenvector accumulation, as in the EISPACK routines tred2 and tql2 (Bowdler,
// Martin, Reinsch, Wilkinson; Handbook for Automatic Computation II) in their
// public-domain JAMA form, n = 3.  The reference's copy is at
// kernels_no_3d_write.cl:1824-2068 (CPU copy eigenanalysisOfHessian.cpp:10-260).
// The operation order below is the one those routines define, so that a float
// instantiation rounds exactly like the reference's float code.
template <typename T>
static void householder3(T V[3][3], T d[3], T e[3]) {
    const int n = 3;
    for (int j = 0; j < n; j++) d[j] = V[n - 1][j];
    for (int i = n - 1; i > 0; i--) {
        T scale = 0, h = 0;
        for (int k = 0; k < i; k++) scale = scale + std::fabs(d[k]);
        if (scale == (T)0) {
            e[i] = d[i - 1];
            for (int j = 0; j < i; j++) {
                d[j] = V[i - 1][j];
                V[i][j] = 0;
                V[j][i] = 0;
            }
        } else {
            for (int k = 0; k < i; k++) {
                d[k] /= scale;
                h += d[k] * d[k];
            }
            T f = d[i - 1];
            T g = std::sqrt(h);
            if (f > 0) g = -g;
            e[i] = scale * g;
            h = h - f * g;
            d[i - 1] = f - g;
            for (int j = 0; j < i; j++) e[j] = 0;
            for (int j = 0; j < i; j++) {
                f = d[j];
                V[j][i] = f;
                g = e[j] + V[j][j] * f;
                for (int k = j + 1; k <= i - 1; k++) {
                    g += V[k][j] * d[k];
                    e[k] += V[k][j] * f;
                }
                e[j] = g;
            }
            f = 0;
            for (int j = 0; j < i; j++) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            T hh = f / (h + h);
            for (int j = 0; j < i; j++) e[j] -= hh * d[j];
            for (int j = 0; j < i; j++) {
                f = d[j];
                g = e[j];
                for (int k = j; k <= i - 1; k++) V[k][j] -= (f * e[k] + g * d[k]);
                d[j] = V[i - 1][j];
                V[i][j] = 0;
            }
        }
        d[i] = h;
    }
    for (int i = 0; i < n - 1; i++) {
        V[n - 1][i] = V[i][i];
        V[i][i] = 1;
        T h = d[i + 1];
        if (h != (T)0) {
            for (int k = 0; k <= i; k++) d[k] = V[k][i + 1] / h;
            for (int j = 0; j <= i; j++) {
                T g = 0;
                for (int k = 0; k <= i; k++) g += V[k][i + 1] * V[k][j];
                for (int k = 0; k <= i; k++) V[k][j] -= g * d[k];
            }
        }
        for (int k = 0; k <= i; k++) V[k][i + 1] = 0;
    }
    for (int j = 0; j < n; j++) {
        d[j] = V[n - 1][j];
        V[n - 1][j] = 0;
    }
    V[n - 1][n - 1] = 1;
    e[0] = 0;
}

template <typename T>
static inline T hyp(T a, T b) { return std::sqrt(a * a + b * b); }

template <typename T>
static void ql_implicit3(T V[3][3], T d[3], T e[3], bool* converged = nullptr) {
    const int n = 3;
    for (int i = 1; i < n; i++) e[i - 1] = e[i];
    e[n - 1] = 0;
    T f = 0, tst1 = 0;
    const T eps = (T)std::ldexp(1.0, -52);
    for (int l = 0; l < n; l++) {
        tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
        int m = l;
        while (m < n - 1 && !(std::fabs(e[m]) <= eps * tst1)) m++;
        if (m > l) {
            int iter = 0;
            do {
                iter++;
                T g = d[l];
                T p = (d[l + 1] - g) / ((T)2 * e[l]);
                T r = hyp(p, (T)1);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                T dl1 = d[l + 1];
                T h = g - d[l];
                for (int i = l + 2; i < n; i++) d[i] -= h;
                f = f + h;
                p = d[m];
                T c = 1, c2 = c, c3 = c;
                T el1 = e[l + 1];
                T s = 0, s2 = 0;
                for (int i = m - 1; i >= l; i--) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    r = hyp(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    for (int k = 0; k < n; k++) {
                        h = V[k][i + 1];
                        V[k][i + 1] = s * V[k][i] + c * h;
                        V[k][i] = c * V[k][i] - s * h;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (std::fabs(e[l]) > eps * tst1 && iter < TSF_QL_MAX_ITER);
            if (converged && std::fabs(e[l]) > eps * tst1) *converged = false;
        }
        d[l] = d[l] + f;
        e[l] = 0;
    }
    // Selection sort by ascending |lambda| with column swaps.
    for (int i = 0; i < n - 1; i++) {
        int k = i;
        T p = d[i];
        for (int j = i + 1; j < n; j++)
            if (std::fabs(d[j]) < std::fabs(p)) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; j++) {
                p = V[j][i];
                V[j][i] = V[j][k];
                V[j][k] = p;
            }
        }
    }
}

template <typename T>
static void eigen3(const T A[3][3], T V[3][3], T d[3], bool* converged = nullptr) {
    T e[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) V[i][j] = A[i][j];
    householder3(V, d, e);
    ql_implicit3(V, d, e, converged);
}

// The precision form's eigensolve (tsf_tdf_circle_mixed): the float
// decomposition, redone in double from the same float matrix and rounded to
// float when two eigenvalues lie within 2^-10 of the largest |eigenvalue| of
// each other or the float QL stopped at its sweep cap.  For T = double the
// plain decomposition.
template <typename T>
static void eigen3_fp64_fallback(const T A[3][3], T V[3][3], T d[3]) {
    bool conv = true;
    eigen3(A, V, d, &conv);
    if (!std::is_same<T, float>::value) return;
    const T m = std::max(std::fabs(d[0]), std::max(std::fabs(d[1]), std::fabs(d[2])));
    const T g = std::min(std::fabs(d[1] - d[0]), std::min(std::fabs(d[2] - d[1]), std::fabs(d[2] - d[0])));
    if (conv && !(g < (T)std::ldexp(1.0, -10) * m)) return;
    double Ad[3][3], Vd[3][3], dd[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ad[i][j] = (double)A[i][j];
    eigen3(Ad, Vd, dd);
    for (int i = 0; i < 3; i++) {
        d[i] = (T)dd[i];
        for (int j = 0; j < 3; j++) V[i][j] = (T)Vd[i][j];
    }
}

// ---------------------------------------------------------------- precision form: 16-bit state, stochastic rounding
// Restatement of tsf_gvf_euler_sr (csrc/gvf_sr.hip), the precision form of
// GVF3DIteration (kernels_no_3d_write.cl:1746-1787): the iterate is stored
// as 16-bit numbers (F16 holding value * scale, or BF16) and every launch of
// two Jacobi steps rounds its result stochastically with random bits from a
// hash of the voxel index and the launch; the last launch writes floats and
// w = |v|.  The arithmetic is in storage units and in the device kernel's
// order, fma included, so the device result is reproduced bit for bit.

static inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// IEEE binary16 <-> binary32 (round to nearest even), by bit manipulation.
static inline float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        const float v = std::ldexp((float)m, -24);  // subnormal (or zero): exact
        return s ? -v : v;
    }
    if (e == 31) return bits_f(s | 0x7f800000u | (m << 13));
    return bits_f(s | ((e + 112u) << 23) | (m << 13));
}
static inline uint16_t float_to_half_rne(float f) {
    const uint32_t x = f_bits(f), s = (x >> 16) & 0x8000u, a = x & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(s | 0x7e00u);   // NaN
    if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00u);  // rounds to infinity (>= 65520)
    if (a < 0x38800000u) {                                  // below 2^-14: a half subnormal
        const float q = bits_f(a) * 16777216.0f;             // in units of 2^-24, exact
        return (uint16_t)(s | (uint32_t)std::nearbyint(q));  // ties to even (default rounding mode)
    }
    uint32_t r = a - 0x38000000u;  // rebias the exponent by 127 - 15
    r = r + 0xfffu + ((r >> 13) & 1u);
    return (uint16_t)(s | (r >> 13));
}
// binary32 -> binary16 rounding toward zero (v_cvt_pkrtz_f16_f32): the
// magnitude truncated, finite values beyond the range to the largest half.
static inline uint16_t float_to_half_rtz(float f) {
    const uint32_t x = f_bits(f), s = (x >> 16) & 0x8000u, a = x & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(s | 0x7e00u);   // NaN
    if (a == 0x7f800000u) return (uint16_t)(s | 0x7c00u);  // infinity
    if (a >= 0x47800000u) return (uint16_t)(s | 0x7bffu);  // >= 65536: largest finite
    if (a < 0x38800000u) {                                  // below 2^-14: a half subnormal
        const float q = bits_f(a) * 16777216.0f;             // in units of 2^-24, exact
        return (uint16_t)(s | (uint32_t)q);                  // truncated
    }
    return (uint16_t)(s | ((a - 0x38000000u) >> 13));
}
static inline uint16_t float_to_bf16_rne(float f) {
    const uint32_t x = f_bits(f);
    return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

// storage 1 = F16, 2 = BF16 (tsf_storage)
static inline float s16_get(uint16_t h, int st) { return st == 1 ? half_to_float(h) : bits_f((uint32_t)h << 16); }
static inline uint16_t s16_rne(float v, int st) { return st == 1 ? float_to_half_rne(v) : float_to_bf16_rne(v); }
// r in [0, 1024) scaled to the bits below the 16-bit mantissa (13 for F16,
// 16 for BF16), added to the float's bits, then rounded toward zero: the
// magnitude rounds up with probability within 2^-10 of (discarded bits) /
// 2^bits.
static inline uint16_t s16_sr(float v, uint32_t r, int st) {
    const uint32_t b = f_bits(v);
    if (st == 1) return float_to_half_rtz(bits_f(b + (r << 3)));
    return (uint16_t)((b + (r << 6)) >> 16);
}
// The lowbias32 mixer with 24-bit multiplies (the device's full-rate
// v_mul_u32_u24), of index ^ seed; bits 0-9, 10-19, 20-29 are the random
// numbers of components x, y, z.
static inline uint32_t umul24(uint32_t a, uint32_t b) {
    return (uint32_t)((uint64_t)(a & 0xffffffu) * (uint64_t)(b & 0xffffffu));
}
static inline uint32_t sr_hash(uint32_t i, uint32_t seed) {
    uint32_t h = i ^ seed;
    h ^= h >> 16;
    h = umul24(h, 0x7feb35u);
    h ^= h >> 15;
    h = umul24(h, 0x846ca7u);
    h ^= h >> 16;
    return h;
}

// One step of a two-step launch at interior voxel (x,y,z):
// v' = v (1 - 6 mu - |f|^2) + f |f|^2 + mu (sum of the six neighbours), in
// the device's order: s = (u(x-1) + u(x+1)) + (u(y-1) + u(y+1)),
// A = fma(mu, fma(zmul, u(z-1), s), fma(u, c1, g)), out = fma(mz, u(z+1), A)
// with c1 = (1 - 6 mu) - |f|^2, g = f |f|^2.  The mirror rule in z for the
// second step: on plane 1 the z-1 term is left out (zmul 0) and the z+1
// term doubled (mz 2 mu), on plane Z-2 the z-1 term is doubled and the z+1
// term left out.
static inline float sr_stage(const float* u, const float* c1, const float* g, Dims d, int x, int y, int z, float mu,
                             bool zmirror) {
    const i64 i = d.at(x, y, z), sx = 1, sy = d.X, sz = (i64)d.X * d.Y;
    const float s = (u[i - sx] + u[i + sx]) + (u[i - sy] + u[i + sy]);
    float zmul = 1.0f, mz = mu;
    if (zmirror && z == 1) {
        zmul = 0.0f;
        mz = 2.0f * mu;
    } else if (zmirror && z == d.Z - 2) {
        zmul = 2.0f;
        mz = 0.0f;
    }
    const float A = std::fma(mu, std::fma(zmul, u[i - sz], s), std::fma(u[i], c1[i], g[i]));
    return std::fma(mz, u[i + sz], A);
}

// f's stored numbers in units of the iterate: F16 / BF16 as stored (value *
// scale), SNORM16 (storage 3; the iterate then F16 at unit 2^14) as
// max(s, -32767) * (unit / 32767).
static inline float sr_f_units(uint16_t c, int st, float unit) {
    if (st == 3) return std::max((float)(int16_t)c, -32767.0f) * (unit / 32767.0f);
    return s16_get(c, st);
}
static inline float sr_unit(int st, float scale) { return st == 3 ? 16384.0f : scale; }

static void gvf_euler_sr16_body(const uint16_t* f16, int st, float scale, float mu, int iters, uint32_t seed,
                                Dims d, float* out4);

// The last `tail` iterations in float (gvf_step, the reference's order) on
// the stored f's values, from the 16-bit part's float result; then w.
static void gvf_euler_sr16(const uint16_t* f16, int st, float scale, float mu, int iters, int tail, uint32_t seed,
                           Dims d, float* out4) {
    tail = std::min(tail, iters);
    gvf_euler_sr16_body(f16, st, scale, mu, iters - tail, seed, d, out4);
    if (tail == 0) return;
    const i64 N = d.n();
    const float unit = sr_unit(st, scale), inv = 1.0f / unit;
    std::vector<float> f(3 * N), v(out4, out4 + 3 * N), o(3 * N);
    for (i64 i = 0; i < 3 * N; i++) f[i] = sr_f_units(f16[i], st, unit) * inv;
    for (int k = 0; k < tail; k++) {
        gvf_step<float>(f.data(), v.data(), o.data(), d, mu);
        v.swap(o);
    }
    for (i64 i = 0; i < N; i++) {
        const float x = v[i], y = v[N + i], z = v[2 * N + i];
        const float l = std::sqrt(x * x + y * y + z * z);
        out4[i] = x;
        out4[N + i] = y;
        out4[2 * N + i] = z;
        out4[3 * N + i] = l > 0.0f ? l : 1.0f;
    }
}

static void gvf_euler_sr16_body(const uint16_t* f16, int st, float scale, float mu, int iters, uint32_t seed,
                                Dims d, float* out4) {
    const i64 N = d.n();
    const float unit = sr_unit(st, scale), inv = 1.0f / unit, m2s = 1.0f / (unit * unit);
    const int sv = st == 3 ? 1 : st;  // the iterate's storage
    const float c0 = 1.0f - 6.0f * mu;
    std::vector<float> f(3 * N), c1(N), g(3 * N);
    for (i64 i = 0; i < 3 * N; i++) f[i] = sr_f_units(f16[i], st, unit);
    for (i64 i = 0; i < N; i++) {
        const float m2 = std::fma(f[2 * N + i], f[2 * N + i], std::fma(f[N + i], f[N + i], f[i] * f[i])) * m2s;
        c1[i] = c0 - m2;
        for (int c = 0; c < 3; c++) g[c * N + i] = f[c * N + i] * m2;
    }
    std::vector<uint16_t> v16(3 * N);
    bool from_f = true;  // the iterate is still f (the first launch reads v0 from f's planes)
    std::vector<float> u(3 * N), v1(3 * N), v2(3 * N);
    auto faces = [&](std::vector<float>& a) {
        for (int c = 0; c < 3; c++)
            for (int z = 0; z < d.Z; z++)
                for (int y = 0; y < d.Y; y++)
                    for (int x = 0; x < d.X; x++)
                        if (x == 0 || y == 0 || z == 0 || x == d.X - 1 || y == d.Y - 1 || z == d.Z - 1)
                            a[c * N + d.at(x, y, z)] = a[c * N + d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z))];
    };
    int j = 0;
    if (iters % 2) {
        // one step in value space, the reference's order, round to nearest
        std::vector<float> fv(3 * N), vv(3 * N), o(3 * N);
        for (i64 i = 0; i < 3 * N; i++) fv[i] = vv[i] = f[i] * inv;
        gvf_step<float>(fv.data(), vv.data(), o.data(), d, mu);
        for (i64 i = 0; i < 3 * N; i++) v16[i] = s16_rne(o[i] * unit, sv);
        from_f = false;
        j = 1;
    }
    const int pairs = iters / 2;
    for (int k = 0; k < pairs; k++, j++) {
        for (i64 i = 0; i < 3 * N; i++) u[i] = from_f ? f[i] : s16_get(v16[i], sv);
        from_f = false;
        for (int c = 0; c < 3; c++)
            for (int z = 1; z < d.Z - 1; z++)
                for (int y = 1; y < d.Y - 1; y++)
                    for (int x = 1; x < d.X - 1; x++)
                        v1[c * N + d.at(x, y, z)] =
                            sr_stage(u.data() + c * N, c1.data(), g.data() + c * N, d, x, y, z, mu, false);
        faces(v1);
        for (int c = 0; c < 3; c++)
            for (int z = 1; z < d.Z - 1; z++)
                for (int y = 1; y < d.Y - 1; y++)
                    for (int x = 1; x < d.X - 1; x++)
                        v2[c * N + d.at(x, y, z)] =
                            sr_stage(v1.data() + c * N, c1.data(), g.data() + c * N, d, x, y, z, mu, true);
        faces(v2);
        if (k == pairs - 1) {
            for (i64 i = 0; i < N; i++) {
                const float x = v2[i] * inv, y = v2[N + i] * inv, z = v2[2 * N + i] * inv;
                const float l = std::sqrt(x * x + y * y + z * z);
                out4[i] = x;
                out4[N + i] = y;
                out4[2 * N + i] = z;
                out4[3 * N + i] = l > 0.0f ? l : 1.0f;
            }
            return;
        }
        // every voxel rounded with its own random bits (the x faces hold their
        // mirror's value here); the y and z faces then copy their mirror's
        // stored value
        const uint32_t sj = seed ^ ((uint32_t)j * 0x9e3779b9u);
        for (i64 i = 0; i < N; i++) {
            const uint32_t h = sr_hash((uint32_t)i, sj);
            const uint32_t r[3] = {h & 0x3ffu, (h >> 10) & 0x3ffu, (h >> 20) & 0x3ffu};
            for (int c = 0; c < 3; c++) v16[c * N + i] = s16_sr(v2[c * N + i], r[c], sv);
        }
        for (int c = 0; c < 3; c++)
            for (int z = 0; z < d.Z; z++)
                for (int y = 0; y < d.Y; y++)
                    for (int x = 0; x < d.X; x++)
                        if (y == 0 || z == 0 || y == d.Y - 1 || z == d.Z - 1)
                            v16[c * N + d.at(x, y, z)] = v16[c * N + d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z))];
    }
    // no two-step launch: the output is the 16-bit iterate (f, or the single step)
    auto val = [&](i64 i) { return (from_f ? f[i] : s16_get(v16[i], sv)) * inv; };
    for (i64 i = 0; i < N; i++) {
        const float x = val(i), y = val(N + i), z = val(2 * N + i);
        const float l = std::sqrt(x * x + y * y + z * z);
        out4[i] = x;
        out4[N + i] = y;
        out4[2 * N + i] = z;
        out4[3 * N + i] = l > 0.0f ? l : 1.0f;
    }
}

// ---------------------------------------------------------------- Hessian (A8)
// gradient / gradientNormalized on the vector field (kernels_no_3d_write.cl:90-200)
// and the matrix layout of circleFittingTDF (1522-1526).  With normalize, each
// sample is divided by the w plane of the sampled voxel.  Clamp-to-edge reads.
template <typename T>
static void hessian_at(const T* F, Dims d, int x, int y, int z, bool normalize, T H[3][3]) {
    const i64 N = d.n();
    auto s = [&](int c, int xx, int yy, int zz) -> T {
        T v = rd_edge(F + c * N, d, xx, yy, zz);
        if (normalize) v = v / rd_edge(F + 3 * N, d, xx, yy, zz);
        return v;
    };
    T fxx = (T)0.5 * (s(0, x + 1, y, z) - s(0, x - 1, y, z));
    T fyx = (T)0.5 * (s(1, x + 1, y, z) - s(1, x - 1, y, z));
    T fyy = (T)0.5 * (s(1, x, y + 1, z) - s(1, x, y - 1, z));
    T fzx = (T)0.5 * (s(2, x + 1, y, z) - s(2, x - 1, y, z));
    T fzy = (T)0.5 * (s(2, x, y + 1, z) - s(2, x, y - 1, z));
    T fzz = (T)0.5 * (s(2, x, y, z + 1) - s(2, x, y, z - 1));
    H[0][0] = fxx; H[0][1] = fyx; H[0][2] = fzx;
    H[1][0] = fyx; H[1][1] = fyy; H[1][2] = fzy;
    H[2][0] = fzx; H[2][1] = fzy; H[2][2] = fzz;
}

template <typename T>
static void hessian_eigen(const T* F, Dims d, int normalize, T* lam, T* vec) {
    const i64 N = d.n();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                T H[3][3], V[3][3], ev[3];
                hessian_at(F, d, x, y, z, normalize != 0, H);
                eigen3(H, V, ev);
                i64 i = d.at(x, y, z);
                for (int k = 0; k < 3; k++) lam[k * N + i] = ev[k];
                if (vec)  // plane 3*j + k = component k of eigenvector j
                    for (int j = 0; j < 3; j++)
                        for (int k = 0; k < 3; k++) vec[(3 * j + k) * N + i] = V[k][j];
            }
}

// ---------------------------------------------------------------- TDF (A11)
// The 32 directions are taken at integer radians j = 0..31 with the tables
// rounded to 6 significant digits ("%g"), as in kernels_no_3d_write.cl:1500-1501.
static float g_cos[32], g_sin[32];
static bool g_tab = false;
static void make_tables() {
    if (g_tab) return;
    char buf[64];
    for (int j = 0; j < 32; j++) {
        snprintf(buf, sizeof buf, "%g", std::cos((double)j));
        g_cos[j] = strtof(buf, nullptr);
        snprintf(buf, sizeof buf, "%g", std::sin((double)j));
        g_sin[j] = strtof(buf, nullptr);
    }
    g_tab = true;
}

// OpenCL CLK_FILTER_LINEAR with unnormalised coordinates and clamp-to-edge
// (OpenCL 1.2 spec 8.2): u' = u - 0.5, i0 = floor(u'), a = u' - i0, indices
// clamped to [0, N-1]; the eight terms are summed in the spec's order.
template <typename T>
static inline void trilinear3(const T* F, Dims d, T u, T v, T w, T out[3]) {
    const i64 N = d.n();
    T uu = u - (T)0.5, vv = v - (T)0.5, ww = w - (T)0.5;
    T fu = std::floor(uu), fv = std::floor(vv), fw = std::floor(ww);
    T a = uu - fu, b = vv - fv, c = ww - fw;
    int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    int i1 = clampi(i0 + 1, 0, d.X - 1), j1 = clampi(j0 + 1, 0, d.Y - 1), k1 = clampi(k0 + 1, 0, d.Z - 1);
    i0 = clampi(i0, 0, d.X - 1); j0 = clampi(j0, 0, d.Y - 1); k0 = clampi(k0, 0, d.Z - 1);
    T wt[8] = {((T)1 - a) * ((T)1 - b) * ((T)1 - c), a * ((T)1 - b) * ((T)1 - c),
               ((T)1 - a) * b * ((T)1 - c),          a * b * ((T)1 - c),
               ((T)1 - a) * ((T)1 - b) * c,          a * ((T)1 - b) * c,
               ((T)1 - a) * b * c,                    a * b * c};
    i64 idx[8] = {d.at(i0, j0, k0), d.at(i1, j0, k0), d.at(i0, j1, k0), d.at(i1, j1, k0),
                  d.at(i0, j0, k1), d.at(i1, j0, k1), d.at(i0, j1, k1), d.at(i1, j1, k1)};
    for (int comp = 0; comp < 3; comp++) {
        const T* p = F + comp * N;
        T s = wt[0] * p[idx[0]];
        for (int k = 1; k < 8; k++) s = s + wt[k] * p[idx[k]];
        out[comp] = s;
    }
}

// The precision form's trilinear sample (tsf_tdf_circle_mixed): the cell and
// weights of trilinear3 at uu = u - 0.5 given directly, the eight taps
// accumulated by fused multiply-adds in the same order.
template <typename T>
static inline void trilinear3_fused(const T* F, Dims d, T uu, T vv, T ww, T out[3]) {
    const i64 N = d.n();
    T fu = std::floor(uu), fv = std::floor(vv), fw = std::floor(ww);
    T a = uu - fu, b = vv - fv, c = ww - fw;
    int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    int i1 = clampi(i0 + 1, 0, d.X - 1), j1 = clampi(j0 + 1, 0, d.Y - 1), k1 = clampi(k0 + 1, 0, d.Z - 1);
    i0 = clampi(i0, 0, d.X - 1); j0 = clampi(j0, 0, d.Y - 1); k0 = clampi(k0, 0, d.Z - 1);
    T wt[8] = {((T)1 - a) * ((T)1 - b) * ((T)1 - c), a * ((T)1 - b) * ((T)1 - c),
               ((T)1 - a) * b * ((T)1 - c),          a * b * ((T)1 - c),
               ((T)1 - a) * ((T)1 - b) * c,          a * ((T)1 - b) * c,
               ((T)1 - a) * b * c,                    a * b * c};
    i64 idx[8] = {d.at(i0, j0, k0), d.at(i1, j0, k0), d.at(i0, j1, k0), d.at(i1, j1, k0),
                  d.at(i0, j0, k1), d.at(i1, j0, k1), d.at(i0, j1, k1), d.at(i1, j1, k1)};
    for (int comp = 0; comp < 3; comp++) {
        const T* p = F + comp * N;
        T s = wt[0] * p[idx[0]];
        for (int k = 1; k < 8; k++) s = std::fma(wt[k], p[idx[k]], s);
        out[comp] = s;
    }
}

// circleFittingTDF, kernels_no_3d_write.cl:1503-1572, with the Hessian taken
// from FH and the trilinear samples from FS.  FH == FS is the reference's
// filter; FS = FH rounded to 16 bits is the precision form's mixed mode
// (the Hessian at full precision, the samples from 16-bit texels).  FUSED:
// the precision form's sample arithmetic -- the direction fma(c, e3, s e2),
// the position fma(r, V, x - 0.5), fused taps, the dot product as
// fma(s2, v2, fma(s1, v1, s0 v0)) subtracted from the sum.
template <typename T, bool FUSED = false>
static void tdf_circle_split(const T* FH, const T* F, Dims d, float rMin, float rMax, float rStep, T* tdf, T* rad) {
    make_tables();
    const bool normalize = !(rMax < 4);
#pragma omp parallel for schedule(dynamic, 4)
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                T H[3][3], V[3][3], ev[3];
                hessian_at(FH, d, x, y, z, normalize, H);
                if (FUSED) eigen3_fp64_fallback(H, V, ev);
                else eigen3(H, V, ev);
                const T e2[3] = {V[0][1], V[1][1], V[2][1]};
                const T e3[3] = {V[0][2], V[1][2], V[2][2]};
                T maxSum = 0, maxRadius = 0;
                for (float radius = rMin; radius <= rMax; radius += rStep) {
                    T sum = 0;
                    for (int j = 0; j < 32; j++) {
                        T va[3], s[3];
                        if (FUSED) {
                            for (int k = 0; k < 3; k++) va[k] = std::fma((T)g_cos[j], e3[k], (T)g_sin[j] * e2[k]);
                            trilinear3_fused(F, d, std::fma((T)radius, va[0], (T)x - (T)0.5),
                                             std::fma((T)radius, va[1], (T)y - (T)0.5),
                                             std::fma((T)radius, va[2], (T)z - (T)0.5), s);
                            sum = sum - std::fma(s[2], va[2], std::fma(s[1], va[1], s[0] * va[0]));
                        } else {
                            for (int k = 0; k < 3; k++) va[k] = (T)g_cos[j] * e3[k] + (T)g_sin[j] * e2[k];
                            trilinear3(F, d, (T)x + (T)radius * va[0], (T)y + (T)radius * va[1],
                                       (T)z + (T)radius * va[2], s);
                            T dp = (-s[0]) * va[0] + (-s[1]) * va[1] + (-s[2]) * va[2];
                            sum += dp;
                        }
                    }
                    sum /= (T)32;
                    if (sum > maxSum) {
                        maxSum = sum;
                        maxRadius = (T)radius;
                    } else {
                        break;
                    }
                }
                i64 i = d.at(x, y, z);
                tdf[i] = maxSum;
                rad[i] = maxRadius;
            }
}

template <typename T>
static void tdf_circle(const T* F, Dims d, float rMin, float rMax, float rStep, T* tdf, T* rad) {
    tdf_circle_split(F, F, d, rMin, rMax, rStep, tdf, rad);
}

// splineTDF, kernels_no_3d_write.cl:1574-1667 (enqueued by runSplineTDF,
// tubeDetectionFilters.cpp:4-55, with 12 arms and minAverageMag 0.1): along
// `arms` rays in the plane of e2, e3 (directions alpha_j = 2 pi j / arms,
// M_PI_F in float), the first radius where the field magnitude, after
// rising, falls and its two-sample mean exceeds the voxel's own by
// minAverageMag.  Radius = sum of the border radii found / arms; TDF = 0
// when an arm finds none or its border field points outward, else
// min(1, mean(1 - |n . e1|) * symmetry + 0.2), symmetry the mean of
// min/max over opposite arms.  The direction tables are cos and sin of the
// float angle, correctly rounded.
static void spline_tables(int arms, float* c, float* s) {
    for (int j = 0; j < arms; j++) {
        const float alpha = 2.0f * 3.14159274f * (float)j / (float)arms;
        c[j] = (float)std::cos((double)alpha);
        s[j] = (float)std::sin((double)alpha);
    }
}
template <typename T>
static void tdf_spline(const T* F, Dims d, float rMin, float rMax, float rStep, int arms, float minAvg, T* tdf,
                       T* rad) {
    float ct[12], st[12];
    spline_tables(arms, ct, st);
    const i64 N = d.n();
    auto len = [](const T v[3]) { return (T)std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
#pragma omp parallel for schedule(dynamic, 4)
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                T H[3][3], V[3][3], ev[3];
                hessian_at(F, d, x, y, z, true, H);
                eigen3(H, V, ev);
                const T e1[3] = {V[0][0], V[1][0], V[2][0]};
                const T e2[3] = {V[0][1], V[1][1], V[2][1]};
                const T e3[3] = {V[0][2], V[1][2], V[2][2]};
                const i64 i = d.at(x, y, z);
                const T f0[3] = {F[i], F[N + i], F[2 * N + i]};
                const T cur = len(f0);
                T maxR[12], avgR = 0, sum = 0;
                bool invalid = false;
                for (int j = 0; j < arms; j++) {
                    maxR[j] = 999;
                    T va[3];
                    for (int k = 0; k < 3; k++) va[k] = (T)ct[j] * e3[k] + (T)st[j] * e2[k];
                    const T prev2 = cur;
                    T s[3];
                    trilinear3(F, d, (T)x + (T)rMin * va[0], (T)y + (T)rMin * va[1], (T)z + (T)rMin * va[2], s);
                    T prev = len(s);
                    int up = prev2 > prev ? 0 : 1;
                    for (float radius = rMin + rStep; radius <= rMax; radius += rStep) {
                        trilinear3(F, d, (T)x + (T)radius * va[0], (T)y + (T)radius * va[1],
                                   (T)z + (T)radius * va[2], s);
                        const T mag = len(s);
                        if (up == 1 && mag < prev && (prev + mag) / (T)2 - cur > (T)minAvg) {
                            maxR[j] = (T)radius;
                            avgR += (T)radius;
                            const T lv = len(va);
                            T nv[3], nva[3];
                            for (int k = 0; k < 3; k++) {
                                nv[k] = s[k] / mag;
                                nva[k] = -(va[k] / lv);
                            }
                            if (nv[0] * nva[0] + nv[1] * nva[1] + nv[2] * nva[2] < (T)0) {
                                invalid = true;
                                sum = 0;
                            }
                            sum += (T)1 - std::fabs(nv[0] * e1[0] + nv[1] * e1[1] + nv[2] * e1[2]);
                            break;
                        }
                        if (mag > prev) up = 1;
                        prev = mag;
                    }
                    if (maxR[j] == (T)999 || invalid) {
                        invalid = true;
                        break;
                    }
                }
                avgR = avgR / (T)arms;
                rad[i] = avgR;
                if (!invalid) {
                    T sym = 0;
                    for (int j = 0; j < arms / 2; j++)
                        sym += std::min(maxR[j], maxR[arms / 2 + j]) / std::max(maxR[j], maxR[arms / 2 + j]);
                    sym /= (T)(arms / 2);
                    tdf[i] = std::fmin((T)1, (sum / (T)arms) * sym + (T)0.2f);
                } else {
                    tdf[i] = 0;
                }
            }
}

// ---------------------------------------------------------------- combine / candidates (A12, A13)
// combine, kernels_no_3d_write.cl:1087-1098.
template <typename T>
static void combine(const T* ts, const T* rs, T* tl, T* rl, i64 n) {
    for (i64 i = 0; i < n; i++)
        if (tl[i] < ts[i]) {
            tl[i] = ts[i];
            rl[i] = rs[i];
        }
}
// findCandidateCenterpoints, kernels_no_3d_write.cl:1006-1017.
template <typename T>
static void candidates(const T* tdf, uint8_t* mask, i64 n, float thr) {
    for (i64 i = 0; i < n; i++) mask[i] = tdf[i] < (T)thr ? 0 : 1;
}

// ---------------------------------------------------------------- NMS (A16)
// findCandidateCenterpoints2, kernels_no_3d_write.cl:1019-1084 (maxD cap 5 on
// the buffer path, 8 on the image path, kernels.cl:1334).
template <typename T>
static void candidate_nms(const T* tdf, const T* radius, const T* F, Dims d, const int32_t* pos, i64 count,
                          int maxDcap, uint8_t* out) {
    (void)tdf;
    const i64 N = d.n();
    const T* W = F + 3 * N;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 t = 0; t < count; t++) {
        const int x = pos[3 * t], y = pos[3 * t + 1], z = pos[3 * t + 2];
        const float rr = (float)rd_edge(radius, d, x, y, z);
        const int maxD = (int)std::max(std::min(std::round(rr), (float)maxDcap), 1.0f);
        T H[3][3], V[3][3], ev[3];
        hessian_at(F, d, x, y, z, true, H);
        eigen3(H, V, ev);
        const T e1[3] = {V[0][0], V[1][0], V[2][0]};
        const T wp = rd_edge(W, d, x, y, z);
        bool invalid = false;
        for (int a = -maxD; a <= maxD; a++)
            for (int b = -maxD; b <= maxD; b++)
                for (int c = -maxD; c <= maxD; c++) {
                    if (a == 0 && b == 0 && c == 0) continue;
                    T r[3] = {(T)a, (T)b, (T)c};
                    T dp = e1[0] * r[0] + e1[1] * r[1] + e1[2] * r[2];
                    T rp[3] = {r[0] - e1[0] * dp, r[1] - e1[1] * dp, r[2] - e1[2] * dp};
                    T lr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                    T lp = std::sqrt(rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2]);
                    T cs = (r[0] / lr) * (rp[0] / lp) + (r[1] / lr) * (rp[1] / lp) + (r[2] / lr) * (rp[2] / lp);
                    T theta = std::acos(cs);
                    if (theta < (T)0.5f && lr < (T)maxD) {
                        if (rd_edge(W, d, x + a, y + b, z + c) < wp) invalid = true;
                    }
                }
        out[d.at(x, y, z)] = invalid ? 0 : 1;
    }
}

// ---------------------------------------------------------------- cube max (A17)
// dd, kernels_no_3d_write.cl:970-1004 (in-range voxels only, see header).
template <typename T>
static void cube_max(const T* tdf, const uint8_t* cand, Dims d, int cube, uint8_t* out) {
    const int cx = (d.X + cube - 1) / cube, cy = (d.Y + cube - 1) / cube, cz = (d.Z + cube - 1) / cube;
    for (int gz = 0; gz < cz; gz++)
        for (int gy = 0; gy < cy; gy++)
            for (int gx = 0; gx < cx; gx++) {
                T best = 0;
                i64 bi = -1;
                for (int a = 0; a < cube; a++)
                    for (int b = 0; b < cube; b++)
                        for (int c = 0; c < cube; c++) {
                            int x = gx * cube + a, y = gy * cube + b, z = gz * cube + c;
                            if (x >= d.X || y >= d.Y || z >= d.Z) continue;
                            i64 i = d.at(x, y, z);
                            if (cand[i] == 1 && tdf[i] > best) {
                                best = tdf[i];
                                bi = i;
                            }
                        }
                if (bi >= 0) out[bi] = 1;
            }
}

// ---------------------------------------------------------------- HP compaction (A14, A15)
// Histogram-pyramid order: at every level the eight children are visited in
// the order (0,0,0),(1,0,0),(0,0,1),(1,0,1),(0,1,0),(1,1,0),(0,1,1),(1,1,1)
// (cubeOffsets, kernels_no_3d_write.cl:78-87), i.e. child k = x + 2z + 4y.
static const int kOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1},
                               {0, 1, 0}, {1, 1, 0}, {0, 1, 1}, {1, 1, 1}};

static inline uint64_t hp_key(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t k = 0;
    for (int b = 0; b < 21; b++) {
        uint64_t dgt = ((x >> b) & 1u) + 2u * ((z >> b) & 1u) + 4u * ((y >> b) & 1u);
        k |= dgt << (3 * b);
    }
    return k;
}

// Method 1: sort the set voxels by the pyramid key.
static i64 hp_compact_sort(const uint8_t* mask, Dims d, int32_t* pos, i64 cap) {
    std::vector<std::pair<uint64_t, i64>> v;
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++)
                if (mask[d.at(x, y, z)]) v.push_back({hp_key(x, y, z), d.at(x, y, z)});
    std::sort(v.begin(), v.end());
    i64 n = (i64)v.size();
    for (i64 t = 0; t < n && t < cap; t++) {
        i64 i = v[t].second;
        pos[3 * t] = (int32_t)(i % d.X);
        pos[3 * t + 1] = (int32_t)((i / d.X) % d.Y);
        pos[3 * t + 2] = (int32_t)(i / ((i64)d.X * d.Y));
    }
    return n;
}

// Method 2: build the pyramid literally (constructHPLevel*, cl:246-347: each
// cell is the sum of its 8 children, zero outside) and descend it for every
// target (scanHPLevel* / traverseHP3DBuffer, cl:393-616; target >= sum -> 0).
static i64 hp_compact_pyramid(const uint8_t* mask, Dims d, int32_t* pos, i64 cap) {
    int size = 8;
    while (size < d.X || size < d.Y || size < d.Z) size *= 2;
    int L = 0;
    while ((1 << L) < size) L++;
    // level l has size >> l cells per axis; level 0 is the mask padded with zeros
    std::vector<std::vector<i64>> lev(L + 1);
    lev[0].assign((i64)size * size * size, 0);
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++)
                lev[0][(i64)x + (i64)size * (y + (i64)size * z)] = mask[d.at(x, y, z)] ? 1 : 0;
    for (int l = 1; l <= L; l++) {
        int s = size >> l, sp = size >> (l - 1);
        lev[l].assign((i64)s * s * s, 0);
        for (int z = 0; z < s; z++)
            for (int y = 0; y < s; y++)
                for (int x = 0; x < s; x++) {
                    i64 acc = 0;
                    for (int k = 0; k < 8; k++)
                        acc += lev[l - 1][(i64)(2 * x + kOff[k][0]) +
                                          (i64)sp * ((2 * y + kOff[k][1]) + (i64)sp * (2 * z + kOff[k][2]))];
                    lev[l][(i64)x + (i64)s * (y + (i64)s * z)] = acc;
                }
    }
    const i64 sum = lev[L][0];
    for (i64 t0 = 0; t0 < sum && t0 < cap; t0++) {
        i64 target = t0;
        int px = 0, py = 0, pz = 0;
        i64 base = 0;
        for (int l = L - 1; l >= 0; l--) {
            int s = size >> l;
            i64 nb[8];
            for (int k = 0; k < 8; k++)
                nb[k] = lev[l][(i64)(px + kOff[k][0]) + (i64)s * ((py + kOff[k][1]) + (i64)s * (pz + kOff[k][2]))];
            i64 acc = base;
            int sel = 0;
            for (int k = 0; k < 7; k++) {
                acc += nb[k];
                if (acc <= target) sel++;
            }
            i64 add = 0;
            for (int k = 0; k < sel; k++) add += nb[k];
            base += add;
            px = 2 * (px + kOff[sel][0]);
            py = 2 * (py + kOff[sel][1]);
            pz = 2 * (pz + kOff[sel][2]);
        }
        pos[3 * t0] = px / 2;
        pos[3 * t0 + 1] = py / 2;
        pos[3 * t0 + 2] = pz / 2;
    }
    return sum;
}

// ---------------------------------------------------------------- doEigen (A10)
// eigenanalysisOfHessian.cpp:264-389: every sample is divided by the magnitude
// sqrt(Fx^2+Fy^2+Fz^2) computed on the host from the sampled voxel; there is no
// clamping (callers keep pos one voxel away from the border).
template <typename T>
static void do_eigen(const T* Fx, const T* Fy, const T* Fz, Dims d, int x, int y, int z, T* lam, T* e1, T* e2,
                     T* e3) {
    auto nv = [&](const T* c, int xx, int yy, int zz) -> T {
        i64 i = d.at(xx, yy, zz);
        return c[i] / std::sqrt(Fx[i] * Fx[i] + Fy[i] * Fy[i] + Fz[i] * Fz[i]);
    };
    T fxx = (T)0.5 * (nv(Fx, x + 1, y, z) - nv(Fx, x - 1, y, z));
    T fyx = (T)0.5 * (nv(Fy, x + 1, y, z) - nv(Fy, x - 1, y, z));
    T fyy = (T)0.5 * (nv(Fy, x, y + 1, z) - nv(Fy, x, y - 1, z));
    T fzx = (T)0.5 * (nv(Fz, x + 1, y, z) - nv(Fz, x - 1, y, z));
    T fzy = (T)0.5 * (nv(Fz, x, y + 1, z) - nv(Fz, x, y - 1, z));
    T fzz = (T)0.5 * (nv(Fz, x, y, z + 1) - nv(Fz, x, y, z - 1));
    T H[3][3] = {{fxx, fyx, fzx}, {fyx, fyy, fzy}, {fzx, fzy, fzz}};
    T V[3][3], ev[3];
    eigen3(H, V, ev);
    for (int k = 0; k < 3; k++) {
        lam[k] = ev[k];
        e1[k] = V[k][0];
        e2[k] = V[k][1];
        e3[k] = V[k][2];
    }
}

// ================================================================ adjoints
// Reverse-mode derivatives of the differentiable chain
//   B = blur_sigma(I), F = VF(B), V = GVF_N(F; mu), W = |V|, T = TDF(V, W)
// for a scalar L = sum_p gT(p) T(p).  Discrete choices are held fixed: the
// blur mask size, the Fmax branch of the vector field, the eigenvalue order,
// and the TDF radius r* (T = 0 voxels carry no gradient).  These functions
// are the FP64 reference gradient; they are pinned by central finite
// differences of the forward chain in tests/test_oracle_adjoint.py.

// Normalised Gaussian weights in precision T and their sigma-derivative.
template <typename T>
static int blur_weights(T sigma, std::vector<T>& w, std::vector<T>* dw) {
    int m = (int)std::ceil((float)sigma / 0.5f);
    m = std::min(5, std::max(1, m));
    const int M = 2 * m + 1;
    std::vector<T> g(M * M * M), gp(M * M * M);
    T G = 0, Gp = 0;
    for (int a = -m; a <= m; a++)
        for (int b = -m; b <= m; b++)
            for (int c = -m; c <= m; c++) {
                const int i = (a + m) + (b + m) * M + (c + m) * M * M;
                const T r2 = (T)(a * a + b * b + c * c);
                g[i] = std::exp(-r2 / (2 * sigma * sigma));
                gp[i] = g[i] * r2 / (sigma * sigma * sigma);
                G += g[i];
                Gp += gp[i];
            }
    w.resize(g.size());
    if (dw) dw->resize(g.size());
    for (size_t i = 0; i < g.size(); i++) {
        w[i] = g[i] / G;
        if (dw) (*dw)[i] = (gp[i] * G - g[i] * Gp) / (G * G);
    }
    return m;
}

template <typename T>
static void blur_w(const T* in, T* out, Dims d, int m, const std::vector<T>& w) {
    const int M = 2 * m + 1;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                T s = 0;
                for (int a = -m; a <= m; a++)
                    for (int b = -m; b <= m; b++)
                        for (int c = -m; c <= m; c++)
                            s += w[(a + m) + (b + m) * M + (c + m) * M * M] * rd_edge(in, d, x + a, y + b, z + c);
                out[d.at(x, y, z)] = s;
            }
}

// dI = B^T dB (clamp-to-edge transpose), dsigma = sum_p dB(p) sum_k dw_k I(clamp(p+k)).
template <typename T>
static void blur_adjoint(const T* dB, const T* I, Dims d, T sigma, T* dI, double* dsigma) {
    std::vector<T> w, dw;
    const int m = blur_weights(sigma, w, &dw), M = 2 * m + 1;
    std::fill(dI, dI + d.n(), (T)0);
    double ds = 0;
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                const T g = dB[d.at(x, y, z)];
                if (g == (T)0) continue;
                T acc = 0;
                for (int a = -m; a <= m; a++)
                    for (int b = -m; b <= m; b++)
                        for (int c = -m; c <= m; c++) {
                            const int k = (a + m) + (b + m) * M + (c + m) * M * M;
                            const i64 j = d.at(clampi(x + a, 0, d.X - 1), clampi(y + b, 0, d.Y - 1),
                                               clampi(z + c, 0, d.Z - 1));
                            dI[j] += w[k] * g;
                            acc += dw[k] * I[j];
                        }
                ds += (double)g * (double)acc;
            }
    *dsigma = ds;
}

// Vector field adjoint (Fmax branch held fixed): dB from dF (planes x, y, z).
template <typename T>
static void vector_field_adjoint(const T* B, const T* dF, Dims d, float Fmax, int sign, T* dB) {
    const i64 N = d.n();
    std::fill(dB, dB + N, (T)0);
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                const int xm = std::max(x - 1, 0), xp = std::min(x + 1, d.X - 1);
                const int ym = std::max(y - 1, 0), yp = std::min(y + 1, d.Y - 1);
                const int zm = std::max(z - 1, 0), zp = std::min(z + 1, d.Z - 1);
                T g[3] = {(T)sign * (T)0.5 * (B[d.at(xp, y, z)] - B[d.at(xm, y, z)]),
                          (T)sign * (T)0.5 * (B[d.at(x, yp, z)] - B[d.at(x, ym, z)]),
                          (T)sign * (T)0.5 * (B[d.at(x, y, zp)] - B[d.at(x, y, zm)])};
                const i64 i = d.at(x, y, z);
                const T df[3] = {dF[i], dF[N + i], dF[2 * N + i]};
                const T l = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                T dg[3];
                if (l < (T)Fmax) {
                    for (int k = 0; k < 3; k++) dg[k] = df[k] / (T)Fmax;
                } else {
                    const T f[3] = {g[0] / l, g[1] / l, g[2] / l};
                    const T fd = f[0] * df[0] + f[1] * df[1] + f[2] * df[2];
                    for (int k = 0; k < 3; k++) dg[k] = (df[k] - f[k] * fd) / l;
                }
                const T h = (T)sign * (T)0.5;
                dB[d.at(xp, y, z)] += h * dg[0];
                dB[d.at(xm, y, z)] -= h * dg[0];
                dB[d.at(x, yp, z)] += h * dg[1];
                dB[d.at(x, ym, z)] -= h * dg[1];
                dB[d.at(x, y, zp)] += h * dg[2];
                dB[d.at(x, y, zm)] -= h * dg[2];
            }
}

// Mirror fold: Lam(q) = sum of lam(w) over all w with mirror(w) == q (zero
// on boundary voxels, which are never the image of the mirror map).
template <typename T>
static void mirror_fold(const T* lam, T* out, Dims d) {
    const i64 N = d.n();
    std::fill(out, out + N, (T)0);
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++)
                out[d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z))] += lam[d.at(x, y, z)];
}

// GVF adjoint: given lam = dL/dv^N (3 planes), the forward iterates
// states[n] (v^n, n = 0..N-1, 3 planes each) and F, returns dF (3 planes,
// including v^0 = F) and dmu.
template <typename T>
static void gvf_adjoint(const T* F, const T* states, Dims d, T mu, int N_it, const T* dV, T* dF, double* dmu) {
    const i64 N = d.n();
    std::vector<T> lam(dV, dV + 3 * N), Lam(3 * N), nl(3 * N);
    std::fill(dF, dF + 3 * N, (T)0);
    double gm = 0;
    const i64 sx = 1, sy = d.X, sz = (i64)d.X * d.Y;
    for (int n = N_it - 1; n >= 0; n--) {
        const T* v = states + (i64)n * 3 * N;
        for (int c = 0; c < 3; c++) mirror_fold(lam.data() + c * N, Lam.data() + c * N, d);
        for (int z = 0; z < d.Z; z++)
            for (int y = 0; y < d.Y; y++)
                for (int x = 0; x < d.X; x++) {
                    const i64 i = d.at(x, y, z);
                    const T f[3] = {F[i], F[N + i], F[2 * N + i]};
                    const T m2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
                    T s = 0;
                    for (int c = 0; c < 3; c++) {
                        const T* L = Lam.data() + c * N;
                        T nb = 0;
                        if (x + 1 < d.X) nb += L[i + sx];
                        if (x > 0) nb += L[i - sx];
                        if (y + 1 < d.Y) nb += L[i + sy];
                        if (y > 0) nb += L[i - sy];
                        if (z + 1 < d.Z) nb += L[i + sz];
                        if (z > 0) nb += L[i - sz];
                        nl[c * N + i] = ((T)1 - (T)6 * mu - m2) * L[i] + mu * nb;
                        s += L[i] * (f[c] - v[c * N + i]);
                    }
                    for (int c = 0; c < 3; c++) dF[c * N + i] += Lam[c * N + i] * m2 + (T)2 * f[c] * s;
                    const bool interior = x > 0 && y > 0 && z > 0 && x < d.X - 1 && y < d.Y - 1 && z < d.Z - 1;
                    if (interior)
                        for (int c = 0; c < 3; c++) {
                            const T* p = v + c * N;
                            const T lap = p[i + sx] + p[i - sx] + p[i + sy] + p[i - sy] + p[i + sz] + p[i - sz] -
                                          (T)6 * p[i];
                            gm += (double)Lam[c * N + i] * (double)lap;
                        }
                }
        lam.swap(nl);
    }
    for (i64 k = 0; k < 3 * N; k++) dF[k] += lam[k];
    *dmu = gm;
}

// Trilinear sample of planes 0..2 and its spatial Jacobian J[c][m] = d s_c / d x_m.
template <typename T>
static void trilinear_grad(const T* F, Dims d, T u, T v, T w, T s[3], T J[3][3], i64 idx[8], T wt[8]) {
    const i64 N = d.n();
    T uu = u - (T)0.5, vv = v - (T)0.5, ww = w - (T)0.5;
    T fu = std::floor(uu), fv = std::floor(vv), fw = std::floor(ww);
    T a = uu - fu, b = vv - fv, c = ww - fw;
    int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    int i1 = clampi(i0 + 1, 0, d.X - 1), j1 = clampi(j0 + 1, 0, d.Y - 1), k1 = clampi(k0 + 1, 0, d.Z - 1);
    i0 = clampi(i0, 0, d.X - 1); j0 = clampi(j0, 0, d.Y - 1); k0 = clampi(k0, 0, d.Z - 1);
    const T A[8] = {1 - a, a, 1 - a, a, 1 - a, a, 1 - a, a}, Bv[8] = {1 - b, 1 - b, b, b, 1 - b, 1 - b, b, b};
    const T C[8] = {1 - c, 1 - c, 1 - c, 1 - c, c, c, c, c};
    const T dA[8] = {-1, 1, -1, 1, -1, 1, -1, 1}, dBv[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const T dC[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    const i64 id[8] = {d.at(i0, j0, k0), d.at(i1, j0, k0), d.at(i0, j1, k0), d.at(i1, j1, k0),
                       d.at(i0, j0, k1), d.at(i1, j0, k1), d.at(i0, j1, k1), d.at(i1, j1, k1)};
    for (int k = 0; k < 8; k++) {
        idx[k] = id[k];
        wt[k] = A[k] * Bv[k] * C[k];
    }
    for (int comp = 0; comp < 3; comp++) {
        const T* p = F + comp * N;
        T sv = 0, g0 = 0, g1 = 0, g2 = 0;
        for (int k = 0; k < 8; k++) {
            const T t = p[id[k]];
            sv += wt[k] * t;
            g0 += dA[k] * Bv[k] * C[k] * t;
            g1 += A[k] * dBv[k] * C[k] * t;
            g2 += A[k] * Bv[k] * dC[k] * t;
        }
        s[comp] = sv;
        J[comp][0] = g0;
        J[comp][1] = g1;
        J[comp][2] = g2;
    }
}

// TDF adjoint: dV (3 planes) and dW (w plane) from gT, radius r* held fixed.
// ABS: the same scatter of |contribution|, a per-voxel scale of what the
// sums add up (for tolerances where contributions cancel).
template <typename T, bool ABS = false>
static void tdf_adjoint(const T* F, Dims d, float rMin, float rMax, float rStep, const T* gT, T* dV, T* dW) {
    auto m = [](T v) { return ABS ? std::fabs(v) : v; };
    make_tables();
    const i64 N = d.n();
    const bool norm = !(rMax < 4);
    std::vector<T> tdf(N), rad(N);
    tdf_circle(F, d, rMin, rMax, rStep, tdf.data(), rad.data());
    std::fill(dV, dV + 3 * N, (T)0);
    std::fill(dW, dW + N, (T)0);
    const T* W = F + 3 * N;
    for (int z = 0; z < d.Z; z++)
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++) {
                const i64 p = d.at(x, y, z);
                const T g = gT[p];
                if (g == (T)0 || !(tdf[p] > (T)0)) continue;
                const T r = rad[p];
                T H[3][3], E[3][3], lam[3];
                hessian_at(F, d, x, y, z, norm, H);
                eigen3(H, E, lam);
                T A2[3] = {0, 0, 0}, A3[3] = {0, 0, 0};
                for (int j = 0; j < 32; j++) {
                    T dj[3];
                    for (int k = 0; k < 3; k++) dj[k] = (T)g_cos[j] * E[k][2] + (T)g_sin[j] * E[k][1];
                    T s[3], J[3][3], wt[8];
                    i64 idx[8];
                    trilinear_grad(F, d, (T)x + r * dj[0], (T)y + r * dj[1], (T)z + r * dj[2], s, J, idx, wt);
                    const T cf = -g / (T)32;
                    for (int k = 0; k < 8; k++)
                        for (int c = 0; c < 3; c++) dV[c * N + idx[k]] += m(cf * dj[c] * wt[k]);
                    T u[3];
                    for (int mm = 0; mm < 3; mm++)
                        u[mm] = s[mm] + r * (dj[0] * J[0][mm] + dj[1] * J[1][mm] + dj[2] * J[2][mm]);
                    for (int mm = 0; mm < 3; mm++) {
                        A3[mm] += cf * (T)g_cos[j] * u[mm];
                        A2[mm] += cf * (T)g_sin[j] * u[mm];
                    }
                }
                // eigenvector perturbation: de_k = sum_{i != k} e_i (e_i^T dH e_k) / (lam_k - lam_i)
                T G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                const T* Ak[2] = {A2, A3};
                for (int kk = 0; kk < 2; kk++) {
                    const int k = kk + 1;
                    for (int i = 0; i < 3; i++) {
                        if (i == k || lam[k] == lam[i]) continue;
                        const T ae = Ak[kk][0] * E[0][i] + Ak[kk][1] * E[1][i] + Ak[kk][2] * E[2][i];
                        const T coef = ae / (lam[k] - lam[i]);
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++) G[a][b] += coef * E[a][i] * E[b][k];
                    }
                }
                const T gxx = G[0][0], gyx = G[0][1] + G[1][0], gzx = G[0][2] + G[2][0];
                const T gyy = G[1][1], gzy = G[1][2] + G[2][1], gzz = G[2][2];
                // H entries are 0.5 (s(+) - s(-)) of component samples, each divided by w when normalised
                auto put = [&](int comp, int xx, int yy, int zz, T coefv) {
                    const i64 q = d.at(clampi(xx, 0, d.X - 1), clampi(yy, 0, d.Y - 1), clampi(zz, 0, d.Z - 1));
                    if (norm) {
                        const T wq = W[q];
                        dV[comp * N + q] += m(coefv / wq);
                        dW[q] += m(-(coefv * F[comp * N + q] / (wq * wq)));
                    } else {
                        dV[comp * N + q] += m(coefv);
                    }
                };
                put(0, x + 1, y, z, (T)0.5 * gxx);
                put(0, x - 1, y, z, -(T)0.5 * gxx);
                put(1, x + 1, y, z, (T)0.5 * gyx);
                put(1, x - 1, y, z, -(T)0.5 * gyx);
                put(1, x, y + 1, z, (T)0.5 * gyy);
                put(1, x, y - 1, z, -(T)0.5 * gyy);
                put(2, x + 1, y, z, (T)0.5 * gzx);
                put(2, x - 1, y, z, -(T)0.5 * gzx);
                put(2, x, y + 1, z, (T)0.5 * gzy);
                put(2, x, y - 1, z, -(T)0.5 * gzy);
                put(2, x, y, z + 1, (T)0.5 * gzz);
                put(2, x, y, z - 1, -(T)0.5 * gzz);
            }
}

// The differentiable chain forward (T-precision mask) and its gradient.
template <typename T>
static void chain_forward(const T* I, Dims d, T sigma, float Fmax, int sign, T mu, int iters, float rMin, float rMax,
                          float rStep, T* tdf, T* rad, std::vector<T>* B_out, std::vector<T>* F_out,
                          std::vector<T>* states_out, std::vector<T>* V_out) {
    const i64 N = d.n();
    std::vector<T> w;
    const int m = blur_weights<T>(sigma, w, (std::vector<T>*)nullptr);
    std::vector<T> B(N), F(4 * N);
    blur_w(I, B.data(), d, m, w);
    vector_field(B.data(), F.data(), d, Fmax, sign);
    std::vector<T> states;
    if (states_out) states.resize((size_t)iters * 3 * N);
    std::vector<T> a(F.begin(), F.begin() + 3 * N), b(3 * N);
    for (int it = 0; it < iters; it++) {
        if (states_out) std::copy(a.begin(), a.end(), states.begin() + (i64)it * 3 * N);
        gvf_step(F.data(), a.data(), b.data(), d, mu);
        a.swap(b);
    }
    std::vector<T> V(4 * N);
    std::copy(a.begin(), a.end(), V.begin());
    gvf_finish(V.data(), d);
    tdf_circle(V.data(), d, rMin, rMax, rStep, tdf, rad);
    if (B_out) B_out->swap(B);
    if (F_out) F_out->swap(F);
    if (states_out) states_out->swap(states);
    if (V_out) V_out->swap(V);
}

template <typename T>
static void chain_gradient(const T* I, Dims d, T sigma, float Fmax, int sign, T mu, int iters, float rMin, float rMax,
                           float rStep, const T* gT, T* tdf, T* dI, double* dmu, double* dsigma) {
    const i64 N = d.n();
    std::vector<T> B, F, states, V, rad(N);
    chain_forward(I, d, sigma, Fmax, sign, mu, iters, rMin, rMax, rStep, tdf, rad.data(), &B, &F, &states, &V);
    std::vector<T> dV(3 * N), dW(N), dF(3 * N), dB(N);
    tdf_adjoint(V.data(), d, rMin, rMax, rStep, gT, dV.data(), dW.data());
    for (i64 i = 0; i < N; i++) {  // w = |v| (1 where |v| == 0)
        const T l = std::sqrt(V[i] * V[i] + V[N + i] * V[N + i] + V[2 * N + i] * V[2 * N + i]);
        if (l > (T)0)
            for (int c = 0; c < 3; c++) dV[c * N + i] += dW[i] * V[c * N + i] / l;
    }
    gvf_adjoint(F.data(), states.data(), d, mu, iters, dV.data(), dF.data(), dmu);
    vector_field_adjoint(B.data(), dF.data(), d, Fmax, sign, dB.data());
    blur_adjoint(dB.data(), I, d, sigma, dI, dsigma);
}

}  // namespace

// ------------------------------------------------------------------ C exports
#define DIMS Dims{X, Y, Z}
#define EXPORT_T(T, S)                                                                                       \
    extern "C" void orc_blur_##S(const T* in, T* out, int X, int Y, int Z, int m, const float* mask) {      \
        blur<T>(in, out, DIMS, m, mask);                                                                     \
    }                                                                                                        \
    extern "C" void orc_to_float_##S(const double* in, T* out, long long n, float mn, float mx) {            \
        to_float<T>(in, out, n, mn, mx);                                                                     \
    }                                                                                                        \
    extern "C" void orc_vector_field_##S(const T* I, T* F, int X, int Y, int Z, float Fmax, int sign) {      \
        vector_field<T>(I, F, DIMS, Fmax, sign);                                                             \
    }                                                                                                        \
    extern "C" void orc_gvf_euler_##S(const T* f4, T* v4, int X, int Y, int Z, float mu, int iters) {       \
        gvf_euler<T>(f4, v4, DIMS, mu, iters);                                                               \
    }                                                                                                        \
    extern "C" void orc_gvf_step_##S(const T* f3, const T* v3, T* o3, int X, int Y, int Z, float mu) {      \
        gvf_step<T>(f3, v3, o3, DIMS, (T)mu);                                                                \
    }                                                                                                        \
    extern "C" void orc_gvf_fmg_##S(const T* F, T* out, int X, int Y, int Z, float mu, int iters, int v0,   \
                                    int v1, int v2, int lmax) {                                              \
        gvf_fmg<T>(F, out, DIMS, mu, iters, v0, v1, v2, lmax);                                               \
    }                                                                                                        \
    extern "C" void orc_eigen3_##S(const T* A, T* V, T* d) {                                                 \
        T a[3][3], v[3][3];                                                                                  \
        for (int i = 0; i < 9; i++) a[i / 3][i % 3] = A[i];                                                  \
        eigen3<T>(a, v, d);                                                                                  \
        for (int i = 0; i < 9; i++) V[i] = v[i / 3][i % 3];                                                  \
    }                                                                                                        \
    extern "C" void orc_hessian_eigen_##S(const T* F, int X, int Y, int Z, int normalize, T* lam, T* vec) { \
        hessian_eigen<T>(F, DIMS, normalize, lam, vec);                                                      \
    }                                                                                                        \
    extern "C" void orc_tdf_circle_##S(const T* F, int X, int Y, int Z, float rmin, float rmax, float rstep, \
                                       T* tdf, T* rad) {                                                     \
        tdf_circle<T>(F, DIMS, rmin, rmax, rstep, tdf, rad);                                                 \
    }                                                                                                        \
    extern "C" void orc_tdf_circle_split_##S(const T* FH, const T* FS, int X, int Y, int Z, float rmin,       \
                                             float rmax, float rstep, T* tdf, T* rad) {                      \
        tdf_circle_split<T>(FH, FS, DIMS, rmin, rmax, rstep, tdf, rad);                                      \
    }                                                                                                        \
    extern "C" void orc_tdf_circle_mixed_##S(const T* FH, const T* FS, int X, int Y, int Z, float rmin,       \
                                             float rmax, float rstep, T* tdf, T* rad) {                      \
        tdf_circle_split<T, true>(FH, FS, DIMS, rmin, rmax, rstep, tdf, rad);                                \
    }                                                                                                        \
    extern "C" void orc_tdf_spline_##S(const T* F, int X, int Y, int Z, float rmin, float rmax, float rstep, \
                                       int arms, float min_avg, T* tdf, T* rad) {                            \
        tdf_spline<T>(F, DIMS, rmin, rmax, rstep, arms, min_avg, tdf, rad);                                  \
    }                                                                                                        \
    extern "C" void orc_combine_##S(const T* ts, const T* rs, T* tl, T* rl, long long n) {                  \
        combine<T>(ts, rs, tl, rl, n);                                                                       \
    }                                                                                                        \
    extern "C" void orc_candidates_##S(const T* tdf, uint8_t* mask, long long n, float thr) {               \
        candidates<T>(tdf, mask, n, thr);                                                                    \
    }                                                                                                        \
    extern "C" void orc_candidate_nms_##S(const T* tdf, const T* rad, const T* F, int X, int Y, int Z,      \
                                          const int32_t* pos, long long count, int cap, uint8_t* out) {      \
        candidate_nms<T>(tdf, rad, F, DIMS, pos, count, cap, out);                                           \
    }                                                                                                        \
    extern "C" void orc_cube_max_##S(const T* tdf, const uint8_t* cand, int X, int Y, int Z, int cube,      \
                                     uint8_t* out) {                                                         \
        cube_max<T>(tdf, cand, DIMS, cube, out);                                                             \
    }                                                                                                        \
    extern "C" void orc_do_eigen_##S(const T* Fx, const T* Fy, const T* Fz, int X, int Y, int Z, int x,     \
                                     int y, int z, T* lam, T* e1, T* e2, T* e3) {                            \
        do_eigen<T>(Fx, Fy, Fz, DIMS, x, y, z, lam, e1, e2, e3);                                             \
    }                                                                                                        \
    extern "C" void orc_trilinear_##S(const T* F, int X, int Y, int Z, T u, T v, T w, T* out) {             \
        trilinear3<T>(F, DIMS, u, v, w, out);                                                                \
    }                                                                                                        \
    extern "C" void orc_blur_sigma_##S(const T* in, T* out, int X, int Y, int Z, double sigma) {            \
        std::vector<T> w;                                                                                    \
        int m = blur_weights<T>((T)sigma, w, (std::vector<T>*)nullptr);                                      \
        blur_w<T>(in, out, DIMS, m, w);                                                                      \
    }                                                                                                        \
    extern "C" void orc_blur_adjoint_##S(const T* dB, const T* I, int X, int Y, int Z, double sigma, T* dI, \
                                         double* dsigma) {                                                   \
        blur_adjoint<T>(dB, I, DIMS, (T)sigma, dI, dsigma);                                                  \
    }                                                                                                        \
    extern "C" void orc_vector_field_adjoint_##S(const T* B, const T* dF, int X, int Y, int Z, float Fmax,  \
                                                 int sign, T* dB) {                                          \
        vector_field_adjoint<T>(B, dF, DIMS, Fmax, sign, dB);                                                \
    }                                                                                                        \
    extern "C" void orc_gvf_adjoint_##S(const T* F, int X, int Y, int Z, double mu, int iters, const T* dV, \
                                        T* dF, double* dmu) {                                                \
        Dims d = DIMS;                                                                                       \
        const i64 N = d.n();                                                                                 \
        std::vector<T> states((size_t)iters * 3 * N), a(F, F + 3 * N), b(3 * N);                             \
        for (int it = 0; it < iters; it++) {                                                                 \
            std::copy(a.begin(), a.end(), states.begin() + (i64)it * 3 * N);                                 \
            gvf_step<T>(F, a.data(), b.data(), d, (T)mu);                                                    \
            a.swap(b);                                                                                       \
        }                                                                                                    \
        gvf_adjoint<T>(F, states.data(), d, (T)mu, iters, dV, dF, dmu);                                      \
    }                                                                                                        \
    extern "C" void orc_tdf_adjoint_##S(const T* F, int X, int Y, int Z, float rmin, float rmax,            \
                                        float rstep, const T* gT, T* dV, T* dW) {                            \
        tdf_adjoint<T>(F, DIMS, rmin, rmax, rstep, gT, dV, dW);                                              \
    }                                                                                                        \
    extern "C" void orc_tdf_adjoint_abs_##S(const T* F, int X, int Y, int Z, float rmin, float rmax,        \
                                            float rstep, const T* gT, T* dV, T* dW) {                        \
        tdf_adjoint<T, true>(F, DIMS, rmin, rmax, rstep, gT, dV, dW);                                        \
    }                                                                                                        \
    extern "C" void orc_chain_forward_##S(const T* I, int X, int Y, int Z, double sigma, float Fmax,        \
                                          int sign, double mu, int iters, float rmin, float rmax,           \
                                          float rstep, T* tdf, T* rad) {                                     \
        chain_forward<T>(I, DIMS, (T)sigma, Fmax, sign, (T)mu, iters, rmin, rmax, rstep, tdf, rad, nullptr, \
                         nullptr, nullptr, nullptr);                                                         \
    }                                                                                                        \
    extern "C" void orc_chain_gradient_##S(const T* I, int X, int Y, int Z, double sigma, float Fmax,       \
                                           int sign, double mu, int iters, float rmin, float rmax,          \
                                           float rstep, const T* gT, T* tdf, T* dI, double* dmu,            \
                                           double* dsigma) {                                                 \
        chain_gradient<T>(I, DIMS, (T)sigma, Fmax, sign, (T)mu, iters, rmin, rmax, rstep, gT, tdf, dI, dmu, \
                          dsigma);                                                                           \
    }

EXPORT_T(float, f32)
EXPORT_T(double, f64)

extern "C" void orc_blur_mask(float sigma, float* mask, int* maskSize) { blur_mask(sigma, mask, maskSize); }
// 16-bit mode: round values to the SNORM16 format in place; store codes; the
// vector field and GVF with every store in the format.
extern "C" void orc_sn16_round(float* a, long long n) {
    for (long long i = 0; i < n; i++) a[i] = sn16_round(a[i]);
}
extern "C" void orc_gvf_euler_sr16(const uint16_t* f16, int storage, float scale, float mu, int iters, int tail,
                                   unsigned seed, int X, int Y, int Z, float* out4) {
    gvf_euler_sr16(f16, storage, scale, mu, iters, tail, seed, Dims{X, Y, Z}, out4);
}
extern "C" void orc_eigen3_fp64_fallback_f32(const float* A, float* V, float* d) {
    float a[3][3], v[3][3];
    for (int i = 0; i < 9; i++) a[i / 3][i % 3] = A[i];
    eigen3_fp64_fallback(a, v, d);
    for (int i = 0; i < 9; i++) V[i] = v[i / 3][i % 3];
}
extern "C" void orc_half_from_float(const float* a, uint16_t* h, long long n) {
    for (long long i = 0; i < n; i++) h[i] = float_to_half_rne(a[i]);
}
extern "C" void orc_half_to_float(const uint16_t* h, float* a, long long n) {
    for (long long i = 0; i < n; i++) a[i] = half_to_float(h[i]);
}
extern "C" void orc_s16_sr(const float* a, const unsigned* r, int storage, uint16_t* h, long long n) {
    for (long long i = 0; i < n; i++) h[i] = s16_sr(a[i], r[i], storage);
}
extern "C" unsigned orc_sr_hash(unsigned i, unsigned seed) { return sr_hash(i, seed); }
extern "C" void orc_sn16_store(const float* a, int16_t* s, long long n) {
    for (long long i = 0; i < n; i++) s[i] = sn16_store(a[i]);
}
extern "C" void orc_sn16_load(const int16_t* s, float* a, long long n) {
    for (long long i = 0; i < n; i++) a[i] = sn16_load(s[i]);
}
// 16-bit TDF buffers (kernels_no_3d_write.cl:19-20): convert_ushort_sat_rte(t *
// 65535) stored, (float)c / 65535.0f read.
extern "C" void orc_unorm16_round(float* a, long long n) {
    for (long long i = 0; i < n; i++) {
        const float x = a[i] * 65535.0f;
        const float c = x != x ? 0.0f : std::nearbyint(std::min(std::max(x, 0.0f), 65535.0f));
        a[i] = c / 65535.0f;
    }
}
extern "C" void orc_vector_field_sn16(const float* I, float* F, int X, int Y, int Z, float Fmax, int sign) {
    vector_field<float>(I, F, DIMS, Fmax, sign);
    orc_sn16_round(F, 4LL * X * Y * Z);
}
extern "C" void orc_gvf_euler_low_memory(const float* f4, float* v4, int X, int Y, int Z, float mu, int iters,
                                         int sn16) {
    gvf_euler_low_memory(f4, v4, DIMS, mu, iters, sn16 != 0);
}
extern "C" void orc_gvf_euler_sn16(const float* f4, float* v4, int X, int Y, int Z, float mu, int iters) {
    gvf_euler_sn16(f4, v4, DIMS, mu, iters);
}
extern "C" void orc_circle_tables(float* c, float* s) {
    make_tables();
    for (int j = 0; j < 32; j++) {
        c[j] = g_cos[j];
        s[j] = g_sin[j];
    }
}
extern "C" long long orc_hp_compact_sort(const uint8_t* mask, int X, int Y, int Z, int32_t* pos, long long cap) {
    return hp_compact_sort(mask, DIMS, pos, cap);
}
extern "C" long long orc_hp_compact_pyramid(const uint8_t* mask, int X, int Y, int Z, int32_t* pos,
                                            long long cap) {
    return hp_compact_pyramid(mask, DIMS, pos, cap);
}
extern "C" unsigned long long orc_hp_key(unsigned x, unsigned y, unsigned z) { return hp_key(x, y, z); }

// ---------------------------------------------------------------- centerline graph
// runNewCenterlineAlg after the vertex list (parallelCenterlineExtraction.cpp:
// 912-1160, 1270-1316), restated with the reference's own data layout: the
// count x count length image, per-vertex compacted rows of (length,
// neighbour) -- filled in ascending neighbour index, the order an in-order
// atomic_inc would give -- a count x count edge image, removeDuplicateEdges,
// and a 2D histogram pyramid (constructHPLevel2D sums, scanHPLevel2D descent
// with the children (0,0), (0,1), (1,0), (1,1)) that lists the edges.

// float distance of two integer positions (OpenCL distance / length)
static float odist(const int32_t* v, int a, int b) {
    const float dx = (float)v[3 * a] - (float)v[3 * b], dy = (float)v[3 * a + 1] - (float)v[3 * b + 1];
    const float dz = (float)v[3 * a + 2] - (float)v[3 * b + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}
// CLK_FILTER_LINEAR, unnormalised, clamp-to-edge (the TDF's sampler) on a
// scalar float volume
static float olinear(const float* T, int X, int Y, int Z, float u, float v, float w) {
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = std::floor(uu), fv = std::floor(vv), fw = std::floor(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    auto cl = [](int q, int n) { return q < 0 ? 0 : (q > n - 1 ? n - 1 : q); };
    const int i0 = cl((int)fu, X), i1 = cl((int)fu + 1, X), j0 = cl((int)fv, Y), j1 = cl((int)fv + 1, Y);
    const int k0 = cl((int)fw, Z), k1 = cl((int)fw + 1, Z);
    auto at = [&](int i, int j, int k) { return T[(size_t)i + (size_t)X * ((size_t)j + (size_t)Y * k)]; };
    float s = (1.0f - a) * (1.0f - b) * (1.0f - c) * at(i0, j0, k0);
    s = s + a * (1.0f - b) * (1.0f - c) * at(i1, j0, k0);
    s = s + (1.0f - a) * b * (1.0f - c) * at(i0, j1, k0);
    s = s + a * b * (1.0f - c) * at(i1, j1, k0);
    s = s + (1.0f - a) * (1.0f - b) * c * at(i0, j0, k1);
    s = s + a * (1.0f - b) * c * at(i1, j0, k1);
    s = s + (1.0f - a) * b * c * at(i0, j1, k1);
    s = s + a * b * c * at(i1, j1, k1);
    return s;
}

extern "C" int orc_link_centerpoints(const int32_t* v, int sum, const float* T, int X, int Y, int Z, float maxDistance,
                                     float minAvgTDF, int32_t* edgesOut, int cap) {
    // linkLengths and compact
    std::vector<float> lengths((size_t)sum * sum);
    for (int j = 0; j < sum; j++)
        for (int i = 0; i < sum; i++) lengths[(size_t)j * sum + i] = odist(v, i, j);
    std::vector<std::vector<std::pair<float, int>>> rows(sum);
    for (int i = 0; i < sum; i++)
        for (int j = 0; j < sum; j++) {
            const float l = lengths[(size_t)j * sum + i];
            if (l < maxDistance && l > 0.0f) rows[i].push_back({l, j});
        }
    // linkCenterpoints, kernels_no_3d_write.cl:753-882
    std::vector<uint8_t> E((size_t)sum * sum, 0);  // E[y * sum + x]: image (x, y)
    for (int id = 0; id < sum; id++) {
        const float xa[3] = {(float)v[3 * id], (float)v[3 * id + 1], (float)v[3 * id + 2]};
        int bx = 0, by = 0;
        float shortestDistance = maxDistance * 2;
        bool valid = false;
        const auto& R = rows[id];
        for (size_t i = 0; i < R.size(); i++) {
            const float clx = R[i].first;
            const int cly = R[i].second;
            const int db = (int)std::round(clx);
            if (db >= shortestDistance) continue;
            for (size_t j = 0; j < i; j++) {
                if (R[j].second == cly) continue;
                const int dc = (int)std::round(R[j].first);
                if (db + dc < shortestDistance) {
                    float ab[3], ac[3];
                    for (int q = 0; q < 3; q++) {
                        ab[q] = (float)v[3 * cly + q] - xa[q];
                        ac[q] = (float)v[3 * R[j].second + q] - xa[q];
                    }
                    const float nb = std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
                    const float nc = std::sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
                    const float d = (ab[0] / nb) * (ac[0] / nc) + (ab[1] / nb) * (ac[1] / nc) + (ab[2] / nb) * (ac[2] / nc);
                    if (std::acos(d) < 2.0f) continue;
                    float avg = 0.0f;
                    for (int k = 0; k <= db; k++) {
                        const float alpha = (float)k / db;
                        avg += olinear(T, X, Y, Z, xa[0] + ab[0] * alpha, xa[1] + ab[1] * alpha, xa[2] + ab[2] * alpha);
                    }
                    avg /= db + 1;
                    if (avg < minAvgTDF) continue;
                    avg = 0.0f;
                    for (int k = 0; k <= dc; k++) {
                        const float alpha = (float)k / dc;
                        avg += olinear(T, X, Y, Z, xa[0] + ac[0] * alpha, xa[1] + ac[1] * alpha, xa[2] + ac[2] * alpha);
                    }
                    avg /= dc + 1;
                    if (avg < minAvgTDF) continue;
                    valid = true;
                    bx = cly;
                    by = R[j].second;
                    shortestDistance = db + dc;
                }
            }
        }
        if (valid) {
            E[(size_t)bx * sum + id] = 1;
            E[(size_t)by * sum + id] = 1;
        }
    }
    // removeDuplicateEdges, kernels_no_3d_write.cl:947-968
    std::vector<uint8_t> E2((size_t)sum * sum, 0);
    for (int y = 0; y < sum; y++)
        for (int x = 0; x < sum; x++) {
            if (E[(size_t)y * sum + x] == 0) continue;
            E2[(size_t)y * sum + x] = (x > y && E[(size_t)x * sum + y] == 1) ? 0 : 1;
        }
    // 2D histogram pyramid: level 0 = the edge image padded to a power of two
    int size = 1;
    while (size < sum) size *= 2;
    std::vector<std::vector<int>> lv;
    lv.push_back(std::vector<int>((size_t)size * size, 0));
    for (int y = 0; y < sum; y++)
        for (int x = 0; x < sum; x++) lv[0][(size_t)y * size + x] = E2[(size_t)y * sum + x];
    for (int s = size; s > 1; s /= 2) {
        const int h = s / 2;
        std::vector<int> up((size_t)h * h);
        const auto& lo = lv.back();
        for (int y = 0; y < h; y++)
            for (int x = 0; x < h; x++)
                up[(size_t)y * h + x] = lo[(size_t)(2 * y) * s + 2 * x] + lo[(size_t)(2 * y) * s + 2 * x + 1] +
                                        lo[(size_t)(2 * y + 1) * s + 2 * x] + lo[(size_t)(2 * y + 1) * s + 2 * x + 1];
        lv.push_back(up);
    }
    const int total = lv.back()[0];
    // descent (scanHPLevel2D): children (0,0), (0,1), (1,0), (1,1)
    const int off[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
    for (int target = 0; target < total && target < cap; target++) {
        int cx = 0, cy = 0, acc0 = 0;
        for (int l = (int)lv.size() - 2; l >= 0; l--) {
            const int s = size >> l;
            const auto& L = lv[l];
            cx *= 2;
            cy *= 2;
            int acc = acc0;
            int k = 0;
            for (; k < 3; k++) {
                const int n = L[(size_t)(cy + off[k][1]) * s + cx + off[k][0]];
                if (acc + n <= target) acc += n;
                else break;
            }
            cx += off[k][0];
            cy += off[k][1];
            acc0 = acc;
        }
        edgesOut[2 * target] = cx;
        edgesOut[2 * target + 1] = cy;
    }
    return total;
}

// graphComponentLabeling to its fixed point, then calculateTreeLength
extern "C" void orc_graph_components(const int32_t* edges, int nEdges, int sum, int32_t* C, int32_t* S) {
    for (int i = 0; i < sum; i++) {
        C[i] = i;
        S[i] = 0;
    }
    bool m = true;
    while (m) {
        m = false;
        for (int e = 0; e < nEdges; e++) {
            const int ca = C[edges[2 * e]], cb = C[edges[2 * e + 1]];
            if (ca == cb) continue;
            if (ca < cb) C[edges[2 * e + 1]] = std::min(C[edges[2 * e + 1]], ca);
            else C[edges[2 * e]] = std::min(C[edges[2 * e]], cb);
            m = true;
        }
    }
    for (int i = 0; i < sum; i++) S[C[i]]++;
}

// removeSmallTrees, kernels_no_3d_write.cl:920-945
extern "C" void orc_rasterize_tree_edges(const int32_t* edges, int nEdges, const int32_t* v, const int32_t* C,
                                         const int32_t* S, int minTreeLength, uint8_t* out, int X, int Y, int Z) {
    for (int e = 0; e < nEdges; e++) {
        const int a = edges[2 * e], b = edges[2 * e + 1];
        if (S[C[a]] < minTreeLength) continue;
        float xa[3], d[3];
        for (int q = 0; q < 3; q++) {
            xa[q] = (float)v[3 * a + q];
            d[q] = (float)v[3 * b + q] - xa[q];
        }
        const int l = (int)std::round(std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
        for (int i = 0; i < l; i++) {
            const float alpha = (float)i / l;
            const int px = (int)std::round(xa[0] + d[0] * alpha), py = (int)std::round(xa[1] + d[1] * alpha);
            const int pz = (int)std::round(xa[2] + d[2] * alpha);
            if (px >= 0 && py >= 0 && pz >= 0 && px < X && py < Y && pz < Z)
                out[(size_t)px + (size_t)X * ((size_t)py + (size_t)Y * pz)] = 1;
        }
    }
}

// ---------------------------------------------------------------- segmentation
// runInverseGradientSegmentation (segmentation.cpp:5-221) with the kernels
// initGrowing / grow / dilate / erode (kernels_no_3d_write.cl:1100-1207,
// 1370-1412) on a float field (x, y, z, w planes).  The reference's next
// state is written by many threads in an unspecified order; this
// restatement evaluates every candidate on the current state, then merges
// with the precedence accepted > nominated > dropped.
static int oclamp(int q, int n) { return q < 0 ? 0 : (q > n - 1 ? n - 1 : q); }

extern "C" int orc_segment_inverse_gradient(const uint8_t* cl, const float* radius, const float* F, int X, int Y,
                                            int Z, uint8_t* seg) {
    const size_t n = (size_t)X * Y * Z;
    auto at = [&](int x, int y, int z) { return (size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z); };
    auto in = [&](int x, int y, int z) { return x >= 0 && y >= 0 && z >= 0 && x < X && y < Y && z < Z; };
    auto cat = [&](int x, int y, int z) { return at(oclamp(x, X), oclamp(y, Y), oclamp(z, Z)); };
    const float* Fx = F;
    const float* Fy = F + n;
    const float* Fz = F + 2 * n;
    const float* Fw = F + 3 * n;
    std::vector<uint8_t> cur(cl, cl + n);
    // initGrowing
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++) {
                if (cl[at(x, y, z)] != 1) continue;
                const int N = std::min(std::max(1, (int)std::round(radius[at(x, y, z)] / 2.0f)), 4);
                for (int a = -N; a < N + 1; a++)
                    for (int b = -N; b < N + 1; b++)
                        for (int c = -N; c < N + 1; c++)
                            if (cl[cat(x + a, y + b, z + c)] == 0 && in(x + a, y + b, z + c))
                                cur[at(x + a, y + b, z + c)] = 2;
            }
    int passes = 0;
    for (;;) {
        std::vector<uint8_t> accepted(n, 0), nominated(n, 0);
        bool grew = false;
        for (int z = 0; z < Z; z++)
            for (int y = 0; y < Y; y++)
                for (int x = 0; x < X; x++) {
                    if (cur[at(x, y, z)] != 2) continue;
                    const float FNXw = Fw[at(x, y, z)];
                    bool cont = false;
                    for (int a = -1; a < 2; a++)
                        for (int b = -1; b < 2; b++)
                            for (int c = -1; c < 2; c++) {
                                if (a == 0 && b == 0 && c == 0) continue;
                                const int yx = x + a, yy = y + b, yz = z + c;
                                const size_t q = cat(yx, yy, yz);
                                if (cur[q] == 1) continue;
                                const float w = Fw[q], fx = Fx[q] / w, fy = Fy[q] / w, fz = Fz[q] / w;
                                if (!(w > FNXw)) continue;
                                int zx = 0, zy = 0, zz = 0;
                                float maxDot = -2.0f;
                                for (int a2 = -1; a2 < 2; a2++)
                                    for (int b2 = -1; b2 < 2; b2++)
                                        for (int c2 = -1; c2 < 2; c2++) {
                                            if (a2 == 0 && b2 == 0 && c2 == 0) continue;
                                            const float l = std::sqrt((float)(a2 * a2 + b2 * b2 + c2 * c2));
                                            const float v = fx * ((float)a2 / l) + fy * ((float)b2 / l) + fz * ((float)c2 / l);
                                            if (v > maxDot) {
                                                maxDot = v;
                                                zx = yx + a2;
                                                zy = yy + b2;
                                                zz = yz + c2;
                                            }
                                        }
                                if (zx == x && zy == y && zz == z && in(yx, yy, yz)) {
                                    nominated[at(yx, yy, yz)] = 1;
                                    cont = true;
                                }
                            }
                    if (cont) {
                        accepted[at(x, y, z)] = 1;
                        grew = true;
                    }
                }
        for (size_t i = 0; i < n; i++) cur[i] = (cur[i] == 1 || accepted[i]) ? 1 : (nominated[i] ? 2 : 0);
        passes++;
        if (passes >= 2 && !grew) break;
    }
    // dilate into an empty volume, then erode (clamp-to-edge reads)
    std::vector<uint8_t> dil(n, 0);
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++) {
                if (cur[at(x, y, z)] != 1) continue;
                for (int a = -1; a < 2; a++)
                    for (int b = -1; b < 2; b++)
                        for (int c = -1; c < 2; c++)
                            if (in(x + a, y + b, z + c)) dil[at(x + a, y + b, z + c)] = 1;
            }
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++) {
                bool keep = dil[at(x, y, z)] == 1;
                for (int a = -1; a < 2; a++)
                    for (int b = -1; b < 2; b++)
                        for (int c = -1; c < 2; c++) keep = keep && dil[cat(x + a, y + b, z + c)] == 1;
                seg[at(x, y, z)] = keep ? 1 : 0;
            }
    return passes;
}

// sphereSegmentation (kernels_no_3d_write.cl:1209-1234)
extern "C" void orc_segment_spheres(const uint8_t* cl, const float* radius, int X, int Y, int Z, uint8_t* seg) {
    const size_t n = (size_t)X * Y * Z;
    std::fill(seg, seg + n, 0);
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++) {
                const size_t i = (size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z);
                if (cl[i] == 0) continue;
                const float r = radius[i];
                const int N = (int)std::ceil(r);
                for (int a = -N; a <= N; a++)
                    for (int b = -N; b <= N; b++)
                        for (int c = -N; c <= N; c++) {
                            if (!(std::sqrt((float)a * a + (float)b * b + (float)c * c) < r)) continue;
                            const int px = x + a, py = y + b, pz = z + c;
                            if (px >= 0 && py >= 0 && pz >= 0 && px < X && py < Y && pz < Z)
                                seg[(size_t)px + (size_t)X * ((size_t)py + (size_t)Y * pz)] = 1;
                        }
            }
}

// ---------------------------------------------------------------- ingest
// getMinimum / getMaximum (tube-segmentation.cpp:1356-1374): std::min /
// std::max from numeric_limits<T>::max() / ::min() in index order.  The
// values arrive as doubles (exact for every element type); `type` 0..4 =
// uint8, int8, uint16, int16, float picks T.
template <class T>
static void limits_t(const double* v, long long n, float* mn, float* mx) {
    T lo = std::numeric_limits<T>::max(), hi = std::numeric_limits<T>::min();
    for (long long i = 0; i < n; i++) {
        const T t = (T)v[i];
        lo = std::min(lo, t);
        hi = std::max(hi, t);
    }
    *mn = (float)lo;
    *mx = (float)hi;
}
extern "C" void orc_volume_limits(const double* v, long long n, int type, float* mn, float* mx) {
    switch (type) {
        case 0: limits_t<uint8_t>(v, n, mn, mx); break;
        case 1: limits_t<int8_t>(v, n, mn, mx); break;
        case 2: limits_t<uint16_t>(v, n, mn, mx); break;
        case 3: limits_t<int16_t>(v, n, mn, mx); break;
        default: limits_t<float>(v, n, mn, mx); break;
    }
}

// cropDatasetThreshold / cropDatasetLung (kernels_no_3d_write.cl:1236-1368):
// one work-item per slice of direction dir (0: x, 1: y, 2: z), scan lines and
// elements in the kernel's order.  type as above; method 0 threshold, 1 lung.
extern "C" void orc_crop_scan_lines(const double* v, int X, int Y, int Z, int type, int method, float threshold,
                                    int32_t* lines_x, int32_t* lines_y, int32_t* lines_z) {
    const int hu_limit = -150 + ((type == 0 || type == 2) ? 1024 : 0);
    for (int dir = 0; dir < 3; dir++) {
        const int slices = dir == 0 ? X : (dir == 1 ? Y : Z);
        const int lines = dir == 1 ? X : Y;
        const int elems = dir == 2 ? X : Z;
        int32_t* out = dir == 0 ? lines_x : (dir == 1 ? lines_y : lines_z);
        for (int s = 0; s < slices; s++) {
            int count = 0;
            for (int l = 0; l < lines; l++) {
                bool found = false;
                int wc = 0, bc = 0, dw = 0, db = 0;
                for (int e = 0; e < elems; e++) {
                    int x, y, z;
                    if (dir == 0) {
                        x = s; y = l; z = e;
                    } else if (dir == 1) {
                        x = l; y = s; z = e;
                    } else {
                        x = e; y = l; z = s;
                    }
                    const double val = v[(size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z)];
                    if (method == 0) {
                        if ((float)val > threshold) found = true;
                    } else {
                        const short hu = (short)(int)val;
                        if (hu > hu_limit) {
                            if (wc == 30) {
                                dw++;
                                bc = 0;
                            }
                            wc++;
                        } else {
                            if (bc == 30) {
                                db++;
                                wc = 0;
                            }
                            bc++;
                        }
                    }
                }
                if (method == 0 ? found : ((dw == 2 && db == 1) || (db > 1 && dw > 1))) count++;
            }
            out[s] = count;
        }
    }
}

// The crop region of readDatasetAndTransfer (tube-segmentation.cpp:1612-1745)
// from the scan-line counts.  Returns 0, or -1 where an extent is empty or
// negative.  offset and size are 3 ints each.
extern "C" int orc_crop_region(const int32_t* lx, const int32_t* ly, const int32_t* lz, int X, int Y, int Z,
                               int min_lines, int start_z_middle, int32_t* offset, int32_t* size) {
    int x1 = 0, x2 = X, y1 = 0, y2 = Y, z1 = 0, z2 = Z;
    const int start = start_z_middle ? Z / 2 : 0, a = start_z_middle ? -1 : 1;
    for (int s = 0; s < X; s++)
        if (lx[s] > min_lines) {
            x1 = s;
            break;
        }
    for (int s = X - 1; s > 0; s--)
        if (lx[s] > min_lines) {
            x2 = s;
            break;
        }
    for (int s = 0; s < Y; s++)
        if (ly[s] > min_lines) {
            y1 = s;
            break;
        }
    for (int s = Y - 1; s > 0; s--)
        if (ly[s] > min_lines) {
            y2 = s;
            break;
        }
    for (int s = start; s < Z; s++)
        if (a * lz[s] > a * min_lines) {
            z2 = s;
            break;
        }
    for (int s = Z - start - 1; s > 0; s--)
        if (a * lz[s] > a * min_lines) {
            z1 = s;
            break;
        }
    if (!start_z_middle) std::swap(z1, z2);
    int SX = x2 - x1, SY = y2 - y1, SZ = z2 - z1;
    if (SX <= 0 || SY <= 0 || SZ <= 0) return -1;
    bool lower = false;
    while (SX % 4 != 0 && SX < X) {
        if (lower && x1 > 0) x1--;
        else if (x2 < X) x2++;
        lower = !lower;
        SX = x2 - x1;
    }
    while (SX % 4 != 0) SX--;
    while (SY % 4 != 0 && SY < Y) {
        if (lower && y1 > 0) y1--;
        else if (y2 < Y) y2++;
        lower = !lower;
        SY = y2 - y1;
    }
    while (SY % 4 != 0) SY--;
    while (SZ % 4 != 0 && SZ < Z) {
        if (lower && z1 > 0) z1--;
        else if (z2 < Z) z2++;
        lower = !lower;
        SZ = z2 - z1;
    }
    while (SZ % 4 != 0) SZ--;
    if (SX <= 0 || SY <= 0 || SZ <= 0) return -1;
    offset[0] = x1;
    offset[1] = y1;
    offset[2] = z1;
    size[0] = SX;
    size[1] = SY;
    size[2] = SZ;
    return 0;
}

// Worker threads of the OpenMP loops (the bench's CPU baseline times one
// thread and all of them); the count in effect.
extern "C" int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
