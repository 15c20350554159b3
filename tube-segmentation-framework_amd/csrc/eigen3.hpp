// Register-resident symmetric 3x3 eigensolver for gfx950.
//
// Householder tridiagonalisation followed by the implicit QL method with
// eigenvector accumulation -- the EISPACK routines tred2/tql2 (Bowdler,
// Martin, Reinsch, Wilkinson, Handbook for Automatic Computation II) in their
// public-domain JAMA formulation, which is what the reference runs per voxel
// (kernels_no_3d_write.cl:1824-2068).  The operation order is the one those
// routines define, so float results round as the reference's float code does.
//
// Host code shares it (the single-point doEigen).  Every loop has a
// compile-time trip count and every array index is a constant after
// unrolling, so V, d and e live in VGPRs (no scratch).  The
// data-dependent QL window [l, m] is expressed with predicated constant-index
// branches.  Two fixed rules where the reference is unsafe: the small
// sub-diagonal search stops at index 2, and the QL iteration is capped at
// TSF_QL_MAX_ITER sweeps.
#pragma once

#include <hip/hip_runtime.h>

#define TSF_QL_MAX_ITER 100

namespace tsf {

__host__ __device__ __forceinline__ float tsqrt(float v) { return sqrtf(v); }
__host__ __device__ __forceinline__ double tsqrt(double v) { return sqrt(v); }
__host__ __device__ __forceinline__ float tabs(float v) { return fabsf(v); }
__host__ __device__ __forceinline__ double tabs(double v) { return fabs(v); }

template <typename T>
__host__ __device__ __forceinline__ void tridiag3(T V[3][3], T d[3], T e[3]) {
#pragma unroll
    for (int j = 0; j < 3; j++) d[j] = V[2][j];
#pragma unroll
    for (int i = 2; i > 0; i--) {
        T scale = 0, h = 0;
#pragma unroll
        for (int k = 0; k < i; k++) scale = scale + tabs(d[k]);
        if (scale == (T)0) {
            e[i] = d[i - 1];
#pragma unroll
            for (int j = 0; j < i; j++) {
                d[j] = V[i - 1][j];
                V[i][j] = 0;
                V[j][i] = 0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < i; k++) {
                d[k] /= scale;
                h += d[k] * d[k];
            }
            T f = d[i - 1];
            T g = tsqrt(h);
            if (f > 0) g = -g;
            e[i] = scale * g;
            h = h - f * g;
            d[i - 1] = f - g;
#pragma unroll
            for (int j = 0; j < i; j++) e[j] = 0;
#pragma unroll
            for (int j = 0; j < i; j++) {
                f = d[j];
                V[j][i] = f;
                g = e[j] + V[j][j] * f;
#pragma unroll
                for (int k = j + 1; k <= i - 1; k++) {
                    g += V[k][j] * d[k];
                    e[k] += V[k][j] * f;
                }
                e[j] = g;
            }
            f = 0;
#pragma unroll
            for (int j = 0; j < i; j++) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            T hh = f / (h + h);
#pragma unroll
            for (int j = 0; j < i; j++) e[j] -= hh * d[j];
#pragma unroll
            for (int j = 0; j < i; j++) {
                f = d[j];
                g = e[j];
#pragma unroll
                for (int k = j; k <= i - 1; k++) V[k][j] -= (f * e[k] + g * d[k]);
                d[j] = V[i - 1][j];
                V[i][j] = 0;
            }
        }
        d[i] = h;
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
        V[2][i] = V[i][i];
        V[i][i] = 1;
        T h = d[i + 1];
        if (h != (T)0) {
#pragma unroll
            for (int k = 0; k <= i; k++) d[k] = V[k][i + 1] / h;
#pragma unroll
            for (int j = 0; j <= i; j++) {
                T g = 0;
#pragma unroll
                for (int k = 0; k <= i; k++) g += V[k][i + 1] * V[k][j];
#pragma unroll
                for (int k = 0; k <= i; k++) V[k][j] -= g * d[k];
            }
        }
#pragma unroll
        for (int k = 0; k <= i; k++) V[k][i + 1] = 0;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        d[j] = V[2][j];
        V[2][j] = 0;
    }
    V[2][2] = 1;
    e[0] = 0;
}

template <typename T>
__host__ __device__ __forceinline__ void ql3(T V[3][3], T d[3], T e[3], bool* converged = nullptr) {
    e[0] = e[1];
    e[1] = e[2];
    e[2] = 0;
    T f = 0, tst1 = 0;
    const T eps = (T)2.220446049250313080847263336181640625e-16;  // 2^-52
#pragma unroll
    for (int l = 0; l < 3; l++) {
        T t = tabs(d[l]) + tabs(e[l]);
        tst1 = tst1 > t ? tst1 : t;
        int m = l;
#pragma unroll
        for (int mm = l; mm < 2; mm++)
            if (m == mm && !(tabs(e[mm]) <= eps * tst1)) m = mm + 1;
        if (l < 2 && m > l) {
            int iter = 0;
            do {
                iter++;
                T g = d[l];
                T p = (d[l + 1] - g) / ((T)2 * e[l]);
                T r = tsqrt(p * p + (T)1 * (T)1);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                T dl1 = d[l + 1];
                T h = g - d[l];
#pragma unroll
                for (int i = l + 2; i < 3; i++) d[i] -= h;
                f = f + h;
                p = (m == 1) ? d[1] : d[2];
                T c = 1, c2 = c, c3 = c;
                T el1 = e[l + 1];
                T s = 0, s2 = 0;
#pragma unroll
                for (int i = 1; i >= 0; i--) {
                    if (i >= l && i <= m - 1) {
                        c3 = c2;
                        c2 = c;
                        s2 = s;
                        g = c * e[i];
                        h = c * p;
                        r = tsqrt(p * p + e[i] * e[i]);
                        e[i + 1] = s * r;
                        s = e[i] / r;
                        c = p / r;
                        p = c * d[i] - s * g;
                        d[i + 1] = h + s * (c * g + s * d[i]);
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            h = V[k][i + 1];
                            V[k][i + 1] = s * V[k][i] + c * h;
                            V[k][i] = c * V[k][i] - s * h;
                        }
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (tabs(e[l]) > eps * tst1 && iter < TSF_QL_MAX_ITER);
            if (converged && tabs(e[l]) > eps * tst1) *converged = false;
        }
        d[l] = d[l] + f;
        e[l] = 0;
    }
    // ascending |lambda|, columns follow
#pragma unroll
    for (int i = 0; i < 2; i++) {
        int k = i;
        T p = d[i];
#pragma unroll
        for (int j = i + 1; j < 3; j++)
            if (tabs(d[j]) < tabs(p)) {
                k = j;
                p = d[j];
            }
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            if (k == j) {
                d[j] = d[i];
                d[i] = p;
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    T q = V[r][i];
                    V[r][i] = V[r][j];
                    V[r][j] = q;
                }
            }
        }
    }
}

template <typename T>
__host__ __device__ __forceinline__ void eigen3(const T A[3][3], T V[3][3], T d[3], bool* converged = nullptr) {
    T e[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) V[i][j] = A[i][j];
    tridiag3(V, d, e);
    ql3(V, d, e, converged);
}

// The precision form's eigensolve: eigen3 in float, redone in FP64 from
// the same float matrix -- and rounded to float -- when two eigenvalues lie
// within 2^-10 of the largest |eigenvalue| of each other (a float
// eigenvector's error grows as eps |H| / gap: about 1e-4 there) or the
// float QL sweeps stopped at their cap.  Same routine and ordering rules in
// both precisions.
__host__ __device__ __forceinline__ void eigen3_fp64_fallback(const float A[3][3], float V[3][3], float d[3]) {
    bool conv = true;
    eigen3<float>(A, V, d, &conv);
    auto mx = [](float a, float b) { return a < b ? b : a; };  // as std::max / std::min
    auto mn = [](float a, float b) { return b < a ? b : a; };
    const float m = mx(tabs(d[0]), mx(tabs(d[1]), tabs(d[2])));
    const float g = mn(tabs(d[1] - d[0]), mn(tabs(d[2] - d[1]), tabs(d[2] - d[0])));
    if (conv && !(g < 0x1p-10f * m)) return;
    double Ad[3][3], Vd[3][3], dd[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Ad[i][j] = (double)A[i][j];
    eigen3<double>(Ad, Vd, dd);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        d[i] = (float)dd[i];
#pragma unroll
        for (int j = 0; j < 3; j++) V[i][j] = (float)Vd[i][j];
    }
}

}  // namespace tsf
