// Vector-field access shared by the Hessian, TDF and candidate kernels.
#pragma once

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "tsf_common.hpp"

namespace tsf {

// SoA field planes as a kernel argument (any storage).
template <class S>
struct Field {
    const typename S::T* x;
    const typename S::T* y;
    const typename S::T* z;
    const typename S::T* w;
    float inv;  // 1 / scale
    __device__ __forceinline__ float c(int comp, i64 i) const {
        return S::ld(comp == 0 ? x : (comp == 1 ? y : z), i, inv);
    }
    __device__ __forceinline__ float mag(i64 i) const { return S::ld(w, i, inv); }
};

// Jacobian "Hessian" of circleFittingTDF / findCandidateCenterpoints2
// (gradient / gradientNormalized, kernels_no_3d_write.cl:90-200; layout
// 1522-1526).  Samples clamp to the edge; with NORM each sample is divided by
// the w plane of its own voxel.  `A` is any accessor with c(comp, i) and
// mag(i).
template <bool NORM, class A>
__device__ __forceinline__ void hessian(const A& F, const Dims& d, int x, int y, int z, float H[3][3]) {
    const int xm = max(x - 1, 0), xp = min(x + 1, d.X - 1);
    const int ym = max(y - 1, 0), yp = min(y + 1, d.Y - 1);
    const int zm = max(z - 1, 0), zp = min(z + 1, d.Z - 1);
    const i64 ixp = d.at(xp, y, z), ixm = d.at(xm, y, z);
    const i64 iyp = d.at(x, yp, z), iym = d.at(x, ym, z);
    const i64 izp = d.at(x, y, zp), izm = d.at(x, y, zm);
    float a0, a1, b0, b1, b2, b3, c0, c1, c2, c3, c4, c5;
    a0 = F.c(0, ixp); a1 = F.c(0, ixm);
    b0 = F.c(1, ixp); b1 = F.c(1, ixm); b2 = F.c(1, iyp); b3 = F.c(1, iym);
    c0 = F.c(2, ixp); c1 = F.c(2, ixm); c2 = F.c(2, iyp); c3 = F.c(2, iym); c4 = F.c(2, izp); c5 = F.c(2, izm);
    if (NORM) {
        const float wxp = F.mag(ixp), wxm = F.mag(ixm), wyp = F.mag(iyp), wym = F.mag(iym);
        const float wzp = F.mag(izp), wzm = F.mag(izm);
        a0 = a0 / wxp; a1 = a1 / wxm;
        b0 = b0 / wxp; b1 = b1 / wxm; b2 = b2 / wyp; b3 = b3 / wym;
        c0 = c0 / wxp; c1 = c1 / wxm; c2 = c2 / wyp; c3 = c3 / wym; c4 = c4 / wzp; c5 = c5 / wzm;
    }
    const float fxx = 0.5f * (a0 - a1);
    const float fyx = 0.5f * (b0 - b1), fyy = 0.5f * (b2 - b3);
    const float fzx = 0.5f * (c0 - c1), fzy = 0.5f * (c2 - c3), fzz = 0.5f * (c4 - c5);
    H[0][0] = fxx; H[0][1] = fyx; H[0][2] = fzx;
    H[1][0] = fyx; H[1][1] = fyy; H[1][2] = fzy;
    H[2][0] = fzx; H[2][1] = fzy; H[2][2] = fzz;
}

// The 32 circle directions at integer radians j = 0..31, rounded to six
// significant digits as "%g" prints cos(j) and sin(j) (the tables of
// kernels_no_3d_write.cl:1500-1501 hold exactly these values).
struct CircleTables {
    float c[32];
    float s[32];
};

inline CircleTables make_circle_tables() {
    CircleTables t;
    char buf[64];
    for (int j = 0; j < 32; j++) {
        snprintf(buf, sizeof buf, "%g", cos((double)j));
        t.c[j] = strtof(buf, nullptr);
        snprintf(buf, sizeof buf, "%g", sin((double)j));
        t.s[j] = strtof(buf, nullptr);
    }
    return t;
}

}  // namespace tsf
