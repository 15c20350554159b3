// Hand-written adjoints of the differentiable chain
//   B = blur_sigma(I), F = VF(B), V = GVF_N(F; mu), W = |V|, T = TDF(V, W)
// for L = sum_p gT(p) T(p).  Discrete choices are held fixed (blur mask
// size, Fmax branch, eigenvalue order, TDF radius r*).  All adjoints are
// float with double accumulation of the scalar parameter gradients.
//
// Kernels and their data movement:
//   k_blur_adjoint     gather B^T dB with clamp-to-edge folding (deterministic)
//   k_blur_dsigma      per-voxel dB * (dw * I) -> block partial sums (double)
//   k_vf_dg / k_vf_gather   the vector-field adjoint as two gathers
//   k_gvf_adj_march    the transposed Jacobi step on the folded adjoint,
//                      z-marching: transposed 7-point stencil and data term,
//                      dF accumulation, d/dmu block partials (double), and
//                      the mirror fold of its output for the next reverse step
//   k_tdf_adjoint      per-voxel recomputation of the circle fit at r*,
//                      then atomic scatter of the sample, direction
//                      (eigenvector perturbation) and Hessian contributions
// The GVF backward runs over checkpoints of the forward taken every k
// iterations; each segment is recomputed into a k-state buffer.
#include <math.h>

#include <algorithm>
#include <vector>

#include "eigen3.hpp"
#include "field.hpp"
#include "texel.hpp"

namespace tsf {

tsf_status gvf_steps_f32(const float* const f[3], const float* const v0[3], float* const out[3], Dims d, float mu,
                         int iters, float* ws, hipStream_t stream);
tsf_status gvf_states_f32(const float* const f[3], const float* v0, float* const* states, int count, Dims d,
                          float mu, hipStream_t stream);

constexpr int RB = 256;     // reduction block
constexpr int NPART = 1024; // partial sums per launch

__device__ __forceinline__ void block_sum_double(double v, double* out) {
    __shared__ double sh[RB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// out[0] += sum of n partials (one block, deterministic order)
__global__ void __launch_bounds__(RB) k_fold_partials(const double* __restrict__ part, int n, double* __restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += RB) s += part[i];
    __shared__ double sh[RB];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += sh[0];
}

__global__ void k_zero_double(double* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// ---------------------------------------------------------------- blur
// Per axis, the positions p with clamp(p + a) == y: one position inside,
// a run of them at the clamped ends.
__device__ __forceinline__ void pre_range(int y, int a, int N, int& lo, int& hi) {
    if (N == 1) {
        lo = hi = 0;
    } else if (y == 0) {
        lo = 0;
        hi = min(-a, N - 1);
    } else if (y == N - 1) {
        lo = max(0, N - 1 - a);
        hi = N - 1;
    } else {
        lo = hi = y - a;
        if (lo < 0 || lo > N - 1) hi = lo - 1;  // empty
    }
}

__global__ void k_blur_adjoint(const float* __restrict__ dB, float* __restrict__ dI, Dims d, int m,
                               const float* __restrict__ w) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= d.X || y >= d.Y) return;
    const int M = 2 * m + 1;
    float acc = 0.0f;
    for (int a = -m; a <= m; a++) {
        int xl, xh;
        pre_range(x, a, d.X, xl, xh);
        if (xl > xh) continue;
        for (int b = -m; b <= m; b++) {
            int yl, yh;
            pre_range(y, b, d.Y, yl, yh);
            if (yl > yh) continue;
            for (int c = -m; c <= m; c++) {
                int zl, zh;
                pre_range(z, c, d.Z, zl, zh);
                if (zl > zh) continue;
                float s = 0.0f;
                for (int pz = zl; pz <= zh; pz++)
                    for (int py = yl; py <= yh; py++)
                        for (int px = xl; px <= xh; px++) s += dB[d.at(px, py, pz)];
                acc += w[(a + m) + (b + m) * M + (c + m) * M * M] * s;
            }
        }
    }
    dI[d.at(x, y, z)] = acc;
}

__global__ void __launch_bounds__(RB) k_blur_dsigma(const float* __restrict__ dB, const float* __restrict__ I, Dims d,
                                                    int m, const double* __restrict__ dw, double* __restrict__ part) {
    const i64 n = d.n();
    const int M = 2 * m + 1;
    double s = 0.0;
    for (i64 i = blockIdx.x * (i64)RB + threadIdx.x; i < n; i += (i64)gridDim.x * RB) {
        const float g = dB[i];
        if (g == 0.0f) continue;
        const int x = (int)(i % d.X), y = (int)((i / d.X) % d.Y), z = (int)(i / ((i64)d.X * d.Y));
        double acc = 0.0;
        for (int a = -m; a <= m; a++)
            for (int b = -m; b <= m; b++)
                for (int c = -m; c <= m; c++)
                    acc += dw[(a + m) + (b + m) * M + (c + m) * M * M] *
                           (double)I[d.at(clampi(x + a, 0, d.X - 1), clampi(y + b, 0, d.Y - 1), clampi(z + c, 0, d.Z - 1))];
        s += (double)g * acc;
    }
    block_sum_double(s, &part[blockIdx.x]);
}

// ---------------------------------------------------------------- vector field
__global__ void k_vf_dg(const float* __restrict__ B, const float* __restrict__ dFx, const float* __restrict__ dFy,
                        const float* __restrict__ dFz, Dims d, float fmax, float sign, float* __restrict__ dg) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= d.X || y >= d.Y) return;
    const i64 N = d.n(), i = d.at(x, y, z);
    const int xm = max(x - 1, 0), xp = min(x + 1, d.X - 1), ym = max(y - 1, 0), yp = min(y + 1, d.Y - 1);
    const int zm = max(z - 1, 0), zp = min(z + 1, d.Z - 1);
    float g0 = sign * (0.5f * (B[d.at(xp, y, z)] - B[d.at(xm, y, z)]));
    float g1 = sign * (0.5f * (B[d.at(x, yp, z)] - B[d.at(x, ym, z)]));
    float g2 = sign * (0.5f * (B[d.at(x, y, zp)] - B[d.at(x, y, zm)]));
    const float l = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
    const float a = dFx[i], b = dFy[i], c = dFz[i];
    float r0, r1, r2;
    if (l < fmax) {
        r0 = a / fmax;
        r1 = b / fmax;
        r2 = c / fmax;
    } else {
        const float f0 = g0 / l, f1 = g1 / l, f2 = g2 / l;
        const float fd = f0 * a + f1 * b + f2 * c;
        r0 = (a - f0 * fd) / l;
        r1 = (b - f1 * fd) / l;
        r2 = (c - f2 * fd) / l;
    }
    const float h = sign * 0.5f;
    dg[i] = h * r0;
    dg[N + i] = h * r1;
    dg[2 * N + i] = h * r2;
}

// dB(y) = sum over axes of dg(y - e)[y >= 1] + dg(y)[y == N-1] - dg(y + e)[y <= N-2] - dg(y)[y == 0]
__global__ void k_vf_gather(const float* __restrict__ dg, Dims d, float* __restrict__ dB) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= d.X || y >= d.Y) return;
    const i64 N = d.n(), i = d.at(x, y, z);
    const int p[3] = {x, y, z}, n[3] = {d.X, d.Y, d.Z};
    const i64 st[3] = {1, (i64)d.X, (i64)d.X * d.Y};
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float* g = dg + a * N;
        if (p[a] >= 1) s += g[i - st[a]];
        if (p[a] == n[a] - 1) s += g[i];
        if (p[a] <= n[a] - 2) s -= g[i + st[a]];
        if (p[a] == 0) s -= g[i];
    }
    dB[i] = s;
}

// ---------------------------------------------------------------- GVF
// Mirror fold: the adjoint that reaches the update at interior position q
// is the sum over every w whose mirror image is q (per axis: q itself, plus
// 0 when q == 2, plus N-1 when q == N-3); boundary positions receive none.
__device__ __forceinline__ int pre_list(int q, int N, int* o) {
    if (q <= 0 || q >= N - 1) return 0;
    int k = 0;
    o[k++] = q;
    if (q == 2) o[k++] = 0;
    if (q == N - 3) o[k++] = N - 1;
    return k;
}
__device__ __forceinline__ float fold(const float* __restrict__ lam, const Dims& d, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= d.X || y >= d.Y || z >= d.Z) return 0.0f;
    int px[3], py[3], pz[3];
    const int nx = pre_list(x, d.X, px), ny = pre_list(y, d.Y, py), nz = pre_list(z, d.Z, pz);
    float s = 0.0f;
    for (int c = 0; c < nz; c++)
        for (int b = 0; b < ny; b++)
            for (int a = 0; a < nx; a++) s += lam[d.at(px[a], py[b], pz[c])];
    return s;
}

struct P3 {
    const float* c[3];
};
struct M3 {
    float* c[3];
};

// Lf = fold(lam) on all three planes (the adjoint that enters the first
// reverse step).
__global__ void k_fold3(P3 lam, Dims d, M3 out) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= d.X || y >= d.Y) return;
    const i64 i = d.at(x, y, z);
#pragma unroll
    for (int c = 0; c < 3; c++) out.c[c][i] = fold(lam.c[c], d, x, y, z);
}

// ---------------------------------------------------------------- GVF reverse step, z-marching
// One reverse Euler step as a z-march.  In: Lf = fold(lam^{n+1}) (zero on every
// boundary plane), v = v^n, F.  lam^n(y) = (1 - 6 mu - |f(y)|^2) Lf(y) + mu sum_e
// Lf(y+e).  FOLD: out = fold(lam^n), the input of the next reverse step;
// otherwise (the step from v^0 = f) dF += lam^n.  Always: dF += data-term
// gradient Lf m2 + 2 f sum_c Lf_c (f_c - v_c); part[block] = the block's
// share of d/dmu = sum over interior q of Lf(q) . lap(v)(q).  Lf is zero off
// the interior and the 7-point Laplacian is symmetric, so that sum equals
// sum over every voxel w of v(w) . (S(w) - 6 Lf(w)), S(w) the sum of Lf over
// w's in-volume neighbours, which the update computes anyway.
// Structure: a workgroup of ABX x ABY threads owns that tile of columns and a
// chunk of AZC planes; per plane (one barrier) Lf of the plane, with a
// one-column ring, goes to LDS for the x/y neighbours, and each thread keeps
// its column's planes in register rings loaded three planes ahead.
// The fold of the output needs no gather: Lf is zero on every boundary
// plane, so lam at a boundary position w is mu * Lf at w's single interior
// neighbour when w has one boundary coordinate, and zero when it has more.
// Hence fold(lam)(q) = lam(q) + mu * sum over axes of [q_a == 2] Lf(q - e_a)
// + [q_a == N_a - 3] Lf(q + e_a) at interior q, and zero on boundary planes.
constexpr int ABX = 64, ABY = 8, AZC = 64;

struct AdjCtx {
    P3 Lf, v, F;
    M3 out, dF;
    Dims d;
    float mu;
    int lane, ty, cx, cy, z1;
    bool act, icol, hrole, hok, fx2, fxN, fy2, fyN;
    i64 P, col, hidx;
    int hly, hlx;
};

template <bool FOLD, int T4>
__device__ __forceinline__ void adj_step(const AdjCtx& m, int t, float (&lf)[4][3], float (&vv)[4][3],
                                         float (&fr)[4][3], float (&dfr)[4][3], float (&hl)[2][3],
                                         float (&LL)[2][3][ABY + 2][ABX + 2], double& gm) {
    constexpr int SM1 = (T4 + 3) & 3, S0 = T4, SP1 = (T4 + 1) & 3, B = T4 & 1;
    const int lane = m.lane, ty = m.ty;
    // out-of-volume cells read as zero
#pragma unroll
    for (int c = 0; c < 3; c++) LL[B][c][ty + 1][lane + 1] = m.act ? lf[S0][c] : 0.0f;
    if (m.hrole) {
#pragma unroll
        for (int c = 0; c < 3; c++) LL[B][c][m.hly][m.hlx] = m.hok ? hl[B][c] : 0.0f;
    }
    __syncthreads();
    if (m.act && t < m.z1) {
        const int z = t;
        const bool zm_ok = z > 0, zp_ok = z < m.d.Z - 1;
        const bool interior = m.icol && zm_ok && zp_ok;
        const bool fz2 = z == 2, fzN = z == m.d.Z - 3;
        const float f0 = fr[S0][0], f1 = fr[S0][1], f2 = fr[S0][2];
        const float fc[3] = {f0, f1, f2};
        const float m2 = f0 * f0 + f1 * f1 + f2 * f2;
        const float diag = 1.0f - 6.0f * m.mu - m2;
        float lam[3], fold[3];
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float l0 = lf[S0][c];
            const float lxp = LL[B][c][ty + 1][lane + 2], lxm = LL[B][c][ty + 1][lane];
            const float lyp = LL[B][c][ty + 2][lane + 1], lym = LL[B][c][ty][lane + 1];
            const float lzp = zp_ok ? lf[SP1][c] : 0.0f, lzm = zm_ok ? lf[SM1][c] : 0.0f;
            const float nb = lxp + lxm + lyp + lym + lzp + lzm;
            lam[c] = diag * l0 + m.mu * nb;
            const float v0 = vv[S0][c];
            s += l0 * (fc[c] - v0);
            gm += (double)v0 * (double)(nb - 6.0f * l0);
            const float e = (m.fx2 ? lxm : 0.0f) + (m.fxN ? lxp : 0.0f) + (m.fy2 ? lym : 0.0f) +
                            (m.fyN ? lyp : 0.0f) + (fz2 ? lzm : 0.0f) + (fzN ? lzp : 0.0f);
            fold[c] = interior ? lam[c] + m.mu * e : 0.0f;
        }
        const i64 i = z * m.P + m.col;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float g = dfr[S0][c] + (lf[S0][c] * m2 + 2.0f * fc[c] * s);
            if (!FOLD) g += lam[c];
            m.dF.c[c][i] = g;
            if (FOLD) m.out.c[c][i] = fold[c];
        }
    }
    // loads: plane t+3 of the rings into the slots of t-1, the ring cells of
    // plane t+2 into this step's halo slot
    const int z3 = min(t + 3, m.d.Z - 1), z2 = min(t + 2, m.d.Z - 1);
    const i64 i3 = m.act ? z3 * m.P + m.col : 0, ih = m.hok ? z2 * m.P + m.hidx : 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        lf[SM1][c] = m.Lf.c[c][i3];
        vv[SM1][c] = m.v.c[c][i3];
        fr[SM1][c] = m.F.c[c][i3];
        dfr[SM1][c] = m.dF.c[c][i3];
        hl[B][c] = m.Lf.c[c][ih];
    }
}

template <bool FOLD>
__global__ void __launch_bounds__(ABX* ABY, 4) k_gvf_adj_march(P3 Lf, P3 v, P3 F, Dims d, float mu, M3 out, M3 dF,
                                                            double* __restrict__ part) {
    __shared__ float LL[2][3][ABY + 2][ABX + 2];
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    AdjCtx m;
    m.Lf = Lf;
    m.v = v;
    m.F = F;
    m.out = out;
    m.dF = dF;
    m.d = d;
    m.mu = mu;
    m.lane = threadIdx.x;
    m.ty = threadIdx.y;
    const int x0 = tbx * ABX, y0 = tby * ABY;
    m.cx = x0 + m.lane;
    m.cy = y0 + m.ty;
    m.act = m.cx < d.X && m.cy < d.Y;
    m.icol = m.cx >= 1 && m.cx <= d.X - 2 && m.cy >= 1 && m.cy <= d.Y - 2;
    m.fx2 = m.cx == 2;
    m.fxN = m.cx == d.X - 3;
    m.fy2 = m.cy == 2;
    m.fyN = m.cy == d.Y - 3;
    m.P = (i64)d.X * d.Y;
    m.col = m.act ? (i64)m.cy * d.X + m.cx : 0;
    // ring of the tile: rows y0-1 (ty 0) and y0+ABY (ty 1) at this thread's
    // column; columns x0-1 and x0+ABX (ty 2, lanes 0..2*ABY-1)
    int hx = 0, hy = 0;
    m.hrole = false;
    m.hly = 0;
    m.hlx = 0;
    if (m.ty == 0 || m.ty == 1) {
        m.hrole = true;
        hx = m.cx;
        hy = m.ty == 0 ? y0 - 1 : y0 + ABY;
        m.hly = m.ty == 0 ? 0 : ABY + 1;
        m.hlx = m.lane + 1;
    } else if (m.ty == 2 && m.lane < 2 * ABY) {
        m.hrole = true;
        const bool left = m.lane < ABY;
        hx = left ? x0 - 1 : x0 + ABX;
        hy = y0 + (left ? m.lane : m.lane - ABY);
        m.hly = (left ? m.lane : m.lane - ABY) + 1;
        m.hlx = left ? 0 : ABX + 1;
    }
    m.hok = m.hrole && hx >= 0 && hx < d.X && hy >= 0 && hy < d.Y;
    m.hidx = m.hok ? (i64)hy * d.X + hx : 0;
    const int z0 = tbz * AZC;
    m.z1 = min(z0 + AZC, d.Z);
    float lf[4][3], vv[4][3], fr[4][3], dfr[4][3], hl[2][3];
    // the loads of the four steps before z0 in the steps' order
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int z = min(max(z0 - 1 + g, 0), d.Z - 1), k = (g + 3) & 3;
        const i64 i = m.act ? z * m.P + m.col : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            lf[k][c] = Lf.c[c][i];
            vv[k][c] = v.c[c][i];
            fr[k][c] = F.c[c][i];
            dfr[k][c] = dF.c[c][i];
            if (g >= 2) {
                const i64 ih = m.hok ? min(z0 + g - 2, d.Z - 1) * m.P + m.hidx : 0;
                hl[g & 1][c] = Lf.c[c][ih];
            }
        }
    }
    double gm = 0.0;
    for (int t = z0; t < m.z1; t += 4) {
        adj_step<FOLD, 0>(m, t, lf, vv, fr, dfr, hl, LL, gm);
        adj_step<FOLD, 1>(m, t + 1, lf, vv, fr, dfr, hl, LL, gm);
        adj_step<FOLD, 2>(m, t + 2, lf, vv, fr, dfr, hl, LL, gm);
        adj_step<FOLD, 3>(m, t + 3, lf, vv, fr, dfr, hl, LL, gm);
    }
    __shared__ double sh[ABX * ABY];
    const int tid = m.ty * ABX + m.lane;
    sh[tid] = gm;
    __syncthreads();
    for (int o = ABX * ABY / 2; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x + (i64)gridDim.x * (blockIdx.y + (i64)gridDim.y * blockIdx.z)] = sh[0];
}

// ---------------------------------------------------------------- two reverse steps per launch
// k_gvf_adj2 runs reverse steps n+1 and n in one launch, so the adjoint, the
// states, f and dF cross HBM once per two steps:
//   in   Lf = fold(lam^{n+2}) (zero on every boundary plane), va = v^{n+1},
//        vb = v^n, F, dF
//   Lf1 = fold(lam^{n+1}) is computed on the tile E and kept on chip
//   out  fold(lam^n) (FOLD), or dF += lam^n (the step from v^0 = f)
//   dF  += the data-term gradients of both steps
//   part[b], part[nblk + b]: block b's share of d/dmu of step n+1 and of
//        step n, in the form of k_gvf_adj_march (v . (S - 6 Lf) per voxel)
// A workgroup of 64 x A2Y threads (one wave per row) covers 64 x A2Y columns
// starting two columns left of and one row above its output tile O (A2OW x
// A2OH columns); Lf1 is computed on every row at lanes 1..62 (the tile E),
// the second step on O.  Waves 0 and 1 also load the Lf rows just above and
// below the workgroup.  Tiles do not overlap, since dF is read and written:
// the last one along x and y ends past the volume's edge.  A chunk of A2ZC
// output planes is marched in z; per step t (one barrier): LDS <- Lf plane t
// (with the ring rows) and Lf1 plane t-1; Lf1(t) on E from the LDS x/y
// neighbours and the register ring of Lf planes t-1..t+1 (with step n+1's
// d/dmu term at plane t on O); then, at plane t-1 on O, lam^n from Lf1
// planes t-2..t, the data terms of both steps, step n's d/dmu term, the dF
// update and the output.  Every value comes from the expressions of
// k_gvf_adj_march in the same order (the build does not contract), so one
// launch of this kernel and two of that one give the same adjoint and dF
// bit for bit.  Loads run ahead into register rings: Lf and f three planes,
// va and the ring rows two, vb and dF one, issued in that order so every
// value is waited for a step after its load.
constexpr int A2X = 64, A2Y = 12;
constexpr int A2OW = A2X - 4, A2OH = A2Y - 2;
constexpr int A2ZC = 64;  // output planes per z chunk

// The arrays are read and written through buffer resources: L, va, vb, dF
// and out each hold three component volumes back to back (one resource each,
// the component and plane offsets in a scalar register), f's components are
// three resources.  The host launches the kernel only when 3 volumes' bytes
// fit 32 bits.
struct Adj2Ctx {
    __amdgpu_buffer_rsrc_t rL, ra, rb, rd, ro, rf[3];
    unsigned cb, pb;     // bytes per component volume, per plane
    unsigned col, hcol;  // byte offsets of this thread's cell and its ring-row cell within a plane
    unsigned ocol;       // col on the rows of O, else past every resource's range (the load reads 0, no memory)
    Dims d;
    float mu;
    int lane, ty, q0, q1, hly;
    bool act, icol, own, hrow, hin, fx2, fxN, fy2, fyN;
};

__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    const float v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
    // keeps the loads in program order, so the waits counted at the loop
    // head are the same on entry and on the back edge
    asm volatile("" ::: "memory");
    return v;
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, 0);
}

template <bool FOLD, int K>
__device__ __forceinline__ void adj2_step(const Adj2Ctx& m, int t, float (&lr)[4][3], float (&ar)[4][3],
                                          float (&br)[4][3], float (&fr)[4][3], float (&dr)[4][3],
                                          float (&l1)[4][3], float (&hl)[2][3], float& mq,
                                          float (&LL)[2][3][A2Y + 2][A2X + 2], float (&L1)[2][3][A2Y][A2X],
                                          double& gmA, double& gmB) {
    constexpr int S0 = K, SM1 = (K + 3) & 3, SM2 = (K + 2) & 3, SP1 = (K + 1) & 3;
    constexpr int B = K & 1, BP = (K + 1) & 1;
    const int lane = m.lane, ty = m.ty, Z = m.d.Z;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        LL[B][c][ty + 1][lane + 1] = m.act ? lr[S0][c] : 0.0f;
        L1[BP][c][ty][lane] = l1[SM1][c];
    }
    if (m.hrow) {
#pragma unroll
        for (int c = 0; c < 3; c++) LL[B][c][m.hly][lane + 1] = m.hin ? hl[B][c] : 0.0f;
    }
    __syncthreads();
    // Lf1 = fold(lam^{n+1}) at plane t (zero off the interior), and step
    // n+1's d/dmu term at plane t on O
    const bool own_t = m.own && t >= m.q0 && t < m.q1;
    const float f0 = fr[S0][0], f1 = fr[S0][1], f2 = fr[S0][2];
    const float m2 = f0 * f0 + f1 * f1 + f2 * f2;
    {
        const bool in = m.icol && t >= 1 && t <= Z - 2;
        const float diag = 1.0f - 6.0f * m.mu - m2;
        const bool fz2 = t == 2, fzN = t == Z - 3;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float l0 = lr[S0][c];
            const float lxp = LL[B][c][ty + 1][lane + 2], lxm = LL[B][c][ty + 1][lane];
            const float lyp = LL[B][c][ty + 2][lane + 1], lym = LL[B][c][ty][lane + 1];
            const float lzp = t < Z - 1 ? lr[SP1][c] : 0.0f, lzm = t > 0 ? lr[SM1][c] : 0.0f;
            const float nb = lxp + lxm + lyp + lym + lzp + lzm;
            const float lam = diag * l0 + m.mu * nb;
            const double q = (double)ar[S0][c] * (double)(nb - 6.0f * l0);
            gmA += own_t ? q : 0.0;
            const float e = (m.fx2 ? lxm : 0.0f) + (m.fxN ? lxp : 0.0f) + (m.fy2 ? lym : 0.0f) +
                            (m.fyN ? lyp : 0.0f) + (fz2 ? lzm : 0.0f) + (fzN ? lzp : 0.0f);
            l1[S0][c] = in ? lam + m.mu * e : 0.0f;
        }
    }
    // the second step at plane p = t - 1 (|f|^2 of plane p carried from the
    // last step)
    const int p = t - 1;
    if (m.own && p >= m.q0 && p < m.q1) {
        const float m2p = mq;
        const bool pin = m.icol && p >= 1 && p <= Z - 2;
        const bool pz2 = p == 2, pzN = p == Z - 3;
        const float diagp = 1.0f - 6.0f * m.mu - m2p;
        const float fc[3] = {fr[SM1][0], fr[SM1][1], fr[SM1][2]};
        float lam0[3], fold0[3];
        float sA = 0.0f, sB = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float k0 = l1[SM1][c];
            const float kxp = L1[BP][c][ty][lane + 1], kxm = L1[BP][c][ty][lane - 1];
            const float kyp = L1[BP][c][ty + 1][lane], kym = L1[BP][c][ty - 1][lane];
            const float kzp = l1[S0][c], kzm = l1[SM2][c];
            const float nb = kxp + kxm + kyp + kym + kzp + kzm;
            lam0[c] = diagp * k0 + m.mu * nb;
            sA += lr[SM1][c] * (fc[c] - ar[SM1][c]);
            sB += k0 * (fc[c] - br[SM1][c]);
            gmB += (double)br[SM1][c] * (double)(nb - 6.0f * k0);
            const float e = (m.fx2 ? kxm : 0.0f) + (m.fxN ? kxp : 0.0f) + (m.fy2 ? kym : 0.0f) +
                            (m.fyN ? kyp : 0.0f) + (pz2 ? kzm : 0.0f) + (pzN ? kzp : 0.0f);
            fold0[c] = pin ? lam0[c] + m.mu * e : 0.0f;
        }
        const unsigned sp = (unsigned)p * m.pb;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float g = dr[SM1][c] + (lr[SM1][c] * m2p + 2.0f * fc[c] * sA);
            g = g + (l1[SM1][c] * m2p + 2.0f * fc[c] * sB);
            if (!FOLD) g += lam0[c];
            bst(m.rd, m.col, c * m.cb + sp, g);
            if (FOLD) bst(m.ro, m.col, c * m.cb + sp, fold0[c]);
        }
    }
    mq = m2;
    // loads: Lf and f of plane t+3, va and the Lf ring-row cells of plane
    // t+2, vb and dF of plane t+1
    const unsigned s3 = (unsigned)min(t + 3, Z - 1) * m.pb, s2 = (unsigned)min(t + 2, Z - 1) * m.pb;
    const unsigned s1 = (unsigned)min(t + 1, Z - 1) * m.pb;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        lr[SM1][c] = bld(m.rL, m.col, c * m.cb + s3);
        fr[SM1][c] = bld(m.rf[c], m.col, s3);
        ar[SM2][c] = bld(m.ra, m.ocol, c * m.cb + s2);
        hl[B][c] = bld(m.rL, m.hcol, c * m.cb + s2);
        br[SP1][c] = bld(m.rb, m.ocol, c * m.cb + s1);
        dr[SP1][c] = bld(m.rd, m.ocol, c * m.cb + s1);
    }
}

template <bool FOLD>
__global__ void __launch_bounds__(A2X* A2Y, 4) k_gvf_adj2(const float* L, const float* va, const float* vb, P3 F,
                                                         Dims d, float mu, float* out, float* dF,
                                                         double* __restrict__ part) {
    __shared__ float LL[2][3][A2Y + 2][A2X + 2];
    __shared__ float L1[2][3][A2Y][A2X];
    __shared__ double red[2][A2Y];
    Adj2Ctx m;
    m.d = d;
    m.mu = mu;
    m.lane = threadIdx.x;
    m.ty = threadIdx.y;
    m.pb = (unsigned)d.X * (unsigned)d.Y * 4u;
    m.cb = m.pb * (unsigned)d.Z;
    const int nr = (int)(3u * m.cb), nf = (int)m.cb;
    m.rL = __builtin_amdgcn_make_buffer_rsrc((void*)L, 0, nr, BUFFER_WORD3);
    m.ra = __builtin_amdgcn_make_buffer_rsrc((void*)va, 0, nr, BUFFER_WORD3);
    m.rb = __builtin_amdgcn_make_buffer_rsrc((void*)vb, 0, nr, BUFFER_WORD3);
    m.rd = __builtin_amdgcn_make_buffer_rsrc((void*)dF, 0, nr, BUFFER_WORD3);
    m.ro = __builtin_amdgcn_make_buffer_rsrc((void*)out, 0, nr, BUFFER_WORD3);
#pragma unroll
    for (int c = 0; c < 3; c++) m.rf[c] = __builtin_amdgcn_make_buffer_rsrc((void*)F.c[c], 0, nf, BUFFER_WORD3);
    // the pad cells left and right of every Lf row hold 0
    for (int i = m.ty * A2X + m.lane; i < 2 * 3 * (A2Y + 2) * 2; i += A2X * A2Y) {
        const int side = i & 1, row = i >> 1;
        (&LL[0][0][0][0])[row * (A2X + 2) + (side ? A2X + 1 : 0)] = 0.0f;
    }
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int ex = tbx * A2OW - 2, ey = tby * A2OH - 1;
    const int cx = ex + m.lane, cy = ey + m.ty;
    m.act = cx >= 0 && cx < d.X && cy >= 0 && cy < d.Y;
    m.icol = m.act && cx >= 1 && cx <= d.X - 2 && cy >= 1 && cy <= d.Y - 2;
    m.own = m.act && m.lane >= 2 && m.lane <= A2X - 3 && m.ty >= 1 && m.ty <= A2Y - 2;
    m.fx2 = cx == 2;
    m.fxN = cx == d.X - 3;
    m.fy2 = cy == 2;
    m.fyN = cy == d.Y - 3;
    m.col = m.act ? (unsigned)(cy * d.X + cx) * 4u : 0u;
    // va, vb and dF are used on O only: the waves of rows 0 and A2Y-1 read
    // none of them (a buffer load at an offset of 2^31 or more is past the
    // range, 3 volumes < 2^31 bytes, and returns 0 without a memory access)
    m.ocol = m.ty >= 1 && m.ty <= A2Y - 2 ? m.col : 0x80000000u;
    m.hrow = m.ty == 0 || m.ty == 1;
    const int hy = m.ty == 0 ? ey - 1 : ey + A2Y;
    m.hly = m.ty == 0 ? 0 : A2Y + 1;
    m.hin = m.hrow && cx >= 0 && cx < d.X && hy >= 0 && hy < d.Y;
    m.hcol = m.hin ? (unsigned)(hy * d.X + cx) * 4u : 0u;
    m.q0 = tbz * A2ZC;
    m.q1 = min(m.q0 + A2ZC, d.Z);
    const int t0 = m.q0 - 1, te = m.q1;
    float lr[4][3], ar[4][3], br[4][3], fr[4][3], dr[4][3], l1[4][3], hl[2][3];
    // the loads of the four steps before t0, in the order the steps issue
    // them; ring slot k holds a plane congruent to t0 + k (mod 4)
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int s = t0 - 4 + g, k3 = (g + 3) & 3, k2 = (g + 2) & 3, k1 = (g + 1) & 3;
        const unsigned s3 = (unsigned)min(max(s + 3, 0), d.Z - 1) * m.pb;
        const unsigned s2 = (unsigned)min(max(s + 2, 0), d.Z - 1) * m.pb;
        const unsigned s1 = (unsigned)min(max(s + 1, 0), d.Z - 1) * m.pb;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            lr[k3][c] = bld(m.rL, m.col, c * m.cb + s3);
            fr[k3][c] = bld(m.rf[c], m.col, s3);
            ar[k2][c] = bld(m.ra, m.ocol, c * m.cb + s2);
            hl[g & 1][c] = bld(m.rL, m.hcol, c * m.cb + s2);
            br[k1][c] = bld(m.rb, m.ocol, c * m.cb + s1);
            dr[k1][c] = bld(m.rd, m.ocol, c * m.cb + s1);
            l1[g][c] = 0.0f;
        }
    }
    float mq = 0.0f;
    double gmA = 0.0, gmB = 0.0;
    // four unconditional steps per iteration (up to three past te: they
    // store nothing, and their loads are clamped into the volume)
    for (int t = t0; t <= te; t += 4) {
        adj2_step<FOLD, 0>(m, t, lr, ar, br, fr, dr, l1, hl, mq, LL, L1, gmA, gmB);
        adj2_step<FOLD, 1>(m, t + 1, lr, ar, br, fr, dr, l1, hl, mq, LL, L1, gmA, gmB);
        adj2_step<FOLD, 2>(m, t + 2, lr, ar, br, fr, dr, l1, hl, mq, LL, L1, gmA, gmB);
        adj2_step<FOLD, 3>(m, t + 3, lr, ar, br, fr, dr, l1, hl, mq, LL, L1, gmA, gmB);
    }
    // block sums in a fixed order: within each wave by xor shuffles, then
    // the waves in order
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        gmA += __shfl_xor(gmA, o);
        gmB += __shfl_xor(gmB, o);
    }
    if (m.lane == 0) {
        red[0][m.ty] = gmA;
        red[1][m.ty] = gmB;
    }
    __syncthreads();
    if (m.lane == 0 && m.ty == 0) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int w = 0; w < A2Y; w++) {
            a += red[0][w];
            b += red[1][w];
        }
        const i64 nblk = (i64)gridDim.x * gridDim.y * gridDim.z;
        const i64 blk = blockIdx.x + (i64)gridDim.x * (blockIdx.y + (i64)gridDim.y * blockIdx.z);
        part[blk] = a;
        part[nblk + blk] = b;
    }
}

// part2[b] = sum of a contiguous chunk of n partials (fixed order)
__global__ void __launch_bounds__(RB) k_reduce_parts(const double* __restrict__ part, i64 n, double* __restrict__ part2) {
    const i64 chunk = (n + gridDim.x - 1) / gridDim.x;
    const i64 lo = blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0;
    for (i64 j = lo + threadIdx.x; j < hi; j += RB) s += part[j];
    block_sum_double(s, &part2[blockIdx.x]);
}

// W = |V| (1 where |V| == 0): dV += dW V / |V|
__global__ void k_finish_adjoint(const float* __restrict__ vx, const float* __restrict__ vy,
                                 const float* __restrict__ vz, const float* __restrict__ dW, float* __restrict__ dx,
                                 float* __restrict__ dy, float* __restrict__ dz, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const float a = vx[i], b = vy[i], c = vz[i];
        const float l = sqrtf(a * a + b * b + c * c);
        if (l > 0.0f) {
            const float g = dW[i] / l;
            dx[i] += g * a;
            dy[i] += g * b;
            dz[i] += g * c;
        }
    }
}

// ---------------------------------------------------------------- TDF
// The filter's trilinear sample (trilinear_tex: identical terms and order,
// so the recomputed radius is the forward's) together with its spatial
// Jacobian and the eight corner indices and weights, in two parts:
// tri_fetch issues the eight texel loads, tri_eval uses them, so a caller
// can have the next sample's loads in flight while it evaluates this one.
struct TriTaps {
    float4 t[8];
    float a, b, c;
    int i0, i1, j0, j1, k0, k1;
};
__device__ __forceinline__ void tri_fetch(const Tex<F32>& T, const Dims& d, float u, float v, float w, TriTaps& q) {
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    q.a = uu - fu;
    q.b = vv - fv;
    q.c = ww - fw;
    const int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    q.i1 = clampi(i0 + 1, 0, d.X - 1);
    q.j1 = clampi(j0 + 1, 0, d.Y - 1);
    q.k1 = clampi(k0 + 1, 0, d.Z - 1);
    q.i0 = clampi(i0, 0, d.X - 1);
    q.j0 = clampi(j0, 0, d.Y - 1);
    q.k0 = clampi(k0, 0, d.Z - 1);
    const auto r00 = T.row(d, q.j0, q.k0), r10 = T.row(d, q.j1, q.k0);
    const auto r01 = T.row(d, q.j0, q.k1), r11 = T.row(d, q.j1, q.k1);
    q.t[0] = T.at(r00, q.i0);
    q.t[1] = T.at(r00, q.i1);
    q.t[2] = T.at(r10, q.i0);
    q.t[3] = T.at(r10, q.i1);
    q.t[4] = T.at(r01, q.i0);
    q.t[5] = T.at(r01, q.i1);
    q.t[6] = T.at(r11, q.i0);
    q.t[7] = T.at(r11, q.i1);
}
__device__ __forceinline__ void tri_eval(const TriTaps& q, float s[3], float J[3][3], int ci[8], int cj[8], int ck[8],
                                         float wt[8]) {
    const float a = q.a, b = q.b, c = q.c;
    const int I[8] = {q.i0, q.i1, q.i0, q.i1, q.i0, q.i1, q.i0, q.i1};
    const int Jj[8] = {q.j0, q.j0, q.j1, q.j1, q.j0, q.j0, q.j1, q.j1};
    const int K[8] = {q.k0, q.k0, q.k0, q.k0, q.k1, q.k1, q.k1, q.k1};
    const float A[8] = {1 - a, a, 1 - a, a, 1 - a, a, 1 - a, a};
    const float Bv[8] = {1 - b, 1 - b, b, b, 1 - b, 1 - b, b, b};
    const float C[8] = {1 - c, 1 - c, 1 - c, 1 - c, c, c, c, c};
    const float dA[8] = {-1, 1, -1, 1, -1, 1, -1, 1}, dB[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const float dC[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        ci[k] = I[k];
        cj[k] = Jj[k];
        ck[k] = K[k];
        wt[k] = A[k] * Bv[k] * C[k];
    }
    // the filter's sum, term by term
    float sx = wt[0] * q.t[0].x, sy = wt[0] * q.t[0].y, sz = wt[0] * q.t[0].z;
#pragma unroll
    for (int k = 1; k < 8; k++) {
        sx = sx + wt[k] * q.t[k].x;
        sy = sy + wt[k] * q.t[k].y;
        sz = sz + wt[k] * q.t[k].z;
    }
    s[0] = sx;
    s[1] = sy;
    s[2] = sz;
#pragma unroll
    for (int m = 0; m < 3; m++) J[m][0] = J[m][1] = J[m][2] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float gx = dA[k] * Bv[k] * C[k], gy = A[k] * dB[k] * C[k], gz = A[k] * Bv[k] * dC[k];
        const float tv[3] = {q.t[k].x, q.t[k].y, q.t[k].z};
#pragma unroll
        for (int m = 0; m < 3; m++) {
            J[m][0] += gx * tv[m];
            J[m][1] += gy * tv[m];
            J[m][2] += gz * tv[m];
        }
    }
}

// Accumulation tile of dV (3 planes) over one brick plus a halo of `h`
// voxels, in LDS; contributions outside it go to global atomics.
// The sample scatter accumulates in an LDS tile over one brick plus a halo
// of `halo` voxels, as 64-bit fixed point: a contribution v adds
// rint(v * S) to its cell, S = 2^(50 - e) for the brick's largest |dL/dT|
// < 2^e.  Every sample contribution is at most |dL/dT| / 32 in magnitude
// (times a unit direction component and a trilinear weight), so a rounded
// contribution is below 2^45 (larger ones, never expected, take the global
// path) and a cell's sum -- at most 8 * 32 contributions from each of the
// brick's 512 voxels, 2^17 in all -- stays below 2^62.  The quantum 2^(e-51)
// is 2^-46 of the largest contribution, far below a float's rounding, so a
// cell fed only by a brick's small gradients (1e-6 beside 1) and sums that
// cancel keep float precision.  Integer adds are exact and order-free, so the tile's sums are
// deterministic, and on this chip an integer LDS atomic costs a third of a
// float one.  Contributions that are not finite, or fall outside the tile,
// or arrive when the tile is off, go to global float atomics.  The tile's
// atomics name the LDS address space explicitly (through a generic pointer
// they would become flat atomics).
constexpr int TAX = 16, TAY = 8, TAZ = 4;  // TDF adjoint brick
constexpr int TA_HALO = 5;                 // the largest halo the tile covers (r_max < 4.5)
constexpr int TA_CELLS = (TAX + 2 * TA_HALO) * (TAY + 2 * TA_HALO) * (TAZ + 2 * TA_HALO);
typedef __attribute__((address_space(3))) long long lds_i64;
struct AdjTile {
    long long* a;  // 3 * cells fixed-point sums in LDS
    float S;       // scale (a power of two)
    bool on;
    int ox, oy, oz, tx, ty, tz, cells;
    __device__ __forceinline__ void add(float* const dV[3], const Dims& d, int c, int x, int y, int z, float v) const {
        const int lx = x - ox, ly = y - oy, lz = z - oz;
        const float q = v * S;  // |q| < 2^45 for a sample's contribution; NaN fails the test
        if (on && fabsf(q) < 70368744177664.0f && (unsigned)lx < (unsigned)tx && (unsigned)ly < (unsigned)ty &&
            (unsigned)lz < (unsigned)tz)
            __hip_atomic_fetch_add((lds_i64*)(a + (c * cells + lx + tx * (ly + ty * lz))),
                                   (long long)__builtin_rintf(q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else if (v != 0.0f)
            unsafeAtomicAdd(dV[c] + d.at(x, y, z), v);
    }
};

// Adjoint of circleFittingTDF at the forward's radius.  Block = one 16x8x4
// brick; halo > 0 (at most TA_HALO) turns the fixed-point tile on, halo == 0
// leaves global atomics only.  The Hessian's terms (12 per voxel, scaled by
// 1 / eigenvalue gaps, so unbounded) always go to global atomics.
template <bool NORM>
__global__ void __launch_bounds__(TAX* TAY* TAZ) k_tdf_adjoint(Tex<F32> T, Dims d, float rmin, float rmax, float rstep,
                                                              CircleTables tab, const float* __restrict__ gT,
                                                              const float* __restrict__ rad,
                                                              float* __restrict__ dVx, float* __restrict__ dVy,
                                                              float* __restrict__ dVz, float* __restrict__ dW,
                                                              int halo) {
    constexpr int NT = TAX * TAY * TAZ;
    __shared__ long long acc[3 * TA_CELLS];
    // the circle's direction tables, indexed per lane below
    __shared__ float tc[32], ts[32];
    __shared__ float wmax[NT / 64];
    const int tid = threadIdx.x + TAX * (threadIdx.y + TAY * threadIdx.z);
    if (tid < 32) {
        tc[tid] = tab.c[tid];
        ts[tid] = tab.s[tid];
    }
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int x = tbx * TAX + threadIdx.x, y = tby * TAY + threadIdx.y;
    const int z = tbz * TAZ + threadIdx.z;
    const float g = (x < d.X && y < d.Y && z < d.Z) ? gT[d.at(x, y, z)] : 0.0f;
    // the brick's largest |dL/dT| (NaN or infinity turn the tile off)
    float m = __builtin_isfinite(g) ? fabsf(g) : __builtin_inff();
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((tid & 63) == 0) wmax[tid >> 6] = m;
    AdjTile tile;
    tile.a = acc;
    tile.tx = TAX + 2 * halo;
    tile.ty = TAY + 2 * halo;
    tile.tz = TAZ + 2 * halo;
    tile.cells = tile.tx * tile.ty * tile.tz;
    tile.ox = tbx * TAX - halo;
    tile.oy = tby * TAY - halo;
    tile.oz = tbz * TAZ - halo;
    const bool lds_on = halo > 0 && halo <= TA_HALO;
    if (lds_on)
        for (int k = tid; k < 3 * tile.cells; k += NT) acc[k] = 0;
    __syncthreads();
    float gmax = 0.0f;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) gmax = fmaxf(gmax, wmax[w]);
    tile.on = lds_on && gmax > 0.0f && __builtin_isfinite(gmax);
    int e = 0;
    if (tile.on) frexpf(gmax, &e);  // gmax < 2^e
    // S = 2^(50 - e) must be a finite float: bricks whose gradients are all
    // below 2^-77 use the global float atomics
    tile.on = tile.on && e >= -76;
    tile.S = tile.on ? ldexpf(1.0f, 50 - e) : 0.0f;
    float* const dV[3] = {dVx, dVy, dVz};
    if (g != 0.0f) {
        float H[3][3], E[3][3], lam[3];
        hessian_tex<NORM>(T, d, x, y, z, H);
        eigen3(H, E, lam);
        const float e2x = E[0][1], e2y = E[1][1], e2z = E[2][1];
        const float e3x = E[0][2], e3y = E[1][2], e3z = E[2][2];
        float maxSum = 0.0f, rstar = 0.0f;
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        if (rad) {
            // the forward's r*: a radius the search can return (>= rmin) is
            // one where the forward's response was its sum there.  With
            // rmin > 0 the forward records such a radius only where that sum
            // is positive (elsewhere it records 0), which is all the
            // gradient needs to know; otherwise the sum is recomputed by the
            // forward's code, so bit-identical
            rstar = rad[d.at(x, y, z)];
            if (rmin > 0.0f && rstar >= rmin && rstar <= rmax) {
                maxSum = 1.0f;
            } else if (rstar >= rmin && rstar <= rmax) {
                float sum = 0.0f;
                for (int j = 0; j < 32; j++) {
                    const float cj = tab.c[j], sj = tab.s[j];
                    const float vx = cj * e3x + sj * e2x, vy = cj * e3y + sj * e2y, vz = cj * e3z + sj * e2z;
                    float s[3];
                    trilinear_tex(T, d, fx + rstar * vx, fy + rstar * vy, fz + rstar * vz, s);
                    sum += (-s[0]) * vx + (-s[1]) * vy + (-s[2]) * vz;
                }
                maxSum = sum / 32.0f;
            }
        }
        // the forward's radius search
        for (float radius = rmin; !rad && radius <= rmax; radius += rstep) {
            float sum = 0.0f;
            for (int j = 0; j < 32; j++) {
                const float cj = tab.c[j], sj = tab.s[j];
                const float vx = cj * e3x + sj * e2x, vy = cj * e3y + sj * e2y, vz = cj * e3z + sj * e2z;
                float s[3];
                trilinear_tex(T, d, fx + radius * vx, fy + radius * vy, fz + radius * vz, s);
                sum += (-s[0]) * vx + (-s[1]) * vy + (-s[2]) * vz;
            }
            sum /= 32.0f;
            if (sum > maxSum) {
                maxSum = sum;
                rstar = radius;
            } else {
                break;
            }
        }
        if (maxSum > 0.0f) {
            const float cf = -g / 32.0f;
            float A2[3] = {0, 0, 0}, A3[3] = {0, 0, 0};
            // the lanes of a wave take the 32 directions in rotated orders:
            // neighbouring voxels then sample far-apart points in the same
            // iteration, and their LDS atomics rarely meet on one cell
            // (the taps of direction jr + 1 are loaded while jr is evaluated)
            const int rot = tid & 31;
            TriTaps cur, nxt;
            {
                const float cc = tc[rot], ss = ts[rot];
                tri_fetch(T, d, fx + rstar * (cc * e3x + ss * e2x), fy + rstar * (cc * e3y + ss * e2y),
                          fz + rstar * (cc * e3z + ss * e2z), cur);
            }
            for (int jr = 0; jr < 32; jr++) {
                const int j = (jr + rot) & 31;
                const float cc = tc[j], ss = ts[j];
                const float dj[3] = {cc * e3x + ss * e2x, cc * e3y + ss * e2y, cc * e3z + ss * e2z};
                if (jr < 31) {
                    const int jn = (j + 1) & 31;
                    const float cn = tc[jn], sn = ts[jn];
                    tri_fetch(T, d, fx + rstar * (cn * e3x + sn * e2x), fy + rstar * (cn * e3y + sn * e2y),
                              fz + rstar * (cn * e3z + sn * e2z), nxt);
                }
                float s[3], J[3][3], wt[8];
                int ci[8], cj[8], ck[8];
                tri_eval(cur, s, J, ci, cj, ck, wt);
                cur = nxt;
#pragma unroll
                for (int k = 0; k < 8; k++)
#pragma unroll
                    for (int c = 0; c < 3; c++) tile.add(dV, d, c, ci[k], cj[k], ck[k], cf * dj[c] * wt[k]);
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const float u = s[m] + rstar * (dj[0] * J[0][m] + dj[1] * J[1][m] + dj[2] * J[2][m]);
                    A3[m] += cf * cc * u;
                    A2[m] += cf * ss * u;
                }
            }
            // eigenvector perturbation: de_k = sum_{i != k} e_i (e_i^T dH e_k) / (lam_k - lam_i)
            float G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                const int k = kk + 1;
                const float* A = kk == 0 ? A2 : A3;
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    if (i == k || lam[k] == lam[i]) continue;
                    const float coef = (A[0] * E[0][i] + A[1] * E[1][i] + A[2] * E[2][i]) / (lam[k] - lam[i]);
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++) G[a][b] += coef * E[a][i] * E[b][k];
                }
            }
            const float gxx = G[0][0], gyx = G[0][1] + G[1][0], gzx = G[0][2] + G[2][0];
            const float gyy = G[1][1], gzy = G[1][2] + G[2][1], gzz = G[2][2];
            // H entries are central differences (clamp-to-edge) of the samples
            // V_c / W (NORM) or V_c
            auto put = [&](int comp, int xx, int yy, int zz, float cv) {
                xx = clampi(xx, 0, d.X - 1);
                yy = clampi(yy, 0, d.Y - 1);
                zz = clampi(zz, 0, d.Z - 1);
                const i64 q = d.at(xx, yy, zz);
                if (NORM) {
                    const float4 tq = T.at(T.row(d, yy, zz), xx);
                    const float vq = comp == 0 ? tq.x : (comp == 1 ? tq.y : tq.z);
                    unsafeAtomicAdd(dV[comp] + q, cv / tq.w);
                    unsafeAtomicAdd(dW + q, -cv * vq / (tq.w * tq.w));
                } else {
                    unsafeAtomicAdd(dV[comp] + q, cv);
                }
            };
            put(0, x + 1, y, z, 0.5f * gxx);
            put(0, x - 1, y, z, -0.5f * gxx);
            put(1, x + 1, y, z, 0.5f * gyx);
            put(1, x - 1, y, z, -0.5f * gyx);
            put(1, x, y + 1, z, 0.5f * gyy);
            put(1, x, y - 1, z, -0.5f * gyy);
            put(2, x + 1, y, z, 0.5f * gzx);
            put(2, x - 1, y, z, -0.5f * gzx);
            put(2, x, y + 1, z, 0.5f * gzy);
            put(2, x, y - 1, z, -0.5f * gzy);
            put(2, x, y, z + 1, 0.5f * gzz);
            put(2, x, y, z - 1, -0.5f * gzz);
        }
    }
    if (tile.on) {
        // one global atomic per touched tile cell inside the volume
        __syncthreads();
        const double inv = 1.0 / (double)tile.S;
        for (int k = tid; k < tile.cells; k += NT) {
            const int lx = k % tile.tx, ly = (k / tile.tx) % tile.ty, lz = k / (tile.tx * tile.ty);
            const int gx = tile.ox + lx, gy = tile.oy + ly, gz = tile.oz + lz;
            if (gx < 0 || gy < 0 || gz < 0 || gx >= d.X || gy >= d.Y || gz >= d.Z) continue;
            const i64 q = d.at(gx, gy, gz);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const long long v = acc[c * tile.cells + k];
                if (v != 0) unsafeAtomicAdd(dV[c] + q, (float)((double)v * inv));
            }
        }
    }
}

// w = |v| where nonzero, else 1 (GVF3DFinish)
__global__ void k_mag_plane(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                            float* __restrict__ w, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const float l = sqrtf(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
        w[i] = l > 0.0f ? l : 1.0f;
    }
}

// L = (1/n) sum (t - target)^2 -> grad = 2 (t - target) / n, block partials of L
__global__ void __launch_bounds__(RB) k_mse(const float* __restrict__ t, const float* __restrict__ target, i64 n,
                                            float* __restrict__ grad, double* __restrict__ part) {
    const float inv = 2.0f / (float)n;
    double s = 0.0;
    for (i64 i = blockIdx.x * (i64)RB + threadIdx.x; i < n; i += (i64)gridDim.x * RB) {
        const float e = t[i] - target[i];
        grad[i] = inv * e;
        s += (double)e * (double)e;
    }
    block_sum_double(s, &part[blockIdx.x]);
}

static unsigned g1d(i64 n) {
    i64 g = (n + 255) / 256;
    return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

static tsf_status sync_read_double(const double* dptr, double* host, hipStream_t s, const char* W) {
    hipError_t e = hipMemcpyAsync(host, dptr, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

#define STENCIL_GRID(d) dim3(((d).X + 63) / 64, ((d).Y + 3) / 4, (d).Z), dim3(64, 4)

extern "C" {

size_t tsf_blur_adjoint_workspace_size(void) { return 1331 * (sizeof(float) + sizeof(double)) + (NPART + 1) * sizeof(double) + 512; }

tsf_status tsf_blur_adjoint(const float* dB, const float* in, float* dI, tsf_dims dims, float sigma, double* dsigma,
                            void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_blur_adjoint";
    if (!dB || !dI || (dsigma && !in)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (dB == dI) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: dB and dI must differ", W);
    TSF_TRY(check_dims(dims, 1, W));
    float mask[1331];
    int32_t m;
    TSF_TRY(tsf_create_blur_mask(sigma, mask, &m));
    const int M = 2 * m + 1, K = M * M * M;
    // d w_k / d sigma of the float mask of tsf_create_blur_mask, in double
    std::vector<double> dw(K);
    {
        double G = 0, Gp = 0;
        std::vector<double> g(K), gp(K);
        for (int a = -m; a <= m; a++)
            for (int b = -m; b <= m; b++)
                for (int c = -m; c <= m; c++) {
                    const int i = (a + m) + (b + m) * M + (c + m) * M * M;
                    const double r2 = a * a + b * b + c * c, s = sigma;
                    g[i] = exp(-r2 / (2 * s * s));
                    gp[i] = g[i] * r2 / (s * s * s);
                    G += g[i];
                    Gp += gp[i];
                }
        for (int i = 0; i < K; i++) dw[i] = (gp[i] * G - g[i] * Gp) / (G * G);
    }
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_blur_adjoint_workspace_size(), W));
    double* part = (double*)sc.ptr;
    double* acc = part + NPART;
    double* ddw = acc + 1;
    float* dmask = (float*)(ddw + 1331);
    hipError_t e = hipMemcpyAsync(dmask, mask, sizeof(float) * K, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ddw, dw.data(), sizeof(double) * K, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    Dims d = D(dims);
    hipLaunchKernelGGL(k_blur_adjoint, STENCIL_GRID(d), 0, stream, dB, dI, d, m, dmask);
    TSF_TRY(check_launch(W));
    if (dsigma) {
        hipLaunchKernelGGL(k_zero_double, 1, 64, 0, stream, acc, 1);
        hipLaunchKernelGGL(k_blur_dsigma, NPART, RB, 0, stream, dB, in, d, m, ddw, part);
        hipLaunchKernelGGL(k_fold_partials, 1, RB, 0, stream, part, NPART, acc);
        TSF_TRY(check_launch(W));
        TSF_TRY(sync_read_double(acc, dsigma, stream, W));
    } else {
        e = hipStreamSynchronize(stream);  // the staged mask lives in scratch
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

size_t tsf_vector_field_adjoint_workspace_size(tsf_dims dims) { return (size_t)3 * D(dims).n() * sizeof(float); }

tsf_status tsf_vector_field_adjoint(const float* volume, const tsf_field* dF, tsf_dims dims, float fmax, int32_t sign,
                                    float* dvolume, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_vector_field_adjoint";
    if (!volume || !dvolume) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL volume", W);
    TSF_TRY(check_field(dF, false, W));
    if (dF->storage != TSF_STORAGE_F32) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: dF must be F32", W);
    TSF_TRY(check_dims(dims, 1, W));
    if (!(fmax > 0.0f)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: fmax must be positive", W);
    if (sign != 1 && sign != -1) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: sign must be +1 or -1", W);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_vector_field_adjoint_workspace_size(dims), W));
    Dims d = D(dims);
    hipLaunchKernelGGL(k_vf_dg, STENCIL_GRID(d), 0, stream, volume, (const float*)dF->x, (const float*)dF->y,
                       (const float*)dF->z, d, fmax, (float)sign, (float*)sc.ptr);
    hipLaunchKernelGGL(k_vf_gather, STENCIL_GRID(d), 0, stream, (const float*)sc.ptr, d, dvolume);
    TSF_TRY(check_launch(W));
    if (sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

int64_t tsf_gvf_checkpoint_count(int32_t iterations, int32_t checkpoint_every) {
    if (iterations <= 0 || checkpoint_every <= 0) return 0;
    return (iterations + checkpoint_every - 1) / checkpoint_every;
}

tsf_status tsf_gvf_euler_checkpointed(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu,
                                      int32_t iterations, int32_t checkpoint_every, float* checkpoints,
                                      size_t checkpoint_bytes, hipStream_t stream) {
    const char* W = "tsf_gvf_euler_checkpointed";
    TSF_TRY(check_field(f, false, W));
    TSF_TRY(check_field(out, true, W));
    if (f->storage != TSF_STORAGE_F32 || out->storage != TSF_STORAGE_F32)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: F32 fields only", W);
    TSF_TRY(check_dims(dims, 4, W));
    if (iterations < 1 || checkpoint_every < 1)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: iterations and checkpoint_every must be >= 1", W);
    if (!(mu >= 0.0f)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: mu must be >= 0", W);
    Dims d = D(dims);
    const i64 N = d.n();
    const i64 C = tsf_gvf_checkpoint_count(iterations, checkpoint_every);
    if (!checkpoints || checkpoint_bytes < (size_t)C * 3 * N * sizeof(float))
        return set_error(TSF_ERR_OUT_OF_MEMORY, "%s: checkpoint buffer too small", W);
    Scratch sc;
    TSF_TRY(sc.get(nullptr, 0, (size_t)3 * N * sizeof(float), W));
    const float* fp[3] = {(const float*)f->x, (const float*)f->y, (const float*)f->z};
    for (i64 s = 0; s < C; s++) {
        float* cp = checkpoints + s * 3 * N;
        float* cps[3] = {cp, cp + N, cp + 2 * N};
        if (s == 0) {
            for (int c = 0; c < 3; c++) {
                hipError_t e = hipMemcpyAsync(cps[c], fp[c], N * sizeof(float), hipMemcpyDeviceToDevice, stream);
                if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
            }
        } else {
            const float* prev = checkpoints + (s - 1) * 3 * N;
            const float* pv[3] = {prev, prev + N, prev + 2 * N};
            TSF_TRY(gvf_steps_f32(fp, pv, cps, d, mu, checkpoint_every, (float*)sc.ptr, stream));
        }
    }
    // last segment into out, then the magnitude plane
    const float* last = checkpoints + (C - 1) * 3 * N;
    const float* lv[3] = {last, last + N, last + 2 * N};
    float* op[3] = {(float*)out->x, (float*)out->y, (float*)out->z};
    const int rem = iterations - (int)((C - 1) * checkpoint_every);
    TSF_TRY(gvf_steps_f32(fp, lv, op, d, mu, rem, (float*)sc.ptr, stream));
    hipLaunchKernelGGL(k_mag_plane, g1d(N), 256, 0, stream, op[0], op[1], op[2], (float*)out->w, N);
    TSF_TRY(check_launch(W));
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    return TSF_OK;
}

static dim3 adj_grid(const Dims& d) { return dim3((d.X + ABX - 1) / ABX, (d.Y + ABY - 1) / ABY, (d.Z + AZC - 1) / AZC); }
static size_t adj_blocks(const Dims& d) {
    const dim3 g = adj_grid(d);
    return (size_t)g.x * g.y * g.z;
}
static dim3 adj2_grid(const Dims& d) {
    return dim3((d.X + A2OW - 1) / A2OW, (d.Y + A2OH - 1) / A2OH, (d.Z + A2ZC - 1) / A2ZC);
}
static size_t adj2_blocks(const Dims& d) {
    const dim3 g = adj2_grid(d);
    return (size_t)g.x * g.y * g.z;
}
// d/dmu block partials: one set per single reverse step, two per pair
static size_t adj_parts(const Dims& d) { return std::max(adj_blocks(d), 2 * adj2_blocks(d)); }

// float states first, then the double partial sums at a 256-byte boundary
static size_t adj_float_bytes(const Dims& d, int k) {
    return (((size_t)k + 2) * 3 * (size_t)d.n() * sizeof(float) + 255) & ~(size_t)255;
}

size_t tsf_gvf_euler_adjoint_workspace_size(tsf_dims dims, int32_t checkpoint_every) {
    const Dims d = D(dims);
    if (checkpoint_every < 1) return 0;
    return adj_float_bytes(d, checkpoint_every) + (NPART + 1 + adj_parts(d)) * sizeof(double);
}

tsf_status tsf_gvf_euler_adjoint(const tsf_field* f, tsf_dims dims, float mu, int32_t iterations,
                                 int32_t checkpoint_every, const float* checkpoints, const float* dV, float* dF,
                                 double* dmu, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_gvf_euler_adjoint";
    TSF_TRY(check_field(f, false, W));
    if (f->storage != TSF_STORAGE_F32) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: F32 field only", W);
    TSF_TRY(check_dims(dims, 4, W));
    if (!checkpoints || !dV || !dF) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", W);
    if (iterations < 1 || checkpoint_every < 1)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: iterations and checkpoint_every must be >= 1", W);
    Dims d = D(dims);
    const i64 N = d.n();
    const int k = checkpoint_every;
    const i64 C = tsf_gvf_checkpoint_count(iterations, k);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_gvf_euler_adjoint_workspace_size(dims, k), W));
    float* seg = (float*)sc.ptr;              // k states of 3 planes
    float* lamA = seg + (i64)k * 3 * N;       // folded adjoint, ping-pong
    float* lamB = lamA + 3 * N;
    double* part2 = (double*)((char*)sc.ptr + adj_float_bytes(d, k));  // NPART
    double* acc = part2 + NPART;
    double* part = acc + 1;                   // one per stencil block
    const i64 nblk = (i64)adj_blocks(d), nblk2 = (i64)adj2_blocks(d);
    // two reverse steps per launch address three volumes with 32-bit byte
    // offsets
    const bool pairs = 3 * N * (i64)sizeof(float) < ((i64)1 << 31);
    const float* fp[3] = {(const float*)f->x, (const float*)f->y, (const float*)f->z};
    hipError_t e = hipMemsetAsync(dF, 0, 3 * N * sizeof(float), stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    hipLaunchKernelGGL(k_zero_double, 1, 64, 0, stream, acc, 1);
    hipLaunchKernelGGL(k_fold3, STENCIL_GRID(d), 0, stream, P3{{dV, dV + N, dV + 2 * N}}, d,
                       M3{{lamA, lamA + N, lamA + 2 * N}});
    TSF_TRY(check_launch(W));
    P3 Fp{{fp[0], fp[1], fp[2]}};
    M3 dFp{{dF, dF + N, dF + 2 * N}};
    float* lin = lamA;
    float* lout = lamB;
    for (i64 s = C - 1; s >= 0; s--) {
        const int start = (int)(s * k), end = (int)std::min<i64>((s + 1) * k, iterations);
        // recompute v^start .. v^(end-1) into seg
        const float* cp = checkpoints + s * 3 * N;
        e = hipMemcpyAsync(seg, cp, 3 * N * sizeof(float), hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
        if (end - start > 1) {
            std::vector<float*> st(end - start - 1);
            for (int t = 1; t < end - start; t++) st[t - 1] = seg + (i64)t * 3 * N;
            TSF_TRY(gvf_states_f32(fp, seg, st.data(), end - start - 1, d, mu, stream));
        }
        // steps n and n-1 in one launch while both lie in the segment; the
        // adjoint, dF and d/dmu come out as from two single steps
        int n = end - 1;
        for (; pairs && n - 1 >= start; n -= 2) {
            const float* va = seg + (i64)(n - start) * 3 * N;
            const float* vb = va - 3 * N;
            if (n - 1 > 0)
                hipLaunchKernelGGL(k_gvf_adj2<true>, adj2_grid(d), dim3(A2X, A2Y), 0, stream, lin, va, vb, Fp, d, mu,
                                   lout, dF, part);
            else
                hipLaunchKernelGGL(k_gvf_adj2<false>, adj2_grid(d), dim3(A2X, A2Y), 0, stream, lin, va, vb, Fp, d,
                                   mu, lout, dF, part);
            for (int h = 0; h < 2; h++) {
                hipLaunchKernelGGL(k_reduce_parts, NPART, RB, 0, stream, part + h * nblk2, nblk2, part2);
                hipLaunchKernelGGL(k_fold_partials, 1, RB, 0, stream, part2, NPART, acc);
            }
            TSF_TRY(check_launch(W));
            std::swap(lin, lout);
        }
        for (; n >= start; n--) {
            const float* vn = seg + (i64)(n - start) * 3 * N;
            P3 L{{lin, lin + N, lin + 2 * N}};
            P3 Vn{{vn, vn + N, vn + 2 * N}};
            M3 Lo{{lout, lout + N, lout + 2 * N}};
            if (n > 0)
                hipLaunchKernelGGL(k_gvf_adj_march<true>, adj_grid(d), dim3(ABX, ABY), 0, stream, L, Vn, Fp, d, mu, Lo,
                                   dFp, part);
            else  // v^0 = f: the last adjoint flows into dF
                hipLaunchKernelGGL(k_gvf_adj_march<false>, adj_grid(d), dim3(ABX, ABY), 0, stream, L, Vn, Fp, d, mu,
                                   Lo, dFp, part);
            hipLaunchKernelGGL(k_reduce_parts, NPART, RB, 0, stream, part, nblk, part2);
            hipLaunchKernelGGL(k_fold_partials, 1, RB, 0, stream, part2, NPART, acc);
            TSF_TRY(check_launch(W));
            std::swap(lin, lout);
        }
    }
    double h = 0.0;
    TSF_TRY(sync_read_double(acc, &h, stream, W));
    if (dmu) *dmu = h;
    return TSF_OK;
}

tsf_status tsf_gvf_finish_adjoint(const tsf_field* v, const tsf_field* dv, tsf_dims dims, hipStream_t stream) {
    const char* W = "tsf_gvf_finish_adjoint";
    TSF_TRY(check_field(v, false, W));
    TSF_TRY(check_field(dv, true, W));
    if (v->storage != TSF_STORAGE_F32 || dv->storage != TSF_STORAGE_F32)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: F32 fields only", W);
    TSF_TRY(check_dims(dims, 1, W));
    const i64 N = D(dims).n();
    hipLaunchKernelGGL(k_finish_adjoint, g1d(N), 256, 0, stream, (const float*)v->x, (const float*)v->y,
                       (const float*)v->z, (const float*)dv->w, (float*)dv->x, (float*)dv->y, (float*)dv->z, N);
    return check_launch(W);
}

size_t tsf_mse_workspace_size(void) { return (NPART + 1) * sizeof(double); }

tsf_status tsf_mse(const float* t, const float* target, int64_t n, float* grad, double* loss, void* workspace,
                   size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_mse";
    if (!t || !target || !grad || !loss) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (n < 1) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: n must be >= 1", W);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_mse_workspace_size(), W));
    double* part = (double*)sc.ptr;
    double* acc = part + NPART;
    hipLaunchKernelGGL(k_zero_double, 1, 64, 0, stream, acc, 1);
    hipLaunchKernelGGL(k_mse, NPART, RB, 0, stream, t, target, (i64)n, grad, part);
    hipLaunchKernelGGL(k_fold_partials, 1, RB, 0, stream, part, NPART, acc);
    TSF_TRY(check_launch(W));
    double h = 0.0;
    TSF_TRY(sync_read_double(acc, &h, stream, W));
    *loss = h / (double)n;
    return TSF_OK;
}

size_t tsf_tdf_adjoint_workspace_size(tsf_dims dims) { return (size_t)D(dims).n() * sizeof(float4); }

tsf_status tsf_tdf_adjoint(const tsf_field* field, tsf_dims dims, float rmin, float rmax, float rstep,
                           const float* tdf_grad, const float* radius, const tsf_field* dfield, void* workspace,
                           size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_tdf_adjoint";
    const bool norm = !(rmax < 4.0f);
    TSF_TRY(check_field(field, norm, W));
    TSF_TRY(check_field(dfield, norm, W));
    if (field->storage != TSF_STORAGE_F32 || dfield->storage != TSF_STORAGE_F32)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: F32 fields only", W);
    TSF_TRY(check_dims(dims, 1, W));
    if (!tdf_grad) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL gradient", W);
    if (!(rstep > 0.0f) || !__builtin_isfinite(rmin) || !__builtin_isfinite(rmax))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: radius range must be finite with rstep > 0", W);
    if ((double)(rmax - rmin) / rstep > 100000.0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: too many radii", W);
    static const CircleTables tab = make_circle_tables();
    Dims d = D(dims);
    const i64 N = d.n();
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_tdf_adjoint_workspace_size(dims), W));
    pack_texels_f32(field, (float4*)sc.ptr, N, stream);
    Tex<F32> T{(const float4*)sc.ptr, 1.0f};
    // LDS tile: samples reach ceil(rmax + 0.5) voxels from the centre (the
    // trilinear corner below u - 0.5); wider halos than the tile's leave
    // global atomics only
    int halo = (int)ceilf(fmaxf(fabsf(rmin), fabsf(rmax)) + 0.5f);
    if (halo < 1) halo = 1;
    if (halo > TA_HALO) halo = 0;
    float* dx = (float*)dfield->x;
    float* dy = (float*)dfield->y;
    float* dz = (float*)dfield->z;
    float* dw = (float*)dfield->w;
    dim3 grid((d.X + TAX - 1) / TAX, (d.Y + TAY - 1) / TAY, (d.Z + TAZ - 1) / TAZ), block(TAX, TAY, TAZ);
    if (norm)
        hipLaunchKernelGGL(k_tdf_adjoint<true>, grid, block, 0, stream, T, d, rmin, rmax, rstep, tab, tdf_grad,
                           radius, dx, dy, dz, dw, halo);
    else
        hipLaunchKernelGGL(k_tdf_adjoint<false>, grid, block, 0, stream, T, d, rmin, rmax, rstep, tab, tdf_grad,
                           radius, dx, dy, dz, dw, halo);
    TSF_TRY(check_launch(W));
    if (sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

}  // extern "C"
