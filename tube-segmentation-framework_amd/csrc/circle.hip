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
#include <atomic>

#include "eigen3.hpp"
#include "field.hpp"
#include "texel.hpp"

namespace tsf {

// Pack SoA planes (raw stored values) into texels.  w == NULL packs 1.
__global__ void k_pack_f32(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                           const float* __restrict__ w, float4* __restrict__ out, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        out[i] = make_float4(x[i], y[i], z[i], w ? w[i] : 1.0f);
}
// sn16: code -32768 (0x8000) is stored as -32767 (0x8001); both read as -1.
__global__ void k_pack_16(const uint16_t* __restrict__ x, const uint16_t* __restrict__ y,
                          const uint16_t* __restrict__ z, const uint16_t* __restrict__ w, uint16_t one, bool sn16,
                          uint2* __restrict__ out, i64 n) {
    auto c = [sn16](unsigned v) { return sn16 && v == 0x8000u ? 0x8001u : v; };
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const unsigned ww = w ? w[i] : one;
        out[i] = make_uint2(c(x[i]) | (c(y[i]) << 16), c(z[i]) | (c(ww) << 16));
    }
}



// The mean of dot(-F(p + r V_j), V_j) over the 32 circle directions, in the
// reference's operation order (kernels_no_3d_write.cl:1539-1556); FUSED: the
// precision form's, with the direction fma(c_j, e3, s_j e2), the position
// fma(r, V_j, p - 0.5), fused taps (trilinear_fused) and the dot product as
// a chain of fused multiply-adds subtracted from the sum with one rounding.
template <bool FUSED, class A>
__device__ __forceinline__ float circle_sum(const A& T, const Dims& d, float fx, float fy, float fz, float radius,
                                            const float (&e2)[3], const float (&e3)[3], const CircleTables& tab) {
    float sum = 0.0f;
    if constexpr (FUSED) {
        const float xm = fx - 0.5f, ym = fy - 0.5f, zm = fz - 0.5f;
#pragma unroll 8
        for (int j = 0; j < 32; j++) {
            const float cj = tab.c[j], sj = tab.s[j];
            const float vx = __builtin_fmaf(cj, e3[0], sj * e2[0]), vy = __builtin_fmaf(cj, e3[1], sj * e2[1]);
            const float vz = __builtin_fmaf(cj, e3[2], sj * e2[2]);
            float s[3];
            trilinear_fused(T, d, __builtin_fmaf(radius, vx, xm), __builtin_fmaf(radius, vy, ym),
                            __builtin_fmaf(radius, vz, zm), s);
            const float dp = __builtin_fmaf(s[2], vz, __builtin_fmaf(s[1], vy, s[0] * vx));
            sum = __builtin_fmaf(-T.inv, dp, sum);
        }
    } else {
        for (int j = 0; j < 32; j++) {
            const float cj = tab.c[j], sj = tab.s[j];
            const float vx = cj * e3[0] + sj * e2[0], vy = cj * e3[1] + sj * e2[1], vz = cj * e3[2] + sj * e2[2];
            float s[3];
            trilinear_tex(T, d, fx + radius * vx, fy + radius * vy, fz + radius * vz, s);
            sum += (-s[0]) * vx + (-s[1]) * vy + (-s[2]) * vz;
        }
    }
    return sum / 32.0f;
}

// splineTDF (kernels_no_3d_write.cl:1574-1667) at one voxel, with the 12
// arms runSplineTDF enqueues (tubeDetectionFilters.cpp:46): normalised
// Hessian, eigenvectors e1..e3, and along each arm cos(a_j) e3 + sin(a_j) e2
// the first radius where the sampled field magnitude, after rising, falls
// with a two-sample mean above the voxel's own by min_avg.  The arms loop is
// unrolled so the border radii stay in registers.
struct SplineTables {
    float c[12], s[12];
};
inline SplineTables make_spline_tables() {
    SplineTables t;
    for (int j = 0; j < 12; j++) {
        const float alpha = 2.0f * 3.14159274f * (float)j / 12.0f;  // 2 * M_PI_F * j / arms, in float
        t.c[j] = (float)cos((double)alpha);
        t.s[j] = (float)sin((double)alpha);
    }
    return t;
}

__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

template <class S>
__global__ void __launch_bounds__(CBX* CBY* CBZ) k_tdf_spline(Tex<S> T, Dims d, float rmin, float rmax, float rstep,
                                                              float min_avg, SplineTables tab,
                                                              float* __restrict__ tdf, float* __restrict__ rad) {
    constexpr int ARMS = 12;
    const int x = blockIdx.x * CBX + threadIdx.x, y = blockIdx.y * CBY + threadIdx.y;
    const int z = blockIdx.z * CBZ + threadIdx.z;
    if (x >= d.X || y >= d.Y || z >= d.Z) return;
    float H[3][3], V[3][3], ev[3];
    hessian_tex<true>(T, d, x, y, z, H);
    eigen3(H, V, ev);
    const float e1x = V[0][0], e1y = V[1][0], e1z = V[2][0];
    const float e2x = V[0][1], e2y = V[1][1], e2z = V[2][1];
    const float e3x = V[0][2], e3y = V[1][2], e3z = V[2][2];
    const float4 f0 = T.at(T.row(d, y, z), x);
    const float cur = len3(f0.x, f0.y, f0.z);
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    float maxR[ARMS];
    float avgR = 0.0f, sum = 0.0f;
    bool invalid = false;
#pragma unroll
    for (int j = 0; j < ARMS; j++) {
        maxR[j] = 999.0f;
        if (invalid) continue;  // the reference leaves the arms loop here
        const float vx = tab.c[j] * e3x + tab.s[j] * e2x, vy = tab.c[j] * e3y + tab.s[j] * e2y;
        const float vz = tab.c[j] * e3z + tab.s[j] * e2z;
        float s[3];
        trilinear_tex(T, d, fx + rmin * vx, fy + rmin * vy, fz + rmin * vz, s);
        float prev = len3(s[0], s[1], s[2]);
        int up = cur > prev ? 0 : 1;
        for (float radius = rmin + rstep; radius <= rmax; radius += rstep) {
            trilinear_tex(T, d, fx + radius * vx, fy + radius * vy, fz + radius * vz, s);
            const float mag = len3(s[0], s[1], s[2]);
            if (up == 1 && mag < prev && (prev + mag) / 2.0f - cur > min_avg) {
                maxR[j] = radius;
                avgR += radius;
                const float lv = len3(vx, vy, vz);
                const float nx = s[0] / mag, ny = s[1] / mag, nz = s[2] / mag;
                if (nx * (-(vx / lv)) + ny * (-(vy / lv)) + nz * (-(vz / lv)) < 0.0f) {
                    invalid = true;
                    sum = 0.0f;
                }
                sum += 1.0f - fabsf(nx * e1x + ny * e1y + nz * e1z);
                break;
            }
            if (mag > prev) up = 1;
            prev = mag;
        }
        if (maxR[j] == 999.0f) invalid = true;
    }
    const i64 i = d.at(x, y, z);
    rad[i] = avgR / (float)ARMS;
    if (invalid) {
        tdf[i] = 0.0f;
        return;
    }
    float sym = 0.0f;
#pragma unroll
    for (int j = 0; j < ARMS / 2; j++) sym += fminf(maxR[j], maxR[ARMS / 2 + j]) / fmaxf(maxR[j], maxR[ARMS / 2 + j]);
    sym /= (float)(ARMS / 2);
    tdf[i] = fminf(1.0f, (sum / (float)ARMS) * sym + 0.2f);
}

// The brick's footprint -- the brick plus a halo of h voxels, clamp-to-edge
// -- staged in LDS once; the Hessian and every trilinear tap read LDS.  h
// covers the Hessian's +-1 and the farthest sample corner,
// floor(rmax + 0.5) + 1.
// The radius search runs in rounds, one radius per round for every voxel
// of the brick still improving: after each round the survivors (voxel, e2,
// e3, best sum) are compacted to the front of the workgroup through LDS, so
// a wave's lanes stay busy and waves with no survivor retire.  The radius of
// round k is the same float accumulation for every voxel, so each voxel's
// arithmetic is that of the reference's sequential loop over radii
// (kernels_no_3d_write.cl:1547-1567: radius += step from rmin, break at the
// first sum that does not exceed the best).  T is the sample source (a
// staged tile, whose fill is the caller's, followed by a barrier, or the
// texels in device memory).  The Hessian reads TH: the same tile through a
// run-time-stride view, or the F32 planes in the mixed mode.
template <bool NORM, int BX, int BY, int BZ, bool FUSED = false, class A, class AH>
__device__ __forceinline__ void tdf_brick(const A& T, const AH& TH, Dims d, int bx0, int by0, int bz0, float rmin, float rmax,
                                          float rstep, const CircleTables& tab, float* __restrict__ tdf,
                                          float* __restrict__ rad) {
    constexpr int NT = BX * BY * BZ, NW = NT / 64;
    __shared__ float st_e[6][NT];
    __shared__ float st_m[NT];
    __shared__ int st_v[NT];
    __shared__ int wcount[NW];
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    // the first radius (rmin) at the thread's own voxel
    int v = tid;  // voxel of the brick this thread works on
    float e2[3], e3[3], best = 0.0f;
    bool act = false;
    {
        const int x = bx0 + threadIdx.x, y = by0 + threadIdx.y, z = bz0 + threadIdx.z;
        if (x < d.X && y < d.Y && z < d.Z) {
            float H[3][3], V[3][3], ev[3];
            hessian_tex<NORM>(TH, d, x, y, z, H);
            if constexpr (FUSED) eigen3_fp64_fallback(H, V, ev);
            else eigen3(H, V, ev);
            e2[0] = V[0][1]; e2[1] = V[1][1]; e2[2] = V[2][1];
            e3[0] = V[0][2]; e3[1] = V[1][2]; e3[2] = V[2][2];
            float r0 = 0.0f;
            if (rmin <= rmax) {
                const float sum = circle_sum<FUSED>(T, d, (float)x, (float)y, (float)z, rmin, e2, e3, tab);
                if (sum > 0.0f) {
                    best = sum;
                    r0 = rmin;
                    act = true;
                }
            }
            if (!act) {
                const i64 i = d.at(x, y, z);
                tdf[i] = best;
                rad[i] = r0;
            }
        }
    }
    const int lane = tid & 63, wid = tid >> 6;
    float rprev = rmin;
    for (float radius = rmin + rstep; ; radius += rstep) {
        // voxel of this thread's state -> coordinates
        auto coords = [&](int vv, int& x, int& y, int& z) {
            x = bx0 + vv % BX;
            y = by0 + (vv / BX) % BY;
            z = bz0 + vv / (BX * BY);
        };
        if (!(radius <= rmax)) {
            if (act) {
                int x, y, z;
                coords(v, x, y, z);
                const i64 i = d.at(x, y, z);
                tdf[i] = best;
                rad[i] = rprev;
            }
            break;
        }
        // compact the survivors: slot = rank among active threads
        const unsigned long long bal = __ballot(act);
        if (lane == 0) wcount[wid] = __popcll(bal);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const int cnt = wcount[w];
            off += w < wid ? cnt : 0;
            total += cnt;
        }
        if (act) {
            const int slot = off + __popcll(bal & ((1ull << lane) - 1ull));
            st_e[0][slot] = e2[0]; st_e[1][slot] = e2[1]; st_e[2][slot] = e2[2];
            st_e[3][slot] = e3[0]; st_e[4][slot] = e3[1]; st_e[5][slot] = e3[2];
            st_m[slot] = best;
            st_v[slot] = v;
        }
        __syncthreads();
        if (total == 0) break;
        act = tid < total;
        if (act) {
            e2[0] = st_e[0][tid]; e2[1] = st_e[1][tid]; e2[2] = st_e[2][tid];
            e3[0] = st_e[3][tid]; e3[1] = st_e[4][tid]; e3[2] = st_e[5][tid];
            best = st_m[tid];
            v = st_v[tid];
        }
        __syncthreads();  // the state area is rewritten next round
        if (act) {
            int x, y, z;
            coords(v, x, y, z);
            const float sum = circle_sum<FUSED>(T, d, (float)x, (float)y, (float)z, radius, e2, e3, tab);
            if (sum > best) {
                best = sum;
            } else {
                const i64 i = d.at(x, y, z);
                tdf[i] = best;
                rad[i] = rprev;
                act = false;
            }
        }
        rprev = radius;
    }
}

// Packed texels (8 bytes per cell for the 16-bit storage types).
template <class S, bool NORM, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX* BY* BZ) k_tdf_lds(const typename Texel<S>::P* __restrict__ tex, float inv,
                                                        Dims d, float rmin, float rmax, float rstep, CircleTables tab,
                                                        float* __restrict__ tdf, float* __restrict__ rad, int h) {
    constexpr int NT = BX * BY * BZ;
    typedef typename Texel<S>::P P;
    extern __shared__ float4 lds_f4[];
    P* tile = reinterpret_cast<P*>(lds_f4);
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int bx0 = tbx * BX, by0 = tby * BY, bz0 = tbz * BZ;
    LdsTex<S> T{tile, inv, bx0 - h, by0 - h, bz0 - h, BX + 2 * h, BY + 2 * h, BZ + 2 * h};
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    const int cells = T.tx * T.ty * T.tz;
    for (int c = tid; c < cells; c += NT) {
        const int lx = c % T.tx, rr = c / T.tx, ly = rr % T.ty, lz = rr / T.ty;
        const int gx = clampi(T.ox + lx, 0, d.X - 1), gy = clampi(T.oy + ly, 0, d.Y - 1);
        const int gz = clampi(T.oz + lz, 0, d.Z - 1);
        tile[c] = tex[d.at(gx, gy, gz)];
    }
    __syncthreads();
    tdf_brick<NORM, BX, BY, BZ>(T, T, d, bx0, by0, bz0, rmin, rmax, rstep, tab, tdf, rad);
}

// Mixed mode: the samples from a tile of 16-bit texels packed from an F32
// field, the Hessian from the F32 planes themselves (device memory), so the
// eigenvectors see the field at full precision.
template <class S, bool NORM, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX* BY* BZ)
    k_tdf_lds_mixed(const typename Texel<S>::P* __restrict__ tex, float inv, PlanesF32 H, Dims d, float rmin,
                    float rmax, float rstep, CircleTables tab, float* __restrict__ tdf, float* __restrict__ rad, int h) {
    constexpr int NT = BX * BY * BZ;
    typedef typename Texel<S>::P P;
    extern __shared__ float4 lds_f4[];
    P* tile = reinterpret_cast<P*>(lds_f4);
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int bx0 = tbx * BX, by0 = tby * BY, bz0 = tbz * BZ;
    LdsTex<S> T{tile, inv, bx0 - h, by0 - h, bz0 - h, BX + 2 * h, BY + 2 * h, BZ + 2 * h};
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    const int cells = T.tx * T.ty * T.tz;
    for (int c = tid; c < cells; c += NT) {
        const int lx = c % T.tx, rr = c / T.tx, ly = rr % T.ty, lz = rr / T.ty;
        const int gx = clampi(T.ox + lx, 0, d.X - 1), gy = clampi(T.oy + ly, 0, d.Y - 1);
        const int gz = clampi(T.oz + lz, 0, d.Z - 1);
        tile[c] = tex[d.at(gx, gy, gz)];
    }
    __syncthreads();
    tdf_brick<NORM, BX, BY, BZ, true>(T, H, d, bx0, by0, bz0, rmin, rmax, rstep, tab, tdf, rad);
}

// Texels gathered from device memory (halos too wide for a workgroup's LDS
// tile), with the radius search compacted round by round as in the tiled
// kernels (tdf_brick), so a wave runs as many radii as its voxels need and
// not as many as its slowest voxel: only the compaction state is in LDS.
// MIXED: the precision form's (Hessian from the F32 planes, fused samples).
template <class S, bool NORM, bool MIXED, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX* BY* BZ)
    k_tdf_gather(Tex<S> T, PlanesF32 H, Dims d, float rmin, float rmax, float rstep, CircleTables tab,
                 float* __restrict__ tdf, float* __restrict__ rad) {
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    if constexpr (MIXED)
        tdf_brick<NORM, BX, BY, BZ, true>(T, H, d, tbx * BX, tby * BY, tbz * BZ, rmin, rmax, rstep, tab, tdf, rad);
    else
        tdf_brick<NORM, BX, BY, BZ>(T, T, d, tbx * BX, tby * BY, tbz * BZ, rmin, rmax, rstep, tab, tdf, rad);
}
constexpr int GBX = 16, GBY = 8, GBZ = 4;  // k_tdf_gather's bricks

// Stage the clamp-to-edge texels of a TX x TY x TZ tile at (ox, oy, oz) in
// LDS, then a barrier: every load is issued before the first LDS store, so
// one memory latency covers them all (the cell's coordinates come from
// divisions by constants).
template <int NT, int TX, int TY, int TZ, class P>
__device__ __forceinline__ void stage_tile(P* tile, const P* __restrict__ tex, const Dims& d, int ox, int oy, int oz,
                                           int tid) {
    constexpr int CELLS = TX * TY * TZ, NIT = (CELLS + NT - 1) / NT;
    const i64 plane = (i64)d.X * d.Y;
    P v[NIT];
#pragma unroll
    for (int k = 0; k < NIT; k++) {
        const int c = tid + NT * k;
        if (CELLS % NT == 0 || c < CELLS) {
            const int lx = c % TX, rr = c / TX, ly = rr % TY, lz = rr / TY;
            const int gx = clampi(ox + lx, 0, d.X - 1), gy = clampi(oy + ly, 0, d.Y - 1);
            const int gz = clampi(oz + lz, 0, d.Z - 1);
            v[k] = tex[plane * gz + ((i64)gy * d.X + gx)];
        }
    }
#pragma unroll
    for (int k = 0; k < NIT; k++) {
        const int c = tid + NT * k;
        if (CELLS % NT == 0 || c < CELLS) tile[c] = v[k];
    }
    __syncthreads();
}

// The same with the halo HC a compile-time constant: the tile's strides are
// immediates, so a sample's cell costs a few float operations and its eight
// taps are reads at constant offsets from one address.
template <class S, bool NORM, int BX, int BY, int BZ, int HC>
__global__ void __launch_bounds__(BX* BY* BZ)
    k_tdf_lds_mixed_h(const typename Texel<S>::P* __restrict__ tex, float inv, PlanesF32 H, Dims d, float rmin,
                      float rmax, float rstep, CircleTables tab, float* __restrict__ tdf, float* __restrict__ rad) {
    constexpr int NT = BX * BY * BZ, TX = BX + 2 * HC, TY = BY + 2 * HC, TZ = BZ + 2 * HC;
    typedef typename Texel<S>::P P;
    extern __shared__ float4 lds_f4[];
    P* tile = reinterpret_cast<P*>(lds_f4);
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int bx0 = tbx * BX, by0 = tby * BY, bz0 = tbz * BZ;
    const int ox = bx0 - HC, oy = by0 - HC, oz = bz0 - HC;
    const int lo = (int)(size_t)(__attribute__((address_space(3))) P*)tile;  // the tile's LDS address
    const LdsTexC<S, TX, TY, TZ> T{inv, (float)(lo - ((ox + oy * TX) + oz * (TX * TY)) * (int)sizeof(P)), (float)lo};
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    stage_tile<NT, TX, TY, TZ>(tile, tex, d, ox, oy, oz, tid);
    tdf_brick<NORM, BX, BY, BZ, true>(T, H, d, bx0, by0, bz0, rmin, rmax, rstep, tab, tdf, rad);
}

// The reference-order TDF on 8-byte texels with the halo HC a compile-time
// constant: the samples address the tile as k_tdf_lds_mixed_h does
// (trilinear_tex on an LdsTexC, the same arithmetic as on an LdsTex), the
// Hessian reads the same tile through an LdsTex.
template <class S, bool NORM, int BX, int BY, int BZ, int HC>
__global__ void __launch_bounds__(BX* BY* BZ)
    k_tdf_lds_h(const typename Texel<S>::P* __restrict__ tex, float inv, Dims d, float rmin, float rmax, float rstep,
                CircleTables tab, float* __restrict__ tdf, float* __restrict__ rad) {
    constexpr int NT = BX * BY * BZ, TX = BX + 2 * HC, TY = BY + 2 * HC, TZ = BZ + 2 * HC;
    typedef typename Texel<S>::P P;
    extern __shared__ float4 lds_f4[];
    P* tile = reinterpret_cast<P*>(lds_f4);
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int bx0 = tbx * BX, by0 = tby * BY, bz0 = tbz * BZ;
    const int ox = bx0 - HC, oy = by0 - HC, oz = bz0 - HC;
    const int lo = (int)(size_t)(__attribute__((address_space(3))) P*)tile;  // the tile's LDS address
    const LdsTexC<S, TX, TY, TZ> T{inv, (float)(lo - ((ox + oy * TX) + oz * (TX * TY)) * (int)sizeof(P)), (float)lo};
    const LdsTex<S> TH{tile, inv, ox, oy, oz, TX, TY, TZ};
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    stage_tile<NT, TX, TY, TZ>(tile, tex, d, ox, oy, oz, tid);
    tdf_brick<NORM, BX, BY, BZ>(T, TH, d, bx0, by0, bz0, rmin, rmax, rstep, tab, tdf, rad);
}

// F32 planes -> 16-bit texels holding value * scale (x, y, z; w = 0, the
// mixed mode's Hessian reads w from its F32 plane).
// (round to nearest even)
template <class S>
__device__ __forceinline__ unsigned bits16(float v);
template <>
__device__ __forceinline__ unsigned bits16<F16>(float v) {
    return __half_as_ushort(__float2half_rn(v));
}
template <>
__device__ __forceinline__ unsigned bits16<BF16>(float v) {
    return __bfloat16_as_ushort(__float2bfloat16(v));
}
template <class S>
__global__ void k_pack_f32_to_16(const float* __restrict__ x, const float* __restrict__ y,
                                 const float* __restrict__ z, float scale, uint2* __restrict__ out, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        out[i] = make_uint2(bits16<S>(x[i] * scale) | (bits16<S>(y[i] * scale) << 16), bits16<S>(z[i] * scale));
}

// F32 fields straight from their planes: the tile holds x, y and z as three
// float planes (12 bytes per cell); w, read only by the normalised Hessian,
// comes from its plane in device memory.
template <bool NORM, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX* BY* BZ)
    k_tdf_lds_f32(const float* __restrict__ fx, const float* __restrict__ fy, const float* __restrict__ fz,
                  const float* __restrict__ fw, Dims d, float rmin, float rmax, float rstep, CircleTables tab,
                  float* __restrict__ tdf, float* __restrict__ rad, int h) {
    constexpr int NT = BX * BY * BZ;
    extern __shared__ float4 lds_f4[];
    float* tile = reinterpret_cast<float*>(lds_f4);
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int bx0 = tbx * BX, by0 = tby * BY, bz0 = tbz * BZ;
    const int tx = BX + 2 * h, ty = BY + 2 * h, tz = BZ + 2 * h, cells = tx * ty * tz;
    LdsSoaTex T{tile, fw, cells, bx0 - h, by0 - h, bz0 - h, tx, ty, tz};
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    for (int c = tid; c < cells; c += NT) {
        const int lx = c % tx, rr = c / tx, ly = rr % ty, lz = rr / ty;
        const int gx = clampi(T.ox + lx, 0, d.X - 1), gy = clampi(T.oy + ly, 0, d.Y - 1);
        const int gz = clampi(T.oz + lz, 0, d.Z - 1);
        const i64 g = d.at(gx, gy, gz);
        tile[c] = fx[g];
        tile[cells + c] = fy[g];
        tile[2 * cells + c] = fz[g];
    }
    __syncthreads();
    tdf_brick<NORM, BX, BY, BZ>(T, T, d, bx0, by0, bz0, rmin, rmax, rstep, tab, tdf, rad);
}

// The same with the halo HC a compile-time constant: the samples address
// the tile through an LdsSoaTexC, the Hessian through an LdsSoaTex (the same
// arithmetic as k_tdf_lds_f32); every load of the tile is issued before the
// first LDS store.
template <bool NORM, int BX, int BY, int BZ, int HC>
__global__ void __launch_bounds__(BX* BY* BZ)
    k_tdf_lds_f32_h(const float* __restrict__ fx, const float* __restrict__ fy, const float* __restrict__ fz,
                    const float* __restrict__ fw, Dims d, float rmin, float rmax, float rstep, CircleTables tab,
                    float* __restrict__ tdf, float* __restrict__ rad) {
    constexpr int NT = BX * BY * BZ, TX = BX + 2 * HC, TY = BY + 2 * HC, TZ = BZ + 2 * HC, CELLS = TX * TY * TZ;
    constexpr int NIT = (CELLS + NT - 1) / NT;
    extern __shared__ float4 lds_f4[];
    float* tile = reinterpret_cast<float*>(lds_f4);
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int bx0 = tbx * BX, by0 = tby * BY, bz0 = tbz * BZ;
    const int ox = bx0 - HC, oy = by0 - HC, oz = bz0 - HC;
    const int lo = (int)(size_t)(__attribute__((address_space(3))) float*)tile;  // the tile's LDS address
    const LdsSoaTexC<TX, TY, TZ> T{(float)(lo - ((ox + oy * TX) + oz * (TX * TY)) * (int)sizeof(float)), (float)lo};
    const LdsSoaTex TH{tile, fw, CELLS, ox, oy, oz, TX, TY, TZ};
    const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
    float vx[NIT], vy[NIT], vz[NIT];
#pragma unroll
    for (int k = 0; k < NIT; k++) {
        const int c = tid + NT * k;
        if (CELLS % NT == 0 || c < CELLS) {
            const int lx = c % TX, rr = c / TX, ly = rr % TY, lz = rr / TY;
            const int gx = clampi(ox + lx, 0, d.X - 1), gy = clampi(oy + ly, 0, d.Y - 1);
            const int gz = clampi(oz + lz, 0, d.Z - 1);
            const i64 g = d.at(gx, gy, gz);
            vx[k] = fx[g];
            vy[k] = fy[g];
            vz[k] = fz[g];
        }
    }
#pragma unroll
    for (int k = 0; k < NIT; k++) {
        const int c = tid + NT * k;
        if (CELLS % NT == 0 || c < CELLS) {
            tile[c] = vx[k];
            tile[CELLS + c] = vy[k];
            tile[2 * CELLS + c] = vz[k];
        }
    }
    __syncthreads();
    tdf_brick<NORM, BX, BY, BZ>(T, TH, d, bx0, by0, bz0, rmin, rmax, rstep, tab, tdf, rad);
}

static unsigned grid1d(i64 n) {
    i64 g = (n + 255) / 256;
    return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

// LDS of the current device, queried once per device: the most a workgroup
// may allocate (static + dynamic) and the size of a CU's LDS (both 160 KiB
// on gfx950).  64 KiB when the query fails.
static size_t device_lds(bool per_cu) {
    static std::atomic<int> cache[2][64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
    int v = cache[per_cu][dev].load(std::memory_order_relaxed);
    if (v <= 0) {
        const hipDeviceAttribute_t a =
            per_cu ? hipDeviceAttributeMaxSharedMemoryPerMultiprocessor : hipDeviceAttributeMaxSharedMemoryPerBlock;
        if (hipDeviceGetAttribute(&v, a, dev) != hipSuccess || v <= 0) v = 65536;
        cache[per_cu][dev].store(v, std::memory_order_relaxed);
    }
    return (size_t)v;
}
static size_t lds_per_block() { return device_lds(false); }

// Waves a CU holds of a TDF kernel whose workgroups of nt threads take
// `bytes` of LDS each (the TDF kernels' registers allow six or more waves
// per SIMD, so their LDS sets the count); 0 when a workgroup does not fit.
static int tdf_waves(size_t bytes, int nt) {
    if (bytes > lds_per_block()) return 0;
    return (int)(device_lds(true) / bytes) * (nt / 64);
}

// Whether the compile-time-stride kernels' float cell addresses are exact
// on this volume: the tile's origin term o8 and every partial sum of a
// sample's address are integers bounded by the workgroup's LDS plus 8 bytes
// times the cell offset of the volume's far corner in the tile's strides
// (tx, tx ty), and a float holds every integer below 2^24.
static bool tdf_addresses_exact(Dims d, int tx, int ty, int hc, size_t cap) {
    const double x = d.X + 2.0 * hc + 2.0, y = d.Y + 2.0 * hc + 2.0, z = d.Z + 2.0 * hc + 2.0;
    return 8.0 * (x + (double)tx * y + (double)tx * ty * z) + (double)cap < 16777216.0;
}

void pack_texels_f32(const tsf_field* field, float4* out, i64 n, hipStream_t stream) {
    hipLaunchKernelGGL(k_pack_f32, grid1d(n), 256, 0, stream, (const float*)field->x, (const float*)field->y,
                       (const float*)field->z, (const float*)field->w, out, n);
}

// The halo of a brick's LDS tile: floor(r_max + 0.5) + 1 cells keep both
// corners of every sample inside it.  Tiles that do not fit a workgroup's
// LDS gather from device memory instead.
static int tdf_halo(float rmin, float rmax) {
    const float r = fmaxf(fabsf(rmin), fabsf(rmax));
    return (int)floorf(r + 0.5f) + 1;
}

// F32 through the three-plane tile: bricks of 16 x F3Y x 4 voxels, F3Y 8 or
// 16, whichever holds more waves per CU (the pass wants every wave it can
// get; 16 x 16 x 4 spends less LDS on a wide halo per voxel: measured at
// 512^3, 43.3 ms at radii 0.5-3 on 16 x 8 x 4 and 45.9 ms at 2.5-4 on
// 16 x 16 x 4, against 51.6 and 58.0 ms on 8 x 8 x 4 bricks).  F3Y and
// the tile's size; F3Y = 0 when neither fits LDS.
static int tdf_f32_brick(float rmin, float rmax, size_t& lds) {
    const int h = tdf_halo(rmin, rmax);
    int best = 0, waves = 0;
    for (int by : {8, 16}) {
        const size_t tile = (size_t)(16 + 2 * h) * (by + 2 * h) * (4 + 2 * h) * 3 * sizeof(float);
        const size_t state = (size_t)16 * by * 4 * 8 * sizeof(float) + 64;  // tdf_brick's static LDS
        // from halo 7 on, k_tdf_gather on F32 texels is faster than the tile
        // (256^3, radii 2.5-6: 11.2 against 12.5 ms; halo 6: 10.8 against 10.4)
        const int w = h <= 6 ? tdf_waves(tile + state, 16 * by * 4) : 0;
        if (w > waves) {
            waves = w;
            best = by;
            lds = tile;
        }
    }
    return best;
}
static bool tdf_f32_lds(float rmin, float rmax, size_t& lds) { return tdf_f32_brick(rmin, rmax, lds) != 0; }

template <bool NORM, int F3Y>
static void launch_tdf_f32_brick(const tsf_field* f, Dims d, float rmin, float rmax, float rstep, float* tdf,
                                 float* radius, size_t lds, hipStream_t stream) {
    static const CircleTables tab = make_circle_tables();
    constexpr int F3X = 16, F3Z = 4;
    const int h = tdf_halo(rmin, rmax);
    dim3 grid((d.X + F3X - 1) / F3X, (d.Y + F3Y - 1) / F3Y, (d.Z + F3Z - 1) / F3Z), block(F3X, F3Y, F3Z);
    const float *x = (const float*)f->x, *y = (const float*)f->y, *z = (const float*)f->z;
    const float* w = NORM ? (const float*)f->w : nullptr;
    // halos 4 and 5 with compile-time strides (the same arithmetic)
    if ((h == 4 || h == 5) && tdf_addresses_exact(d, F3X + 2 * h, F3Y + 2 * h, h, lds_per_block())) {
        if (h == 4)
            hipLaunchKernelGGL((k_tdf_lds_f32_h<NORM, F3X, F3Y, F3Z, 4>), grid, block, lds, stream, x, y, z, w, d,
                               rmin, rmax, rstep, tab, tdf, radius);
        else
            hipLaunchKernelGGL((k_tdf_lds_f32_h<NORM, F3X, F3Y, F3Z, 5>), grid, block, lds, stream, x, y, z, w, d,
                               rmin, rmax, rstep, tab, tdf, radius);
        return;
    }
    hipLaunchKernelGGL((k_tdf_lds_f32<NORM, F3X, F3Y, F3Z>), grid, block, lds, stream, x, y, z, w, d, rmin, rmax,
                       rstep, tab, tdf, radius, h);
}
// lds: the tile's size from tdf_f32_brick for the same radii
template <bool NORM>
static void launch_tdf_f32(const tsf_field* f, Dims d, float rmin, float rmax, float rstep, float* tdf,
                           float* radius, size_t lds, hipStream_t stream) {
    size_t t = 0;
    if (tdf_f32_brick(rmin, rmax, t) == 16)
        launch_tdf_f32_brick<NORM, 16>(f, d, rmin, rmax, rstep, tdf, radius, lds, stream);
    else
        launch_tdf_f32_brick<NORM, 8>(f, d, rmin, rmax, rstep, tdf, radius, lds, stream);
}


template <class S, bool NORM>
static void launch_tdf(const void* tex, float inv, Dims d, float rmin, float rmax, float rstep, float* tdf,
                       float* radius, hipStream_t stream) {
    static const CircleTables tab = make_circle_tables();
    typedef typename Texel<S>::P P;
    // 16 x 8 x 4 bricks (measured at 512^3 with SNORM16 texels: 43.0 and
    // 47.0 ms at radii 0.5-3 and 2.5-4, against 53.5 and 51.4 ms on
    // 16 x 4 x 4 and 59.8 and 48.1 ms on 16 x 16 x 4)
    constexpr int BX = sizeof(P) == 16 ? 8 : 16, BY = 8, BZ = 4;
    const int h = tdf_halo(rmin, rmax);
    const size_t lds = (size_t)(BX + 2 * h) * (BY + 2 * h) * (BZ + 2 * h) * sizeof(P);
    const size_t state = (size_t)BX * BY * BZ * 8 * sizeof(float) + 64;  // tdf_brick's static LDS
    // the tile up to halo 5; wider, the tile holds a CU to 8-16 waves and
    // k_tdf_gather is faster (256^3, SNORM16, halos 6, 7 and 8: 9.2, 10.0 and
    // 10.2 ms against 11.8, 14.4 and 15.3 ms)
    if (sizeof(P) == 8 && h <= 5 && lds + state <= lds_per_block()) {
        dim3 grid((d.X + BX - 1) / BX, (d.Y + BY - 1) / BY, (d.Z + BZ - 1) / BZ), block(BX, BY, BZ);
        // halos 4 and 5 with compile-time strides (the same arithmetic)
        if constexpr (sizeof(P) == 8) {
            if ((h == 4 || h == 5) && tdf_addresses_exact(d, BX + 2 * h, BY + 2 * h, h, lds_per_block())) {
                if (h == 4)
                    hipLaunchKernelGGL((k_tdf_lds_h<S, NORM, 16, 8, 4, 4>), grid, block, lds, stream, (const P*)tex,
                                       inv, d, rmin, rmax, rstep, tab, tdf, radius);
                else
                    hipLaunchKernelGGL((k_tdf_lds_h<S, NORM, 16, 8, 4, 5>), grid, block, lds, stream, (const P*)tex,
                                       inv, d, rmin, rmax, rstep, tab, tdf, radius);
                return;
            }
        }
        hipLaunchKernelGGL((k_tdf_lds<S, NORM, BX, BY, BZ>), grid, block, lds, stream, (const P*)tex, inv, d, rmin,
                           rmax, rstep, tab, tdf, radius, h);
        return;
    }
    // halos too wide for LDS: texels gathered from device memory (measured at
    // 256^3, radii 2.5-9 and 2.5-15: 15.2 -> 11.4 ms for F32 texels and
    // 14.7 -> 10.3 ms for SNORM16 against a per-voxel kernel without the
    // compaction)
    const dim3 grid((d.X + GBX - 1) / GBX, (d.Y + GBY - 1) / GBY, (d.Z + GBZ - 1) / GBZ), block(GBX, GBY, GBZ);
    hipLaunchKernelGGL((k_tdf_gather<S, NORM, false, GBX, GBY, GBZ>), grid, block, 0, stream,
                       Tex<S>{(const P*)tex, inv}, PlanesF32{}, d, rmin, rmax, rstep, tab, tdf, radius);
}

// The compile-time-halo kernels on BX x BY x BZ bricks, when the float cell
// addresses are exact and the tile and the brick's state fit the
// workgroup's LDS; false when they do not run.
template <class S, bool NORM, int BX, int BY, int BZ, int HC>
static bool launch_tdf_mixed_h(const void* tex, float inv, const PlanesF32& H, Dims d, float rmin, float rmax,
                               float rstep, const CircleTables& tab, float* tdf, float* radius, hipStream_t stream) {
    typedef typename Texel<S>::P P;
    constexpr int TX = BX + 2 * HC, TY = BY + 2 * HC;
    const size_t lds = (size_t)TX * TY * (BZ + 2 * HC) * sizeof(P);
    const size_t state = (size_t)BX * BY * BZ * 8 * sizeof(float) + 64;  // tdf_brick's static LDS
    const size_t cap = lds_per_block();
    if (!tdf_addresses_exact(d, TX, TY, HC, cap) || lds + state > cap) return false;
    const dim3 grid((d.X + BX - 1) / BX, (d.Y + BY - 1) / BY, (d.Z + BZ - 1) / BZ), block(BX, BY, BZ);
    hipLaunchKernelGGL((k_tdf_lds_mixed_h<S, NORM, BX, BY, BZ, HC>), grid, block, lds, stream, (const P*)tex, inv, H,
                       d, rmin, rmax, rstep, tab, tdf, radius);
    return true;
}

template <class S, bool NORM>
static tsf_status launch_tdf_mixed(const tsf_field* f, const void* tex, float inv, Dims d, float rmin, float rmax,
                                   float rstep, float* tdf, float* radius, hipStream_t stream, const char* W) {
    static const CircleTables tab = make_circle_tables();
    typedef typename Texel<S>::P P;
    constexpr int BX = 16, BY = 8, BZ = 4;  // other halos: the 8-byte-texel kernel's bricks
    const int h = tdf_halo(rmin, rmax);
    const size_t lds = (size_t)(BX + 2 * h) * (BY + 2 * h) * (BZ + 2 * h) * sizeof(P);
    const size_t state = (size_t)BX * BY * BZ * 8 * sizeof(float) + 64;  // tdf_brick's static LDS
    const PlanesF32 H{(const float*)f->x, (const float*)f->y, (const float*)f->z, NORM ? (const float*)f->w : nullptr};
    if (h > 64 || lds + state > lds_per_block()) {
        (void)W;
        const dim3 grid((d.X + GBX - 1) / GBX, (d.Y + GBY - 1) / GBY, (d.Z + GBZ - 1) / GBZ), block(GBX, GBY, GBZ);
        hipLaunchKernelGGL((k_tdf_gather<S, NORM, true, GBX, GBY, GBZ>), grid, block, 0, stream,
                           Tex<S>{(const P*)tex, inv}, H, d, rmin, rmax, rstep, tab, tdf, radius);
        return TSF_OK;
    }
    // halos 4 and 5 (radii up to 3.5 and 4.5) with compile-time strides, on
    // 512-voxel bricks: the pass's speed follows the waves a CU holds (its
    // registers allow six per SIMD, its LDS sets the count), and a larger
    // brick spends less LDS on its halo per wave.  Halo 4 on 16 x 8 x 4 (tile 24 x 16 x 12, three
    // workgroups, 24 waves per CU); halo 5 on 8 x 8 x 8 (tile 18^3, two
    // workgroups, 16 waves).  Measured at 512^3: 25.9 -> 20.9 ms and
    // 26.4 -> 24.3 ms against 256-voxel bricks (16 x 4 x 4 and 8 x 2 x 16,
    // 16 and 12 waves); 16-column bricks read the Hessian's F32 rows in
    // longer runs (8 x 8 x 8 at halo 4: 25.2 ms).
    if (h == 4 && launch_tdf_mixed_h<S, NORM, 16, 8, 4, 4>(tex, inv, H, d, rmin, rmax, rstep, tab, tdf, radius, stream))
        return TSF_OK;
    if (h == 5 && launch_tdf_mixed_h<S, NORM, 8, 8, 8, 5>(tex, inv, H, d, rmin, rmax, rstep, tab, tdf, radius, stream))
        return TSF_OK;
    dim3 grid((d.X + BX - 1) / BX, (d.Y + BY - 1) / BY, (d.Z + BZ - 1) / BZ), block(BX, BY, BZ);
    hipLaunchKernelGGL((k_tdf_lds_mixed<S, NORM, BX, BY, BZ>), grid, block, lds, stream, (const P*)tex, inv, H, d,
                       rmin, rmax, rstep, tab, tdf, radius, h);
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_tdf_circle_mixed_workspace_size(tsf_dims dims) { return (size_t)D(dims).n() * 8; }

tsf_status tsf_tdf_circle_mixed(const tsf_field* field, int32_t sample_storage, float sample_scale, tsf_dims dims,
                                float rmin, float rmax, float rstep, float* tdf, float* radius, void* workspace,
                                size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_tdf_circle_mixed";
    const bool norm = !(rmax < 4.0f);
    TSF_TRY(check_field(field, norm, W));
    TSF_TRY(check_dims(dims, 1, W));
    if (field->storage != TSF_STORAGE_F32) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: field must be F32", W);
    if (sample_storage != TSF_STORAGE_F16 && sample_storage != TSF_STORAGE_BF16)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: sample storage must be F16 or BF16", W);
    int ex;
    if (!(sample_scale > 0.0f) || frexpf(sample_scale, &ex) != 0.5f)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: sample scale must be a positive power of two", W);
    if (!tdf || !radius) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL output", W);
    if (!(rstep > 0.0f) || !__builtin_isfinite(rmin) || !__builtin_isfinite(rmax))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: radius range must be finite with rstep > 0", W);
    if ((double)(rmax - rmin) / rstep > 100000.0)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: more than 1e5 radii", W);
    Dims d = D(dims);
    const i64 n = d.n();
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_tdf_circle_mixed_workspace_size(dims), W));
    const float inv = 1.0f / sample_scale;
    if (sample_storage == TSF_STORAGE_F16) {
        hipLaunchKernelGGL(k_pack_f32_to_16<F16>, grid1d(n), 256, 0, stream, (const float*)field->x,
                           (const float*)field->y, (const float*)field->z, sample_scale, (uint2*)sc.ptr, n);
        const tsf_status st =
            norm ? launch_tdf_mixed<F16, true>(field, sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream, W)
                 : launch_tdf_mixed<F16, false>(field, sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream, W);
        if (st != TSF_OK) return st;
    } else {
        hipLaunchKernelGGL(k_pack_f32_to_16<BF16>, grid1d(n), 256, 0, stream, (const float*)field->x,
                           (const float*)field->y, (const float*)field->z, sample_scale, (uint2*)sc.ptr, n);
        const tsf_status st =
            norm ? launch_tdf_mixed<BF16, true>(field, sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream, W)
                 : launch_tdf_mixed<BF16, false>(field, sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream, W);
        if (st != TSF_OK) return st;
    }
    TSF_TRY(check_launch(W));
    if (sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

size_t tsf_tdf_circle_workspace_size(tsf_dims dims, int32_t storage) {
    if (storage < 0 || storage > 3) return 0;
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
    // F32 fields whose tile fits LDS are read from their planes: no packed
    // texels, no workspace
    size_t lds_f32 = 0;
    const bool planes = field->storage == TSF_STORAGE_F32 && tdf_f32_lds(rmin, rmax, lds_f32);
    Scratch sc;
    if (!planes)
        TSF_TRY(sc.get(workspace, workspace_bytes, tsf_tdf_circle_workspace_size(dims, field->storage), W));
    // texel units: F16/BF16 hold value * scale, SN16 texels are read as
    // value * 2^15
    const float inv = field->storage == TSF_STORAGE_SN16 ? 1.0f / 32768.0f : 1.0f / field->scale;
    switch (field->storage) {
        case TSF_STORAGE_F32:
            if (planes) {
                if (norm) launch_tdf_f32<true>(field, d, rmin, rmax, rstep, tdf, radius, lds_f32, stream);
                else launch_tdf_f32<false>(field, d, rmin, rmax, rstep, tdf, radius, lds_f32, stream);
                break;
            }
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
                               field->storage == TSF_STORAGE_SN16, (uint2*)sc.ptr, n);
            if (norm) launch_tdf<F16, true>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            else launch_tdf<F16, false>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            break;
        }
        case TSF_STORAGE_SN16: {
            const uint16_t one = 32767;
            hipLaunchKernelGGL(k_pack_16, grid1d(n), 256, 0, stream, (const uint16_t*)field->x,
                               (const uint16_t*)field->y, (const uint16_t*)field->z, (const uint16_t*)field->w, one,
                               field->storage == TSF_STORAGE_SN16, (uint2*)sc.ptr, n);
            if (norm) launch_tdf<SN16, true>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            else launch_tdf<SN16, false>(sc.ptr, inv, d, rmin, rmax, rstep, tdf, radius, stream);
            break;
        }
        default: {
            const uint16_t one = 0x3f80;
            hipLaunchKernelGGL(k_pack_16, grid1d(n), 256, 0, stream, (const uint16_t*)field->x,
                               (const uint16_t*)field->y, (const uint16_t*)field->z, (const uint16_t*)field->w, one,
                               field->storage == TSF_STORAGE_SN16, (uint2*)sc.ptr, n);
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

size_t tsf_tdf_spline_workspace_size(tsf_dims dims, int32_t storage) {
    return tsf_tdf_circle_workspace_size(dims, storage);
}

tsf_status tsf_tdf_spline(const tsf_field* field, tsf_dims dims, float rmin, float rmax, float rstep,
                          float min_average_mag, float* tdf, float* radius, void* workspace, size_t workspace_bytes,
                          hipStream_t stream) {
    const char* W = "tsf_tdf_spline";
    TSF_TRY(check_field(field, true, W));
    TSF_TRY(check_dims(dims, 1, W));
    if (!tdf || !radius) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL output", W);
    if (!(rstep > 0.0f) || !__builtin_isfinite(rmin) || !__builtin_isfinite(rmax) ||
        !__builtin_isfinite(min_average_mag))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: radius range must be finite with rstep > 0", W);
    if ((double)(rmax - rmin) / rstep > 100000.0)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: more than 1e5 radii", W);
    const Dims d = D(dims);
    const i64 n = d.n();
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_tdf_spline_workspace_size(dims, field->storage), W));
    static const SplineTables tab = make_spline_tables();
    const float inv = field->storage == TSF_STORAGE_SN16 ? 1.0f / 32768.0f : 1.0f / field->scale;
    const dim3 grid((d.X + CBX - 1) / CBX, (d.Y + CBY - 1) / CBY, (d.Z + CBZ - 1) / CBZ), block(CBX, CBY, CBZ);
    if (field->storage == TSF_STORAGE_F32) {
        hipLaunchKernelGGL(k_pack_f32, grid1d(n), 256, 0, stream, (const float*)field->x, (const float*)field->y,
                           (const float*)field->z, (const float*)field->w, (float4*)sc.ptr, n);
        Tex<F32> T{(const float4*)sc.ptr, inv};
        hipLaunchKernelGGL(k_tdf_spline<F32>, grid, block, 0, stream, T, d, rmin, rmax, rstep, min_average_mag, tab,
                           tdf, radius);
    } else {
        const bool sn16 = field->storage == TSF_STORAGE_SN16;
        hipLaunchKernelGGL(k_pack_16, grid1d(n), 256, 0, stream, (const uint16_t*)field->x, (const uint16_t*)field->y,
                           (const uint16_t*)field->z, (const uint16_t*)field->w, (uint16_t)0, sn16, (uint2*)sc.ptr, n);
        const uint2* p = (const uint2*)sc.ptr;
        if (sn16)
            hipLaunchKernelGGL(k_tdf_spline<SN16>, grid, block, 0, stream, Tex<SN16>{p, inv}, d, rmin, rmax, rstep,
                               min_average_mag, tab, tdf, radius);
        else if (field->storage == TSF_STORAGE_F16)
            hipLaunchKernelGGL(k_tdf_spline<F16>, grid, block, 0, stream, Tex<F16>{p, inv}, d, rmin, rmax, rstep,
                               min_average_mag, tab, tdf, radius);
        else
            hipLaunchKernelGGL(k_tdf_spline<BF16>, grid, block, 0, stream, Tex<BF16>{p, inv}, d, rmin, rmax, rstep,
                               min_average_mag, tab, tdf, radius);
    }
    TSF_TRY(check_launch(W));
    if (sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return TSF_OK;
}

}  // extern "C"
