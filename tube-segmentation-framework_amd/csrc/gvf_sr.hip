// Gradient vector flow, precision form: the iterate streams through HBM as
// 16-bit planes (F16 holding value * scale, or BF16) and is rounded to them
// stochastically; every step computes in float.
//
// Why stochastic rounding: near convergence a step changes v by less than
// half a 16-bit ulp in most voxels, so round-to-nearest returns the stored
// value unchanged and the iteration stalls short of the fixed point.  A
// rounding that goes up with probability equal to the discarded fraction
// keeps every update in expectation.  The random bits come from a hash of
// the voxel index and the launch, so results are reproducible.
//
// k_gvf_sr2 runs two Jacobi steps per launch (the middle iterate stays in
// float), so the 16-bit v and f planes cross HBM once per two iterations:
// 18 bytes per voxel per launch.  A workgroup of 16 waves covers a tile of
// 64 x 32 columns: lane li of half hf of wave w owns columns x0 + 2 li,
// +1 of row y0 + 2w + hf, so every load is one 32-bit word (two 16-bit
// values) per lane, a whole cache line per half wave.  Two voxels per lane
// keep the registers low enough for four waves per SIMD.  The workgroup marches a chunk of planes in z:
//   - x neighbours come from the lane's other column or, across lanes, by
//     DPP row shifts (wave_shr / wave_shl) folded into the adds;
//   - y neighbours through LDS (each wave publishes its two rows, one
//     barrier per plane);
//   - z neighbours from registers: each stage keeps the previous plane of
//     its input and a partial update A(p) = v + mu (v(p-1) + sum_xy - 6 v)
//     - (v - f) |f|^2, completed with mu v(p+1) when that plane arrives, so
//     no three-plane ring is needed.
// Stage 1 (v1 = step(v0)) is valid on the tile less one column / row at
// each side, stage 2 on the tile less two: the output tile is 60 x 28.  (A
// 64-column tile, the two halves of a wave stacked in y, wastes less than a
// 128-column one on 512-wide volumes: 9 x 64 columns computed for 512
// instead of 5 x 128.)
// The reference's mirror rule (kernels_no_3d_write.cl:1757-1760: the value
// at a face voxel is the update computed at its mirror, 0 -> 2 and N-1 ->
// N-3) is applied to v1 in x (columns 0 and X-1 take columns 2 and X-3
// across lanes), in y (the LDS rows of the mirrors) and in z (scalar
// multipliers of the z terms on planes 1 and Z-2), and to the output in x
// (DPP across lanes); its y and z faces are copied from their mirrors by
// k_gvf_sr_faces.
// The last launch of a run can write F32 planes and the magnitude plane w
// instead (the final iterate at full precision for the Hessian).
#include <algorithm>
#include <type_traits>

#include "tsf_common.hpp"

namespace tsf {

constexpr int SRW = 16;                          // waves per workgroup, two rows each
constexpr int SRTY = 2 * SRW;                    // tile rows
constexpr int SRTX = 64;                         // tile columns: 32 lanes x 2
constexpr int SROX = SRTX - 4, SROY = SRTY - 4;  // output tile
constexpr int SRZ = 32;                          // output planes per z chunk
constexpr unsigned SR_OOB = 0x80000000u;         // buffer offset past every resource: the access is dropped

// The two 16-bit numbers of a 32-bit word as floats in units (the stored
// numbers: value * scale for F16, the value for BF16; for SNORM16 f, read
// as the integer times k = unit / 32767, at least -unit), one number alone,
// and the stochastic rounding of two floats (in units) to 16 bits with 10
// random bits each.
template <class S>
struct H16;
template <>
struct H16<F16> {
    static __device__ __forceinline__ float one(unsigned short c, float) { return __half2float(__ushort_as_half(c)); }
    static __device__ __forceinline__ float lo(unsigned u, float) {
        return __half2float(__ushort_as_half((unsigned short)(u & 0xffffu)));
    }
    static __device__ __forceinline__ float hi(unsigned u, float) {
        return __half2float(__ushort_as_half((unsigned short)(u >> 16)));
    }
    // Two values with random numbers ra, rb in [0, 1024), scaled to the 13
    // bits below the half mantissa and added to the floats' bits, converted
    // to halves rounding toward zero (one v_cvt_pkrtz): each magnitude rounds
    // up with probability within 2^-10 of (discarded bits) / 2^13.
    static __device__ __forceinline__ unsigned sr2(float a, unsigned ra, float b, unsigned rb) {
        const auto h = __builtin_amdgcn_cvt_pkrtz(__uint_as_float(__float_as_uint(a) + (ra << 3)),
                                                  __uint_as_float(__float_as_uint(b) + (rb << 3)));
        return __builtin_bit_cast(unsigned, h);
    }
};
template <>
struct H16<BF16> {
    static __device__ __forceinline__ float one(unsigned short c, float) { return __uint_as_float((unsigned)c << 16); }
    static __device__ __forceinline__ float lo(unsigned u, float) { return __uint_as_float(u << 16); }
    static __device__ __forceinline__ float hi(unsigned u, float) { return __uint_as_float(u & 0xffff0000u); }
    // the high halves of the two sums (bf16 truncation), packed by v_perm
    static __device__ __forceinline__ unsigned sr2(float a, unsigned ra, float b, unsigned rb) {
        return __builtin_amdgcn_perm(__float_as_uint(b) + (rb << 6), __float_as_uint(a) + (ra << 6), 0x07060302u);
    }
};
template <>
struct H16<SN16> {  // f only (the reference's 16-bit vector format)
    static __device__ __forceinline__ float one(unsigned short c, float k) {
        return fmaxf((float)(int)(short)c, -32767.0f) * k;
    }
    static __device__ __forceinline__ float lo(unsigned u, float k) {
        return fmaxf((float)(int)(short)(u & 0xffffu), -32767.0f) * k;
    }
    static __device__ __forceinline__ float hi(unsigned u, float k) { return fmaxf((float)((int)u >> 16), -32767.0f) * k; }
};

// Hash of the voxel index and the launch seed, after the lowbias32 mixer
// (shift-xor-multiply rounds) with 24-bit multiplies, which run at full rate
// on the vector ALU; its bits 0-9, 10-19, 20-29 are the random numbers of
// components x, y, z.
__device__ __forceinline__ unsigned sr_hash(unsigned i, unsigned seed) {
    unsigned h = i ^ seed;
    h ^= h >> 16;
    h = __umul24(h, 0x7feb35u);
    h ^= h >> 15;
    h = __umul24(h, 0x846ca7u);
    h ^= h >> 16;
    return h;
}

// lane i gets lane i-1's value (lane 0 gets 0), lane i gets lane i+1's
// (lane 63 gets 0): DPP wave_shr:1 / wave_shl:1
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

struct SrArgs {
    const void* f[3];  // f planes (16-bit)
    const void* v[3];  // input iterate planes (16-bit: the iterate's storage, or f's on the first launch)
    void* o[4];        // output planes: 16-bit x, y, z, or F32 x, y, z, w
    Dims d;
    float mu, m2s, inv;  // m2s = 1 / unit^2, inv = 1 / unit (unit = the 16-bit scale)
    float c0;            // 1 - 6 mu
    float k;             // SNORM16 integer -> units (unit / 32767)
    unsigned seed;
    int zspan;  // planes covered by the load resources of a chunk
};

// The update v' = v (1 - 6 mu - |f|^2) + f |f|^2 + mu (sum of the six
// neighbours), in units, as a partial A = fma(mu, s + zmul z-1, fma(v, c1, g))
// with c1 = (1 - 6 mu) - |f|^2 and g = f |f|^2 (per voxel and plane, shared
// by both steps), completed by fma(mz, z+1, A).  One component, the lane's
// two voxels (columns c = 0, 1 of its row): u = centre values, up and dn =
// the rows above and below, s = (x-1 + x+1) + (y-1 + y+1), the x-1 of
// column 0 and the x+1 of column 1 from the neighbouring lanes (DPP folded
// into the adds).
__device__ __forceinline__ void sr_partial(const float (&u)[2], float2 up, float2 dn, const float (&zm)[2],
                                           float zmul, const float (&c1)[2], const float (&g)[2], float mu,
                                           float (&A)[2]) {
    float s[2];
    s[0] = (from_left(u[1]) + u[1]) + (up.x + dn.x);
    s[1] = (u[0] + from_right(u[0])) + (up.y + dn.y);
#pragma unroll
    for (int k = 0; k < 2; k++)
        A[k] = __builtin_fmaf(mu, __builtin_fmaf(zmul, zm[k], s[k]), __builtin_fmaf(u[k], c1[k], g[k]));
}

// One z chunk of one tile: XT for tiles at an x face of the volume (the
// mirror columns), a separate instance so the interior tiles' code has no
// branch inside its stencil.
template <class S, class SV, class SF, bool OUT32, bool XT>
__device__ __forceinline__ void sr_march(const SrArgs& a, float2 (&L0)[2][3][SRTY][32], float2 (&L1)[2][3][SRTY][32],
                                         int x0, int y0, int qs, int qe) {
    typedef H16<S> H;
    typedef H16<SV> HV;
    typedef H16<SF> HF;
    const Dims d = a.d;
    // lane li of half hf of wave w: columns x0 + 2 li, +1 of row y0 + 2 w + hf
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 31, hf = lane >> 5;
    const int xa = x0 + 2 * li, rw = 2 * w + hf, ya = y0 + rw;
    // x faces (column 0 of a tile at x = 0, X-1 of a tile ending at X-1): v1
    // and the output there take the values of columns 2 and X-3
    const bool xl = XT && x0 == 0 && li == 0, xr = XT && x0 + SRTX == d.X && li == 31;
    // LDS rows of the y neighbours; stage 2 reads v1's mirror rows at the y
    // faces (row 1's upper neighbour, row 0, holds row 2; row Y-2's lower one
    // holds row Y-3)
    const int ru = max(rw - 1, 0), rd = min(rw + 1, SRTY - 1);
    const int ru2 = ya - 1 == 0 ? 2 - y0 : ru, rd2 = ya + 1 == d.Y - 1 ? d.Y - 3 - y0 : rd;
    // output columns and rows (interior tiles: 60 x 28); the y and z faces are
    // k_gvf_sr_faces'
    const int ylo = y0 == 0 ? 1 : y0 + 2, yhi = y0 + SRTY == d.Y ? d.Y - 2 : y0 + SRTY - 3;
    const int xlo = x0 == 0 ? 0 : x0 + 2, xhi = x0 + SRTX == d.X ? d.X - 1 : x0 + SRTX - 3;
    // tiles start at even columns and X is even: a lane's two columns are
    // both output columns or neither
    const bool ok = xa >= xlo && xa <= xhi && ya >= ylo && ya <= yhi;
    const float mu = a.mu, m2s = a.m2s, c0 = a.c0;

    // loads: a resource per plane array over the chunk's planes zb .. zb +
    // zspan - 1; this lane's word at byte offset voff within a plane
    const int q0 = max(qs - 1, 1), q1 = qe + 1;  // steps
    const int zb = q0 - 1;
    const unsigned pb = (unsigned)((i64)d.X * d.Y * 2);  // bytes per 16-bit plane
    const int zspan = min(a.zspan, d.Z - zb);
    __amdgpu_buffer_rsrc_t rv[3], rf[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        rv[c] = __builtin_amdgcn_make_buffer_rsrc((void*)((const char*)a.v[c] + (i64)zb * pb), 0,
                                                  (int)((unsigned)zspan * pb), BUFFER_WORD3);
        rf[c] = __builtin_amdgcn_make_buffer_rsrc((void*)((const char*)a.f[c] + (i64)zb * pb), 0,
                                                  (int)((unsigned)zspan * pb), BUFFER_WORD3);
    }
    const unsigned voff = (unsigned)(((i64)ya * d.X + xa) * 2);
    auto plane_off = [&](int z) -> int { return (int)((unsigned)(min(z, d.Z - 1) - zb) * pb); };
    // stores: resources over the chunk's output planes qs .. qe-1 (a store at
    // SR_OOB is dropped)
    const i64 P = (i64)d.X * d.Y;
    constexpr int OB = OUT32 ? 4 : 2;  // bytes per output element
    __amdgpu_buffer_rsrc_t ro[4];
#pragma unroll
    for (int c = 0; c < (OUT32 ? 4 : 3); c++)
        ro[c] = __builtin_amdgcn_make_buffer_rsrc((void*)((char*)a.o[c] + (i64)qs * P * OB), 0,
                                                  (int)((i64)(qe - qs) * P * OB), BUFFER_WORD3);
    const unsigned orow = (unsigned)((i64)ya * d.X + xa);  // element index within an output plane

    // state carried from plane to plane (units): v0 of the previous plane,
    // the partial updates of stage 1 (previous plane) and stage 2 (the one
    // before), v1 of the previous plane, c1 and g of the previous plane
    float v0p[3][2], a1[3][2], v1p[3][2], a2[3][2], gp[3][2], c1p[2];
    unsigned pv[2][3], pf[2][3];
    {
        unsigned t[3];
#pragma unroll
        for (int c = 0; c < 3; c++) t[c] = __builtin_amdgcn_raw_buffer_load_b32(rv[c], (int)voff, plane_off(q0 - 1), 0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v0p[c][0] = HV::lo(t[c], a.k);
            v0p[c][1] = HV::hi(t[c], a.k);
#pragma unroll
            for (int k = 0; k < 2; k++) a1[c][k] = v1p[c][k] = a2[c][k] = gp[c][k] = 0.0f;
        }
        c1p[0] = c1p[1] = 0.0f;
    }
    // the loads of planes q0 and q0 + 1, each after the stores a step issues
    // (dropped here: offset past the resource), so the memory operations in
    // flight at the loop head are the same on entry as on the back edge and
    // the compiler's wait counts stay partial
#pragma unroll
    for (int s = 0; s < 2; s++) {
#pragma unroll
        for (int k = 0; k < (OUT32 ? 8 : 3); k++)
            __builtin_amdgcn_raw_buffer_store_b32(0u, ro[k % (OUT32 ? 4 : 3)], (int)SR_OOB, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            pv[s][c] = __builtin_amdgcn_raw_buffer_load_b32(rv[c], (int)voff, plane_off(q0 + s), 0);
            pf[s][c] = __builtin_amdgcn_raw_buffer_load_b32(rf[c], (int)voff, plane_off(q0 + s), 0);
        }
    }

    auto step = [&](const int q, auto slot) {
        constexpr int SL = decltype(slot)::value;
        // plane q of v0 and f; c1 and g of plane q
        float u0[3][2], gq[3][2], c1q[2];
        {
            float fq[3][2];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                u0[c][0] = HV::lo(pv[SL][c], a.k);
                u0[c][1] = HV::hi(pv[SL][c], a.k);
                fq[c][0] = HF::lo(pf[SL][c], a.k);
                fq[c][1] = HF::hi(pf[SL][c], a.k);
            }
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const float m2 =
                    __builtin_fmaf(fq[2][k], fq[2][k], __builtin_fmaf(fq[1][k], fq[1][k], fq[0][k] * fq[0][k])) * m2s;
                c1q[k] = c0 - m2;
#pragma unroll
                for (int c = 0; c < 3; c++) gq[c][k] = fq[c][k] * m2;
            }
        }
        // v1(q-1) = A1(q-1) + mu v0(q); its x faces from their mirrors
        float v1c[3][2];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int k = 0; k < 2; k++) v1c[c][k] = __builtin_fmaf(mu, u0[c][k], a1[c][k]);
        if (XT) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float t0 = from_right(v1c[c][0]), t1 = from_left(v1c[c][1]);
                v1c[c][0] = xl ? t0 : v1c[c][0];
                v1c[c][1] = xr ? t1 : v1c[c][1];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            L0[SL][c][rw][li] = make_float2(u0[c][0], u0[c][1]);
            L1[SL][c][rw][li] = make_float2(v1c[c][0], v1c[c][1]);
        }
        __syncthreads();
        // stage 1: A1(q) (v0 is real on every plane and face: no mirrors)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float A[2];
            sr_partial(u0[c], L0[SL][c][ru][li], L0[SL][c][rd][li], v0p[c], 1.0f, c1q, gq[c], mu, A);
#pragma unroll
            for (int k = 0; k < 2; k++) {
                v0p[c][k] = u0[c][k];
                a1[c][k] = A[k];
            }
        }
        // stage 2: v2(q-2) = A2(q-2) + mu v1(q-1) (plane 1: 2 mu, its z-1
        // neighbour is the mirror of its z+1; plane Z-2: already complete)
        {
            const int p = q - 2;
            const float mz = p == 1 ? 2.0f * mu : (p == d.Z - 2 ? 0.0f : mu);
            const bool pok = p >= qs && p < qe;
            float v2[3][2];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int k = 0; k < 2; k++) v2[c][k] = __builtin_fmaf(mz, v1c[c][k], a2[c][k]);
            if (XT) {
                // x faces: column 0 takes column 2, X-1 takes X-3
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float t0 = from_right(v2[c][0]), t1 = from_left(v2[c][1]);
                    v2[c][0] = xl ? t0 : v2[c][0];
                    v2[c][1] = xr ? t1 : v2[c][1];
                }
            }
            const bool st = pok && ok;
            const unsigned ic = (unsigned)(pok ? p - qs : 0) * (unsigned)P + orow;  // index within the chunk
            if (OUT32) {
                float o[4][2];
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const float x = v2[0][k] * a.inv, y = v2[1][k] * a.inv, z = v2[2][k] * a.inv;
                    const float l = sqrtf(x * x + y * y + z * z);
                    o[0][k] = x;
                    o[1][k] = y;
                    o[2][k] = z;
                    o[3][k] = l > 0.0f ? l : 1.0f;
                }
                const unsigned off = st ? ic * 4u : SR_OOB;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(o[c][0]), ro[c], (int)off, 0, 0);
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(o[c][1]), ro[c], (int)off, 4, 0);
                }
            } else {
                const unsigned gi = (unsigned)((i64)xa + (i64)d.X * ((i64)ya + (i64)d.Y * p));
                const unsigned h0 = sr_hash(gi, a.seed), h1 = sr_hash(gi + 1, a.seed);
                const unsigned word[3] = {H::sr2(v2[0][0], h0 & 0x3ffu, v2[0][1], h1 & 0x3ffu),
                                          H::sr2(v2[1][0], (h0 >> 10) & 0x3ffu, v2[1][1], (h1 >> 10) & 0x3ffu),
                                          H::sr2(v2[2][0], (h0 >> 20) & 0x3ffu, v2[2][1], (h1 >> 20) & 0x3ffu)};
                const unsigned off = st ? ic * 2u : SR_OOB;
#pragma unroll
                for (int c = 0; c < 3; c++) __builtin_amdgcn_raw_buffer_store_b32(word[c], ro[c], (int)off, 0, 0);
            }
        }
        // stage 2: A2(q-1) from v1(q-1) (its LDS rows, mirrored at the y faces),
        // v1(q-2), c1 and g of q-1
        {
            const int p = q - 1;
            const float zmul = p == 1 ? 0.0f : (p == d.Z - 2 ? 2.0f : 1.0f);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float A[2];
                sr_partial(v1c[c], L1[SL][c][ru2][li], L1[SL][c][rd2][li], v1p[c], zmul, c1p, gp[c], mu, A);
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    v1p[c][k] = v1c[c][k];
                    a2[c][k] = A[k];
                    gp[c][k] = gq[c][k];
                }
            }
            c1p[0] = c1q[0];
            c1p[1] = c1q[1];
        }
        // loads of plane q + 2 into this slot; the scheduling barriers keep
        // them after this step's work and before the next step's, so each
        // plane is loaded two steps before its use
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            pv[SL][c] = __builtin_amdgcn_raw_buffer_load_b32(rv[c], (int)voff, plane_off(q + 2), 0);
            pf[SL][c] = __builtin_amdgcn_raw_buffer_load_b32(rf[c], (int)voff, plane_off(q + 2), 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    for (int q = q0; q <= q1; q += 2) {
        step(q, std::integral_constant<int, 0>{});
        step(q + 1, std::integral_constant<int, 1>{});
    }
}

template <class S, class SV, class SF, bool OUT32>
__global__ void __launch_bounds__(64 * SRW) k_gvf_sr2(SrArgs a) {
    __shared__ float2 L0[2][3][SRTY][32];  // v0 rows of the current plane
    __shared__ float2 L1[2][3][SRTY][32];  // v1 rows of the previous plane
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const Dims d = a.d;
    const int x0 = min(tbx * SROX, d.X - SRTX), y0 = min(tby * SROY, d.Y - SRTY);
    const int qs = 1 + tbz * SRZ, qe = min(qs + SRZ, d.Z - 1);  // output planes [qs, qe)
    if (x0 == 0 || x0 + SRTX == d.X)
        sr_march<S, SV, SF, OUT32, true>(a, L0, L1, x0, y0, qs, qe);
    else
        sr_march<S, SV, SF, OUT32, false>(a, L0, L1, x0, y0, qs, qe);
}

// The y and z faces of an output from their mirrors (written by k_gvf_sr2,
// x faces included): z faces (whole planes), y faces (rows of the other
// planes).  Element type T, NP planes.
template <class T, int NP>
__global__ void k_gvf_sr_faces(T* p0, T* p1, T* p2, T* p3, Dims d) {
    const i64 XY = (i64)d.X * d.Y, nz = 2 * XY, ny = 2 * (i64)d.X * (d.Z - 2);
    T* P[4] = {p0, p1, p2, p3};
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nz + ny; i += (i64)gridDim.x * blockDim.x) {
        int x, y, z;
        if (i < nz) {
            const i64 r = i % XY;
            x = (int)(r % d.X);
            y = (int)(r / d.X);
            z = i < XY ? 0 : d.Z - 1;
        } else {
            const i64 j = i - nz, row = j / d.X;
            x = (int)(j % d.X);
            y = (row & 1) ? d.Y - 1 : 0;
            z = 1 + (int)(row / 2);
        }
        const i64 wi = d.at(x, y, z), q = d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z));
#pragma unroll
        for (int c = 0; c < NP; c++) P[c][wi] = P[c][q];
    }
}

}  // namespace tsf

namespace tsf {

// A single Euler step from f (16-bit) into 16-bit planes, round to nearest
// (gvf.hip); used for the first step of an odd count.
tsf_status gvf_single_step16(const tsf_field* f, void* const out[3], Dims d, float mu, hipStream_t stream);

// x0 of the last tile is X - 128: X even and at least one tile wide; the
// load spans of a chunk (SRZ + 6 planes of 16 bits) and the store spans (SRZ
// planes of floats) addressable by a buffer resource.
static bool sr_shape_ok(Dims d) {
    const double plane = (double)d.X * d.Y;
    return d.X % 2 == 0 && d.X >= SRTX && d.Y >= SRTY && d.Z >= 4 && plane * 2 * (SRZ + 6) < 2147483648.0 &&
           plane * 4 * SRZ < 2147483648.0;
}

// One two-step launch: iterate S out, SV in (f's storage on the first
// launch, which reads v0 = f), f in SF; unit = the iterate's scale.
template <class S, class SV, class SF>
static void launch_sr2(bool out32, const void* const f[3], const void* const v[3], void* const o[4], Dims d,
                       float mu, float unit, unsigned seed, hipStream_t stream) {
    SrArgs a;
    for (int c = 0; c < 3; c++) {
        a.f[c] = f[c];
        a.v[c] = v[c];
    }
    for (int c = 0; c < 4; c++) a.o[c] = o[c];
    a.d = d;
    a.mu = mu;
    a.m2s = 1.0f / (unit * unit);
    a.inv = 1.0f / unit;
    a.c0 = 1.0f - 6.0f * mu;
    a.k = unit / 32767.0f;
    a.seed = seed;
    // a chunk uses planes qs-2 .. qe+1: SRZ + 4 of them; the look-ahead loads
    // past those fall outside the resources and fetch nothing
    a.zspan = SRZ + 4;
    const dim3 grid((unsigned)std::max(1, (d.X - 4 + SROX - 1) / SROX), (unsigned)std::max(1, (d.Y - 4 + SROY - 1) / SROY),
                    (unsigned)((d.Z - 2 + SRZ - 1) / SRZ));
    if (out32) hipLaunchKernelGGL((k_gvf_sr2<S, SV, SF, true>), grid, dim3(64 * SRW), 0, stream, a);
    else hipLaunchKernelGGL((k_gvf_sr2<S, SV, SF, false>), grid, dim3(64 * SRW), 0, stream, a);
    const i64 nf = 2 * ((i64)d.X * d.Y + (i64)d.X * (d.Z - 2));
    const unsigned g = (unsigned)std::min<i64>((nf + 255) / 256, 4096);
    if (out32)
        hipLaunchKernelGGL((k_gvf_sr_faces<float, 4>), g, 256, 0, stream, (float*)o[0], (float*)o[1], (float*)o[2],
                           (float*)o[3], d);
    else
        hipLaunchKernelGGL((k_gvf_sr_faces<unsigned short, 3>), g, 256, 0, stream, (unsigned short*)o[0],
                           (unsigned short*)o[1], (unsigned short*)o[2], (unsigned short*)nullptr, d);
}

// 16-bit planes (codec C, units) as F32 values, and optionally w = |v| (1
// where 0): the output of a run with no two-step launch, and f for the
// float tail.
template <class C, bool W>
__global__ void k_sr_values(const void* px, const void* py, const void* pz, float k, float inv, float* ox, float* oy,
                            float* oz, float* ow, i64 n) {
    const unsigned short* P[3] = {(const unsigned short*)px, (const unsigned short*)py, (const unsigned short*)pz};
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const float x = H16<C>::one(P[0][i], k) * inv, y = H16<C>::one(P[1][i], k) * inv;
        const float z = H16<C>::one(P[2][i], k) * inv;
        ox[i] = x;
        oy[i] = y;
        oz[i] = z;
        if (W) {
            const float l = sqrtf(x * x + y * y + z * z);
            ow[i] = l > 0.0f ? l : 1.0f;
        }
    }
}

// One Euler step from v0 = f (codec SF) in values, the reference's order
// (kernels_no_3d_write.cl:1746-1787, the mirror rule at the faces), rounded
// to nearest into the iterate's 16-bit planes (S, value * unit): the first
// step of an odd count when f is SNORM16.  Volumes of a few hundred million
// voxels: one pass, read through the caches.
template <class SF, class S>
__global__ void k_sr_step1(const void* fx, const void* fy, const void* fz, Dims d, float mu, float k, float inv,
                           float unit, void* ox, void* oy, void* oz) {
    const unsigned short* F[3] = {(const unsigned short*)fx, (const unsigned short*)fy, (const unsigned short*)fz};
    typename S::T* O[3] = {(typename S::T*)ox, (typename S::T*)oy, (typename S::T*)oz};
    const i64 n = d.n(), sy = d.X, sz = (i64)d.X * d.Y;
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const int x = (int)(i % d.X), y = (int)((i / d.X) % d.Y), z = (int)(i / sz);
        const i64 q = d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z));
        auto val = [&](int c, i64 j) { return H16<SF>::one(F[c][j], k) * inv; };
        const float ix = val(0, q), iy = val(1, q), iz = val(2, q);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float vc = c == 0 ? ix : (c == 1 ? iy : iz);
            float lap = -6.0f * vc;
            lap = lap + val(c, q + 1);
            lap = lap + val(c, q - 1);
            lap = lap + val(c, q + sy);
            lap = lap + val(c, q - sy);
            lap = lap + val(c, q + sz);
            lap = lap + val(c, q - sz);
            S::st(O[c], i, vc + mu * lap, unit);  // v0 = f: the data term (v - f) |f|^2 is zero
        }
    }
}

// The 16-bit part of a run: f in SF (scale `unit` for F16/BF16; SNORM16 read
// as value * unit), iterate in S at `unit`; `iters` Euler steps from v0 = f,
// the last launch's output as F32 x, y, z, w in O.
template <class S, class SF>
static tsf_status run_sr(const tsf_field* f, float unit, float* const O[4], Dims d, float mu, int iters, unsigned seed,
                         void* ws, hipStream_t stream) {
    const i64 n = d.n();
    const float k = unit / 32767.0f, inv = 1.0f / unit;
    const void* F[3] = {f->x, f->y, f->z};
    void* const OV[4] = {O[0], O[1], O[2], O[3]};
    const unsigned G = 4096;
    if (iters == 0) {
        hipLaunchKernelGGL((k_sr_values<SF, true>), G, 256, 0, stream, f->x, f->y, f->z, k, inv, O[0], O[1], O[2], O[3],
                           n);
        return check_launch("tsf_gvf_euler_sr");
    }
    char* A = (char*)ws;
    void* WA[4] = {A, A + 2 * n, A + 4 * n, nullptr};
    void* WB[4] = {A + 6 * n, A + 8 * n, A + 10 * n, nullptr};
    const void* src[3] = {F[0], F[1], F[2]};
    bool from_f = true;  // src holds f (codec SF), not the iterate
    int j = 0;
    if (iters % 2) {
        if (std::is_same<S, SF>::value) {
            TSF_TRY(gvf_single_step16(f, WA, d, mu, stream));
        } else {
            hipLaunchKernelGGL((k_sr_step1<SF, S>), G, 256, 0, stream, f->x, f->y, f->z, d, mu, k, inv, unit, WA[0],
                               WA[1], WA[2]);
            TSF_TRY(check_launch("tsf_gvf_euler_sr"));
        }
        for (int c = 0; c < 3; c++) src[c] = WA[c];
        from_f = false;
        j = 1;
    }
    const int pairs = iters / 2;
    for (int p = 0; p < pairs; p++, j++) {
        const bool last = p == pairs - 1;
        void* const* dst = last ? OV : ((j % 2) ? WB : WA);
        const unsigned sj = seed ^ ((unsigned)j * 0x9e3779b9u);
        if (from_f) launch_sr2<S, SF, SF>(last, F, src, dst, d, mu, unit, sj, stream);
        else launch_sr2<S, S, SF>(last, F, src, dst, d, mu, unit, sj, stream);
        for (int c = 0; c < 3; c++) src[c] = dst[c];
        from_f = false;
        TSF_TRY(check_launch("tsf_gvf_euler_sr"));
    }
    if (pairs == 0)  // one step: its 16-bit result as the F32 output
        hipLaunchKernelGGL((k_sr_values<S, true>), G, 256, 0, stream, WA[0], WA[1], WA[2], k, inv, O[0], O[1], O[2],
                           O[3], n);
    return check_launch("tsf_gvf_euler_sr");
}

// w = |v| (1 where 0) of F32 planes.
__global__ void k_sr_finish(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                            float* __restrict__ w, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const float a = x[i], b = y[i], c = z[i];
        const float l = sqrtf(a * a + b * b + c * c);
        w[i] = l > 0.0f ? l : 1.0f;
    }
}

// float `iters` Euler steps from v0 (gvf.hip; the reference's operation order).
tsf_status gvf_steps_f32(const float* const f[3], const float* const v0[3], float* const out[3], Dims d, float mu,
                         int iters, float* ws, hipStream_t stream);

template <class S, class SF>
static tsf_status run_sr_tail(const tsf_field* f, float unit, const tsf_field* out, Dims d, float mu, int body,
                              int tail, unsigned seed, void* ws, hipStream_t stream) {
    const i64 n = d.n();
    float* T = tail > 0 ? (float*)((char*)ws + 12 * n) : nullptr;  // f32 f, body result (4), ping-pong (3)
    float* const Obody[4] = {tail > 0 ? T + 3 * n : (float*)out->x, tail > 0 ? T + 4 * n : (float*)out->y,
                             tail > 0 ? T + 5 * n : (float*)out->z, tail > 0 ? T + 6 * n : (float*)out->w};
    TSF_TRY((run_sr<S, SF>(f, unit, Obody, d, mu, body, seed, ws, stream)));
    if (tail == 0) return TSF_OK;
    hipLaunchKernelGGL((k_sr_values<SF, false>), 4096, 256, 0, stream, f->x, f->y, f->z, unit / 32767.0f, 1.0f / unit,
                       T, T + n, T + 2 * n, nullptr, n);
    const float* const ff[3] = {T, T + n, T + 2 * n};
    const float* const v0[3] = {Obody[0], Obody[1], Obody[2]};
    float* const o[3] = {(float*)out->x, (float*)out->y, (float*)out->z};
    TSF_TRY(gvf_steps_f32(ff, v0, o, d, mu, tail, T + 7 * n, stream));
    hipLaunchKernelGGL(k_sr_finish, 4096, 256, 0, stream, o[0], o[1], o[2], (float*)out->w, n);
    return check_launch("tsf_gvf_euler_sr");
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_gvf_euler_sr_workspace_size(tsf_dims dims, int32_t f32_tail) {
    const size_t n = (size_t)D(dims).n();
    // two 16-bit iterates; with a float tail also f, the 16-bit part's result
    // (x, y, z, w) and the float steps' ping-pong as F32 planes
    return 12 * n + (f32_tail > 0 ? 40 * n : 0);
}

int32_t tsf_gvf_euler_sr_supported(tsf_dims dims) { return sr_shape_ok(D(dims)) ? 1 : 0; }

tsf_status tsf_gvf_euler_sr(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu, int32_t iterations,
                            int32_t f32_tail, uint32_t seed, void* workspace, size_t workspace_bytes,
                            hipStream_t stream) {
    const char* W = "tsf_gvf_euler_sr";
    TSF_TRY(check_field(f, false, W));
    TSF_TRY(check_field(out, true, W));
    TSF_TRY(check_dims(dims, 4, W));
    if (f->storage != TSF_STORAGE_F16 && f->storage != TSF_STORAGE_BF16 && f->storage != TSF_STORAGE_SN16)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f must be F16, BF16 or SNORM16", W);
    if (out->storage != TSF_STORAGE_F32) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: out must be F32", W);
    if (iterations < 0 || f32_tail < 0)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative iteration count", W);
    if (!(mu >= 0.0f) || !__builtin_isfinite(mu)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: mu must be >= 0", W);
    if (f->storage != TSF_STORAGE_SN16 && !(f->scale >= 0x1p-30f && f->scale <= 0x1p30f))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: scale out of range", W);
    Dims d = D(dims);
    if (!sr_shape_ok(d))
        return set_error(TSF_ERR_DIMENSIONS, "%s: needs X even and >= %d, Y >= %d, Z >= 4 (got %dx%dx%d)", W, SRTX,
                         SRTY, d.X, d.Y, d.Z);
    for (const void* p : {f->x, f->y, f->z})
        if ((uintptr_t)p % 4) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f planes must be 4-byte aligned", W);
    const int tail = f32_tail < iterations ? f32_tail : iterations, body = iterations - tail;
    Scratch sc;
    if (iterations > 2 || iterations % 2 || tail > 0)
        TSF_TRY(sc.get(workspace, workspace_bytes, tsf_gvf_euler_sr_workspace_size(dims, tail), W));
    tsf_status st;
    switch (f->storage) {
        case TSF_STORAGE_F16:
            st = run_sr_tail<F16, F16>(f, f->scale, out, d, mu, body, tail, seed, sc.ptr, stream);
            break;
        case TSF_STORAGE_BF16:
            st = run_sr_tail<BF16, BF16>(f, f->scale, out, d, mu, body, tail, seed, sc.ptr, stream);
            break;
        default:  // SNORM16 f, the iterate in F16
            st = run_sr_tail<F16, SN16>(f, TSF_SR_SN16_UNIT, out, d, mu, body, tail, seed, sc.ptr, stream);
            break;
    }
    if (st == TSF_OK && sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return st;
}

}  // extern "C"
