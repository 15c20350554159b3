// Packed field texels and the TDF's sampling primitives, shared by the
// forward filter (circle.hip) and its adjoint (adjoint.hip).
#pragma once

#include "tsf_common.hpp"

namespace tsf {

template <class S>
struct Texel;
// mulx/muly/mulz(w, p): w times the x / y / z component of texel p, as
// stored (one rounding, as w * raw(p).x).
template <>
struct Texel<F32> {
    typedef float4 P;
    static __device__ __forceinline__ float4 unpack(P p, float) { return p; }
    static __device__ __forceinline__ float4 raw(P p) { return p; }
    static __device__ __forceinline__ float mulx(float w, P p) { return w * p.x; }
    static __device__ __forceinline__ float muly(float w, P p) { return w * p.y; }
    static __device__ __forceinline__ float mulz(float w, P p) { return w * p.z; }
    static __device__ __forceinline__ float fmx(float w, P p, float acc) { return __builtin_fmaf(w, p.x, acc); }
    static __device__ __forceinline__ float fmy(float w, P p, float acc) { return __builtin_fmaf(w, p.y, acc); }
    static __device__ __forceinline__ float fmz(float w, P p, float acc) { return __builtin_fmaf(w, p.z, acc); }
};
template <>
struct Texel<F16> {
    typedef uint2 P;
    static __device__ __forceinline__ float h(unsigned b) { return __half2float(__ushort_as_half((unsigned short)b)); }
    static __device__ __forceinline__ float4 unpack(P p, float inv) {
        return make_float4(h(p.x & 0xffffu) * inv, h(p.x >> 16) * inv, h(p.y & 0xffffu) * inv, h(p.y >> 16) * inv);
    }
    static __device__ __forceinline__ float4 raw(P p) {
        return make_float4(h(p.x & 0xffffu), h(p.x >> 16), h(p.y & 0xffffu), h(p.y >> 16));
    }
    // v_fma_mix_f32 with a half operand and addend -0: the product w * h of
    // the converted half, rounded once (x * y + (-0) == x * y exactly, signed
    // zeros included), without a separate conversion instruction
    static __device__ __forceinline__ float mulx(float w, P p) {
        float r;
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel_hi:[0,1,0]" : "=v"(r) : "v"(w), "v"(p.x));
        return r;
    }
    static __device__ __forceinline__ float muly(float w, P p) {
        float r;
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel:[0,1,0] op_sel_hi:[0,1,0]" : "=v"(r) : "v"(w), "v"(p.x));
        return r;
    }
    static __device__ __forceinline__ float mulz(float w, P p) {
        float r;
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel_hi:[0,1,0]" : "=v"(r) : "v"(w), "v"(p.y));
        return r;
    }
    // fma(w, h, acc), one rounding (v_fma_mix_f32 reads the half directly)
    static __device__ __forceinline__ float fmx(float w, P p, float acc) {
        return __builtin_fmaf(w, h(p.x & 0xffffu), acc);
    }
    static __device__ __forceinline__ float fmy(float w, P p, float acc) { return __builtin_fmaf(w, h(p.x >> 16), acc); }
    static __device__ __forceinline__ float fmz(float w, P p, float acc) {
        return __builtin_fmaf(w, h(p.y & 0xffffu), acc);
    }
};
template <>
struct Texel<BF16> {
    typedef uint2 P;
    static __device__ __forceinline__ float b(unsigned v) { return __uint_as_float(v << 16); }
    static __device__ __forceinline__ float4 unpack(P p, float inv) {
        return make_float4(b(p.x & 0xffffu) * inv, b(p.x >> 16) * inv, b(p.y & 0xffffu) * inv, b(p.y >> 16) * inv);
    }
    static __device__ __forceinline__ float4 raw(P p) {
        return make_float4(b(p.x & 0xffffu), b(p.x >> 16), b(p.y & 0xffffu), b(p.y >> 16));
    }
    static __device__ __forceinline__ float mulx(float w, P p) { return w * b(p.x & 0xffffu); }
    static __device__ __forceinline__ float muly(float w, P p) { return w * b(p.x >> 16); }
    static __device__ __forceinline__ float mulz(float w, P p) { return w * b(p.y & 0xffffu); }
    static __device__ __forceinline__ float fmx(float w, P p, float acc) {
        return __builtin_fmaf(w, b(p.x & 0xffffu), acc);
    }
    static __device__ __forceinline__ float fmy(float w, P p, float acc) { return __builtin_fmaf(w, b(p.x >> 16), acc); }
    static __device__ __forceinline__ float fmz(float w, P p, float acc) {
        return __builtin_fmaf(w, b(p.y & 0xffffu), acc);
    }
};

// SN16 texels read in units of 2^-15 (32768 * value, exactly); the samplers
// take inv = 2^-15, so unpack() and the scaled trilinear sums give the
// decoded values bit for bit.  The packing kernel stores code -32768 as
// -32767 (both read as -1), so the decode needs no clamp: fma(s, 1/32767, s)
// is the correctly rounded 32768 * s / 32767.
template <>
struct Texel<SN16> {
    typedef uint2 P;
    static __device__ __forceinline__ float u(unsigned b) {
        const float sf = (float)(short)(b & 0xffffu);
        return __builtin_fmaf(sf, 1.0f / 32767.0f, sf);
    }
    static __device__ __forceinline__ float4 unpack(P p, float inv) {
        return make_float4(u(p.x) * inv, u(p.x >> 16) * inv, u(p.y) * inv, u(p.y >> 16) * inv);
    }
    static __device__ __forceinline__ float4 raw(P p) {
        return make_float4(u(p.x), u(p.x >> 16), u(p.y), u(p.y >> 16));
    }
    static __device__ __forceinline__ float mulx(float w, P p) { return w * u(p.x); }
    static __device__ __forceinline__ float muly(float w, P p) { return w * u(p.x >> 16); }
    static __device__ __forceinline__ float mulz(float w, P p) { return w * u(p.y); }
};

// Texel sources of the sampling functions below: row(d, y, z) locates row
// (y, z) and at(row, x) is the texel at (x, y, z), for coordinates already
// clamped into the volume.
// Tex: the packed texel array in device memory.
template <class S>
struct Tex {
    const typename Texel<S>::P* __restrict__ p;
    float inv;
    __device__ __forceinline__ i64 row(const Dims& d, int y, int z) const {
        return (i64)d.X * ((i64)y + (i64)d.Y * z);
    }
    __device__ __forceinline__ float4 at(i64 row, int x) const { return Texel<S>::unpack(p[row + x], inv); }
    __device__ __forceinline__ float4 at_raw(i64 row, int x) const { return Texel<S>::raw(p[row + x]); }
};

// LdsTex: a tile of texels staged in LDS covering [ox, ox+tx) x [oy, oy+ty) x
// [oz, ...).  Local coordinates are clamped into the tile, so a coordinate
// outside it (only possible for non-finite sample positions, whose weights
// are non-finite too) still reads LDS in bounds.
template <class S>
struct LdsTex {
    const typename Texel<S>::P* p;
    float inv;
    int ox, oy, oz, tx, ty, tz;
    __device__ __forceinline__ int row(const Dims&, int y, int z) const {
        return tx * (clampi(y - oy, 0, ty - 1) + ty * clampi(z - oz, 0, tz - 1));
    }
    __device__ __forceinline__ float4 at(int row, int x) const {
        return Texel<S>::unpack(p[row + clampi(x - ox, 0, tx - 1)], inv);
    }
    __device__ __forceinline__ float4 at_raw(int row, int x) const {
        return Texel<S>::raw(p[row + clampi(x - ox, 0, tx - 1)]);
    }
};

// LdsSoaTex: an F32 field's x, y and z staged in LDS as three float planes
// of `cells` floats over the same kind of tile; w (1 when absent) is read
// from its plane in device memory, at the global cell.
struct LdsSoaTex {
    struct Row {
        int l;  // local row offset in the tile
        i64 g;  // global row index
    };
    const float* p;
    const float* __restrict__ w;
    int cells, ox, oy, oz, tx, ty, tz;
    static constexpr float inv = 1.0f;
    __device__ __forceinline__ Row row(const Dims& d, int y, int z) const {
        return {tx * (clampi(y - oy, 0, ty - 1) + ty * clampi(z - oz, 0, tz - 1)), (i64)d.X * ((i64)y + (i64)d.Y * z)};
    }
    __device__ __forceinline__ float4 at(Row r, int x) const {
        const int l = r.l + clampi(x - ox, 0, tx - 1);
        return make_float4(p[l], p[cells + l], p[2 * cells + l], w ? w[r.g + x] : 1.0f);
    }
};

// PlanesF32: an F32 field's x, y, z and w planes in device memory as a texel
// source (w = 1 when absent): the Hessian's source when the samples come from
// 16-bit texels of the same field.
struct PlanesF32 {
    const float* __restrict__ x;
    const float* __restrict__ y;
    const float* __restrict__ z;
    const float* __restrict__ w;
    __device__ __forceinline__ i64 row(const Dims& d, int yy, int zz) const {
        return (i64)d.X * ((i64)yy + (i64)d.Y * zz);
    }
    __device__ __forceinline__ float4 at(i64 r, int xx) const {
        const i64 i = r + xx;
        return make_float4(x[i], y[i], z[i], w ? w[i] : 1.0f);
    }
};

template <bool NORM, class A>
__device__ __forceinline__ void hessian_tex(const A& T, const Dims& d, int x, int y, int z, float H[3][3]) {
    const int xm = max(x - 1, 0), xp = min(x + 1, d.X - 1);
    const int ym = max(y - 1, 0), yp = min(y + 1, d.Y - 1);
    const int zm = max(z - 1, 0), zp = min(z + 1, d.Z - 1);
    const auto r0 = T.row(d, y, z);
    const float4 XP = T.at(r0, xp), XM = T.at(r0, xm);
    const float4 YP = T.at(T.row(d, yp, z), x), YM = T.at(T.row(d, ym, z), x);
    const float4 ZP = T.at(T.row(d, y, zp), x), ZM = T.at(T.row(d, y, zm), x);
    float a0 = XP.x, a1 = XM.x;
    float b0 = XP.y, b1 = XM.y, b2 = YP.y, b3 = YM.y;
    float c0 = XP.z, c1 = XM.z, c2 = YP.z, c3 = YM.z, c4 = ZP.z, c5 = ZM.z;
    if (NORM) {
        a0 = a0 / XP.w; a1 = a1 / XM.w;
        b0 = b0 / XP.w; b1 = b1 / XM.w; b2 = b2 / YP.w; b3 = b3 / YM.w;
        c0 = c0 / XP.w; c1 = c1 / XM.w; c2 = c2 / YP.w; c3 = c3 / YM.w; c4 = c4 / ZP.w; c5 = c5 / ZM.w;
    }
    const float fxx = 0.5f * (a0 - a1);
    const float fyx = 0.5f * (b0 - b1), fyy = 0.5f * (b2 - b3);
    const float fzx = 0.5f * (c0 - c1), fzy = 0.5f * (c2 - c3), fzz = 0.5f * (c4 - c5);
    H[0][0] = fxx; H[0][1] = fyx; H[0][2] = fzx;
    H[1][0] = fyx; H[1][1] = fyy; H[1][2] = fzy;
    H[2][0] = fzx; H[2][1] = fzy; H[2][2] = fzz;
}

// OpenCL CLK_FILTER_LINEAR, unnormalised coordinates, clamp-to-edge: sample
// at u - 0.5, the eight terms in the OpenCL specification's order.  The
// texels' storage scale (a power of two) is applied once to each sum: with
// no overflow or underflow, scaling by a power of two commutes exactly with
// every product and sum, so this equals scaling each tap.
template <class A>
__device__ __forceinline__ void trilinear_tex(const A& T, const Dims& d, float u, float v, float w, float out[3]) {
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    int i1 = clampi(i0 + 1, 0, d.X - 1);
    const int j1 = clampi(j0 + 1, 0, d.Y - 1), k1 = clampi(k0 + 1, 0, d.Z - 1);
    i0 = clampi(i0, 0, d.X - 1);
    j0 = clampi(j0, 0, d.Y - 1);
    k0 = clampi(k0, 0, d.Z - 1);
    const auto r00 = T.row(d, j0, k0), r10 = T.row(d, j1, k0), r01 = T.row(d, j0, k1), r11 = T.row(d, j1, k1);
    const float4 t0 = T.at_raw(r00, i0), t1 = T.at_raw(r00, i1), t2 = T.at_raw(r10, i0), t3 = T.at_raw(r10, i1);
    const float4 t4 = T.at_raw(r01, i0), t5 = T.at_raw(r01, i1), t6 = T.at_raw(r11, i0), t7 = T.at_raw(r11, i1);
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    float sx = wt0 * t0.x, sy = wt0 * t0.y, sz = wt0 * t0.z;
    sx = sx + wt1 * t1.x; sy = sy + wt1 * t1.y; sz = sz + wt1 * t1.z;
    sx = sx + wt2 * t2.x; sy = sy + wt2 * t2.y; sz = sz + wt2 * t2.z;
    sx = sx + wt3 * t3.x; sy = sy + wt3 * t3.y; sz = sz + wt3 * t3.z;
    sx = sx + wt4 * t4.x; sy = sy + wt4 * t4.y; sz = sz + wt4 * t4.z;
    sx = sx + wt5 * t5.x; sy = sy + wt5 * t5.y; sz = sz + wt5 * t5.z;
    sx = sx + wt6 * t6.x; sy = sy + wt6 * t6.y; sz = sz + wt6 * t6.z;
    sx = sx + wt7 * t7.x; sy = sy + wt7 * t7.y; sz = sz + wt7 * t7.z;
    out[0] = sx * T.inv;
    out[1] = sy * T.inv;
    out[2] = sz * T.inv;
}

// The same sampling from an LDS tile.  The tile holds the clamp-to-edge
// texel of every cell and its halo, floor(r + 0.5) + 1 around the brick,
// keeps both corners of every finite sample inside it; so the global clamps
// above change no value here and only the lower corner's local cell is
// computed (and clamped, which matters only for non-finite positions).  The
// taps and the sums are those of trilinear_tex, in the same order.
template <class S>
__device__ __forceinline__ void trilinear_tex(const LdsTex<S>& T, const Dims&, float u, float v, float w,
                                              float out[3]) {
    typedef Texel<S> X;
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    const int sy = T.tx, sz = __mul24(T.tx, T.ty);
    const int li = min(max((int)fu - T.ox, 0), T.tx - 2), lj = min(max((int)fv - T.oy, 0), T.ty - 2);
    const int lk = min(max((int)fw - T.oz, 0), T.tz - 2);
    const typename X::P* q = T.p + (li + __mul24(lj, sy) + __mul24(lk, sz));
    const typename X::P p0 = q[0], p1 = q[1], p2 = q[sy], p3 = q[sy + 1];
    const typename X::P p4 = q[sz], p5 = q[sz + 1], p6 = q[sz + sy], p7 = q[sz + sy + 1];
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    float sx = X::mulx(wt0, p0), sy_ = X::muly(wt0, p0), sz_ = X::mulz(wt0, p0);
    sx = sx + X::mulx(wt1, p1); sy_ = sy_ + X::muly(wt1, p1); sz_ = sz_ + X::mulz(wt1, p1);
    sx = sx + X::mulx(wt2, p2); sy_ = sy_ + X::muly(wt2, p2); sz_ = sz_ + X::mulz(wt2, p2);
    sx = sx + X::mulx(wt3, p3); sy_ = sy_ + X::muly(wt3, p3); sz_ = sz_ + X::mulz(wt3, p3);
    sx = sx + X::mulx(wt4, p4); sy_ = sy_ + X::muly(wt4, p4); sz_ = sz_ + X::mulz(wt4, p4);
    sx = sx + X::mulx(wt5, p5); sy_ = sy_ + X::muly(wt5, p5); sz_ = sz_ + X::mulz(wt5, p5);
    sx = sx + X::mulx(wt6, p6); sy_ = sy_ + X::muly(wt6, p6); sz_ = sz_ + X::mulz(wt6, p6);
    sx = sx + X::mulx(wt7, p7); sy_ = sy_ + X::muly(wt7, p7); sz_ = sz_ + X::mulz(wt7, p7);
    out[0] = sx * T.inv;
    out[1] = sy_ * T.inv;
    out[2] = sz_ * T.inv;
}

// The precision form's sampling (tsf_tdf_circle_mixed): the position
// arrives as uu = u - 0.5 (formed by one fused multiply-add), and the eight
// taps accumulate by fused multiply-adds, one rounding each, in the same
// order; the sums stay in the texels' raw units (the caller applies the
// scale once).  The cell and weight arithmetic is trilinear_tex's.
template <class S>
__device__ __forceinline__ void trilinear_fused(const LdsTex<S>& T, const Dims&, float uu, float vv, float ww,
                                                float out[3]) {
    typedef Texel<S> X;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    const int sy = T.tx, sz = __mul24(T.tx, T.ty);
    // the cell offset in the tile: in range for every finite sample; an
    // unsigned min keeps any other in bounds (its weights are NaN, so the
    // cell read does not matter)
    const int li = (int)min((unsigned)((int)fu - T.ox), (unsigned)(T.tx - 2));
    const int lj = (int)min((unsigned)((int)fv - T.oy), (unsigned)(T.ty - 2));
    const int lk = (int)min((unsigned)((int)fw - T.oz), (unsigned)(T.tz - 2));
    const typename X::P* q = T.p + (li + __mul24(lj, sy) + __mul24(lk, sz));
    const typename X::P p0 = q[0], p1 = q[1], p2 = q[sy], p3 = q[sy + 1];
    const typename X::P p4 = q[sz], p5 = q[sz + 1], p6 = q[sz + sy], p7 = q[sz + sy + 1];
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    float sx = X::mulx(wt0, p0), sy_ = X::muly(wt0, p0), sz_ = X::mulz(wt0, p0);
    sx = X::fmx(wt1, p1, sx); sy_ = X::fmy(wt1, p1, sy_); sz_ = X::fmz(wt1, p1, sz_);
    sx = X::fmx(wt2, p2, sx); sy_ = X::fmy(wt2, p2, sy_); sz_ = X::fmz(wt2, p2, sz_);
    sx = X::fmx(wt3, p3, sx); sy_ = X::fmy(wt3, p3, sy_); sz_ = X::fmz(wt3, p3, sz_);
    sx = X::fmx(wt4, p4, sx); sy_ = X::fmy(wt4, p4, sy_); sz_ = X::fmz(wt4, p4, sz_);
    sx = X::fmx(wt5, p5, sx); sy_ = X::fmy(wt5, p5, sy_); sz_ = X::fmz(wt5, p5, sz_);
    sx = X::fmx(wt6, p6, sx); sy_ = X::fmy(wt6, p6, sy_); sz_ = X::fmz(wt6, p6, sz_);
    sx = X::fmx(wt7, p7, sx); sy_ = X::fmy(wt7, p7, sy_); sz_ = X::fmz(wt7, p7, sz_);
    out[0] = sx;
    out[1] = sy_;
    out[2] = sz_;
}

// LdsTexC: an LdsTex whose tile dimensions are compile-time constants, for
// the fused sampler below, addressed by LDS byte addresses held as
// integer-valued floats.
template <class S, int TX, int TY, int TZ>
struct LdsTexC {
    float inv;
    float o8;  // LDS address of the tile (lo) minus the byte offset of its origin
    float lo;  // LDS address of the tile
};

// trilinear_fused on an LdsTexC: the same cell, weights, taps and order.
// The cell's LDS address is formed in float, fma(fw, 8 TX TY, o8) then y
// then x, every term an integer: exact while the partial sums stay below
// 2^24 in magnitude (the launcher checks 8 (TX Y + X) + 64 KiB < 2^23).  A
// median with the tile's first and last cell keeps every address inside
// the workgroup's LDS, whatever it makes of a NaN (the weights of a
// non-finite position are NaN, so the cell read does not matter).
template <class S, int TX, int TY, int TZ>
__device__ __forceinline__ void trilinear_fused(const LdsTexC<S, TX, TY, TZ>& T, const Dims&, float uu, float vv,
                                                float ww, float out[3]) {
    typedef Texel<S> X;
    typedef typename X::P P;
    constexpr int SY = TX, SZ = TX * TY, B = (int)sizeof(P);
    constexpr float LAST = (float)((TX * TY * TZ - 1 - SZ - SY - 1) * B);
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    const float fo =
        __builtin_fmaf(fu, (float)B, __builtin_fmaf(fv, (float)(SY * B), __builtin_fmaf(fw, (float)(SZ * B), T.o8)));
    const unsigned off = (unsigned)__builtin_amdgcn_fmed3f(fo, T.lo, T.lo + LAST);
    static_assert(sizeof(P) == 8, "8-byte texels");
    typedef __attribute__((address_space(3))) const unsigned long long LW;
    LW* q = (LW*)(size_t)off;
    auto tx = [](unsigned long long w) { P t; __builtin_memcpy(&t, &w, 8); return t; };
    const P p0 = tx(q[0]), p1 = tx(q[1]), p2 = tx(q[SY]), p3 = tx(q[SY + 1]);
    const P p4 = tx(q[SZ]), p5 = tx(q[SZ + 1]), p6 = tx(q[SZ + SY]), p7 = tx(q[SZ + SY + 1]);
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    float sx = X::mulx(wt0, p0), sy_ = X::muly(wt0, p0), sz_ = X::mulz(wt0, p0);
    sx = X::fmx(wt1, p1, sx); sy_ = X::fmy(wt1, p1, sy_); sz_ = X::fmz(wt1, p1, sz_);
    sx = X::fmx(wt2, p2, sx); sy_ = X::fmy(wt2, p2, sy_); sz_ = X::fmz(wt2, p2, sz_);
    sx = X::fmx(wt3, p3, sx); sy_ = X::fmy(wt3, p3, sy_); sz_ = X::fmz(wt3, p3, sz_);
    sx = X::fmx(wt4, p4, sx); sy_ = X::fmy(wt4, p4, sy_); sz_ = X::fmz(wt4, p4, sz_);
    sx = X::fmx(wt5, p5, sx); sy_ = X::fmy(wt5, p5, sy_); sz_ = X::fmz(wt5, p5, sz_);
    sx = X::fmx(wt6, p6, sx); sy_ = X::fmy(wt6, p6, sy_); sz_ = X::fmz(wt6, p6, sz_);
    sx = X::fmx(wt7, p7, sx); sy_ = X::fmy(wt7, p7, sy_); sz_ = X::fmz(wt7, p7, sz_);
    out[0] = sx;
    out[1] = sy_;
    out[2] = sz_;
}

// trilinear_tex on an LdsTexC: the reference-order sample of the LdsTex
// version above (u - 0.5, the same cell, weights, terms and sums), with the
// cell's LDS address formed as trilinear_fused forms it.
template <class S, int TX, int TY, int TZ>
__device__ __forceinline__ void trilinear_tex(const LdsTexC<S, TX, TY, TZ>& T, const Dims&, float u, float v, float w,
                                              float out[3]) {
    typedef Texel<S> X;
    typedef typename X::P P;
    constexpr int SY = TX, SZ = TX * TY, B = (int)sizeof(P);
    constexpr float LAST = (float)((TX * TY * TZ - 1 - SZ - SY - 1) * B);
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    const float fo =
        __builtin_fmaf(fu, (float)B, __builtin_fmaf(fv, (float)(SY * B), __builtin_fmaf(fw, (float)(SZ * B), T.o8)));
    const unsigned off = (unsigned)__builtin_amdgcn_fmed3f(fo, T.lo, T.lo + LAST);
    static_assert(sizeof(P) == 8, "8-byte texels");
    typedef __attribute__((address_space(3))) const unsigned long long LW;
    LW* q = (LW*)(size_t)off;
    auto tx = [](unsigned long long x) { P t; __builtin_memcpy(&t, &x, 8); return t; };
    const P p0 = tx(q[0]), p1 = tx(q[1]), p2 = tx(q[SY]), p3 = tx(q[SY + 1]);
    const P p4 = tx(q[SZ]), p5 = tx(q[SZ + 1]), p6 = tx(q[SZ + SY]), p7 = tx(q[SZ + SY + 1]);
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    float sx = X::mulx(wt0, p0), sy_ = X::muly(wt0, p0), sz_ = X::mulz(wt0, p0);
    sx = sx + X::mulx(wt1, p1); sy_ = sy_ + X::muly(wt1, p1); sz_ = sz_ + X::mulz(wt1, p1);
    sx = sx + X::mulx(wt2, p2); sy_ = sy_ + X::muly(wt2, p2); sz_ = sz_ + X::mulz(wt2, p2);
    sx = sx + X::mulx(wt3, p3); sy_ = sy_ + X::muly(wt3, p3); sz_ = sz_ + X::mulz(wt3, p3);
    sx = sx + X::mulx(wt4, p4); sy_ = sy_ + X::muly(wt4, p4); sz_ = sz_ + X::mulz(wt4, p4);
    sx = sx + X::mulx(wt5, p5); sy_ = sy_ + X::muly(wt5, p5); sz_ = sz_ + X::mulz(wt5, p5);
    sx = sx + X::mulx(wt6, p6); sy_ = sy_ + X::muly(wt6, p6); sz_ = sz_ + X::mulz(wt6, p6);
    sx = sx + X::mulx(wt7, p7); sy_ = sy_ + X::muly(wt7, p7); sz_ = sz_ + X::mulz(wt7, p7);
    out[0] = sx * T.inv;
    out[1] = sy_ * T.inv;
    out[2] = sz_ * T.inv;
}

// The same from texels in device memory (clamp-to-edge at the volume).
template <class S>
__device__ __forceinline__ void trilinear_fused(const Tex<S>& T, const Dims& d, float uu, float vv, float ww,
                                                float out[3]) {
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    int i1 = clampi(i0 + 1, 0, d.X - 1);
    const int j1 = clampi(j0 + 1, 0, d.Y - 1), k1 = clampi(k0 + 1, 0, d.Z - 1);
    i0 = clampi(i0, 0, d.X - 1);
    j0 = clampi(j0, 0, d.Y - 1);
    k0 = clampi(k0, 0, d.Z - 1);
    const auto r00 = T.row(d, j0, k0), r10 = T.row(d, j1, k0), r01 = T.row(d, j0, k1), r11 = T.row(d, j1, k1);
    const float4 t[8] = {T.at_raw(r00, i0), T.at_raw(r00, i1), T.at_raw(r10, i0), T.at_raw(r10, i1),
                         T.at_raw(r01, i0), T.at_raw(r01, i1), T.at_raw(r11, i0), T.at_raw(r11, i1)};
    const float wt[8] = {(1.0f - a) * (1.0f - b) * (1.0f - c), a * (1.0f - b) * (1.0f - c),
                         (1.0f - a) * b * (1.0f - c),          a * b * (1.0f - c),
                         (1.0f - a) * (1.0f - b) * c,          a * (1.0f - b) * c,
                         (1.0f - a) * b * c,                    a * b * c};
    float sx = wt[0] * t[0].x, sy = wt[0] * t[0].y, sz = wt[0] * t[0].z;
#pragma unroll
    for (int k = 1; k < 8; k++) {
        sx = __builtin_fmaf(wt[k], t[k].x, sx);
        sy = __builtin_fmaf(wt[k], t[k].y, sy);
        sz = __builtin_fmaf(wt[k], t[k].z, sz);
    }
    out[0] = sx;
    out[1] = sy;
    out[2] = sz;
}

// The same sampling from the three-plane F32 tile (same taps, terms and
// order; the planes hold the values, so no scale is applied).
__device__ __forceinline__ void trilinear_tex(const LdsSoaTex& T, const Dims&, float u, float v, float w,
                                              float out[3]) {
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    const int sy = T.tx, sz = __mul24(T.tx, T.ty);
    const int li = min(max((int)fu - T.ox, 0), T.tx - 2), lj = min(max((int)fv - T.oy, 0), T.ty - 2);
    const int lk = min(max((int)fw - T.oz, 0), T.tz - 2);
    const int q = li + __mul24(lj, sy) + __mul24(lk, sz);
    const int o[8] = {q, q + 1, q + sy, q + sy + 1, q + sz, q + sz + 1, q + sz + sy, q + sz + sy + 1};
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    const float wt[8] = {wt0, wt1, wt2, wt3, wt4, wt5, wt6, wt7};
    const float* px = T.p;
    const float* py = T.p + T.cells;
    const float* pz = T.p + 2 * T.cells;
    float sx = wt[0] * px[o[0]], sy_ = wt[0] * py[o[0]], sz_ = wt[0] * pz[o[0]];
#pragma unroll
    for (int k = 1; k < 8; k++) {
        sx = sx + wt[k] * px[o[k]];
        sy_ = sy_ + wt[k] * py[o[k]];
        sz_ = sz_ + wt[k] * pz[o[k]];
    }
    out[0] = sx;
    out[1] = sy_;
    out[2] = sz_;
}

// LdsSoaTexC: a three-plane tile whose dimensions are compile-time
// constants, addressed for the samples by LDS byte addresses held as
// integer-valued floats (as LdsTexC; the Hessian keeps an LdsSoaTex).
template <int TX, int TY, int TZ>
struct LdsSoaTexC {
    float o4;  // LDS address of the tile (lo) minus the byte offset of its origin in the x plane
    float lo;  // LDS address of the tile
};

// trilinear_tex on an LdsSoaTexC: the LdsSoaTex version's cell, weights,
// taps and sums, the cell's LDS address formed as trilinear_fused forms it;
// the y and z planes follow the x plane at constant offsets.
template <int TX, int TY, int TZ>
__device__ __forceinline__ void trilinear_tex(const LdsSoaTexC<TX, TY, TZ>& T, const Dims&, float u, float v, float w,
                                              float out[3]) {
    constexpr int SY = TX, SZ = TX * TY, CELLS = TX * TY * TZ, B = (int)sizeof(float);
    constexpr float LAST = (float)((CELLS - 1 - SZ - SY - 1) * B);
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    const float fo =
        __builtin_fmaf(fu, (float)B, __builtin_fmaf(fv, (float)(SY * B), __builtin_fmaf(fw, (float)(SZ * B), T.o4)));
    const unsigned off = (unsigned)__builtin_amdgcn_fmed3f(fo, T.lo, T.lo + LAST);
    typedef __attribute__((address_space(3))) const float LF;
    LF* px = (LF*)(size_t)off;
    LF* py = px + CELLS;
    LF* pz = px + 2 * CELLS;
    constexpr int o[8] = {0, 1, SY, SY + 1, SZ, SZ + 1, SZ + SY, SZ + SY + 1};
    const float wt0 = (1.0f - a) * (1.0f - b) * (1.0f - c), wt1 = a * (1.0f - b) * (1.0f - c);
    const float wt2 = (1.0f - a) * b * (1.0f - c), wt3 = a * b * (1.0f - c);
    const float wt4 = (1.0f - a) * (1.0f - b) * c, wt5 = a * (1.0f - b) * c;
    const float wt6 = (1.0f - a) * b * c, wt7 = a * b * c;
    const float wt[8] = {wt0, wt1, wt2, wt3, wt4, wt5, wt6, wt7};
    float sx = wt[0] * px[o[0]], sy_ = wt[0] * py[o[0]], sz_ = wt[0] * pz[o[0]];
#pragma unroll
    for (int k = 1; k < 8; k++) {
        sx = sx + wt[k] * px[o[k]];
        sy_ = sy_ + wt[k] * py[o[k]];
        sz_ = sz_ + wt[k] * pz[o[k]];
    }
    out[0] = sx;
    out[1] = sy_;
    out[2] = sz_;
}

constexpr int CBX = 16, CBY = 4, CBZ = 4;  // TDF brick

// Packs the field's planes into float4 texels (x, y, z, w; w = 1 when the
// field has no w plane).  Defined in circle.hip.
void pack_texels_f32(const tsf_field* field, float4* out, i64 n, hipStream_t stream);

}  // namespace tsf
