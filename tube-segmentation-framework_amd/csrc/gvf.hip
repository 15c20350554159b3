// Gradient vector flow, explicit Euler (Jacobi) iterations.
//
// Layout: f (the initial field) and v (the iterate) are SoA planes in HBM.
//
// Kernel k_gvf_march (one Jacobi step per launch): a workgroup of 64 x GTY
// threads owns a (64*VEC) x GTY tile of (x, y) columns, each thread VEC
// x-adjacent columns, and marches a chunk of ZC planes in z.  Per column the
// thread keeps planes q-1 .. q+2 of v (a four-slot register ring, the loop
// unrolled by four so the ring never needs register moves) and planes q,
// q+1 of f; every load is issued two planes before its use and is
// unconditional (out-of-range lanes read a safe address and discard it), so
// no wave waits on a load it has just issued.  The current plane of the tile
// plus a one-voxel x/y apron is staged in a double-buffered LDS plane: one
// barrier per plane.  Every v and f value is read from HBM about once per
// step.
//
// Boundary rule of the reference (kernels_no_3d_write.cl:1757-1760): the
// value written at w is the update computed at q = mirror(w), 0 -> 2 and
// N-1 -> N-3.  The kernel computes each interior plane q once and also writes
// it to plane 0 (q == 2) or plane Z-1 (q == Z-3); boundary columns in x/y copy
// the result of their mirror column through LDS.  Tiles are placed so that a
// boundary column and its mirror column always share a tile.
#include "tsf_common.hpp"

namespace tsf {

template <class S>
struct Planes3 {
    const typename S::T* x;
    const typename S::T* y;
    const typename S::T* z;
};
template <class S>
struct Out4 {
    typename S::T* x;
    typename S::T* y;
    typename S::T* z;
    typename S::T* w;
};

constexpr int GTY = 4, ZC = 64;

__device__ __forceinline__ int tile_origin(int b, int T, int N) {
    int o = b * T;
    return N >= T ? min(o, N - T) : 0;
}

// The update of GVF3DIteration (kernels_no_3d_write.cl:1746-1787) in the
// reference's arithmetic order:
//   lap = -6 v + v(x+1) + v(x-1) + v(y+1) + v(y-1) + v(z+1) + v(z-1)
//   v'  = v + (mu lap - (v - f) |f|^2)
__device__ __forceinline__ float gvf_update(float vc, float xp, float xm, float yp, float ym, float zp, float zm,
                                            float fc, float m2, float mu) {
    float lap = -6.0f * vc;
    lap = lap + xp;
    lap = lap + xm;
    lap = lap + yp;
    lap = lap + ym;
    lap = lap + zp;
    lap = lap + zm;
    return vc + (mu * lap - (vc - fc) * m2);
}

// Unconditional load of VEC consecutive elements at i (i even when VEC == 2);
// `ok` false loads element 0 instead and returns zeros.
template <class S, int VEC>
__device__ __forceinline__ void ldv(const typename S::T* p, i64 i, bool ok, float inv, float* o) {
    const i64 j = ok ? i : 0;
    if (VEC == 2) {
        S::ld2(p, j, inv, o[0], o[1]);
    } else {
        o[0] = S::ld(p, j, inv);
    }
#pragma unroll
    for (int k = 0; k < VEC; k++) o[k] = ok ? o[k] : 0.0f;
}
template <class S, int VEC>
__device__ __forceinline__ void stv(typename S::T* p, i64 i, const float* v, float s) {
    if (VEC == 2) S::st2(p, i, v[0], v[1], s);
    else S::st(p, i, v[0], s);
}

template <class S, int VEC>
struct MarchCtx {
    Planes3<S> f, vin;
    Out4<S> vout;
    Dims d;
    float mu, inv_f, inv_v, sc_v;
    int lane, ty, x0, y0, xb, y, qe;
    bool act, blk_edge;
    i64 P, col;
    // apron: this thread loads VEC elements at plane offset hidx and writes
    // element hsel (or all of them for a row apron) to LDS (hly, hlx)
    i64 hidx;
    bool hok, hrow, hcolcell;
    int hsel, hly, hlx;
};

template <class S, int VEC, bool FINISH, int A, int B, int C, int E>
__device__ __forceinline__ void march_plane(const MarchCtx<S, VEC>& m, int q, float (&ring)[4][3][VEC],
                                            float (&fr)[2][3][VEC], float (&hr)[2][3][VEC],
                                            float (&lds)[2][3][GTY + 2][64 * VEC + 2],
                                            float (&res)[3][GTY][64 * VEC]) {
    constexpr int TXV = 64 * VEC;
    const int b = q & 1;
    const int lane = m.lane, ty = m.ty;
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int k = 0; k < VEC; k++) lds[b][c][ty + 1][1 + lane * VEC + k] = ring[B][c][k];
        if (m.hrow) {
#pragma unroll
            for (int k = 0; k < VEC; k++) lds[b][c][m.hly][m.hlx + k] = hr[E][c][k];
        } else if (m.hcolcell) {
            // bitwise select: a select between two elements of the register
            // array would be folded into a dynamic index and sent to scratch
            const unsigned lo = __float_as_uint(hr[E][c][0]), hi = __float_as_uint(hr[E][c][VEC - 1]);
            const unsigned msk = (VEC == 2 && m.hsel) ? 0xffffffffu : 0u;
            lds[b][c][m.hly][m.hlx] = __uint_as_float(lo ^ ((lo ^ hi) & msk));
        }
    }
    __syncthreads();
    float r[3][VEC];
    float m2[VEC];
#pragma unroll
    for (int k = 0; k < VEC; k++)
        m2[k] = fr[E][0][k] * fr[E][0][k] + fr[E][1][k] * fr[E][1][k] + fr[E][2][k] * fr[E][2][k];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            const int lx = 1 + lane * VEC + k;
            r[c][k] = gvf_update(ring[B][c][k], lds[b][c][ty + 1][lx + 1], lds[b][c][ty + 1][lx - 1],
                                 lds[b][c][ty + 2][lx], lds[b][c][ty][lx], ring[C][c][k], ring[A][c][k],
                                 fr[E][c][k], m2[k], m.mu);
        }
    if (m.blk_edge) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int k = 0; k < VEC; k++) res[c][ty][lane * VEC + k] = r[c][k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            const int x = m.xb + k;
            if (m.act && (x == 0 || x == m.d.X - 1 || m.y == 0 || m.y == m.d.Y - 1)) {
                const int mx = remap(x, m.d.X) - m.x0, my = remap(m.y, m.d.Y) - m.y0;
#pragma unroll
                for (int c = 0; c < 3; c++) r[c][k] = res[c][my][mx];
            }
        }
    }
    if (m.act) {
        float wv[VEC];
        if (FINISH) {
            // magnitude of the stored (rounded) components
#pragma unroll
            for (int k = 0; k < VEC; k++) {
                typename S::T t0, t1, t2;
                S::st(&t0, 0, r[0][k], m.sc_v);
                S::st(&t1, 0, r[1][k], m.sc_v);
                S::st(&t2, 0, r[2][k], m.sc_v);
                const float ax = S::ld(&t0, 0, m.inv_v), ay = S::ld(&t1, 0, m.inv_v), az = S::ld(&t2, 0, m.inv_v);
                const float l = sqrtf(ax * ax + ay * ay + az * az);
                wv[k] = l > 0.0f ? l : 1.0f;
            }
        }
        const int nout = 1 + (q == 2 ? 1 : 0) + (q == m.d.Z - 3 ? 1 : 0);
        const int planes[3] = {q, q == 2 ? 0 : m.d.Z - 1, m.d.Z - 1};
        typename S::T* OP[3] = {m.vout.x, m.vout.y, m.vout.z};
#pragma unroll
        for (int o = 0; o < 3; o++) {
            if (o >= nout) break;
            const i64 i = planes[o] * m.P + m.col;
#pragma unroll
            for (int c = 0; c < 3; c++) stv<S, VEC>(OP[c], i, r[c], m.sc_v);
            if (FINISH) stv<S, VEC>(m.vout.w, i, wv, m.sc_v);
        }
    }
    // issue the loads two planes ahead: v(q+3) into the slot of v(q-1),
    // f(q+2) and the apron of q+2 into the current f/apron slots
    const typename S::T* VP[3] = {m.vin.x, m.vin.y, m.vin.z};
    const typename S::T* FP[3] = {m.f.x, m.f.y, m.f.z};
    const int zl = min(q + 3, m.d.Z - 1), zf = min(q + 2, m.d.Z - 1);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ldv<S, VEC>(VP[c], zl * m.P + m.col, m.act, m.inv_v, ring[A][c]);
        ldv<S, VEC>(FP[c], zf * m.P + m.col, m.act, m.inv_f, fr[E][c]);
        ldv<S, VEC>(VP[c], zf * m.P + m.hidx, m.hok, m.inv_v, hr[E][c]);
    }
    (void)TXV;
}

template <class S, int VEC, bool FINISH>
__global__ void __launch_bounds__(64 * GTY) k_gvf_march(Planes3<S> f, Planes3<S> vin, Out4<S> vout, Dims d,
                                                        float mu, float inv_f, float inv_v, float sc_v) {
    constexpr int TXV = 64 * VEC;
    __shared__ float lds[2][3][GTY + 2][TXV + 2];
    __shared__ float res[3][GTY][TXV];
    MarchCtx<S, VEC> m;
    m.f = f;
    m.vin = vin;
    m.vout = vout;
    m.d = d;
    m.mu = mu;
    m.inv_f = inv_f;
    m.inv_v = inv_v;
    m.sc_v = sc_v;
    m.lane = threadIdx.x;
    m.ty = threadIdx.y;
    m.x0 = tile_origin(blockIdx.x, TXV, d.X);
    m.y0 = tile_origin(blockIdx.y, GTY, d.Y);
    m.xb = m.x0 + m.lane * VEC;
    m.y = m.y0 + m.ty;
    m.act = m.y < d.Y && m.xb < d.X;  // VEC == 2 only runs with X even
    m.blk_edge = m.x0 == 0 || m.y0 == 0 || m.x0 + TXV >= d.X || m.y0 + GTY >= d.Y;
    const int qs = 1 + blockIdx.z * ZC;
    m.qe = min(qs + ZC, d.Z - 1);
    m.P = (i64)d.X * d.Y;
    m.col = m.act ? (i64)m.y * d.X + m.xb : 0;
    // apron roles: rows y0-1 (ty 0) and y0+GTY (ty 1) for the thread's own
    // columns; columns x0-1 (ty 2) and x0+TXV (ty 3) for row y0 + lane
    m.hrow = m.ty < 2;
    m.hcolcell = !m.hrow && m.lane < GTY;
    m.hsel = 0;
    int hx, hy;
    if (m.hrow) {
        hy = m.ty == 0 ? m.y0 - 1 : m.y0 + GTY;
        hx = m.xb;
        m.hly = m.ty == 0 ? 0 : GTY + 1;
        m.hlx = 1 + m.lane * VEC;
        m.hok = hy >= 0 && hy < d.Y && hx < d.X;
    } else {
        const int cx = m.ty == 2 ? m.x0 - 1 : m.x0 + TXV;
        hy = m.y0 + m.lane;
        hx = VEC == 2 ? (cx & ~1) : cx;  // aligned pair holding the apron cell
        m.hsel = VEC == 2 ? (cx & 1) : 0;
        m.hly = m.lane + 1;
        m.hlx = m.ty == 2 ? 0 : TXV + 1;
        m.hok = m.hcolcell && cx >= 0 && cx < d.X && hy < d.Y;
    }
    m.hidx = m.hok ? (i64)hy * d.X + hx : 0;

    const typename S::T* VP[3] = {vin.x, vin.y, vin.z};
    const typename S::T* FP[3] = {f.x, f.y, f.z};
    float ring[4][3][VEC], fr[2][3][VEC], hr[2][3][VEC];
    const int z3 = min(qs + 2, d.Z - 1), zf1 = min(qs + 1, d.Z - 1);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ldv<S, VEC>(VP[c], (qs - 1) * m.P + m.col, m.act, inv_v, ring[0][c]);
        ldv<S, VEC>(VP[c], qs * m.P + m.col, m.act, inv_v, ring[1][c]);
        ldv<S, VEC>(VP[c], (qs + 1) * m.P + m.col, m.act, inv_v, ring[2][c]);
        ldv<S, VEC>(VP[c], z3 * m.P + m.col, m.act, inv_v, ring[3][c]);
        ldv<S, VEC>(FP[c], qs * m.P + m.col, m.act, inv_f, fr[0][c]);
        ldv<S, VEC>(FP[c], zf1 * m.P + m.col, m.act, inv_f, fr[1][c]);
        ldv<S, VEC>(VP[c], qs * m.P + m.hidx, m.hok, inv_v, hr[0][c]);
        ldv<S, VEC>(VP[c], zf1 * m.P + m.hidx, m.hok, inv_v, hr[1][c]);
    }
    for (int q = qs; q < m.qe; q += 4) {
        march_plane<S, VEC, FINISH, 0, 1, 2, 0>(m, q, ring, fr, hr, lds, res);
        if (q + 1 < m.qe) march_plane<S, VEC, FINISH, 1, 2, 3, 1>(m, q + 1, ring, fr, hr, lds, res);
        if (q + 2 < m.qe) march_plane<S, VEC, FINISH, 2, 3, 0, 0>(m, q + 2, ring, fr, hr, lds, res);
        if (q + 3 < m.qe) march_plane<S, VEC, FINISH, 3, 0, 1, 1>(m, q + 3, ring, fr, hr, lds, res);
    }
}

// ---------------------------------------------------------------- two steps per launch
// k_gvf_fused2 runs two Jacobi steps per launch, so v and f cross HBM once
// per two iterations.  A workgroup of FBX x FBY threads (each VEC adjacent
// columns) owns an extended column tile E of (FBX*VEC) x FBY columns; it
// computes v1 = step(v0) on E and v2 = step(v1) on the inner tile O (E less
// VEC columns at each x edge and one row at each y edge), marching z.  Per
// step t (one barrier): LDS <- v0 plane t (E plus a one-column ring) and v1
// plane t-1; v1(t) from LDS x/y neighbours and the thread's register ring
// of v0 planes; v2(t-1) from LDS v1 plane t-1 and the register ring of v1
// planes t-2..t.  Loads run three planes ahead into rings of raw stored
// values that are converted at the top of the step that uses them.  v1 is
// kept in float.
// Mirror rule: v1 at a boundary column / plane equals v1 at its mirror
// (0 -> 2, N-1 -> N-3), so v1 reads there are redirected; v2 computed at q
// is also written to every w with mirror(w) = q.  Tiles are placed so a
// boundary column and its mirror lie in the same O (the last tile along x and
// y is shifted to end at the volume's edge).
constexpr int FBX = 64, FBY = 8;
constexpr int FZC = 64;  // output planes per z chunk, a multiple of 4

// raw stored values of VEC consecutive elements; get() converts element k
template <class S, int VEC>
struct RawV;
template <>
struct RawV<F32, 1> {
    typedef float T;
    static __device__ __forceinline__ T load(const float* p, i64 i) { return p[i]; }
    static __device__ __forceinline__ float get(T r, int, float) { return r; }
};
template <>
struct RawV<F32, 2> {
    typedef float2 T;
    static __device__ __forceinline__ T load(const float* p, i64 i) { return *reinterpret_cast<const float2*>(p + i); }
    static __device__ __forceinline__ float get(T r, int k, float) { return k ? r.y : r.x; }
};
template <>
struct RawV<F16, 1> {
    typedef unsigned short T;
    static __device__ __forceinline__ T load(const __half* p, i64 i) {
        return reinterpret_cast<const unsigned short*>(p)[i];
    }
    static __device__ __forceinline__ float get(T r, int, float inv) { return __half2float(__ushort_as_half(r)) * inv; }
};
template <>
struct RawV<F16, 2> {
    typedef unsigned T;
    static __device__ __forceinline__ T load(const __half* p, i64 i) { return *reinterpret_cast<const unsigned*>(p + i); }
    static __device__ __forceinline__ float get(T r, int k, float inv) {
        return __half2float(__ushort_as_half((unsigned short)(k ? (r >> 16) : (r & 0xffffu)))) * inv;
    }
};
template <>
struct RawV<BF16, 1> {
    typedef unsigned short T;
    static __device__ __forceinline__ T load(const __hip_bfloat16* p, i64 i) {
        return reinterpret_cast<const unsigned short*>(p)[i];
    }
    static __device__ __forceinline__ float get(T r, int, float inv) { return __uint_as_float((unsigned)r << 16) * inv; }
};
template <>
struct RawV<BF16, 2> {
    typedef unsigned T;
    static __device__ __forceinline__ T load(const __hip_bfloat16* p, i64 i) {
        return *reinterpret_cast<const unsigned*>(p + i);
    }
    static __device__ __forceinline__ float get(T r, int k, float inv) {
        return __uint_as_float(k ? (r & 0xffff0000u) : (r << 16)) * inv;
    }
};

template <>
struct RawV<SN16, 1> {
    typedef short T;
    static __device__ __forceinline__ T load(const short* p, i64 i) { return p[i]; }
    static __device__ __forceinline__ float get(T r, int, float) { return sn16_to_f(r); }
};
template <>
struct RawV<SN16, 2> {
    typedef unsigned T;
    static __device__ __forceinline__ T load(const short* p, i64 i) { return *reinterpret_cast<const unsigned*>(p + i); }
    static __device__ __forceinline__ float get(T r, int k, float) {
        return sn16_to_f((short)(k ? (r >> 16) : (r & 0xffffu)));
    }
};

// Whether the intermediate iterate of a two-step launch is rounded to the
// storage format: yes for the reference's SNORM16 format, which stores every
// iterate; the FP16/BF16 precision forms keep it in float.
template <class S>
struct RoundMid {
    static constexpr bool value = false;
};
template <>
struct RoundMid<SN16> {
    static constexpr bool value = true;
};

template <int VEC>
struct FGeo {
    static constexpr int EW = FBX * VEC;    // E width (columns)
    static constexpr int RX = VEC;          // E columns beyond O at each x edge (keeps pairs aligned)
    static constexpr int OW = EW - 2 * RX;  // O width
    static constexpr int OH = FBY - 2;      // O height
    static constexpr int LW = EW + 2 * VEC; // LDS row of v0: own columns at VEC + x, ring at VEC-1 and VEC+EW
};

template <class S, int VEC>
struct Fused2 {
    Planes3<S> f, vin;
    Out4<S> vout;
    Dims d;
    float mu, inv_f, inv_v, sc_v;
    int lane, ty, ex, ey, cy, qs, qe;
    int cx[VEC];
    bool act, icol[VEC], inO, hrow, hcol, hok, hsel;
    i64 P, col, hidx;
    int lxp[VEC], lxm[VEC], lyp, lym;  // LDS coordinates of v1 neighbours (mirror-redirected)
    int xw1[VEC], xw2[VEC], yw1, yw2;  // mirror images that receive this column's v2 (-1 = none)
    int hly, hlx;
};

__device__ __forceinline__ int mirror_of(int p, int N) { return p == 0 ? 2 : (p == N - 1 ? N - 3 : p); }

template <class S, int VEC, bool FINISH>
__device__ __forceinline__ void f2_store(const Fused2<S, VEC>& m, int p, const float (&r)[VEC][3],
                                         const bool (&ok)[VEC]) {
    float wv[VEC];
#pragma unroll
    for (int k = 0; k < VEC; k++) {
        wv[k] = 1.0f;
        if (FINISH) {
            // magnitude of the stored (rounded) components
            typename S::T t0, t1, t2;
            S::st(&t0, 0, r[k][0], m.sc_v);
            S::st(&t1, 0, r[k][1], m.sc_v);
            S::st(&t2, 0, r[k][2], m.sc_v);
            const float ax = S::ld(&t0, 0, m.inv_v), ay = S::ld(&t1, 0, m.inv_v), az = S::ld(&t2, 0, m.inv_v);
            const float l = sqrtf(ax * ax + ay * ay + az * az);
            wv[k] = l > 0.0f ? l : 1.0f;
        }
    }
    typename S::T* OP[3] = {m.vout.x, m.vout.y, m.vout.z};
    // this position and its mirror images (-1 = none); fully unrolled with
    // constant indices so the candidates stay in registers
    const int zw[3] = {p, p == 2 ? 0 : -1, p == m.d.Z - 3 ? m.d.Z - 1 : -1};
    const int yw[3] = {m.cy, m.yw1, m.yw2};
#pragma unroll
    for (int k = 0; k < VEC; k++) {
        if (!ok[k]) continue;
        const int xw[3] = {m.cx[k], m.xw1[k], m.xw2[k]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (zw[a] < 0) continue;
#pragma unroll
            for (int bb = 0; bb < 3; bb++) {
                if (yw[bb] < 0) continue;
#pragma unroll
                for (int cc = 0; cc < 3; cc++) {
                    if (xw[cc] < 0) continue;
                    const i64 i = zw[a] * m.P + (i64)yw[bb] * m.d.X + xw[cc];
#pragma unroll
                    for (int c = 0; c < 3; c++) S::st(OP[c], i, r[k][c], m.sc_v);
                    if (FINISH) S::st(m.vout.w, i, wv[k], m.sc_v);
                }
            }
        }
    }
}

template <class S, int VEC, bool FINISH, int T4>
__device__ __forceinline__ void f2_step(const Fused2<S, VEC>& m, int t, typename RawV<S, VEC>::T (&v0)[4][3],
                                        float (&v1)[4][VEC][3], typename RawV<S, VEC>::T (&fr)[4][3],
                                        typename RawV<S, VEC>::T (&hr)[2][3],
                                        float (&L0)[2][3][FBY + 2][FGeo<VEC>::LW],
                                        float (&L1)[2][3][FBY][FGeo<VEC>::EW]) {
    typedef RawV<S, VEC> R;
    constexpr int SM1 = (T4 + 3) & 3, S0 = T4, SP1 = (T4 + 1) & 3, SM2 = (T4 + 2) & 3;
    constexpr int B = T4 & 1, BP = (T4 + 1) & 1;
    const int lane = m.lane, ty = m.ty;
    const float iv = m.inv_v, ifl = m.inv_f;
    // every conversion in straight-line code, before any branch, so the
    // waits on the loads issued two to four planes ago stay exact
    float c0[VEC][3], cp[VEC][3], cm[VEC][3], f1[VEC][3], f2[VEC][3], ch[VEC][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            c0[k][c] = R::get(v0[S0][c], k, iv);
            cp[k][c] = R::get(v0[SP1][c], k, iv);
            cm[k][c] = R::get(v0[SM1][c], k, iv);
            f1[k][c] = R::get(fr[S0][c], k, ifl);
            f2[k][c] = R::get(fr[SM1][c], k, ifl);
            ch[k][c] = R::get(hr[B][c], k, iv);
            // left alone, the compiler sinks these into the branches that
            // use them, where the wait becomes a full drain
            asm volatile("" : "+v"(ch[k][c]), "+v"(f2[k][c]));
        }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            L0[B][c][ty + 1][VEC + lane * VEC + k] = c0[k][c];
            L1[BP][c][ty][lane * VEC + k] = v1[SM1][k][c];
        }
    if (m.hrow) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int k = 0; k < VEC; k++) L0[B][c][m.hly][VEC + lane * VEC + k] = ch[k][c];
    } else if (m.hcol) {
#pragma unroll
        for (int c = 0; c < 3; c++) L0[B][c][m.hly][m.hlx] = m.hsel ? ch[VEC - 1][c] : ch[0][c];
    }
    __syncthreads();
    // v1(t) at this thread's columns
    const bool tin = t >= 1 && t <= m.d.Z - 2;
#pragma unroll
    for (int k = 0; k < VEC; k++) {
        const int lx = VEC + lane * VEC + k;
        const float m2 = f1[k][0] * f1[k][0] + f1[k][1] * f1[k][1] + f1[k][2] * f1[k][2];
        const bool do1 = m.act && m.icol[k] && tin;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float r = gvf_update(c0[k][c], L0[B][c][ty + 1][lx + 1], L0[B][c][ty + 1][lx - 1],
                                       L0[B][c][ty + 2][lx], L0[B][c][ty][lx], cp[k][c], cm[k][c], f1[k][c], m2, m.mu);
            const float rm = RoundMid<S>::value ? sn16_to_f(f_to_sn16(r)) : r;
            v1[S0][k][c] = do1 ? rm : 0.0f;
        }
    }
    // v2(t - 1) at this thread's columns
    const int p = t - 1;
    if (m.inO && p >= m.qs && p < m.qe) {
        const bool lo = (p - 1) == 0, hi = t == m.d.Z - 1;  // plane 0 / Z-1 alias plane 2 / Z-3
        float r[VEC][3];
        bool ok[VEC];
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            ok[k] = m.icol[k];
            const float m2 = f2[k][0] * f2[k][0] + f2[k][1] * f2[k][1] + f2[k][2] * f2[k][2];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float zm = lo ? v1[S0][k][c] : v1[SM2][k][c];
                const float zp = hi ? v1[SM2][k][c] : v1[S0][k][c];
                r[k][c] = gvf_update(v1[SM1][k][c], L1[BP][c][ty][m.lxp[k]], L1[BP][c][ty][m.lxm[k]],
                                     L1[BP][c][m.lyp][lane * VEC + k], L1[BP][c][m.lym][lane * VEC + k], zp, zm,
                                     f2[k][c], m2, m.mu);
            }
        }
        f2_store<S, VEC, FINISH>(m, p, r, ok);
    }
    // loads: v0 and f of plane t+3 into the slots of t-1, the ring cells of
    // plane t+2 into this step's halo slot
    const int z3 = min(t + 3, m.d.Z - 1), z2 = min(t + 2, m.d.Z - 1);
    const i64 iv3 = m.act ? z3 * m.P + m.col : 0, ih2 = m.hok ? z2 * m.P + m.hidx : 0;
    const typename S::T* VP[3] = {m.vin.x, m.vin.y, m.vin.z};
    const typename S::T* FP[3] = {m.f.x, m.f.y, m.f.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v0[SM1][c] = R::load(VP[c], iv3);
        fr[SM1][c] = R::load(FP[c], iv3);
        hr[B][c] = R::load(VP[c], ih2);
    }
}

template <class S, int VEC, bool FINISH>
__global__ void __launch_bounds__(FBX* FBY, FINISH ? 2 : 4) k_gvf_fused2(Planes3<S> f, Planes3<S> vin, Out4<S> vout, Dims d,
                                                            float mu, float inv_f, float inv_v, float sc_v) {
    typedef FGeo<VEC> G;
    typedef RawV<S, VEC> R;
    __shared__ float L0[2][3][FBY + 2][G::LW];
    __shared__ float L1[2][3][FBY][G::EW];
    Fused2<S, VEC> m;
    m.f = f;
    m.vin = vin;
    m.vout = vout;
    m.d = d;
    m.mu = mu;
    m.inv_f = inv_f;
    m.inv_v = inv_v;
    m.sc_v = sc_v;
    m.lane = threadIdx.x;
    m.ty = threadIdx.y;
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int ox = tile_origin(tbx, G::OW, d.X), oy = tile_origin(tby, G::OH, d.Y);
    m.ex = ox - G::RX;
    m.ey = oy - 1;
    m.cy = m.ey + m.ty;
    const int x0 = m.ex + m.lane * VEC;  // even when VEC == 2: ox, X and OW are even
#pragma unroll
    for (int k = 0; k < VEC; k++) m.cx[k] = x0 + k;
    // VEC == 2 runs only with X even, so a pair is wholly inside or outside
    m.act = x0 >= 0 && x0 < d.X && m.cy >= 0 && m.cy < d.Y;
    const bool yint = m.cy >= 1 && m.cy <= d.Y - 2;
#pragma unroll
    for (int k = 0; k < VEC; k++) m.icol[k] = m.act && yint && m.cx[k] >= 1 && m.cx[k] <= d.X - 2;
    m.inO = m.act && m.lane * VEC >= G::RX && m.lane * VEC < G::RX + G::OW && m.ty >= 1 && m.ty <= G::OH;
    m.P = (i64)d.X * d.Y;
    m.col = m.act ? (i64)m.cy * d.X + x0 : 0;
    // v1 neighbours of an O column, redirected at the volume's boundary
    // columns (their mirror lies inside E for every O column that needs it)
#pragma unroll
    for (int k = 0; k < VEC; k++) {
        m.lxp[k] = clampi(mirror_of(m.cx[k] + 1, d.X) - m.ex, 0, G::EW - 1);
        m.lxm[k] = clampi(mirror_of(m.cx[k] - 1, d.X) - m.ex, 0, G::EW - 1);
        m.xw1[k] = m.cx[k] == 2 ? 0 : -1;
        m.xw2[k] = m.cx[k] == d.X - 3 ? d.X - 1 : -1;
    }
    m.lyp = clampi(mirror_of(m.cy + 1, d.Y) - m.ey, 0, FBY - 1);
    m.lym = clampi(mirror_of(m.cy - 1, d.Y) - m.ey, 0, FBY - 1);
    m.yw1 = m.cy == 2 ? 0 : -1;
    m.yw2 = m.cy == d.Y - 3 ? d.Y - 1 : -1;
    // ring of E: rows ey-1 (ty 0) and ey+FBY (ty 1) at this thread's
    // columns; columns ex-1 and ex+EW (ty 2, lanes 0..2*FBY-1), read as the
    // aligned VEC-group holding the cell
    m.hrow = m.ty == 0 || m.ty == 1;
    m.hcol = m.ty == 2 && m.lane < 2 * FBY;
    m.hsel = false;
    m.hly = 0;
    m.hlx = 0;
    int hx = 0, hy = 0;
    bool hin = false;
    if (m.hrow) {
        hy = m.ty == 0 ? m.ey - 1 : m.ey + FBY;
        hx = x0;
        m.hly = m.ty == 0 ? 0 : FBY + 1;
        hin = x0 >= 0 && x0 < d.X && hy >= 0 && hy < d.Y;
    } else if (m.hcol) {
        const bool left = m.lane < FBY;
        const int cx = left ? m.ex - 1 : m.ex + G::EW;
        hy = m.ey + (left ? m.lane : m.lane - FBY);
        hx = VEC == 2 ? (cx & ~1) : cx;
        m.hsel = VEC == 2 && (cx & 1);
        m.hly = (left ? m.lane : m.lane - FBY) + 1;
        m.hlx = left ? VEC - 1 : VEC + G::EW;
        hin = cx >= 0 && cx < d.X && hy >= 0 && hy < d.Y;
    }
    m.hok = hin;
    m.hidx = hin ? (i64)hy * d.X + hx : 0;
    m.qs = 1 + tbz * FZC;
    m.qe = min(m.qs + FZC, d.Z - 1);
    const int t0 = m.qs - 1, te = m.qe;

    typename R::T v0[4][3], fr[4][3], hr[2][3];
    float v1[4][VEC][3];
    const typename S::T* VP[3] = {vin.x, vin.y, vin.z};
    const typename S::T* FP[3] = {f.x, f.y, f.z};
    // the loads of the four steps before t0, in the order the steps issue
    // them (plane t0-1, t0, t0+1 with the ring cells of t0, t0+2 with those
    // of t0+1), so the loop is entered with the steady state's load order
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int z = min(max(t0 - 1 + g, 0), d.Z - 1), k = (g + 3) & 3;
        const i64 i = m.act ? z * m.P + m.col : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v0[k][c] = R::load(VP[c], i);
            fr[k][c] = R::load(FP[c], i);
#pragma unroll
            for (int q = 0; q < VEC; q++) v1[g][q][c] = 0.0f;
            if (g >= 2) hr[g & 1][c] = R::load(VP[c], m.hok ? min(t0 + g - 2, d.Z - 1) * m.P + m.hidx : 0);
        }
    }
    // four unconditional steps per iteration (up to three past te, which
    // store nothing: their v2 plane is >= qe and their loads are clamped) so
    // the loop body has no merge points that would blur the load counters
    for (int t = t0; t <= te; t += 4) {
        f2_step<S, VEC, FINISH, 0>(m, t, v0, v1, fr, hr, L0, L1);
        f2_step<S, VEC, FINISH, 1>(m, t + 1, v0, v1, fr, hr, L0, L1);
        f2_step<S, VEC, FINISH, 2>(m, t + 2, v0, v1, fr, hr, L0, L1);
        f2_step<S, VEC, FINISH, 3>(m, t + 3, v0, v1, fr, hr, L0, L1);
    }
}

template <class S, int VEC>
static void launch_fused2(bool finish, Planes3<S> f, Planes3<S> src, Out4<S> dst, Dims d, float mu, float inv_f,
                          float inv_src, float sc_v, hipStream_t stream) {
    typedef FGeo<VEC> G;
    dim3 grid((d.X + G::OW - 1) / G::OW, (d.Y + G::OH - 1) / G::OH, (d.Z - 2 + FZC - 1) / FZC);
    dim3 block(FBX, FBY);
    if (finish)
        hipLaunchKernelGGL((k_gvf_fused2<S, VEC, true>), grid, block, 0, stream, f, src, dst, d, mu, inv_f, inv_src,
                           sc_v);
    else
        hipLaunchKernelGGL((k_gvf_fused2<S, VEC, false>), grid, block, 0, stream, f, src, dst, d, mu, inv_f, inv_src,
                           sc_v);
}

// iterations == 0: v = f, w = |f| or 1.
template <class S>
__global__ void k_gvf_copy_finish(Planes3<S> f, Out4<S> vout, i64 n, float inv_f, float inv_v, float sc_v) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        S::st(vout.x, i, S::ld(f.x, i, inv_f), sc_v);
        S::st(vout.y, i, S::ld(f.y, i, inv_f), sc_v);
        S::st(vout.z, i, S::ld(f.z, i, inv_f), sc_v);
        float a = S::ld(vout.x, i, inv_v), b = S::ld(vout.y, i, inv_v), c = S::ld(vout.z, i, inv_v);
        float l = sqrtf(a * a + b * b + c * c);
        S::st(vout.w, i, l > 0.0f ? l : 1.0f, sc_v);
    }
}

template <class S, int VEC>
static void launch_step(bool finish, Planes3<S> f, Planes3<S> src, Out4<S> dst, Dims d, float mu, float inv_f,
                        float inv_src, float sc_v, hipStream_t stream) {
    constexpr int TXV = 64 * VEC;
    dim3 grid((d.X + TXV - 1) / TXV, (d.Y + GTY - 1) / GTY, (d.Z - 2 + ZC - 1) / ZC);
    dim3 block(64, GTY);
    if (finish)
        hipLaunchKernelGGL((k_gvf_march<S, VEC, true>), grid, block, 0, stream, f, src, dst, d, mu, inv_f, inv_src, sc_v);
    else
        hipLaunchKernelGGL((k_gvf_march<S, VEC, false>), grid, block, 0, stream, f, src, dst, d, mu, inv_f, inv_src, sc_v);
}

static bool aligned(const void* p, size_t a) { return ((uintptr_t)p % a) == 0; }

template <class S>
static tsf_status run_euler(const tsf_field* f, const tsf_field* out, Dims d, float mu, int iters, int fuse, void* ws,
                            hipStream_t stream) {
    typedef typename S::T T;
    const i64 N = d.n();
    const float inv_f = 1.0f / f->scale, inv_v = 1.0f / out->scale, sc_v = out->scale;
    Planes3<S> F{(const T*)f->x, (const T*)f->y, (const T*)f->z};
    Out4<S> O{(T*)out->x, (T*)out->y, (T*)out->z, (T*)out->w};
    if (iters == 0) {
        hipLaunchKernelGGL(k_gvf_copy_finish<S>, 4096, 256, 0, stream, F, O, N, inv_f, inv_v, sc_v);
        return check_launch("tsf_gvf_euler");
    }
    T* A = (T*)ws;
    Out4<S> Wk{A, A + N, A + 2 * N, nullptr};
    Planes3<S> Wr{A, A + N, A + 2 * N};
    Planes3<S> Or{O.x, O.y, O.z};
    // two-column vector accesses need even rows and aligned planes
    const size_t al = 2 * sizeof(T);
    bool vec2 = (d.X % 2) == 0;
    const void* ptrs[8] = {f->x, f->y, f->z, out->x, out->y, out->z, out->w, A};
    for (const void* p : ptrs) vec2 = vec2 && (p == nullptr || aligned(p, al));
    // launches of one or two steps: with fuse >= 2 an odd count starts with
    // a single step, then pairs
    const int pairs = fuse >= 2 ? iters / 2 : 0;
    const int singles = iters - 2 * pairs;
    const int launches = pairs + singles;
    for (int j = 0; j < launches; j++) {
        // ping-pong so that the last launch lands in `out`
        const bool to_out = ((launches - 1 - j) % 2) == 0;
        Planes3<S> src = (j == 0) ? F : (to_out ? Wr : Or);
        const float inv_src = (j == 0) ? inv_f : inv_v;
        Out4<S> dst = to_out ? O : Wk;
        const bool fin = j == launches - 1;
        if (j < singles) {
            if (vec2) launch_step<S, 2>(fin, F, src, dst, d, mu, inv_f, inv_src, sc_v, stream);
            else launch_step<S, 1>(fin, F, src, dst, d, mu, inv_f, inv_src, sc_v, stream);
        } else {
            launch_fused2<S, 1>(fin, F, src, dst, d, mu, inv_f, inv_src, sc_v, stream);
        }
        TSF_TRY(check_launch("tsf_gvf_euler"));
    }
    return TSF_OK;
}

// `iters` float Euler steps from v0 (which need not be f) into out (planes
// x, y, z; w untouched).  ws: 3 planes, used when iters > 1.  For the
// checkpointed forward and the adjoint's segment recomputation.
tsf_status gvf_steps_f32(const float* const f[3], const float* const v0[3], float* const out[3], Dims d, float mu,
                         int iters, float* ws, hipStream_t stream) {
    const i64 N = d.n();
    Planes3<F32> F{f[0], f[1], f[2]};
    Planes3<F32> V0{v0[0], v0[1], v0[2]};
    Out4<F32> O{out[0], out[1], out[2], nullptr};
    Out4<F32> Wk{ws, ws + N, ws + 2 * N, nullptr};
    Planes3<F32> Wr{ws, ws + N, ws + 2 * N};
    Planes3<F32> Or{out[0], out[1], out[2]};
    bool vec2 = (d.X % 2) == 0;
    const void* ptrs[10] = {f[0], f[1], f[2], v0[0], v0[1], v0[2], out[0], out[1], out[2], ws};
    for (const void* p : ptrs) vec2 = vec2 && (p == nullptr || aligned(p, 8));
    // two steps per launch (the same float operations as single steps), an
    // odd count starting with one single step
    const int pairs = iters / 2, singles = iters - 2 * pairs, launches = pairs + singles;
    for (int j = 0; j < launches; j++) {
        const bool to_out = ((launches - 1 - j) % 2) == 0;
        Planes3<F32> src = (j == 0) ? V0 : (to_out ? Wr : Or);
        Out4<F32> dst = to_out ? O : Wk;
        if (j < singles) {
            if (vec2) launch_step<F32, 2>(false, F, src, dst, d, mu, 1.0f, 1.0f, 1.0f, stream);
            else launch_step<F32, 1>(false, F, src, dst, d, mu, 1.0f, 1.0f, 1.0f, stream);
        } else {
            launch_fused2<F32, 1>(false, F, src, dst, d, mu, 1.0f, 1.0f, 1.0f, stream);
        }
        TSF_TRY(check_launch("gvf_steps_f32"));
    }
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_gvf_euler_workspace_size(tsf_dims dims, int32_t storage) {
    if (storage < 0 || storage > 3) return 0;
    return (size_t)3 * (size_t)D(dims).n() * storage_bytes(storage);
}

tsf_status tsf_gvf_euler(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu, int32_t iterations,
                         int32_t fuse, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_gvf_euler";
    TSF_TRY(check_field(f, false, W));
    TSF_TRY(check_field(out, true, W));
    TSF_TRY(check_dims(dims, 4, W));
    if (f->storage != out->storage)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f and out must use the same storage type", W);
    if (iterations < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: negative iteration count", W);
    if (fuse < 1) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: fuse must be >= 1", W);
    if (!(mu >= 0.0f) || !__builtin_isfinite(mu)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: mu must be >= 0", W);
    if (f->x == out->x || f->y == out->y || f->z == out->z)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: f and out planes must not alias", W);
    Dims d = D(dims);
    Scratch sc;
    if (iterations > 1) TSF_TRY(sc.get(workspace, workspace_bytes, tsf_gvf_euler_workspace_size(dims, out->storage), W));
    tsf_status st;
    switch (out->storage) {
        case TSF_STORAGE_F32: st = run_euler<F32>(f, out, d, mu, iterations, fuse, sc.ptr, stream); break;
        case TSF_STORAGE_F16: st = run_euler<F16>(f, out, d, mu, iterations, fuse, sc.ptr, stream); break;
        case TSF_STORAGE_SN16: st = run_euler<SN16>(f, out, d, mu, iterations, fuse, sc.ptr, stream); break;
        default: st = run_euler<BF16>(f, out, d, mu, iterations, fuse, sc.ptr, stream); break;
    }
    if (st == TSF_OK && sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return st;
}

}  // extern "C"
