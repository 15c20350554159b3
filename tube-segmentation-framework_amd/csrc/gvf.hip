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
#include <algorithm>
#include <type_traits>

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
// per two iterations.  A workgroup of FBX x FBY threads (one wave per row)
// covers FBX x FBY columns, x starting two columns left of its output tile:
// lanes 1..FBX-2 of every row form the tile E where v1 = step(v0) is
// computed, lanes 2..FBX-3 of rows 1..FBY-2 the output tile O (FOW x FOH
// columns) where v2 = step(v1) is computed, and lanes 0 and FBX-1 only supply
// v0 as the x neighbours of E.  The waves of rows 0 and 1 also load the rows
// just above and below.  Every load is one whole row of the wave, so no load
// serves a lone column.  The workgroup marches a chunk of FZC planes in z;
// per step t (one barrier): LDS <- v0 plane t (its rows plus the two ring
// rows) and v1 plane t-1; v1(t) from LDS x/y neighbours and the thread's
// register ring of v0 planes; v2(t-1) from LDS v1 plane t-1 and the register
// ring of v1 planes t-2..t.  Loads run three planes ahead into rings of raw
// stored values (buffer loads: a scalar plane offset plus one constant
// per-thread offset) that are converted at the top of the step that uses
// them.  The arithmetic runs on stored numbers (value * unit, unit a power of
// two), so it rounds exactly as on the values; v1 is kept in float.
// Mirror rule: v1 at a boundary column / row / plane equals v1 at its mirror
// (0 -> 2, N-1 -> N-3), so v1 reads there are redirected.  v2 is stored at
// interior voxels and, on interior rows, at columns 0 and X-1 (the value of
// columns 2 and X-3, taken across lanes); k_gvf_faces then writes the y and
// z faces from their mirrors.  Tiles are placed so a boundary column and its
// mirror lie in the same O (the last tile along x and y is shifted to end at
// the volume's edge).
constexpr int FBX = 64, FBY = 12;
constexpr int FOW = FBX - 4, FOH = FBY - 2;
constexpr int FZC = 64;  // output planes per z chunk, a multiple of 4

// The stored number of one element: getu() reads it as a float in storage
// units, put() stores a float given in units, bload() is a buffer load.
template <class S>
struct Raw;
template <>
struct Raw<F32> {
    typedef float T;
    static __device__ __forceinline__ float getu(T r) { return r; }
    static __device__ __forceinline__ T put(float u) { return u; }
    static __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t b, unsigned v, unsigned s) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(b, (int)v, (int)s, 0));
    }
};
template <>
struct Raw<F16> {
    typedef unsigned short T;
    static __device__ __forceinline__ float getu(T r) { return __half2float(__ushort_as_half(r)); }
    static __device__ __forceinline__ T put(float u) { return __half_as_ushort(__float2half_rn(u)); }
    static __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t b, unsigned v, unsigned s) {
        return __builtin_amdgcn_raw_buffer_load_b16(b, (int)v, (int)s, 0);
    }
};
template <>
struct Raw<BF16> {
    typedef unsigned short T;
    static __device__ __forceinline__ float getu(T r) { return __uint_as_float((unsigned)r << 16); }
    static __device__ __forceinline__ T put(float u) { return __bfloat16_as_ushort(__float2bfloat16(u)); }
    static __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t b, unsigned v, unsigned s) {
        return __builtin_amdgcn_raw_buffer_load_b16(b, (int)v, (int)s, 0);
    }
};
template <>
struct Raw<SN16> {
    typedef short T;
    static __device__ __forceinline__ float getu(T r) { return sn16u_to_f(r); }
    static __device__ __forceinline__ T put(float u) { return f_to_sn16u(u); }
    static __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t b, unsigned v, unsigned s) {
        return (short)__builtin_amdgcn_raw_buffer_load_b16(b, (int)v, (int)s, 0);
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

// Whether values in storage units differ from the values (m2 = |f|^2 then
// needs the factor 1/unit^2): all but F32.
template <class S>
struct UnitScaled {
    static constexpr bool value = true;
};
template <>
struct UnitScaled<F32> {
    static constexpr bool value = false;
};

template <class S>
struct Fused2 {
    Planes3<S> f, vin;
    Out4<S> vout;
    Dims d;
    float mu, m2s, unit;
    int lane, ty, cy, qs, qe;
    bool icol, inO, hrow;
    i64 P;
    unsigned coff, hoff;  // byte offsets within a plane: this thread's column, its ring-row cell
    unsigned pb;          // bytes per plane
    int zb;               // first plane of the chunk's buffer resources
    __amdgpu_buffer_rsrc_t rv[3], rf[3];  // v0 and f planes zb.. of this chunk
    int lxp, lxm, lyp, lym;  // LDS coordinates of v1 neighbours (mirror-redirected)
    int hly;                 // LDS row of this thread's ring-row cell
    bool xedge;              // the tile holds column 0 or X-1
    bool xbnd;               // this thread's column is 0 or X-1, on an interior row
    int xsrc;                // lane holding the mirror column (2 or X-3) of a boundary column, else own lane
};

// Plane z of p (z uniform, so the address stays scalar).
template <class T>
__device__ __forceinline__ T* plane_of(T* p, int z, i64 P) {
    return p + (i64)z * P;
}

__device__ __forceinline__ int mirror_of(int p, int N) { return p == 0 ? 2 : (p == N - 1 ? N - 3 : p); }

// Stores v2 at this thread's column of plane p (an interior voxel).
template <class S, bool FINISH>
__device__ __forceinline__ void f2_store(const Fused2<S>& m, int p, const float (&r)[3]) {
    typedef Raw<S> R;
    typename R::T q[4];
#pragma unroll
    for (int c = 0; c < 3; c++) q[c] = R::put(r[c]);
    if (FINISH) {
        // magnitude of the stored (rounded) components, 1 (in units) where 0
        const float ax = R::getu(q[0]), ay = R::getu(q[1]), az = R::getu(q[2]);
        const float l = sqrtf(ax * ax + ay * ay + az * az);
        q[3] = R::put(l > 0.0f ? l : m.unit);
    }
    typename S::T* OP[4] = {m.vout.x, m.vout.y, m.vout.z, m.vout.w};
#pragma unroll
    for (int c = 0; c < (FINISH ? 4 : 3); c++)
        *reinterpret_cast<typename R::T*>(reinterpret_cast<char*>(plane_of(OP[c], p, m.P)) + m.coff) = q[c];
}

// cv: v0 of planes t-1 and t, fv / mv: f and |f|^2 of plane t-1, all
// converted at an earlier step and carried in registers (each stored value
// is converted once).
template <class S, bool FINISH, int T4>
__device__ __forceinline__ void f2_step(const Fused2<S>& m, int t, typename Raw<S>::T (&v0)[4][3], float (&v1)[4][3],
                                        typename Raw<S>::T (&fr)[4][3], typename Raw<S>::T (&hr)[2][3],
                                        float (&cv)[2][3], float (&fv)[3], float& mv,
                                        float (&L0)[2][3][FBY + 2][FBX + 2], float (&L1)[2][3][FBY][FBX]) {
    typedef Raw<S> R;
    constexpr int SM1 = (T4 + 3) & 3, S0 = T4, SP1 = (T4 + 1) & 3, SM2 = (T4 + 2) & 3;
    constexpr int B = T4 & 1, BP = (T4 + 1) & 1;
    const int lane = m.lane, ty = m.ty, lx = lane + 1;
    // every conversion in straight-line code, before any branch, so the
    // waits on the loads issued two to four planes ago stay exact
    float c0[3], cp[3], cm[3], f1[3], f2[3], ch[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cm[c] = cv[0][c];
        c0[c] = cv[1][c];
        cp[c] = R::getu(v0[SP1][c]);
        f1[c] = R::getu(fr[S0][c]);
        f2[c] = fv[c];
        ch[c] = R::getu(hr[B][c]);
        // left alone, the compiler sinks this into the branch that uses it,
        // where the wait becomes a full drain
        asm volatile("" : "+v"(ch[c]));
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        L0[B][c][ty + 1][lx] = c0[c];
        L1[BP][c][ty][lane] = v1[SM1][c];
    }
    if (m.hrow) {
#pragma unroll
        for (int c = 0; c < 3; c++) L0[B][c][m.hly][lx] = ch[c];
    }
    __syncthreads();
    // v1(t) at this thread's column (lanes 1..FBX-2; the others read the
    // zeroed pad cells and keep 0)
    const bool do1 = m.icol && t >= 1 && t <= m.d.Z - 2;
    float m1 = f1[0] * f1[0] + f1[1] * f1[1] + f1[2] * f1[2];
    if (UnitScaled<S>::value) m1 = m1 * m.m2s;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float r = gvf_update(c0[c], L0[B][c][ty + 1][lx + 1], L0[B][c][ty + 1][lx - 1], L0[B][c][ty + 2][lx],
                                   L0[B][c][ty][lx], cp[c], cm[c], f1[c], m1, m.mu);
        const float rm = RoundMid<S>::value ? R::getu(R::put(r)) : r;
        v1[S0][c] = do1 ? rm : 0.0f;
    }
    // plane 0 / Z-1 of v1 equal planes 2 / Z-3 (the mirror rule): the ring
    // slot of plane 0 takes plane 2 when that is computed (t == 2), the slot
    // of plane Z-1 takes plane Z-3 (t == Z-1); neither slot is read otherwise
    if (t == 2) {
#pragma unroll
        for (int c = 0; c < 3; c++) v1[SM2][c] = v1[S0][c];
    }
    if (t == m.d.Z - 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) v1[S0][c] = v1[SM2][c];
    }
    // v2(t - 1) at this thread's column
    const int p = t - 1;
    if (m.inO && p >= m.qs && p < m.qe) {
        const float m2 = mv;
        float r[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            r[c] = gvf_update(v1[SM1][c], L1[BP][c][ty][m.lxp], L1[BP][c][ty][m.lxm], L1[BP][c][m.lyp][lane],
                              L1[BP][c][m.lym][lane], v1[S0][c], v1[SM2][c], f2[c], m2, m.mu);
        if (m.xedge) {
#pragma unroll
            for (int c = 0; c < 3; c++) r[c] = __shfl(r[c], m.xsrc);
        }
        if (m.icol || m.xbnd) f2_store<S, FINISH>(m, p, r);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cv[0][c] = c0[c];
        cv[1][c] = cp[c];
        fv[c] = f1[c];
    }
    mv = m1;
    // loads: v0 and f of plane t+3 into the slots of t-1, the ring-row cells
    // of plane t+2 into this step's ring slot; issued in a fixed order (the
    // prologue repeats it), so the wait counts at the loop head are the same
    // on entry and on the back edge
    const unsigned s3 = (unsigned)(min(t + 3, m.d.Z - 1) - m.zb) * m.pb;
    const unsigned s2 = (unsigned)(min(t + 2, m.d.Z - 1) - m.zb) * m.pb;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v0[SM1][c] = R::bload(m.rv[c], m.coff, s3);
        asm volatile("" ::: "memory");
        fr[SM1][c] = R::bload(m.rf[c], m.coff, s3);
        asm volatile("" ::: "memory");
        hr[B][c] = R::bload(m.rv[c], m.hoff, s2);
        asm volatile("" ::: "memory");
    }
}

template <class S, bool FINISH>
__global__ void __launch_bounds__(FBX* FBY, FINISH ? 2 : 4)
    k_gvf_fused2(Planes3<S> f, Planes3<S> vin, Out4<S> vout, Dims d, float mu, float m2s, float unit) {
    typedef Raw<S> R;
    typedef typename R::T RT;
    __shared__ float L0[2][3][FBY + 2][FBX + 2];
    __shared__ float L1[2][3][FBY][FBX];
    Fused2<S> m;
    m.f = f;
    m.vin = vin;
    m.vout = vout;
    m.d = d;
    m.mu = mu;
    m.m2s = m2s;
    m.unit = unit;
    m.lane = threadIdx.x;
    m.ty = threadIdx.y;
    // the pad cells left and right of every LDS row of v0 hold 0 (lanes 0
    // and FBX-1 read them and discard the result)
    for (int i = m.ty * FBX + m.lane; i < 2 * 3 * (FBY + 2) * 2; i += FBX * FBY) {
        const int side = i & 1, row = i >> 1;
        (&L0[0][0][0][0])[row * (FBX + 2) + (side ? FBX + 1 : 0)] = 0.0f;
    }
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int ox = tile_origin(tbx, FOW, d.X), oy = tile_origin(tby, FOH, d.Y);
    const int ex = ox - 2, ey = oy - 1;
    const int cx = ex + m.lane;
    m.cy = ey + m.ty;
    const bool act = cx >= 0 && cx < d.X && m.cy >= 0 && m.cy < d.Y;
    m.icol = act && m.lane >= 1 && m.lane <= FBX - 2 && cx >= 1 && cx <= d.X - 2 && m.cy >= 1 && m.cy <= d.Y - 2;
    m.inO = act && m.lane >= 2 && m.lane <= FBX - 3 && m.ty >= 1 && m.ty <= FBY - 2;
    // columns 0 and X-1 of the interior rows take v2 of columns 2 and X-3
    // (the x mirror) from the lanes that compute it, two lanes over; the y
    // and z faces are written by k_gvf_faces
    m.xedge = ex <= 0 || ex + FBX - 1 >= d.X - 1;
    m.xbnd = m.inO && (cx == 0 || cx == d.X - 1) && m.cy >= 1 && m.cy <= d.Y - 2;
    m.xsrc = cx == 0 ? m.lane + 2 : (cx == d.X - 1 ? m.lane - 2 : m.lane);
    m.P = (i64)d.X * d.Y;
    // the host launches this kernel only when a chunk's plane bytes fit 31 bits
    m.coff = act ? (unsigned)(((i64)m.cy * d.X + cx) * (i64)sizeof(typename S::T)) : 0u;
    // v1 neighbours of an O column, redirected at the volume's boundary
    // columns and rows (their mirror lies inside E for every O column)
    m.lxp = clampi(mirror_of(cx + 1, d.X) - ex, 0, FBX - 1);
    m.lxm = clampi(mirror_of(cx - 1, d.X) - ex, 0, FBX - 1);
    m.lyp = clampi(mirror_of(m.cy + 1, d.Y) - ey, 0, FBY - 1);
    m.lym = clampi(mirror_of(m.cy - 1, d.Y) - ey, 0, FBY - 1);
    // ring rows ey-1 (wave 0) and ey+FBY (wave 1) at this thread's column
    m.hrow = m.ty == 0 || m.ty == 1;
    const int hy = m.ty == 0 ? ey - 1 : ey + FBY;
    m.hly = m.ty == 0 ? 0 : FBY + 1;
    const bool hin = m.hrow && cx >= 0 && cx < d.X && hy >= 0 && hy < d.Y;
    m.hoff = hin ? (unsigned)(((i64)hy * d.X + cx) * (i64)sizeof(typename S::T)) : 0u;
    m.qs = 1 + tbz * FZC;
    m.qe = min(m.qs + FZC, d.Z - 1);
    const int t0 = m.qs - 1, te = m.qe;
    // buffer resources over the planes the chunk's outputs depend on, zb ..
    // qe + 1 (two steps from planes below qe); the loop runs up to three
    // steps past te and loads three planes ahead, and those look-ahead loads
    // past the resource fetch nothing (they feed no stored plane)
    m.pb = (unsigned)(m.P * (i64)sizeof(typename S::T));
    m.zb = max(t0 - 1, 0);
    const int zspan = min(d.Z, m.qe + 2) - m.zb;
    const typename S::T* VP[3] = {vin.x, vin.y, vin.z};
    const typename S::T* FP[3] = {f.x, f.y, f.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        m.rv[c] = __builtin_amdgcn_make_buffer_rsrc((void*)plane_of(VP[c], m.zb, m.P), 0, (int)(zspan * m.pb),
                                                    BUFFER_WORD3);
        m.rf[c] = __builtin_amdgcn_make_buffer_rsrc((void*)plane_of(FP[c], m.zb, m.P), 0, (int)(zspan * m.pb),
                                                    BUFFER_WORD3);
    }

    RT v0[4][3], fr[4][3], hr[2][3];
    float v1[4][3];
    // the loads of the four steps before t0, in the order the steps issue
    // them (plane t0-1, t0, t0+1 with the ring cells of t0, t0+2 with those
    // of t0+1), so the loop is entered with the steady state's load order
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int z = min(max(t0 - 1 + g, 0), d.Z - 1), k = (g + 3) & 3;
        const unsigned sz = (unsigned)(z - m.zb) * m.pb, sh = (unsigned)(min(t0 + g - 2, d.Z - 1) - m.zb) * m.pb;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v0[k][c] = R::bload(m.rv[c], m.coff, sz);
            asm volatile("" ::: "memory");
            fr[k][c] = R::bload(m.rf[c], m.coff, sz);
            asm volatile("" ::: "memory");
            v1[g][c] = 0.0f;
            if (g >= 2) {
                hr[g & 1][c] = R::bload(m.rv[c], m.hoff, sh);
                asm volatile("" ::: "memory");
            }
        }
    }
    // planes t0-1 and t0 of v0 and plane t0-1 of f, converted once here
    float cv[2][3], fv[3], mv;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cv[0][c] = R::getu(v0[3][c]);
        cv[1][c] = R::getu(v0[0][c]);
        fv[c] = R::getu(fr[3][c]);
    }
    mv = fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2];
    if (UnitScaled<S>::value) mv = mv * m2s;
    // four unconditional steps per iteration (up to three past te, which
    // store nothing: their v2 plane is >= qe and their loads are clamped) so
    // the loop body has no merge points that would blur the load counters
    for (int t = t0; t <= te; t += 4) {
        f2_step<S, FINISH, 0>(m, t, v0, v1, fr, hr, cv, fv, mv, L0, L1);
        f2_step<S, FINISH, 1>(m, t + 1, v0, v1, fr, hr, cv, fv, mv, L0, L1);
        f2_step<S, FINISH, 2>(m, t + 2, v0, v1, fr, hr, cv, fv, mv, L0, L1);
        f2_step<S, FINISH, 3>(m, t + 3, v0, v1, fr, hr, cv, fv, mv, L0, L1);
    }
}

// ---------------------------------------------------------------- four steps per launch
// k_gvf_fused4 extends the two-step scheme to four steps: v and f cross HBM
// once per four iterations.  A workgroup of 64 x F4Y threads (one wave per
// row) covers 64 x F4Y columns starting four columns left and three rows
// above its output tile; stage k (k = 1..4) computes v_k = step(v_{k-1}) on
// lanes k..63-k of rows k-1..F4Y-k, so stage 4 produces the output tile O of
// (64-8) x (F4Y-6) columns.  Per step t (one barrier): LDS gets v0 plane t
// (with the ring rows) and v_k plane t-k from each stage k < 4; stage k then
// computes plane t-k+1 from LDS x/y neighbours and its z neighbours in the
// register ring of stage k-1 (planes t-k..t-k+2, the last just computed).
// f and |f|^2 of plane t-k+1 pass from stage to stage in registers.  Loads,
// units, the mirror rule and the boundary faces are those of k_gvf_fused2.
constexpr int F4Y = 16;
constexpr int F4OW = FBX - 8, F4OH = F4Y - 6;

// The three intermediate iterates of a four-step launch, stored when the
// caller needs every state (the adjoint's segment recomputation): each one
// block of three component volumes of n elements.
template <class S>
struct Mids3 {
    typename S::T* s[3];
    i64 n;
};

template <class S>
struct Fused4 {
    Planes3<S> f, vin;
    Out4<S> vout;
    Mids3<S> mids;
    Dims d;
    float mu, m2s, unit;
    int lane, ty, cy, qs, qe;
    bool inner, hrow, xedge, xbnd, inO;
    i64 P;
    unsigned coff, hoff, pb;
    int zb, ex;
    __amdgpu_buffer_rsrc_t rv[3], rf[3];
    int lxp, lxm, lyp, lym, hly, xsrc;
};

// Stage k's iterate (k = 1..3) at this thread's column of plane p, for the
// threads of the output tile: interior voxels, and the x-boundary columns
// of interior rows from their mirror lanes (k_gvf_faces writes the rest).
template <class S>
__device__ __forceinline__ void f4_store_mid(const Fused4<S>& m, int k, int p, const float (&v)[3]) {
    typedef Raw<S> R;
    float r[3] = {v[0], v[1], v[2]};
    if (m.xedge) {
#pragma unroll
        for (int c = 0; c < 3; c++) r[c] = __shfl(r[c], m.xsrc);
    }
    if ((m.inner && m.lane >= 4 && m.lane <= FBX - 5) || m.xbnd) {
#pragma unroll
        for (int c = 0; c < 3; c++)
            *reinterpret_cast<typename R::T*>(reinterpret_cast<char*>(plane_of(m.mids.s[k - 1], p, m.P) + c * m.mids.n) +
                                              m.coff) = R::put(r[c]);
    }
}

template <class S, bool FINISH, bool MIDS, int T4>
__device__ __forceinline__ void f4_step(const Fused4<S>& m, int t, typename Raw<S>::T (&v0)[4][3],
                                        typename Raw<S>::T (&fr)[4][3], typename Raw<S>::T (&hr)[2][3],
                                        float (&fk)[4][3], float (&mk)[4], float (&vk)[3][4][3],
                                        float (&L0)[2][3][F4Y + 2][FBX + 2], float (&LK)[3][2][3][F4Y][FBX]) {
    typedef Raw<S> R;
    constexpr int S0 = T4, SP1 = (T4 + 1) & 3, SM1 = (T4 + 3) & 3;
    constexpr int B = T4 & 1, BP = (T4 + 1) & 1;
    const int lane = m.lane, ty = m.ty, lx = lane + 1;
    float c0[3], cp[3], cm[3], fn[3], ch[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        c0[c] = R::getu(v0[S0][c]);
        cp[c] = R::getu(v0[SP1][c]);
        cm[c] = R::getu(v0[SM1][c]);
        fn[c] = R::getu(fr[S0][c]);
        ch[c] = R::getu(hr[B][c]);
        asm volatile("" : "+v"(ch[c]));
    }
    // f and |f|^2 of plane t-k+1 for stage k
#pragma unroll
    for (int k = 3; k >= 1; k--) {
#pragma unroll
        for (int c = 0; c < 3; c++) fk[k][c] = fk[k - 1][c];
        mk[k] = mk[k - 1];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) fk[0][c] = fn[c];
    {
        float m2 = fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2];
        if (UnitScaled<S>::value) m2 = m2 * m.m2s;
        mk[0] = m2;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        L0[B][c][ty + 1][lx] = c0[c];
#pragma unroll
        for (int k = 1; k <= 3; k++) LK[k - 1][BP][c][ty][lane] = vk[k - 1][(T4 - k) & 3][c];
    }
    if (m.hrow) {
#pragma unroll
        for (int c = 0; c < 3; c++) L0[B][c][m.hly][lx] = ch[c];
    }
    __syncthreads();
    // stage 1: v1(t)
    {
        const bool ok = m.inner && lane >= 1 && lane <= FBX - 2 && t >= 1 && t <= m.d.Z - 2;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float r = gvf_update(c0[c], L0[B][c][ty + 1][lx + 1], L0[B][c][ty + 1][lx - 1], L0[B][c][ty + 2][lx],
                                       L0[B][c][ty][lx], cp[c], cm[c], fk[0][c], mk[0], m.mu);
            const float rm = RoundMid<S>::value ? R::getu(R::put(r)) : r;
            vk[0][S0][c] = ok ? rm : 0.0f;
        }
        if (MIDS && m.inO && t >= m.qs && t < m.qe) f4_store_mid(m, 1, t, vk[0][S0]);
        if (t == 2) {
#pragma unroll
            for (int c = 0; c < 3; c++) vk[0][(T4 - 2) & 3][c] = vk[0][S0][c];
        }
        if (t == m.d.Z - 1) {
#pragma unroll
            for (int c = 0; c < 3; c++) vk[0][S0][c] = vk[0][(T4 - 2) & 3][c];
        }
    }
    // stages 2 and 3: v_k(t-k+1) on rows k-1..F4Y-k
#pragma unroll
    for (int k = 2; k <= 3; k++) {
        const int p = t - k + 1;
        if (ty >= k - 1 && ty <= F4Y - k) {
            const int sc = (T4 - k + 1) & 3, sp = (T4 - k + 2) & 3, sm = (T4 - k) & 3;
            const bool ok = m.inner && lane >= k && lane <= FBX - 1 - k && p >= 1 && p <= m.d.Z - 2;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float r = gvf_update(vk[k - 2][sc][c], LK[k - 2][BP][c][ty][m.lxp], LK[k - 2][BP][c][ty][m.lxm],
                                           LK[k - 2][BP][c][m.lyp][lane], LK[k - 2][BP][c][m.lym][lane],
                                           vk[k - 2][sp][c], vk[k - 2][sm][c], fk[k - 1][c], mk[k - 1], m.mu);
                const float rm = RoundMid<S>::value ? R::getu(R::put(r)) : r;
                vk[k - 1][sc][c] = ok ? rm : 0.0f;
            }
        }
        if (MIDS && m.inO && p >= m.qs && p < m.qe) f4_store_mid(m, k, p, vk[k - 1][(T4 - k + 1) & 3]);
        if (p == 2) {
#pragma unroll
            for (int c = 0; c < 3; c++) vk[k - 1][(T4 - k - 1) & 3][c] = vk[k - 1][(T4 - k + 1) & 3][c];
        }
        if (p == m.d.Z - 1) {
#pragma unroll
            for (int c = 0; c < 3; c++) vk[k - 1][(T4 - k + 1) & 3][c] = vk[k - 1][(T4 - k - 1) & 3][c];
        }
    }
    // stage 4: the output plane t-3
    {
        const int p = t - 3;
        if (m.inO && p >= m.qs && p < m.qe) {
            const int sc = (T4 - 3) & 3, sp = (T4 - 2) & 3, sm = (T4 - 4) & 3;
            float r[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
                r[c] = gvf_update(vk[2][sc][c], LK[2][BP][c][ty][m.lxp], LK[2][BP][c][ty][m.lxm],
                                  LK[2][BP][c][m.lyp][lane], LK[2][BP][c][m.lym][lane], vk[2][sp][c], vk[2][sm][c],
                                  fk[3][c], mk[3], m.mu);
            if (m.xedge) {
#pragma unroll
                for (int c = 0; c < 3; c++) r[c] = __shfl(r[c], m.xsrc);
            }
            if ((m.inner && lane >= 4 && lane <= FBX - 5) || m.xbnd) {
                typename R::T q[4];
#pragma unroll
                for (int c = 0; c < 3; c++) q[c] = R::put(r[c]);
                if (FINISH) {
                    const float ax = R::getu(q[0]), ay = R::getu(q[1]), az = R::getu(q[2]);
                    const float l = sqrtf(ax * ax + ay * ay + az * az);
                    q[3] = R::put(l > 0.0f ? l : m.unit);
                }
                typename S::T* OP[4] = {m.vout.x, m.vout.y, m.vout.z, m.vout.w};
#pragma unroll
                for (int c = 0; c < (FINISH ? 4 : 3); c++)
                    *reinterpret_cast<typename R::T*>(reinterpret_cast<char*>(plane_of(OP[c], p, m.P)) + m.coff) =
                        q[c];
            }
        }
    }
    // loads of plane t+3 and the ring rows of plane t+2, fixed order
    const unsigned s3 = (unsigned)(min(t + 3, m.d.Z - 1) - m.zb) * m.pb;
    const unsigned s2 = (unsigned)(min(t + 2, m.d.Z - 1) - m.zb) * m.pb;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v0[SM1][c] = R::bload(m.rv[c], m.coff, s3);
        asm volatile("" ::: "memory");
        fr[SM1][c] = R::bload(m.rf[c], m.coff, s3);
        asm volatile("" ::: "memory");
        hr[B][c] = R::bload(m.rv[c], m.hoff, s2);
        asm volatile("" ::: "memory");
    }
}

template <class S, bool FINISH, bool MIDS = false>
__global__ void __launch_bounds__(FBX* F4Y, 4)
    k_gvf_fused4(Planes3<S> f, Planes3<S> vin, Out4<S> vout, Dims d, float mu, float m2s, float unit,
                 Mids3<S> mids = Mids3<S>{}) {
    typedef Raw<S> R;
    typedef typename R::T RT;
    __shared__ float L0[2][3][F4Y + 2][FBX + 2];
    __shared__ float LK[3][2][3][F4Y][FBX];
    Fused4<S> m;
    m.f = f;
    m.vin = vin;
    m.vout = vout;
    m.mids = mids;
    m.d = d;
    m.mu = mu;
    m.m2s = m2s;
    m.unit = unit;
    m.lane = threadIdx.x;
    m.ty = threadIdx.y;
    for (int i = m.ty * FBX + m.lane; i < 2 * 3 * (F4Y + 2) * 2; i += FBX * F4Y) {
        const int side = i & 1, row = i >> 1;
        (&L0[0][0][0][0])[row * (FBX + 2) + (side ? FBX + 1 : 0)] = 0.0f;
    }
    int tbx, tby, tbz;
    xcd_tile(tbx, tby, tbz);
    const int ox = tile_origin(tbx, F4OW, d.X), oy = tile_origin(tby, F4OH, d.Y);
    const int ex = ox - 4, ey = oy - 3;
    m.ex = ex;
    const int cx = ex + m.lane;
    m.cy = ey + m.ty;
    const bool act = cx >= 0 && cx < d.X && m.cy >= 0 && m.cy < d.Y;
    // interior column: stage outputs exist there
    m.inner = act && cx >= 1 && cx <= d.X - 2 && m.cy >= 1 && m.cy <= d.Y - 2;
    m.inO = act && m.lane >= 4 && m.lane <= FBX - 5 && m.ty >= 3 && m.ty <= F4Y - 4;
    m.P = (i64)d.X * d.Y;
    m.coff = act ? (unsigned)(((i64)m.cy * d.X + cx) * (i64)sizeof(typename S::T)) : 0u;
    m.lxp = clampi(mirror_of(cx + 1, d.X) - ex, 0, FBX - 1);
    m.lxm = clampi(mirror_of(cx - 1, d.X) - ex, 0, FBX - 1);
    m.lyp = clampi(mirror_of(m.cy + 1, d.Y) - ey, 0, F4Y - 1);
    m.lym = clampi(mirror_of(m.cy - 1, d.Y) - ey, 0, F4Y - 1);
    m.xedge = ex <= 0 || ex + FBX - 1 >= d.X - 1;
    m.xbnd = m.inO && (cx == 0 || cx == d.X - 1) && m.cy >= 1 && m.cy <= d.Y - 2;
    m.xsrc = cx == 0 ? m.lane + 2 : (cx == d.X - 1 ? m.lane - 2 : m.lane);
    m.hrow = m.ty == 0 || m.ty == 1;
    const int hy = m.ty == 0 ? ey - 1 : ey + F4Y;
    m.hly = m.ty == 0 ? 0 : F4Y + 1;
    const bool hin = m.hrow && cx >= 0 && cx < d.X && hy >= 0 && hy < d.Y;
    m.hoff = hin ? (unsigned)(((i64)hy * d.X + cx) * (i64)sizeof(typename S::T)) : 0u;
    m.qs = 1 + tbz * FZC;
    m.qe = min(m.qs + FZC, d.Z - 1);
    const int t0 = m.qs - 3, te = m.qe + 2;
    // resources over the planes the chunk's outputs depend on, max(t0-1, 0)
    // .. qe + 3 (four steps from planes below qe); the look-ahead loads past
    // them (the loop reads up to te+6) fetch nothing and feed no stored plane
    m.pb = (unsigned)(m.P * (i64)sizeof(typename S::T));
    m.zb = max(t0 - 1, 0);
    const int zspan = min(d.Z, m.qe + 4) - m.zb;
    const typename S::T* VP[3] = {vin.x, vin.y, vin.z};
    const typename S::T* FP[3] = {f.x, f.y, f.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        m.rv[c] = __builtin_amdgcn_make_buffer_rsrc((void*)plane_of(VP[c], m.zb, m.P), 0, (int)(zspan * m.pb),
                                                    BUFFER_WORD3);
        m.rf[c] = __builtin_amdgcn_make_buffer_rsrc((void*)plane_of(FP[c], m.zb, m.P), 0, (int)(zspan * m.pb),
                                                    BUFFER_WORD3);
    }
    RT v0[4][3], fr[4][3], hr[2][3];
    float fk[4][3], mk[4], vk[3][4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        mk[k] = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            fk[k][c] = 0.0f;
            if (k < 3)
#pragma unroll
                for (int s = 0; s < 4; s++) vk[k][s][c] = 0.0f;
        }
    }
    // the loads of the four steps before t0, in the steps' order
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int z = min(max(t0 - 1 + g, 0), d.Z - 1), k = (g + 3) & 3;
        const unsigned sz = (unsigned)(z - m.zb) * m.pb;
        const unsigned sh = (unsigned)(min(max(t0 + g - 2, 0), d.Z - 1) - m.zb) * m.pb;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v0[k][c] = R::bload(m.rv[c], m.coff, sz);
            asm volatile("" ::: "memory");
            fr[k][c] = R::bload(m.rf[c], m.coff, sz);
            asm volatile("" ::: "memory");
            if (g >= 2) {
                hr[g & 1][c] = R::bload(m.rv[c], m.hoff, sh);
                asm volatile("" ::: "memory");
            }
        }
    }
    for (int t = t0; t <= te; t += 4) {
        f4_step<S, FINISH, MIDS, 0>(m, t, v0, fr, hr, fk, mk, vk, L0, LK);
        f4_step<S, FINISH, MIDS, 1>(m, t + 1, v0, fr, hr, fk, mk, vk, L0, LK);
        f4_step<S, FINISH, MIDS, 2>(m, t + 2, v0, fr, hr, fk, mk, vk, L0, LK);
        f4_step<S, FINISH, MIDS, 3>(m, t + 3, v0, fr, hr, fk, mk, vk, L0, LK);
    }
}

// The volume's boundary voxels after a two-step launch: w takes the stored
// value at mirror(w) (each coordinate 0 -> 2, N-1 -> N-3), which the
// two-step kernel computed (an interior voxel, or an x-boundary column of an
// interior row, which that kernel also writes).  One thread per voxel of
// both z faces, then of the y faces without their z rows.
template <class S, bool W>
__global__ void k_gvf_faces(Out4<S> v, Dims d) {
    const i64 XY = (i64)d.X * d.Y, nz = 2 * XY, ny = 2 * (i64)d.X * (d.Z - 2);
    typename S::T* P[4] = {v.x, v.y, v.z, v.w};
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
        const i64 w = d.at(x, y, z), q = d.at(remap(x, d.X), remap(y, d.Y), remap(z, d.Z));
#pragma unroll
        for (int c = 0; c < (W ? 4 : 3); c++) P[c][w] = P[c][q];
    }
}

template <class S>
static void launch_faces(bool with_w, Out4<S> v, Dims d, hipStream_t stream) {
    const i64 n = 2 * ((i64)d.X * d.Y + (i64)d.X * (d.Z - 2));
    const unsigned g = (unsigned)std::min<i64>((n + 255) / 256, 4096);
    if (with_w) hipLaunchKernelGGL((k_gvf_faces<S, true>), g, 256, 0, stream, v, d);
    else hipLaunchKernelGGL((k_gvf_faces<S, false>), g, 256, 0, stream, v, d);
}

// Units of the two-step kernel: it computes on stored numbers (value *
// unit) and converts only at the ends.  unit is a power of two (the field
// scale; 2^15 for SN16), so every rounding is the one the values would get.
template <class S>
static float storage_unit(float scale) {
    return std::is_same<S, SN16>::value ? 32768.0f : scale;
}

// Whether the two-step kernel can run: f and v in the same units, and the
// bytes of the planes one z chunk reads addressable by a buffer resource.
template <class S>
static bool fused_ok(float scale_f, float scale_v, Dims d, int steps = 2) {
    const double span = (double)d.X * d.Y * sizeof(typename S::T) * (FZC + (steps == 4 ? 13 : 9));
    return scale_f == scale_v && scale_v >= 0x1p-30f && scale_v <= 0x1p30f && span < 2147483648.0;
}

template <class S>
static void launch_fused2(bool finish, Planes3<S> f, Planes3<S> src, Out4<S> dst, Dims d, float mu, float scale,
                          hipStream_t stream) {
    dim3 grid(d.X >= FOW ? (d.X + FOW - 1) / FOW : 1, d.Y >= FOH ? (d.Y + FOH - 1) / FOH : 1, (d.Z - 2 + FZC - 1) / FZC);
    dim3 block(FBX, FBY);
    const float u = storage_unit<S>(scale), m2s = 1.0f / (u * u);
    if (finish)
        hipLaunchKernelGGL((k_gvf_fused2<S, true>), grid, block, 0, stream, f, src, dst, d, mu, m2s, u);
    else
        hipLaunchKernelGGL((k_gvf_fused2<S, false>), grid, block, 0, stream, f, src, dst, d, mu, m2s, u);
}

template <class S>
static void launch_fused4(bool finish, Planes3<S> f, Planes3<S> src, Out4<S> dst, Dims d, float mu, float scale,
                          hipStream_t stream) {
    dim3 grid(d.X >= F4OW ? (d.X + F4OW - 1) / F4OW : 1, d.Y >= F4OH ? (d.Y + F4OH - 1) / F4OH : 1,
              (d.Z - 2 + FZC - 1) / FZC);
    dim3 block(FBX, F4Y);
    const float u = storage_unit<S>(scale), m2s = 1.0f / (u * u);
    if (finish)
        hipLaunchKernelGGL((k_gvf_fused4<S, true>), grid, block, 0, stream, f, src, dst, d, mu, m2s, u);
    else
        hipLaunchKernelGGL((k_gvf_fused4<S, false>), grid, block, 0, stream, f, src, dst, d, mu, m2s, u);
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
    // launches of one, two or four steps: with fuse >= 4 the count is
    // covered by fours, then at most one two-step and one single launch (run
    // first); with fuse 2 or 3 by pairs and at most one single
    const bool f2 = fuse >= 2 && fused_ok<S>(f->scale, out->scale, d);
    const bool f4 = fuse >= 4 && fused_ok<S>(f->scale, out->scale, d, 4);
    const int quads = f4 ? iters / 4 : 0;
    const int pairs = f2 ? (iters - 4 * quads) / 2 : 0;
    const int singles = iters - 4 * quads - 2 * pairs;
    const int launches = quads + pairs + singles;
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
        } else if (j < singles + pairs) {
            launch_fused2<S>(fin, F, src, dst, d, mu, sc_v, stream);
            launch_faces<S>(fin, dst, d, stream);
        } else {
            launch_fused4<S>(fin, F, src, dst, d, mu, sc_v, stream);
            launch_faces<S>(fin, dst, d, stream);
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
    // four steps per launch (the fastest form for float planes; the same
    // float operations as single steps), then at most one two-step and one
    // single launch, run first
    const int quads = fused_ok<F32>(1.0f, 1.0f, d, 4) ? iters / 4 : 0;
    const int pairs = fused_ok<F32>(1.0f, 1.0f, d) ? (iters - 4 * quads) / 2 : 0;
    const int singles = iters - 4 * quads - 2 * pairs, launches = quads + pairs + singles;
    for (int j = 0; j < launches; j++) {
        const bool to_out = ((launches - 1 - j) % 2) == 0;
        Planes3<F32> src = (j == 0) ? V0 : (to_out ? Wr : Or);
        Out4<F32> dst = to_out ? O : Wk;
        if (j < singles) {
            if (vec2) launch_step<F32, 2>(false, F, src, dst, d, mu, 1.0f, 1.0f, 1.0f, stream);
            else launch_step<F32, 1>(false, F, src, dst, d, mu, 1.0f, 1.0f, 1.0f, stream);
        } else if (j < singles + pairs) {
            launch_fused2<F32>(false, F, src, dst, d, mu, 1.0f, stream);
            launch_faces<F32>(false, dst, d, stream);
        } else {
            launch_fused4<F32>(false, F, src, dst, d, mu, 1.0f, stream);
            launch_faces<F32>(false, dst, d, stream);
        }
        TSF_TRY(check_launch("gvf_steps_f32"));
    }
    return TSF_OK;
}

// v^1 .. v^count from v^0 into states[0 .. count-1] (each three planes),
// every state stored with its faces: four steps per launch storing all four
// iterates (bit-identical to single steps), the remainder as single steps.
tsf_status gvf_states_f32(const float* const f[3], const float* v0, float* const* states, int count, Dims d,
                          float mu, hipStream_t stream) {
    const i64 N = d.n();
    Planes3<F32> F{f[0], f[1], f[2]};
    auto planes = [N](const float* p) { return Planes3<F32>{p, p + N, p + 2 * N}; };
    auto out = [N](float* p) { return Out4<F32>{p, p + N, p + 2 * N, nullptr}; };
    bool vec2 = (d.X % 2) == 0;
    for (const void* p : {(const void*)f[0], (const void*)f[1], (const void*)f[2], (const void*)v0})
        vec2 = vec2 && aligned(p, 8);
    for (int i = 0; i < count; i++) vec2 = vec2 && aligned(states[i], 8);
    const bool quads = fused_ok<F32>(1.0f, 1.0f, d, 4);
    const float* src = v0;
    int i = 0;
    while (i < count) {
        if (quads && count - i >= 4) {
            dim3 grid(d.X >= F4OW ? (d.X + F4OW - 1) / F4OW : 1, d.Y >= F4OH ? (d.Y + F4OH - 1) / F4OH : 1,
                      (d.Z - 2 + FZC - 1) / FZC);
            Mids3<F32> mids{{states[i], states[i + 1], states[i + 2]}, N};
            hipLaunchKernelGGL((k_gvf_fused4<F32, false, true>), grid, dim3(FBX, F4Y), 0, stream, F, planes(src),
                               out(states[i + 3]), d, mu, 1.0f, 1.0f, mids);
            for (int k = 0; k < 4; k++) launch_faces<F32>(false, out(states[i + k]), d, stream);
            src = states[i + 3];
            i += 4;
        } else {
            if (vec2) launch_step<F32, 2>(false, F, planes(src), out(states[i]), d, mu, 1.0f, 1.0f, 1.0f, stream);
            else launch_step<F32, 1>(false, F, planes(src), out(states[i]), d, mu, 1.0f, 1.0f, 1.0f, stream);
            src = states[i];
            i += 1;
        }
        TSF_TRY(check_launch("gvf_states_f32"));
    }
    return TSF_OK;
}

// One Euler step from f (F16 or BF16 planes) into the 16-bit planes out[0..2]
// (round to nearest), for the stochastic-rounding entry point's odd counts.
tsf_status gvf_single_step16(const tsf_field* f, void* const out[3], Dims d, float mu, hipStream_t stream) {
    const float inv = 1.0f / f->scale, sc = f->scale;
    if (f->storage == TSF_STORAGE_F16) {
        Planes3<F16> F{(const __half*)f->x, (const __half*)f->y, (const __half*)f->z};
        Out4<F16> O{(__half*)out[0], (__half*)out[1], (__half*)out[2], nullptr};
        launch_step<F16, 1>(false, F, F, O, d, mu, inv, inv, sc, stream);
    } else {
        Planes3<BF16> F{(const __hip_bfloat16*)f->x, (const __hip_bfloat16*)f->y, (const __hip_bfloat16*)f->z};
        Out4<BF16> O{(__hip_bfloat16*)out[0], (__hip_bfloat16*)out[1], (__hip_bfloat16*)out[2], nullptr};
        launch_step<BF16, 1>(false, F, F, O, d, mu, inv, inv, sc, stream);
    }
    return check_launch("gvf_single_step16");
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
