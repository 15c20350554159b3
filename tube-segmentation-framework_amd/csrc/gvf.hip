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
static tsf_status run_euler(const tsf_field* f, const tsf_field* out, Dims d, float mu, int iters, void* ws,
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
    for (int i = 1; i <= iters; i++) {
        // ping-pong so that the last step lands in `out`
        const bool to_out = ((iters - i) % 2) == 0;
        Planes3<S> src = (i == 1) ? F : (to_out ? Wr : Or);
        const float inv_src = (i == 1) ? inv_f : inv_v;
        Out4<S> dst = to_out ? O : Wk;
        const bool fin = to_out && i == iters;
        if (vec2) launch_step<S, 2>(fin, F, src, dst, d, mu, inv_f, inv_src, sc_v, stream);
        else launch_step<S, 1>(fin, F, src, dst, d, mu, inv_f, inv_src, sc_v, stream);
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
    for (int i = 1; i <= iters; i++) {
        const bool to_out = ((iters - i) % 2) == 0;
        Planes3<F32> src = (i == 1) ? V0 : (to_out ? Wr : Or);
        Out4<F32> dst = to_out ? O : Wk;
        if (vec2) launch_step<F32, 2>(false, F, src, dst, d, mu, 1.0f, 1.0f, 1.0f, stream);
        else launch_step<F32, 1>(false, F, src, dst, d, mu, 1.0f, 1.0f, 1.0f, stream);
        TSF_TRY(check_launch("gvf_steps_f32"));
    }
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_gvf_euler_workspace_size(tsf_dims dims, int32_t storage) {
    if (storage < 0 || storage > 2) return 0;
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
        case TSF_STORAGE_F32: st = run_euler<F32>(f, out, d, mu, iterations, sc.ptr, stream); break;
        case TSF_STORAGE_F16: st = run_euler<F16>(f, out, d, mu, iterations, sc.ptr, stream); break;
        default: st = run_euler<BF16>(f, out, d, mu, iterations, sc.ptr, stream); break;
    }
    if (st == TSF_OK && sc.owned) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    }
    return st;
}

}  // extern "C"
