"""ctypes front end of the CPU oracle (oracle/tsf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, as the checker.  The product library never
imports it.

Volumes are numpy arrays of shape (Z, Y, X) (C order), so the flat index is
x + X*(y + Y*z) as in the reference's LPOS macro.  Vector fields are arrays of
shape (4, Z, Y, X): planes x, y, z and w.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libtsf_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


def _fn(name, prec=None):
    return getattr(lib(), f"orc_{name}_{prec}" if prec else f"orc_{name}")


def _dims(vol):
    Z, Y, X = vol.shape[-3:]
    return int(X), int(Y), int(Z)


def set_threads(n=0):
    """Worker threads of the oracle's OpenMP loops (n <= 0 leaves them);
    returns the count in effect."""
    f = lib().orc_set_threads
    f.restype = ctypes.c_int
    return int(f(ctypes.c_int(int(n))))


def blur_mask(sigma):
    mask = np.zeros(11 ** 3, np.float32)
    m = ctypes.c_int(0)
    _fn("blur_mask")(ctypes.c_float(sigma), _p(mask), ctypes.byref(m))
    M = 2 * m.value + 1
    return m.value, mask[: M ** 3].copy()


def circle_tables():
    c = np.zeros(32, np.float32)
    s = np.zeros(32, np.float32)
    _fn("circle_tables")(_p(c), _p(s))
    return c, s


def blur(vol, sigma, prec="f32"):
    vol = np.ascontiguousarray(vol, _dt(prec))
    m, mask = blur_mask(sigma)
    out = np.empty_like(vol)
    X, Y, Z = _dims(vol)
    _fn("blur", prec)(_p(vol), _p(out), X, Y, Z, m, _p(mask))
    return out


def to_float(data, minimum, maximum, prec="f32"):
    d = np.ascontiguousarray(data, np.float64)
    out = np.empty(d.shape, _dt(prec))
    _fn("to_float", prec)(_p(d), _p(out), ctypes.c_longlong(d.size), ctypes.c_float(minimum),
                          ctypes.c_float(maximum))
    return out


def vector_field(vol, fmax, sign=1, prec="f32"):
    vol = np.ascontiguousarray(vol, _dt(prec))
    X, Y, Z = _dims(vol)
    F = np.empty((4,) + vol.shape, _dt(prec))
    _fn("vector_field", prec)(_p(vol), _p(F), X, Y, Z, ctypes.c_float(fmax), int(sign))
    return F


def gvf_euler(F, mu, iters, prec="f32"):
    F = np.ascontiguousarray(F, _dt(prec))
    X, Y, Z = _dims(F)
    V = np.empty((4,) + F.shape[1:], _dt(prec))
    _fn("gvf_euler", prec)(_p(F), _p(V), X, Y, Z, ctypes.c_float(mu), int(iters))
    return V


def gvf_step(f3, v3, mu, prec="f32"):
    f3 = np.ascontiguousarray(f3[:3], _dt(prec))
    v3 = np.ascontiguousarray(v3[:3], _dt(prec))
    X, Y, Z = _dims(f3)
    o = np.empty_like(v3)
    _fn("gvf_step", prec)(_p(f3), _p(v3), _p(o), X, Y, Z, ctypes.c_float(mu))
    return o


def gvf_fmg(F, mu, iters, v0=1, v1=2, v2=2, lmax=2, prec="f32"):
    F = np.ascontiguousarray(F, _dt(prec))
    X, Y, Z = _dims(F)
    V = np.empty((4,) + F.shape[1:], _dt(prec))
    _fn("gvf_fmg", prec)(_p(F), _p(V), X, Y, Z, ctypes.c_float(mu), int(iters), v0, v1, v2, lmax)
    return V


def eigen3(A, prec="f32"):
    A = np.ascontiguousarray(A, _dt(prec)).reshape(9)
    V = np.empty(9, _dt(prec))
    d = np.empty(3, _dt(prec))
    _fn("eigen3", prec)(_p(A), _p(V), _p(d))
    return d, V.reshape(3, 3)


def eigen3_fp64_fallback(A):
    """The precision form's float eigensolve with its FP64 fallback for
    near-degenerate matrices (tsf_tdf_circle_mixed)."""
    A = np.ascontiguousarray(A, np.float32).reshape(9)
    V = np.empty(9, np.float32)
    d = np.empty(3, np.float32)
    lib().orc_eigen3_fp64_fallback_f32(_p(A), _p(V), _p(d))
    return d, V.reshape(3, 3)


def hessian_eigen(F, normalize, vectors=True, prec="f32"):
    F = np.ascontiguousarray(F, _dt(prec))
    X, Y, Z = _dims(F)
    lam = np.empty((3,) + F.shape[1:], _dt(prec))
    vec = np.empty((9,) + F.shape[1:], _dt(prec)) if vectors else None
    _fn("hessian_eigen", prec)(_p(F), X, Y, Z, int(normalize), _p(lam),
                               _p(vec) if vectors else None)
    return lam, vec


def tdf_circle(F, rmin, rmax, rstep, prec="f32"):
    F = np.ascontiguousarray(F, _dt(prec))
    X, Y, Z = _dims(F)
    tdf = np.empty(F.shape[1:], _dt(prec))
    rad = np.empty(F.shape[1:], _dt(prec))
    _fn("tdf_circle", prec)(_p(F), X, Y, Z, ctypes.c_float(rmin), ctypes.c_float(rmax),
                            ctypes.c_float(rstep), _p(tdf), _p(rad))
    return tdf, rad


def tdf_circle_split(FH, FS, rmin, rmax, rstep, prec="f32"):
    """circleFittingTDF with the Hessian from FH and the samples from FS."""
    FH = np.ascontiguousarray(FH, _dt(prec))
    FS = np.ascontiguousarray(FS, _dt(prec))
    assert FH.shape == FS.shape
    X, Y, Z = _dims(FH)
    tdf = np.empty(FH.shape[1:], _dt(prec))
    rad = np.empty(FH.shape[1:], _dt(prec))
    _fn("tdf_circle_split", prec)(_p(FH), _p(FS), X, Y, Z, ctypes.c_float(rmin), ctypes.c_float(rmax),
                                  ctypes.c_float(rstep), _p(tdf), _p(rad))
    return tdf, rad


def tdf_circle_mixed(FH, FS, rmin, rmax, rstep, prec="f32"):
    """The precision form's TDF (tsf_tdf_circle_mixed): the Hessian from FH,
    the samples from FS (FH's 16-bit rounding) with fused multiply-adds."""
    FH = np.ascontiguousarray(FH, _dt(prec))
    FS = np.ascontiguousarray(FS, _dt(prec))
    assert FH.shape == FS.shape
    X, Y, Z = _dims(FH)
    tdf = np.empty(FH.shape[1:], _dt(prec))
    rad = np.empty(FH.shape[1:], _dt(prec))
    _fn("tdf_circle_mixed", prec)(_p(FH), _p(FS), X, Y, Z, ctypes.c_float(rmin), ctypes.c_float(rmax),
                                  ctypes.c_float(rstep), _p(tdf), _p(rad))
    return tdf, rad


def gvf_euler_sr16(f16, storage, scale, mu, iters, seed=0, f32_tail=0):
    """The precision form's GVF (tsf_gvf_euler_sr): f16 = (3, Z, Y, X) uint16
    planes (F16 holding value * scale when storage == 1, BF16 bits when 2,
    SNORM16 integers when 3 -- the iterate then F16 at unit 2^14, scale
    unused), the last f32_tail iterations in float; returns the
    (4, Z, Y, X) float32 output (x, y, z, w)."""
    f16 = np.ascontiguousarray(f16, np.uint16)
    X, Y, Z = _dims(f16)
    out = np.empty((4,) + f16.shape[1:], np.float32)
    lib().orc_gvf_euler_sr16(_p(f16), int(storage), ctypes.c_float(scale), ctypes.c_float(mu), int(iters),
                             int(f32_tail), ctypes.c_uint(seed), X, Y, Z, _p(out))
    return out


def half_from_float(a):
    a = np.ascontiguousarray(a, np.float32)
    h = np.empty(a.shape, np.uint16)
    lib().orc_half_from_float(_p(a), _p(h), ctypes.c_longlong(a.size))
    return h


def half_to_float(h):
    h = np.ascontiguousarray(h, np.uint16)
    a = np.empty(h.shape, np.float32)
    lib().orc_half_to_float(_p(h), _p(a), ctypes.c_longlong(h.size))
    return a


def s16_sr(a, r, storage):
    """Stochastic rounding of float32 values (storage units) with random words r."""
    a = np.ascontiguousarray(a, np.float32)
    r = np.ascontiguousarray(r, np.uint32)
    h = np.empty(a.shape, np.uint16)
    lib().orc_s16_sr(_p(a), _p(r), int(storage), _p(h), ctypes.c_longlong(a.size))
    return h


def sr_hash(i, seed):
    lib().orc_sr_hash.restype = ctypes.c_uint
    return int(lib().orc_sr_hash(ctypes.c_uint(i), ctypes.c_uint(seed)))


def f16_round(a, scale=2.0 ** 14):
    """Values as 16-bit texels holding value * scale store them (round to nearest even)."""
    a = np.asarray(a, np.float32)
    return ((a * np.float32(scale)).astype(np.float16).astype(np.float32) / np.float32(scale)).astype(np.float32)


def tdf_spline(F, rmin, rmax, rstep, arms=12, min_average_mag=0.1, prec="f32"):
    """splineTDF: (tdf, radius) of a 4-plane field (w read by the normalised Hessian)."""
    F = np.ascontiguousarray(F, _dt(prec))
    X, Y, Z = _dims(F)
    tdf = np.empty(F.shape[1:], _dt(prec))
    rad = np.empty(F.shape[1:], _dt(prec))
    _fn("tdf_spline", prec)(_p(F), X, Y, Z, ctypes.c_float(rmin), ctypes.c_float(rmax), ctypes.c_float(rstep),
                            int(arms), ctypes.c_float(min_average_mag), _p(tdf), _p(rad))
    return tdf, rad


def combine(ts, rs, tl, rl, prec="f32"):
    ts, rs = np.ascontiguousarray(ts, _dt(prec)), np.ascontiguousarray(rs, _dt(prec))
    tl, rl = np.array(tl, _dt(prec)), np.array(rl, _dt(prec))
    _fn("combine", prec)(_p(ts), _p(rs), _p(tl), _p(rl), ctypes.c_longlong(ts.size))
    return tl, rl


def candidates(tdf, thr, prec="f32"):
    tdf = np.ascontiguousarray(tdf, _dt(prec))
    m = np.empty(tdf.shape, np.uint8)
    _fn("candidates", prec)(_p(tdf), _p(m), ctypes.c_longlong(tdf.size), ctypes.c_float(thr))
    return m


def hp_compact(mask, method="sort"):
    mask = np.ascontiguousarray(mask, np.uint8)
    X, Y, Z = _dims(mask)
    cap = int(np.count_nonzero(mask))
    pos = np.zeros((max(cap, 1), 3), np.int32)
    n = _fn(f"hp_compact_{method}")(_p(mask), X, Y, Z, _p(pos), ctypes.c_longlong(cap))
    return pos[:n]


def link_centerpoints(vertices, tdf, max_distance, min_mean_tdf):
    """The edge list (n, 2) int32 of runNewCenterlineAlg's linking, in the 2D
    pyramid's order (tsf_oracle.cpp orc_link_centerpoints)."""
    v = np.ascontiguousarray(vertices, np.int32)
    T = np.ascontiguousarray(tdf, np.float32)
    X, Y, Z = _dims(T)
    cap = 2 * len(v)
    out = np.zeros((max(cap, 1), 2), np.int32)
    n = _fn("link_centerpoints")(_p(v), len(v), _p(T), X, Y, Z, ctypes.c_float(max_distance),
                                 ctypes.c_float(min_mean_tdf), _p(out), cap)
    return out[:n]


def graph_components(edges, count):
    """(labels, sizes) int32 of the fixed point of graphComponentLabeling."""
    e = np.ascontiguousarray(edges, np.int32).reshape(-1, 2)
    C = np.zeros(count, np.int32)
    S = np.zeros(count, np.int32)
    _fn("graph_components")(_p(e), len(e), int(count), _p(C), _p(S))
    return C, S


def rasterize_tree_edges(edges, vertices, labels, sizes, min_tree_length, shape, out=None):
    e = np.ascontiguousarray(edges, np.int32).reshape(-1, 2)
    v = np.ascontiguousarray(vertices, np.int32)
    Z, Y, X = shape
    o = np.zeros(shape, np.uint8) if out is None else out
    _fn("rasterize_tree_edges")(_p(e), len(e), _p(v), _p(np.ascontiguousarray(labels, np.int32)),
                                _p(np.ascontiguousarray(sizes, np.int32)), int(min_tree_length), _p(o), X, Y, Z)
    return o


def segment_inverse_gradient(centerlines, radius, F):
    """(segmentation uint8, passes) of runInverseGradientSegmentation;
    F: (4, Z, Y, X) float32 field with w."""
    cl = np.ascontiguousarray(centerlines, np.uint8)
    r = np.ascontiguousarray(radius, np.float32)
    F = np.ascontiguousarray(F, np.float32)
    X, Y, Z = _dims(cl)
    seg = np.zeros(cl.shape, np.uint8)
    passes = _fn("segment_inverse_gradient")(_p(cl), _p(r), _p(F), X, Y, Z, _p(seg))
    return seg, passes


def segment_spheres(centerlines, radius):
    cl = np.ascontiguousarray(centerlines, np.uint8)
    r = np.ascontiguousarray(radius, np.float32)
    X, Y, Z = _dims(cl)
    seg = np.zeros(cl.shape, np.uint8)
    _fn("segment_spheres")(_p(cl), _p(r), X, Y, Z, _p(seg))
    return seg


_TYPES = {np.dtype(np.uint8): 0, np.dtype(np.int8): 1, np.dtype(np.uint16): 2, np.dtype(np.int16): 3,
          np.dtype(np.float32): 4}


def volume_limits(raw):
    """getMinimum / getMaximum of a raw volume (its dtype picks T)."""
    v = np.ascontiguousarray(raw, np.float64)
    mn, mx = ctypes.c_float(0), ctypes.c_float(0)
    _fn("volume_limits")(_p(v), ctypes.c_longlong(v.size), _TYPES[raw.dtype], ctypes.byref(mn), ctypes.byref(mx))
    return mn.value, mx.value


def crop_scan_lines(raw, method, threshold=0.0):
    """(lines_x, lines_y, lines_z) of cropDatasetThreshold (0) / cropDatasetLung (1)."""
    v = np.ascontiguousarray(raw, np.float64)
    X, Y, Z = _dims(raw)
    lx, ly, lz = np.zeros(X, np.int32), np.zeros(Y, np.int32), np.zeros(Z, np.int32)
    _fn("crop_scan_lines")(_p(v), X, Y, Z, _TYPES[raw.dtype], int(method), ctypes.c_float(threshold), _p(lx),
                           _p(ly), _p(lz))
    return lx, ly, lz


def crop_region(lines_x, lines_y, lines_z, min_scan_lines, start_z_middle):
    """((ox, oy, oz), (sx, sy, sz)) or None where the reference cannot crop."""
    lx, ly, lz = (np.ascontiguousarray(a, np.int32) for a in (lines_x, lines_y, lines_z))
    off, size = np.zeros(3, np.int32), np.zeros(3, np.int32)
    r = _fn("crop_region")(_p(lx), _p(ly), _p(lz), lx.size, ly.size, lz.size, int(min_scan_lines),
                           int(bool(start_z_middle)), _p(off), _p(size))
    return None if r != 0 else (tuple(int(a) for a in off), tuple(int(a) for a in size))


def hp_key(x, y, z):
    f = _fn("hp_key")
    f.restype = ctypes.c_ulonglong
    return f(int(x), int(y), int(z))


def candidate_nms(tdf, radius, F, positions, maxd_cap, prec="f32"):
    tdf = np.ascontiguousarray(tdf, _dt(prec))
    radius = np.ascontiguousarray(radius, _dt(prec))
    F = np.ascontiguousarray(F, _dt(prec))
    pos = np.ascontiguousarray(positions, np.int32).reshape(-1, 3)
    X, Y, Z = _dims(F)
    out = np.zeros(tdf.shape, np.uint8)
    _fn("candidate_nms", prec)(_p(tdf), _p(radius), _p(F), X, Y, Z, _p(pos),
                               ctypes.c_longlong(pos.shape[0]), int(maxd_cap), _p(out))
    return out


def cube_max(tdf, cand, cube, prec="f32"):
    tdf = np.ascontiguousarray(tdf, _dt(prec))
    cand = np.ascontiguousarray(cand, np.uint8)
    X, Y, Z = _dims(tdf)
    out = np.zeros(tdf.shape, np.uint8)
    _fn("cube_max", prec)(_p(tdf), _p(cand), X, Y, Z, int(cube), _p(out))
    return out


def do_eigen(Fx, Fy, Fz, pos, prec="f32"):
    Fx, Fy, Fz = (np.ascontiguousarray(a, _dt(prec)) for a in (Fx, Fy, Fz))
    X, Y, Z = _dims(Fx)
    outs = [np.empty(3, _dt(prec)) for _ in range(4)]
    _fn("do_eigen", prec)(_p(Fx), _p(Fy), _p(Fz), X, Y, Z, int(pos[0]), int(pos[1]), int(pos[2]),
                          *[_p(o) for o in outs])
    return tuple(outs)


def trilinear(F, u, v, w, prec="f64"):
    """Sample planes 0..2 of F at unnormalised (u, v, w) with the linear-filter rule."""
    F = np.ascontiguousarray(F, _dt(prec))
    X, Y, Z = _dims(F)
    out = np.empty(3, _dt(prec))
    ct = ctypes.c_float if prec == "f32" else ctypes.c_double
    _fn("trilinear", prec)(_p(F), X, Y, Z, ct(u), ct(v), ct(w), _p(out))
    return out


# ---------------------------------------------------------------- adjoints
_d = ctypes.c_double


def blur_sigma(vol, sigma, prec="f64"):
    """Blur with the mask computed in `prec` (the differentiable chain's blur)."""
    vol = np.ascontiguousarray(vol, _dt(prec))
    out = np.empty_like(vol)
    X, Y, Z = _dims(vol)
    _fn("blur_sigma", prec)(_p(vol), _p(out), X, Y, Z, _d(sigma))
    return out


def blur_adjoint(dB, I, sigma, prec="f64"):
    dB = np.ascontiguousarray(dB, _dt(prec))
    I = np.ascontiguousarray(I, _dt(prec))
    X, Y, Z = _dims(dB)
    dI = np.empty_like(dB)
    ds = _d(0)
    _fn("blur_adjoint", prec)(_p(dB), _p(I), X, Y, Z, _d(sigma), _p(dI), ctypes.byref(ds))
    return dI, ds.value


def vector_field_adjoint(B, dF, fmax, sign=1, prec="f64"):
    B = np.ascontiguousarray(B, _dt(prec))
    dF = np.ascontiguousarray(dF[:3], _dt(prec))
    X, Y, Z = _dims(B)
    dB = np.empty_like(B)
    _fn("vector_field_adjoint", prec)(_p(B), _p(dF), X, Y, Z, ctypes.c_float(fmax), int(sign), _p(dB))
    return dB


def gvf_adjoint(F, mu, iters, dV, prec="f64"):
    F = np.ascontiguousarray(F, _dt(prec))
    dV = np.ascontiguousarray(dV[:3], _dt(prec))
    X, Y, Z = _dims(F)
    dF = np.empty((3,) + F.shape[1:], _dt(prec))
    dm = _d(0)
    _fn("gvf_adjoint", prec)(_p(F), X, Y, Z, _d(mu), int(iters), _p(dV), _p(dF), ctypes.byref(dm))
    return dF, dm.value


def tdf_adjoint(F, rmin, rmax, rstep, gT, prec="f64"):
    F = np.ascontiguousarray(F, _dt(prec))
    gT = np.ascontiguousarray(gT, _dt(prec))
    X, Y, Z = _dims(F)
    dV = np.empty((3,) + F.shape[1:], _dt(prec))
    dW = np.empty(F.shape[1:], _dt(prec))
    _fn("tdf_adjoint", prec)(_p(F), X, Y, Z, ctypes.c_float(rmin), ctypes.c_float(rmax), ctypes.c_float(rstep),
                             _p(gT), _p(dV), _p(dW))
    return dV, dW


def tdf_adjoint_magnitude(F, rmin, rmax, rstep, gT, prec="f64"):
    """The TDF adjoint's scatter of |contribution|: per voxel, the sum of the
    magnitudes that dV (and dW) add up -- the scale of their rounding."""
    F = np.ascontiguousarray(F, _dt(prec))
    gT = np.ascontiguousarray(gT, _dt(prec))
    X, Y, Z = _dims(F)
    dV = np.empty((3,) + F.shape[1:], _dt(prec))
    dW = np.empty(F.shape[1:], _dt(prec))
    _fn("tdf_adjoint_abs", prec)(_p(F), X, Y, Z, ctypes.c_float(rmin), ctypes.c_float(rmax), ctypes.c_float(rstep),
                                 _p(gT), _p(dV), _p(dW))
    return dV, dW


def chain_forward(I, sigma, fmax, sign, mu, iters, rmin, rmax, rstep, prec="f64"):
    I = np.ascontiguousarray(I, _dt(prec))
    X, Y, Z = _dims(I)
    t = np.empty_like(I)
    r = np.empty_like(I)
    _fn("chain_forward", prec)(_p(I), X, Y, Z, _d(sigma), ctypes.c_float(fmax), int(sign), _d(mu), int(iters),
                               ctypes.c_float(rmin), ctypes.c_float(rmax), ctypes.c_float(rstep), _p(t), _p(r))
    return t, r


def chain_gradient(I, sigma, fmax, sign, mu, iters, rmin, rmax, rstep, gT, prec="f64"):
    I = np.ascontiguousarray(I, _dt(prec))
    gT = np.ascontiguousarray(gT, _dt(prec))
    X, Y, Z = _dims(I)
    t = np.empty_like(I)
    dI = np.empty_like(I)
    dm, ds = _d(0), _d(0)
    _fn("chain_gradient", prec)(_p(I), X, Y, Z, _d(sigma), ctypes.c_float(fmax), int(sign), _d(mu), int(iters),
                                ctypes.c_float(rmin), ctypes.c_float(rmax), ctypes.c_float(rstep), _p(gT), _p(t),
                                _p(dI), ctypes.byref(dm), ctypes.byref(ds))
    return t, dI, dm.value, ds.value


# ------------------------------------------------------------- 16-bit mode
def sn16_round(a):
    """Values rounded to the reference's SNORM16 vector format."""
    a = np.array(a, np.float32, copy=True, order="C")
    lib().orc_sn16_round(_p(a), ctypes.c_longlong(a.size))
    return a


def unorm16_round(a):
    """Values rounded to the reference's 16-bit TDF format."""
    a = np.array(a, np.float32, copy=True, order="C")
    lib().orc_unorm16_round(_p(a), ctypes.c_longlong(a.size))
    return a


def sn16_store(a):
    a = np.ascontiguousarray(a, np.float32)
    s = np.empty(a.shape, np.int16)
    lib().orc_sn16_store(_p(a), _p(s), ctypes.c_longlong(a.size))
    return s


def sn16_load(s):
    s = np.ascontiguousarray(s, np.int16)
    a = np.empty(s.shape, np.float32)
    lib().orc_sn16_load(_p(s), _p(a), ctypes.c_longlong(s.size))
    return a


def vector_field_sn16(vol, fmax, sign=1):
    vol = np.ascontiguousarray(vol, np.float32)
    X, Y, Z = _dims(vol)
    F = np.empty((4,) + vol.shape, np.float32)
    lib().orc_vector_field_sn16(_p(vol), _p(F), X, Y, Z, ctypes.c_float(fmax), int(sign))
    return F


def gvf_euler_low_memory(F, mu, iters, sn16=False):
    """runLowMemoryGVF: each component alone with (f_c, |f|^2), |f|^2 and every
    iterate in the field's format (SNORM16 when sn16; F then holds SNORM16
    values)."""
    F = np.ascontiguousarray(F, np.float32)
    X, Y, Z = _dims(F)
    V = np.empty((4,) + F.shape[1:], np.float32)
    lib().orc_gvf_euler_low_memory(_p(F), _p(V), X, Y, Z, ctypes.c_float(mu), int(iters), int(bool(sn16)))
    return V


def gvf_euler_sn16(F, mu, iters):
    """F: (4 or 3, Z, Y, X) float32 holding SNORM16 values."""
    F = np.ascontiguousarray(F, np.float32)
    X, Y, Z = _dims(F)
    V = np.empty((4,) + F.shape[1:], np.float32)
    lib().orc_gvf_euler_sn16(_p(F), _p(V), X, Y, Z, ctypes.c_float(mu), int(iters))
    return V
