"""ctypes declarations of include/tsf.h and small numpy-facing helpers."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(os.path.dirname(_HERE), "lib", "libtsf_hip.so")

STORAGE_F32, STORAGE_F16, STORAGE_BF16, STORAGE_SN16 = 0, 1, 2, 3
_STATUS = {0: "OK", 1: "INVALID_ARGUMENT", 2: "DIMENSIONS", 3: "HIP", 4: "OUT_OF_MEMORY", 5: "CAPACITY",
           6: "RANGE"}


class TSFError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class Dims(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32)]


class FieldDesc(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("w", ctypes.c_void_p), ("storage", ctypes.c_int32), ("scale", ctypes.c_float)]


_lib = None
_vp, _i32, _i64, _f32, _sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_size_t
_PF = ctypes.POINTER(FieldDesc)

_SIGS = {
    "tsf_last_error": ([], ctypes.c_char_p),
    "tsf_version": ([], ctypes.c_char_p),
    "tsf_malloc": ([ctypes.POINTER(_vp), _sz], _i32),
    "tsf_free": ([_vp], _i32),
    "tsf_memcpy_htod": ([_vp, _vp, _sz, _vp], _i32),
    "tsf_memcpy_dtoh": ([_vp, _vp, _sz, _vp], _i32),
    "tsf_memcpy_dtod": ([_vp, _vp, _sz, _vp], _i32),
    "tsf_memset": ([_vp, ctypes.c_int, _sz, _vp], _i32),
    "tsf_stream_create": ([ctypes.POINTER(_vp)], _i32),
    "tsf_stream_destroy": ([_vp], _i32),
    "tsf_stream_synchronize": ([_vp], _i32),
    "tsf_event_create": ([ctypes.POINTER(_vp)], _i32),
    "tsf_event_destroy": ([_vp], _i32),
    "tsf_event_record": ([_vp, _vp], _i32),
    "tsf_event_elapsed_ms": ([_vp, _vp, ctypes.POINTER(_f32)], _i32),
    "tsf_stream_wait_event": ([_vp, _vp], _i32),
    "tsf_device_count": ([ctypes.POINTER(ctypes.c_int)], _i32),
    "tsf_mem_info": ([ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)], _i32),
    "tsf_create_blur_mask": ([_f32, _vp, ctypes.POINTER(_i32)], _i32),
    "tsf_blur_gaussian": ([_vp, _vp, Dims, _f32, _vp], _i32),
    "tsf_to_float": ([_vp, _i32, _vp, Dims, _f32, _f32, _vp], _i32),
    "tsf_create_vector_field": ([_vp, _PF, Dims, _f32, _i32, _vp], _i32),
    "tsf_gvf_euler": ([_PF, _PF, Dims, _f32, _i32, _i32, _vp, _sz, _vp], _i32),
    "tsf_gvf_euler_workspace_size": ([Dims, _i32], _sz),
    "tsf_gvf_euler_low_memory": ([_PF, _PF, Dims, _f32, _i32, _vp, _sz, _vp], _i32),
    "tsf_gvf_euler_low_memory_workspace_size": ([Dims, _i32], _sz),
    "tsf_gvf_fmg": ([_PF, _PF, Dims, _f32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _sz, _vp], _i32),
    "tsf_gvf_fmg_workspace_size": ([Dims, _i32], _sz),
    "tsf_hessian_eigen": ([_PF, Dims, _i32, _vp, _vp, _vp], _i32),
    "tsf_do_eigen_points": ([_vp, _vp, _vp, Dims, _vp, _i64, _vp, _vp], _i32),
    "tsf_tdf_circle": ([_PF, Dims, _f32, _f32, _f32, _vp, _vp, _vp, _sz, _vp], _i32),
    "tsf_tdf_circle_workspace_size": ([Dims, _i32], _sz),
    "tsf_combine": ([_vp, _vp, _vp, _vp, _i64, _vp], _i32),
    "tsf_candidates": ([_vp, _vp, _i64, _f32, _vp], _i32),
    "tsf_tdf_round_unorm16": ([_vp, _i64, _vp], _i32),
    "tsf_hp_compact": ([_vp, Dims, _vp, _i64, ctypes.POINTER(_i64), _vp, _sz, _vp], _i32),
    "tsf_hp_compact_workspace_size": ([Dims], _sz),
    "tsf_gvf_euler_sr": ([_PF, _PF, Dims, _f32, _i32, _i32, ctypes.c_uint32, _vp, _sz, _vp], _i32),
    "tsf_gvf_euler_sr_workspace_size": ([Dims, _i32], _sz),
    "tsf_gvf_euler_sr_supported": ([Dims], _i32),
    "tsf_tdf_circle_mixed": ([_PF, _i32, _f32, Dims, _f32, _f32, _f32, _vp, _vp, _vp, _sz, _vp], _i32),
    "tsf_tdf_circle_mixed_workspace_size": ([Dims], _sz),
    "tsf_hp_compact_noalloc": ([_vp, Dims, _vp, _i64, ctypes.POINTER(_i64), _vp], _i32),
    "tsf_candidate_nms": ([_vp, _vp, _PF, Dims, _vp, _i64, _i32, _vp, _vp], _i32),
    "tsf_cube_max": ([_vp, _vp, Dims, _i32, _vp, _vp], _i32),
    "tsf_link_centerpoints": ([_vp, _i32, _vp, Dims, _f32, _f32, _vp, _i32, ctypes.POINTER(_i32), _vp, _sz, _vp],
                              _i32),
    "tsf_link_centerpoints_workspace_size": ([_i32], _sz),
    "tsf_graph_components": ([_vp, _i32, _i32, _vp, _vp, _vp], _i32),
    "tsf_rasterize_tree_edges": ([_vp, _i32, _vp, _vp, _vp, _i32, _vp, Dims, _vp], _i32),
    "tsf_segment_inverse_gradient": ([_vp, _vp, _PF, Dims, _vp, ctypes.POINTER(_i32), _vp, _sz, _vp], _i32),
    "tsf_segment_inverse_gradient_workspace_size": ([Dims], _sz),
    "tsf_segment_spheres": ([_vp, _vp, Dims, _vp, _vp], _i32),
    "tsf_tdf_spline_workspace_size": ([Dims, _i32], _sz),
    "tsf_tdf_spline": ([_PF, Dims, _f32, _f32, _f32, _f32, _vp, _vp, _vp, _sz, _vp], _i32),
    "tsf_volume_limits": ([_vp, _i32, Dims, ctypes.POINTER(_f32), ctypes.POINTER(_f32), _vp], _i32),
    "tsf_crop_scan_lines": ([_vp, _i32, Dims, _i32, _f32, _vp, _vp, _vp, _vp], _i32),
    "tsf_crop_region": ([_vp, _vp, _vp, Dims, _i32, _i32, ctypes.POINTER(Dims), ctypes.POINTER(Dims)], _i32),
    "tsf_to_float_region": ([_vp, _i32, Dims, Dims, _vp, Dims, _f32, _f32, _vp], _i32),
    "tsf_blur_adjoint": ([_vp, _vp, _vp, Dims, _f32, ctypes.POINTER(ctypes.c_double), _vp, _sz, _vp], _i32),
    "tsf_blur_adjoint_workspace_size": ([], _sz),
    "tsf_vector_field_adjoint": ([_vp, _PF, Dims, _f32, _i32, _vp, _vp, _sz, _vp], _i32),
    "tsf_vector_field_adjoint_workspace_size": ([Dims], _sz),
    "tsf_gvf_checkpoint_count": ([_i32, _i32], _i64),
    "tsf_gvf_euler_checkpointed": ([_PF, _PF, Dims, _f32, _i32, _i32, _vp, _sz, _vp], _i32),
    "tsf_gvf_euler_adjoint": ([_PF, Dims, _f32, _i32, _i32, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_double), _vp,
                               _sz, _vp], _i32),
    "tsf_gvf_euler_adjoint_workspace_size": ([Dims, _i32], _sz),
    "tsf_gvf_finish_adjoint": ([_PF, _PF, Dims, _vp], _i32),
    "tsf_mse": ([_vp, _vp, ctypes.c_int64, _vp, ctypes.POINTER(ctypes.c_double), _vp, ctypes.c_size_t, _vp], _i32),
    "tsf_mse_workspace_size": ([], ctypes.c_size_t),
    "tsf_tdf_adjoint": ([_PF, Dims, _f32, _f32, _f32, _vp, _vp, _PF, _vp, ctypes.c_size_t, _vp], _i32),
    "tsf_tdf_adjoint_workspace_size": ([Dims], ctypes.c_size_t),
}


def lib():
    """Load libtsf_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise RuntimeError(f"{library_path} is missing: build it with "
                               f"`python tube-segmentation-framework_amd/build.py` (no CPU fallback exists)")
        L = ctypes.CDLL(library_path)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def _check(status):
    if status != 0:
        msg = lib().tsf_last_error().decode(errors="replace")
        raise TSFError(status, msg)


def device_count():
    n = ctypes.c_int(0)
    st = lib().tsf_device_count(ctypes.byref(n))
    return n.value if st == 0 else 0


def mem_info():
    """(free, total) bytes of device memory."""
    f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _check(lib().tsf_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def storage_dtype(storage):
    return {STORAGE_F32: np.float32, STORAGE_F16: np.float16, STORAGE_BF16: np.uint16, STORAGE_SN16: np.int16}[storage]


def _to_bf16_bits(a):
    """float32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + 0x7FFF
    out = ((u + r) >> 16).astype(np.uint16)
    nan = np.isnan(a)
    out[nan] = 0x7FC0
    return out


def _from_bf16_bits(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


class DeviceArray:
    """A device allocation owned by Python (freed on garbage collection)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        _check(lib().tsf_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None, stream=None):
        a = np.ascontiguousarray(a, dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().tsf_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes, _sp(stream)))

    def download(self, stream=None):
        a = np.empty(self.shape, self.dtype)
        _check(lib().tsf_memcpy_dtoh(a.ctypes.data, self.ptr, self.nbytes, _sp(stream)))
        return a

    def zero(self, stream=None):
        _check(lib().tsf_memset(self.ptr, 0, self.nbytes, _sp(stream)))

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.tsf_free(self.ptr)
            self.ptr = None


def _sp(stream):
    return None if stream is None else stream.handle


class Stream:
    def __init__(self):
        h = _vp()
        _check(lib().tsf_stream_create(ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self):
        _check(lib().tsf_stream_synchronize(self.handle))

    def wait(self, event):
        """Work queued after this waits for `event`."""
        _check(lib().tsf_stream_wait_event(self.handle, event.handle))

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.tsf_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        h = _vp()
        _check(lib().tsf_event_create(ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        _check(lib().tsf_event_record(self.handle, _sp(stream)))

    def elapsed_ms(self, later):
        ms = _f32(0)
        _check(lib().tsf_event_elapsed_ms(self.handle, later.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.tsf_event_destroy(self.handle)
            self.handle = None


class _PlaneView:
    """Plane c of a multi-plane F32 DeviceArray (keeps the base alive)."""

    def __init__(self, base, c):
        self.base = base
        self.shape = base.shape[1:]
        self.dtype = base.dtype
        self.nbytes = base.nbytes // base.shape[0]
        self.ptr = base.ptr + c * self.nbytes

    def download(self, stream=None):
        a = np.empty(self.shape, self.dtype)
        _check(lib().tsf_memcpy_dtoh(a.ctypes.data, self.ptr, self.nbytes, _sp(stream)))
        return a


class Field:
    """Four device planes (x, y, z, w) of one storage type."""

    @classmethod
    def over(cls, base):
        """F32 field over the planes of a (3 or 4, Z, Y, X) DeviceArray."""
        assert base.dtype == np.float32 and base.shape[0] in (3, 4)
        f = cls.__new__(cls)
        f.shape, f.storage, f.scale = tuple(base.shape[1:]), STORAGE_F32, 1.0
        f.planes = [_PlaneView(base, c) for c in range(base.shape[0])]
        return f

    def __init__(self, shape, storage=STORAGE_F32, scale=1.0, with_w=True):
        self.shape = tuple(shape)
        self.storage = storage
        self.scale = float(scale)
        dt = storage_dtype(storage)
        self.planes = [DeviceArray(shape, dt) for _ in range(4 if with_w else 3)]

    def desc(self):
        p = [a.ptr for a in self.planes] + [None] * (4 - len(self.planes))
        return FieldDesc(p[0], p[1], p[2], p[3], self.storage, self.scale)

    @classmethod
    def from_host(cls, F, storage=STORAGE_F32, scale=1.0, stream=None):
        """F: array (4 or 3, Z, Y, X) of float values (scaled/rounded here)."""
        F = np.asarray(F, np.float32)
        f = cls(F.shape[1:], storage, scale, with_w=F.shape[0] == 4)
        for k in range(F.shape[0]):
            f.planes[k].upload(_encode(F[k], storage, scale), stream)
        return f

    def download(self, stream=None):
        return np.stack([_decode(p.download(stream), self.storage, self.scale) for p in self.planes])


def sn16_encode(a):
    """The reference's 16-bit store: convert_short_sat_rte(v * 32767)."""
    x = np.asarray(a, np.float32) * np.float32(32767.0)
    r = np.rint(np.nan_to_num(x, nan=0.0))
    return np.clip(r, -32768, 32767).astype(np.int16)


def sn16_decode(s):
    """The reference's 16-bit load: max(-1, s / 32767)."""
    return np.maximum(np.float32(-1.0), np.asarray(s, np.int16).astype(np.float32) / np.float32(32767.0))


def _encode(a, storage, scale):
    if storage == STORAGE_SN16:
        return sn16_encode(a)
    a = np.asarray(a, np.float32) * np.float32(scale)
    if storage == STORAGE_F32:
        return a
    if storage == STORAGE_F16:
        return a.astype(np.float16)
    return _to_bf16_bits(a)


def _decode(a, storage, scale):
    if storage == STORAGE_SN16:
        return sn16_decode(a)
    if storage == STORAGE_F32:
        v = a.astype(np.float32)
    elif storage == STORAGE_F16:
        v = a.astype(np.float32)
    else:
        v = _from_bf16_bits(a)
    return v / np.float32(scale)


def _dims(shape):
    Z, Y, X = shape[-3:]
    return Dims(int(X), int(Y), int(Z))


def _ws(ws):
    return (None, 0) if ws is None else (ws.ptr, ws.nbytes)


# ------------------------------------------------------------------- operators
def create_blur_mask(sigma):
    mask = np.zeros(1331, np.float32)
    m = _i32(0)
    _check(lib().tsf_create_blur_mask(sigma, mask.ctypes.data, ctypes.byref(m)))
    M = 2 * m.value + 1
    return m.value, mask[:M ** 3].copy()


def blur_gaussian(src, dst, sigma, stream=None):
    _check(lib().tsf_blur_gaussian(src.ptr, dst.ptr, _dims(src.shape), sigma, _sp(stream)))


_TYPE_CODES = {np.dtype(np.uint8): 0, np.dtype(np.int8): 1, np.dtype(np.uint16): 2, np.dtype(np.int16): 3,
               np.dtype(np.float32): 4}


def to_float(src, dst, minimum, maximum, stream=None):
    _check(lib().tsf_to_float(src.ptr, _TYPE_CODES[src.dtype], dst.ptr, _dims(src.shape), minimum, maximum,
                              _sp(stream)))


def volume_limits(src, stream=None):
    """(minimum, maximum) of a raw volume as getMinimum / getMaximum find them."""
    mn, mx = _f32(0.0), _f32(0.0)
    _check(lib().tsf_volume_limits(src.ptr, _TYPE_CODES[src.dtype], _dims(src.shape), ctypes.byref(mn),
                                   ctypes.byref(mx), _sp(stream)))
    return mn.value, mx.value


CROP_THRESHOLD, CROP_LUNG = 0, 1


def crop_scan_lines(src, method, threshold=0.0, stream=None):
    """Scan-line counts (lines_x, lines_y, lines_z) as int32 numpy arrays."""
    Z, Y, X = src.shape
    lx, ly, lz = DeviceArray((X,), np.int32), DeviceArray((Y,), np.int32), DeviceArray((Z,), np.int32)
    _check(lib().tsf_crop_scan_lines(src.ptr, _TYPE_CODES[src.dtype], _dims(src.shape), int(method), threshold,
                                     lx.ptr, ly.ptr, lz.ptr, _sp(stream)))
    return lx.download(stream), ly.download(stream), lz.download(stream)


def crop_region(lines_x, lines_y, lines_z, min_scan_lines, start_z_middle):
    """Host only: ((ox, oy, oz), (sx, sy, sz)) of the crop."""
    lx, ly, lz = (np.ascontiguousarray(a, np.int32) for a in (lines_x, lines_y, lines_z))
    off, size = Dims(0, 0, 0), Dims(0, 0, 0)
    _check(lib().tsf_crop_region(lx.ctypes.data, ly.ctypes.data, lz.ctypes.data, Dims(lx.size, ly.size, lz.size),
                                 int(min_scan_lines), int(bool(start_z_middle)), ctypes.byref(off),
                                 ctypes.byref(size)))
    return (off.x, off.y, off.z), (size.x, size.y, size.z)


def to_float_region(src, offset, dst, minimum, maximum, stream=None):
    """dst (float32, shape (sz, sy, sx)) = toFloat of src from offset (ox, oy, oz)."""
    _check(lib().tsf_to_float_region(src.ptr, _TYPE_CODES[src.dtype], _dims(src.shape), Dims(*offset), dst.ptr,
                                     _dims(dst.shape), minimum, maximum, _sp(stream)))


def create_vector_field(vol, field, fmax, sign=1, stream=None):
    d = field.desc()
    _check(lib().tsf_create_vector_field(vol.ptr, ctypes.byref(d), _dims(vol.shape), fmax, sign, _sp(stream)))


def gvf_euler_workspace_size(shape, storage):
    return lib().tsf_gvf_euler_workspace_size(_dims(shape), storage)


def gvf_euler(f, out, mu, iterations, fuse=1, workspace=None, stream=None):
    fd, od = f.desc(), out.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_gvf_euler(ctypes.byref(fd), ctypes.byref(od), _dims(f.shape), mu, iterations, fuse, w, wb,
                               _sp(stream)))


def gvf_euler_low_memory_workspace_size(shape, storage):
    return lib().tsf_gvf_euler_low_memory_workspace_size(_dims(shape), storage)


def gvf_euler_low_memory(f, out, mu, iterations, workspace=None, stream=None):
    """runLowMemoryGVF: one component at a time (F32 or SNORM16 fields)."""
    fd, od = f.desc(), out.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_gvf_euler_low_memory(ctypes.byref(fd), ctypes.byref(od), _dims(f.shape), mu, iterations, w, wb,
                                          _sp(stream)))


def gvf_fmg_workspace_size(shape, levels):
    return lib().tsf_gvf_fmg_workspace_size(_dims(shape), levels)


def gvf_fmg(f, out, mu, iterations, levels=2, v0=1, v1=2, v2=2, norms=False, workspace=None, stream=None):
    fd, od = f.desc(), out.desc()
    w, wb = _ws(workspace)
    nr = np.zeros(3 * max(iterations, 1), np.float64) if norms else None
    _check(lib().tsf_gvf_fmg(ctypes.byref(fd), ctypes.byref(od), _dims(f.shape), mu, iterations, levels, v0, v1, v2,
                             nr.ctypes.data if norms else None, w, wb, _sp(stream)))
    return nr[:3 * iterations].reshape(3, iterations) if norms else None


def hessian_eigen(field, normalize, lam, vec=None, stream=None):
    d = field.desc()
    _check(lib().tsf_hessian_eigen(ctypes.byref(d), _dims(field.shape), int(normalize), lam.ptr,
                                   vec.ptr if vec is not None else None, _sp(stream)))


def do_eigen_points(fx, fy, fz, positions, out, stream=None):
    n = positions.shape[0]
    _check(lib().tsf_do_eigen_points(fx.ptr, fy.ptr, fz.ptr, _dims(fx.shape), positions.ptr, n, out.ptr,
                                     _sp(stream)))


def tdf_circle_workspace_size(shape, storage):
    return lib().tsf_tdf_circle_workspace_size(_dims(shape), storage)


def tdf_circle(field, rmin, rmax, rstep, tdf, radius, workspace=None, stream=None):
    d = field.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_tdf_circle(ctypes.byref(d), _dims(field.shape), rmin, rmax, rstep, tdf.ptr, radius.ptr, w, wb,
                                _sp(stream)))


def gvf_euler_sr_workspace_size(shape, f32_tail=0):
    return lib().tsf_gvf_euler_sr_workspace_size(_dims(shape), f32_tail)


def gvf_euler_sr_supported(shape):
    """Whether gvf_euler_sr runs on a volume of this (Z, Y, X) shape."""
    return bool(lib().tsf_gvf_euler_sr_supported(_dims(shape)))


def gvf_euler_sr(f, out, mu, iterations, seed=0, f32_tail=0, workspace=None, stream=None):
    """Precision-form GVF: f F16/BF16, iterate streamed in f's storage with
    stochastic rounding, the last f32_tail iterations in float, out (F32
    field with w) the final iterate."""
    df, do = f.desc(), out.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_gvf_euler_sr(ctypes.byref(df), ctypes.byref(do), _dims(f.shape), mu, iterations, f32_tail, seed,
                                  w, wb, _sp(stream)))


def tdf_circle_mixed_workspace_size(shape):
    return lib().tsf_tdf_circle_mixed_workspace_size(_dims(shape))


def tdf_circle_mixed(field, rmin, rmax, rstep, tdf, radius, sample_storage=STORAGE_F16, sample_scale=2.0 ** 14,
                     workspace=None, stream=None):
    """Circle TDF of an F32 field: Hessian from the F32 planes, samples from 16-bit texels."""
    d = field.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_tdf_circle_mixed(ctypes.byref(d), sample_storage, sample_scale, _dims(field.shape), rmin, rmax,
                                      rstep, tdf.ptr, radius.ptr, w, wb, _sp(stream)))


def tdf_spline(field, rmin, rmax, rstep, tdf, radius, min_average_mag=0.1, workspace=None, stream=None):
    d = field.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_tdf_spline(ctypes.byref(d), _dims(field.shape), rmin, rmax, rstep, min_average_mag, tdf.ptr,
                                radius.ptr, w, wb, _sp(stream)))


def combine(ts, rs, tl, rl, stream=None):
    n = int(np.prod(ts.shape))
    _check(lib().tsf_combine(ts.ptr, rs.ptr, tl.ptr, rl.ptr, n, _sp(stream)))


def tdf_round_unorm16(tdf, stream=None):
    """Rounds TDF values in place to the reference's 16-bit TDF format."""
    _check(lib().tsf_tdf_round_unorm16(tdf.ptr, int(np.prod(tdf.shape)), _sp(stream)))


def candidates(tdf, mask, threshold, stream=None):
    n = int(np.prod(tdf.shape))
    _check(lib().tsf_candidates(tdf.ptr, mask.ptr, n, threshold, _sp(stream)))


def hp_compact_workspace_size(shape):
    return lib().tsf_hp_compact_workspace_size(_dims(shape))


def hp_compact(mask, positions=None, workspace=None, stream=None):
    """Returns (count, positions DeviceArray of shape (capacity, 3)); raises on overflow."""
    cnt = _i64(0)
    cap = positions.shape[0] if positions is not None else 0
    w, wb = _ws(workspace)
    _check(lib().tsf_hp_compact(mask.ptr, _dims(mask.shape), positions.ptr if positions is not None else None, cap,
                                ctypes.byref(cnt), w, wb, _sp(stream)))
    return cnt.value


def hp_compact_noalloc(mask, positions=None, stream=None):
    """tsf_hp_compact with the library's static scratch (no device allocation);
    returns the count, raises on overflow."""
    cnt = _i64(0)
    cap = positions.shape[0] if positions is not None else 0
    _check(lib().tsf_hp_compact_noalloc(mask.ptr, _dims(mask.shape), positions.ptr if positions is not None else None,
                                        cap, ctypes.byref(cnt), _sp(stream)))
    return cnt.value


def candidate_nms(tdf, radius, field, positions, count, maxd_cap, out, stream=None):
    d = field.desc()
    _check(lib().tsf_candidate_nms(tdf.ptr, radius.ptr, ctypes.byref(d), _dims(field.shape), positions.ptr, count,
                                   maxd_cap, out.ptr, _sp(stream)))


def cube_max(tdf, cand, cube, out, stream=None):
    _check(lib().tsf_cube_max(tdf.ptr, cand.ptr, _dims(tdf.shape), cube, out.ptr, _sp(stream)))


# ------------------------------------------------------------------- centerline graph
def link_centerpoints_workspace_size(count):
    return lib().tsf_link_centerpoints_workspace_size(int(count))


def link_centerpoints(vertices, count, tdf, max_distance, min_mean_tdf, edges=None, workspace=None, stream=None):
    """vertices: DeviceArray (>= count, 3) int32; tdf: DeviceArray (Z, Y, X) float32.
    Returns the edge count; edges (DeviceArray (capacity, 2) int32) receives them."""
    n = _i32(0)
    cap = edges.shape[0] if edges is not None else 0
    w, wb = _ws(workspace)
    _check(lib().tsf_link_centerpoints(vertices.ptr, int(count), tdf.ptr, _dims(tdf.shape), max_distance,
                                       min_mean_tdf, edges.ptr if edges is not None else None, cap, ctypes.byref(n),
                                       w, wb, _sp(stream)))
    return n.value


def graph_components(edges, n_edges, count, labels, sizes, stream=None):
    _check(lib().tsf_graph_components(edges.ptr if edges is not None else None, int(n_edges), int(count), labels.ptr,
                                      sizes.ptr, _sp(stream)))


def rasterize_tree_edges(edges, n_edges, vertices, labels, sizes, min_tree_length, centerlines, stream=None):
    _check(lib().tsf_rasterize_tree_edges(edges.ptr, int(n_edges), vertices.ptr, labels.ptr, sizes.ptr,
                                          int(min_tree_length), centerlines.ptr, _dims(centerlines.shape),
                                          _sp(stream)))


# ------------------------------------------------------------------- segmentation
def segment_inverse_gradient_workspace_size(shape):
    return lib().tsf_segment_inverse_gradient_workspace_size(_dims(shape))


def segment_inverse_gradient(centerlines, radius, field, out, workspace=None, stream=None):
    """Returns the number of growing passes."""
    it = _i32(0)
    d = field.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_segment_inverse_gradient(centerlines.ptr, radius.ptr, ctypes.byref(d), _dims(out.shape), out.ptr,
                                              ctypes.byref(it), w, wb, _sp(stream)))
    return it.value


def segment_spheres(centerlines, radius, out, stream=None):
    _check(lib().tsf_segment_spheres(centerlines.ptr, radius.ptr, _dims(out.shape), out.ptr, _sp(stream)))


# ------------------------------------------------------------------- adjoints
def blur_adjoint(dB, src, dI, sigma, want_dsigma=True, stream=None):
    ds = ctypes.c_double(0.0)
    _check(lib().tsf_blur_adjoint(dB.ptr, src.ptr if src is not None else None, dI.ptr, _dims(dB.shape), sigma,
                                  ctypes.byref(ds) if want_dsigma else None, None, 0, _sp(stream)))
    return ds.value if want_dsigma else None


def vector_field_adjoint(vol, dF, fmax, sign, dvol, workspace=None, stream=None):
    d = dF.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_vector_field_adjoint(vol.ptr, ctypes.byref(d), _dims(vol.shape), fmax, sign, dvol.ptr, w, wb,
                                          _sp(stream)))


def gvf_checkpoint_count(iterations, every):
    return lib().tsf_gvf_checkpoint_count(iterations, every)


def gvf_euler_checkpointed(f, out, mu, iterations, every, checkpoints, stream=None):
    fd, od = f.desc(), out.desc()
    _check(lib().tsf_gvf_euler_checkpointed(ctypes.byref(fd), ctypes.byref(od), _dims(f.shape), mu, iterations, every,
                                            checkpoints.ptr, checkpoints.nbytes, _sp(stream)))


def gvf_euler_adjoint_workspace_size(shape, every):
    return lib().tsf_gvf_euler_adjoint_workspace_size(_dims(shape), every)


def gvf_euler_adjoint(f, mu, iterations, every, checkpoints, dV, dF, workspace=None, stream=None):
    """dV, dF: DeviceArray (3, Z, Y, X) float32.  Returns dmu."""
    fd = f.desc()
    dm = ctypes.c_double(0.0)
    w, wb = _ws(workspace)
    _check(lib().tsf_gvf_euler_adjoint(ctypes.byref(fd), _dims(f.shape), mu, iterations, every, checkpoints.ptr,
                                       dV.ptr, dF.ptr, ctypes.byref(dm), w, wb, _sp(stream)))
    return dm.value


def gvf_finish_adjoint(v, dv, stream=None):
    a, b = v.desc(), dv.desc()
    _check(lib().tsf_gvf_finish_adjoint(ctypes.byref(a), ctypes.byref(b), _dims(v.shape), _sp(stream)))


def tdf_adjoint_workspace_size(shape):
    return lib().tsf_tdf_adjoint_workspace_size(_dims(shape))


def tdf_adjoint(field, rmin, rmax, rstep, gT, dfield, stream=None, workspace=None, radius=None):
    """radius: the forward's radius output (r* is read from it), or None to
    repeat the forward's radius search."""
    a, b = field.desc(), dfield.desc()
    w, wb = _ws(workspace)
    _check(lib().tsf_tdf_adjoint(ctypes.byref(a), _dims(field.shape), rmin, rmax, rstep, gT.ptr,
                                 radius.ptr if radius is not None else None, ctypes.byref(b), w, wb, _sp(stream)))


def mse(t, target, grad, workspace=None, stream=None):
    """Returns (1/n) sum (t - target)^2 and writes grad = 2 (t - target) / n."""
    loss = ctypes.c_double(0.0)
    w, wb = _ws(workspace)
    _check(lib().tsf_mse(t.ptr, target.ptr, int(np.prod(t.shape)), grad.ptr, ctypes.byref(loss), w, wb,
                         _sp(stream)))
    return loss.value
