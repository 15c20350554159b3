"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/tsf.h declares, and rejects bad arguments before touching a device.
No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "tsf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tsf_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = header_functions()
    for must in ("tsf_gvf_euler", "tsf_gvf_fmg", "tsf_tdf_circle", "tsf_hessian_eigen", "tsf_hp_compact",
                 "tsf_candidate_nms", "tsf_cube_max", "tsf_create_vector_field", "tsf_blur_gaussian"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import tsf
    L = tsf.lib()
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares a signature for every exported entry point
    assert set(header_functions()) <= set(tsf.exported_symbols())


def test_version_and_error_string():
    import tsf
    assert b"gfx950" in tsf.lib().tsf_version()


@pytest.mark.parametrize("sigma", [0.5, 1.0, 1.5, 2.6, 7.0])
def test_blur_mask_host_matches_oracle(orc, sigma):
    import tsf
    m, mask = tsf.create_blur_mask(sigma)
    mo, masko = orc.blur_mask(sigma)
    assert m == mo
    assert np.array_equal(mask, masko)  # same float expression, same libm expf


def test_invalid_arguments_rejected_on_host():
    import tsf
    L = tsf.lib()
    d = tsf.Dims(8, 8, 8)
    st = L.tsf_gvf_euler(None, None, d, 0.1, 10, 1, None, 0, None)
    assert st == 1 and b"NULL" in L.tsf_last_error()
    f = tsf.FieldDesc(1, 2, 3, 4, 0, 1.0)
    st = L.tsf_gvf_euler(ctypes.byref(f), ctypes.byref(f), tsf.Dims(3, 8, 8), 0.1, 10, 1, None, 0, None)
    assert st == 2  # extent below 4
    g = tsf.FieldDesc(1, 2, 3, 4, 1, 3.0)  # scale not a power of two
    st = L.tsf_gvf_euler(ctypes.byref(g), ctypes.byref(f), d, 0.1, 10, 1, None, 0, None)
    assert st == 1 and b"power of two" in L.tsf_last_error()
    st = L.tsf_gvf_euler(ctypes.byref(f), ctypes.byref(f), d, 0.1, -1, 1, None, 0, None)
    assert st == 1
    st = L.tsf_tdf_circle(ctypes.byref(f), d, 1.0, 5.0, 0.0, 1, 1, None, 0, None)
    assert st == 1  # rstep must be > 0
    cnt = ctypes.c_int64(0)
    st = L.tsf_hp_compact(None, d, None, 0, ctypes.byref(cnt), None, 0, None)
    assert st == 1
    st = L.tsf_cube_max(1, 1, d, 0, 1, None)
    assert st == 1
    st = L.tsf_create_vector_field(1, ctypes.byref(f), d, 0.2, 2, None)
    assert st == 1  # sign must be +-1
    st = L.tsf_gvf_fmg(ctypes.byref(g), ctypes.byref(f), d, 0.1, 1, 2, 1, 2, 2, None, None, 0, None)
    assert st == 1


def test_workspace_sizes():
    import tsf
    assert tsf.gvf_euler_workspace_size((64, 64, 64), 0) == 3 * 64 ** 3 * 4
    assert tsf.gvf_euler_workspace_size((64, 64, 64), 1) == 3 * 64 ** 3 * 2
    assert tsf.hp_compact_workspace_size((64, 64, 64)) > 0
    assert tsf.gvf_fmg_workspace_size((64, 64, 64), 2) > 5 * 64 ** 3 * 4


def test_hp_compact_noalloc_rejects_extent_above_scratch():
    """The no-allocation compaction refuses volumes whose largest extent
    exceeds its static scratch (1024) on the host, before any device work."""
    import tsf
    L = tsf.lib()
    cnt = ctypes.c_int64(-1)
    st = L.tsf_hp_compact_noalloc(ctypes.c_void_p(256), tsf.Dims(1025, 8, 8), None, 0, ctypes.byref(cnt), None)
    assert st == 2  # TSF_ERR_DIMENSIONS
    assert b"1024" in L.tsf_last_error()
    st = L.tsf_hp_compact_noalloc(None, tsf.Dims(8, 8, 8), None, 0, ctypes.byref(cnt), None)
    assert st == 1  # NULL mask
