"""The spline TDF on the device against the oracle: ideal tubes in the three
orientations (FP32), a GVF field of rendered tubes, and 16-bit storage of
the field (the oracle reads the same decoded values)."""
import numpy as np
import pytest

from test_oracle import _spline_tube

pytestmark = pytest.mark.gpu


def run_spline(tsf, fld, rmin, rmax, rstep):
    shape = fld.shape
    t, r = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    tsf.tdf_spline(fld, rmin, rmax, rstep, t, r)
    return t.download(), r.download()


def check(got, exp, what):
    (gt, gr), (et, er) = got, exp
    dt = np.abs(gt.astype(np.float64) - et)
    print(f"{what}: max |dT| {dt.max():.3g}, radii differing {(gr != er).sum()}, T>0 {(et > 0).sum()}")
    assert (dt <= 1e-4 * np.abs(et) + 1e-7).all(), what
    assert np.array_equal(gr, er), what


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_spline_tdf_tubes_vs_oracle(tsf, orc, axis):
    shape = [33, 31, 35]
    shape[axis] = 12
    F = _spline_tube(tuple(shape), 6.0, axis)
    check(run_spline(tsf, tsf.Field.from_host(F), 1.0, 10.0, 0.5), orc.tdf_spline(F, 1.0, 10.0, 0.5), f"axis {axis}")


def _gvf_of_tubes(tsf, shape, iters=40):
    Z, Y, X = shape
    z, y, x = np.mgrid[:Z, :Y, :X].astype(np.float32)
    vol = np.zeros(shape, np.float32)
    for (cy, cz, r) in ((12.0, 14.0, 4.0), (27.0, 20.0, 2.5)):
        vol = np.maximum(vol, (np.sqrt((y - cy) ** 2 + (z - cz) ** 2) < r).astype(np.float32))
    vol += 0.05 * np.random.default_rng(1).random(shape).astype(np.float32)
    d = tsf.DeviceArray.from_host(vol)
    b = tsf.DeviceArray(shape, np.float32)
    tsf.blur_gaussian(d, b, 1.0)
    Fv, V = tsf.Field(shape), tsf.Field(shape)
    tsf.create_vector_field(b, Fv, 0.1, 1)
    tsf.gvf_euler(Fv, V, 0.05, iters)
    return V.download()


def test_spline_tdf_gvf_field_vs_oracle(tsf, orc):
    F = _gvf_of_tubes(tsf, (34, 37, 29))
    check(run_spline(tsf, tsf.Field.from_host(F), 1.5, 8.0, 0.5), orc.tdf_spline(F, 1.5, 8.0, 0.5), "GVF of tubes")


@pytest.mark.parametrize("storage", ["sn16", "f16", "bf16"])
def test_spline_tdf_16bit_fields(tsf, orc, storage):
    S, scale = {"sn16": (tsf.STORAGE_SN16, 1.0), "f16": (tsf.STORAGE_F16, 2.0 ** 14),
                "bf16": (tsf.STORAGE_BF16, 1.0)}[storage]
    F = _spline_tube((33, 31, 12), 6.0, 2) / np.float32(8.0)
    F[3] = np.where(F[3] > 0, F[3], 1.0)
    fld = tsf.Field.from_host(F, S, scale)
    Fd = fld.download()
    check(run_spline(tsf, fld, 1.0, 10.0, 0.5), orc.tdf_spline(Fd, 1.0, 10.0, 0.5), storage)
