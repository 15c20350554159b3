"""Volume ingest against the oracle: the intensity limits, the scan-line
counts of threshold and lung cropping (every element type, ragged shapes),
the crop region (host code, checked here without a GPU) and toFloat of the
cropped region, bit for bit."""
import numpy as np
import pytest

SHAPES = [(37, 45, 70), (8, 130, 9), (64, 64, 64), (1, 1, 1), (100, 3, 129)]
TYPES = [np.uint8, np.int8, np.uint16, np.int16, np.float32]


def _random(rng, shape, dt):
    if dt == np.float32:
        return rng.normal(0.0, 300.0, size=shape).astype(np.float32)
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dt)


def _lungish(rng, shape, dt):
    """Runs of 33 to 41 values alternating between white and black along every
    axis.  The 16-bit types straddle the lung limit (-150, 874 unsigned); 8-bit
    values all lie on one side of it (int8 white, uint8 black)."""
    Z, Y, X = shape
    z, y, x = np.mgrid[:Z, :Y, :X]
    white = ((x // 37 + y // 41 + z // 33) % 2 == 0) ^ (rng.random(shape) < 0.01)
    hi, lo = {np.int16: (0, -1000), np.uint16: (1024, 24)}.get(dt, (np.iinfo(dt).max, np.iinfo(dt).min))
    return np.where(white, hi, lo).astype(dt)


# ------------------------------------------------------------------ CPU: host code
def test_crop_region_host_vs_oracle(orc):
    import tsf
    rng = np.random.default_rng(11)
    checked = errors = 0
    for trial in range(400):
        X, Y, Z = (int(a) for a in rng.integers(1, 40, 3))
        m = int(rng.integers(0, 6))
        dens = rng.random()
        lines = [np.where(rng.random(n) < dens, rng.integers(0, 12, n), 0).astype(np.int32) for n in (X, Y, Z)]
        for middle in (False, True):
            exp = orc.crop_region(*lines, m, middle)
            if exp is None:
                with pytest.raises(tsf.TSFError) as e:
                    tsf.crop_region(*lines, m, middle)
                assert "Invalid cropping" in str(e.value)
                errors += 1
            else:
                assert tsf.crop_region(*lines, m, middle) == exp, (trial, middle)
                checked += 1
    assert checked > 200 and errors > 20


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("dt", TYPES, ids=lambda d: np.dtype(d).name)
def test_volume_limits_vs_oracle(tsf, orc, dt):
    rng = np.random.default_rng(3)
    for shape in SHAPES + [(512, 512, 17)]:
        v = _random(rng, shape, dt)
        got = tsf.volume_limits(tsf.DeviceArray.from_host(v))
        assert got == orc.volume_limits(v), shape
    if dt == np.float32:
        specials = [np.full((3, 5, 7), -2.0, np.float32), np.full((2, 2, 2), np.nan, np.float32)]
        z = np.zeros((4, 100, 300), np.float32)
        z[1, 50, 7] = -0.0       # the first zero is -0: the minimum keeps its sign
        z[0, 0, 0] = -0.0
        z[2:, :, :] = np.nan
        specials.append(z)
        z2 = z.copy()
        z2[0, 0, 0] = 0.0
        specials.append(z2)
        for v in specials:
            got = tsf.volume_limits(tsf.DeviceArray.from_host(v))
            exp = orc.volume_limits(v)
            assert got == exp and np.signbit(got[0]) == np.signbit(exp[0])


@pytest.mark.gpu
@pytest.mark.parametrize("dt", TYPES, ids=lambda d: np.dtype(d).name)
def test_crop_scan_lines_vs_oracle(tsf, orc, dt):
    rng = np.random.default_rng(7)
    for shape in SHAPES:
        v = _random(rng, shape, dt)
        thr = float(np.quantile(v.astype(np.float64), 0.999)) if v.size > 1 else 0.0
        d = tsf.DeviceArray.from_host(v)
        for got, exp in zip(tsf.crop_scan_lines(d, tsf.CROP_THRESHOLD, thr), orc.crop_scan_lines(v, 0, thr)):
            assert (got == exp).all(), ("threshold", shape)
        if dt == np.float32:
            with pytest.raises(tsf.TSFError):
                tsf.crop_scan_lines(d, tsf.CROP_LUNG)
            continue
        v = _lungish(rng, shape, dt)
        d = tsf.DeviceArray.from_host(v)
        exp = orc.crop_scan_lines(v, 1)
        for got, e in zip(tsf.crop_scan_lines(d, tsf.CROP_LUNG), exp):
            assert (got == e).all(), ("lung", shape)
    print(f"{np.dtype(dt).name}: lung lines inside on the last shape {[int(e.sum()) for e in exp]}")


@pytest.mark.gpu
@pytest.mark.parametrize("dt", TYPES, ids=lambda d: np.dtype(d).name)
def test_to_float_region_vs_oracle(tsf, orc, dt):
    rng = np.random.default_rng(9)
    v = _random(rng, (37, 45, 70), dt)
    mn, mx = orc.volume_limits(v)
    for off, size in [((0, 0, 0), (70, 45, 37)), ((3, 5, 1), (64, 40, 36)), ((69, 44, 36), (1, 1, 1))]:
        ox, oy, oz = off
        sx, sy, sz = size
        out = tsf.DeviceArray((sz, sy, sx), np.float32)
        tsf.to_float_region(tsf.DeviceArray.from_host(v), off, out, mn * 0.5, mx * 0.5)
        exp = orc.to_float(v[oz:oz + sz, oy:oy + sy, ox:ox + sx], mn * 0.5, mx * 0.5)
        assert np.array_equal(out.download().view(np.uint32), exp.view(np.uint32)), off
    # X and the x offset multiples of 4: four voxels per thread
    v4 = _random(rng, (21, 19, 72), dt)
    for off, size in [((0, 0, 0), (72, 19, 21)), ((4, 3, 2), (64, 16, 19)), ((1, 0, 0), (68, 19, 21))]:
        ox, oy, oz = off
        sx, sy, sz = size
        out = tsf.DeviceArray((sz, sy, sx), np.float32)
        tsf.to_float_region(tsf.DeviceArray.from_host(v4), off, out, mn * 0.5, mx * 0.5)
        exp = orc.to_float(v4[oz:oz + sz, oy:oy + sy, ox:ox + sx], mn * 0.5, mx * 0.5)
        assert np.array_equal(out.download().view(np.uint32), exp.view(np.uint32)), off
    out = tsf.DeviceArray((37, 45, 70), np.float32)
    with pytest.raises(tsf.TSFError):
        tsf.to_float_region(tsf.DeviceArray.from_host(v), (1, 0, 0), out, 0.0, 1.0)


def test_ingest_argument_errors_on_host():
    """The ingest entry points reject bad arguments before any device work:
    NULL buffers, unknown element types and methods, lung cropping of a
    float volume, a region outside the volume, empty dimensions."""
    import ctypes
    import tsf
    from tsf._capi import Dims, lib
    L = lib()
    d = Dims(8, 8, 8)
    fake = ctypes.c_void_p(256)  # never dereferenced: every call below fails its checks first
    f = ctypes.c_float()
    assert L.tsf_volume_limits(None, 0, d, ctypes.byref(f), ctypes.byref(f), None) == 1
    assert L.tsf_volume_limits(fake, 9, d, ctypes.byref(f), ctypes.byref(f), None) == 1
    assert L.tsf_volume_limits(fake, 0, Dims(0, 8, 8), ctypes.byref(f), ctypes.byref(f), None) == 2
    assert L.tsf_crop_scan_lines(fake, 0, d, 7, 0.0, fake, fake, fake, None) == 1
    assert L.tsf_crop_scan_lines(fake, 4, d, tsf.CROP_LUNG, 0.0, fake, fake, fake, None) == 1
    assert L.tsf_crop_scan_lines(fake, 0, d, tsf.CROP_THRESHOLD, 0.0, None, fake, fake, None) == 1
    assert L.tsf_to_float_region(fake, 0, d, Dims(1, 0, 0), fake, d, 0.0, 1.0, None) == 2
    assert L.tsf_to_float_region(fake, 0, d, Dims(0, 0, 0), fake, d, 1.0, 1.0, None) == 1
    assert L.tsf_crop_region(None, None, None, d, 0, 0, None, None) == 1
    assert "NULL" in L.tsf_last_error().decode()
