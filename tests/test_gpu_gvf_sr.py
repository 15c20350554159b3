"""Precision-form GVF with stochastic rounding (tsf_gvf_euler_sr) against
the oracle's restatement of the same arithmetic: equal values, on shapes
that exercise several tiles along x and y (overlapping last tiles), several
z chunks, the volume faces, odd counts (a first single step rounded to
nearest), float tails of the last iterations, both 16-bit iterate formats
and f in SNORM16 (storage 3: the iterate in F16 at unit 2^14, the first
launch reading v0 from f's planes)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _f(orc, shape, seed):
    from tsf import synth
    vol, _ = synth.tube_volume(shape, 5, 1.0, 3.0, seed, noise=0.05)
    return orc.vector_field(orc.blur(vol, 0.5, "f32"), 0.2, 1, "f32")


@pytest.mark.parametrize("shape,iters,storage,tail", [
    ((6, 48, 64), 2, 1, 0), ((6, 48, 64), 10, 1, 0), ((4, 50, 66), 7, 1, 0), ((70, 48, 64), 12, 1, 0),
    ((10, 100, 130), 9, 1, 0), ((9, 96, 256), 0, 1, 0), ((9, 96, 256), 1, 1, 0), ((5, 48, 128), 3, 2, 0),
    ((40, 64, 192), 16, 2, 0), ((7, 93, 190), 6, 1, 0), ((9, 96, 256), 10, 1, 4), ((6, 48, 64), 7, 1, 3),
    ((5, 48, 128), 4, 1, 4), ((12, 40, 70), 13, 2, 8), ((6, 48, 64), 2, 3, 0), ((70, 48, 64), 12, 3, 0),
    ((4, 50, 66), 7, 3, 0), ((9, 96, 256), 1, 3, 0), ((9, 96, 256), 0, 3, 0), ((10, 100, 130), 11, 3, 4),
    ((7, 93, 190), 6, 3, 2)])
def test_gvf_sr_matches_oracle(tsf, orc, shape, iters, storage, tail):
    F = _f(orc, shape, 50 + iters)
    scale = 2.0 ** 14 if storage == 1 else 1.0  # ignored for SNORM16
    f = tsf.Field.from_host(F[:3], storage, scale)
    out = tsf.Field(shape)
    tsf.gvf_euler_sr(f, out, 0.05, iters, seed=1234, f32_tail=tail)
    got = out.download()
    codes = np.stack([p.download().view(np.uint16) for p in f.planes])
    exp = orc.gvf_euler_sr16(codes, storage, scale, 0.05, iters, seed=1234, f32_tail=tail)
    bad = got != exp
    print(f"{shape} x{iters} (float tail {tail}) storage {storage}: {int(bad.sum())} of {bad.size} differ")
    assert not bad.any(), np.argwhere(bad)[:5]


def test_gvf_sr_rejects_unsupported_shapes(tsf):
    f = tsf.Field((8, 48, 127), tsf.STORAGE_F16, 2.0 ** 14, with_w=False)
    out = tsf.Field((8, 48, 127))
    with pytest.raises(tsf.TSFError) as e:
        tsf.gvf_euler_sr(f, out, 0.05, 2)
    assert e.value.status == 2
