"""The one-component GVF (tsf_gvf_euler_low_memory, runLowMemoryGVF of
gradientVectorFlow.cpp:1173-1445) against the oracle's restatement: equal
values in F32 (where it is also the full-field GVF's result) and in the
reference's 16-bit mode (|f|^2 stored as SNORM16), on shapes with several
tiles and z chunks, odd sizes and counts, and no iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _f(orc, shape, seed):
    from tsf import synth
    vol, _ = synth.tube_volume(shape, 5, 1.0, 3.0, seed, noise=0.05)
    return orc.vector_field(orc.blur(vol, 0.5, "f32"), 0.2, 1, "f32")


@pytest.mark.parametrize("shape,iters", [((6, 9, 70), 5), ((40, 33, 65), 8), ((5, 5, 5), 0), ((37, 6, 130), 3)])
def test_low_memory_f32(tsf, orc, shape, iters):
    F = _f(orc, shape, 70 + iters)
    f = tsf.Field.from_host(F)
    out, full = tsf.Field(shape), tsf.Field(shape)
    tsf.gvf_euler_low_memory(f, out, 0.05, iters)
    tsf.gvf_euler(f, full, 0.05, iters, fuse=1)
    got = out.download()
    assert np.array_equal(got, orc.gvf_euler_low_memory(F, 0.05, iters))
    assert np.array_equal(got, full.download())


@pytest.mark.parametrize("shape,iters", [((6, 9, 70), 5), ((40, 33, 65), 8), ((5, 5, 5), 0)])
def test_low_memory_sn16(tsf, orc, shape, iters):
    F = orc.sn16_round(_f(orc, shape, 80 + iters))
    f = tsf.Field.from_host(F, tsf.STORAGE_SN16)
    out = tsf.Field(shape, tsf.STORAGE_SN16)
    tsf.gvf_euler_low_memory(f, out, 0.05, iters)
    assert np.array_equal(out.download(), orc.gvf_euler_low_memory(F, 0.05, iters, sn16=True))


def test_low_memory_rejects_16bit_floats(tsf):
    f = tsf.Field((8, 8, 8), tsf.STORAGE_F16, 2.0 ** 14)
    out = tsf.Field((8, 8, 8), tsf.STORAGE_F16, 2.0 ** 14)
    with pytest.raises(tsf.TSFError) as e:
        tsf.gvf_euler_low_memory(f, out, 0.05, 2)
    assert e.value.status == 1
