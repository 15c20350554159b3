"""The benchmark's configurations as GPU parity cases.

Each synthetic workload (tsf.synth.config1 .. config4) runs through the path
on the device and is checked
against the oracle on the same input where the oracle finishes in seconds,
and through size-independent properties at the full size otherwise.

  C1  64^3 tree, blur 0.5, GVF mu 0.1 x 250, TDF radii 1-5: in the
      reference's default 16-bit mode (SNORM16 field and iterates, UNORM16
      TDF), which is where mu = 0.1 is meaningful: the float scheme is
      unstable there (12 mu + |f|^2 > 2) and the 16-bit stores saturate.
  C2  airway tree (dark lumen: vector field sign -1), FMG with 3 levels,
      mu 0.15, 10 iterations (the AAA-Vessels-CT preset), TDF, candidate
      compaction:
      oracle parity on a 64^3 airway, properties on the 256^3 one.
  C3  and C5 (the benchmark's 512^3 forward): tests/test_gpu_precision.py,
      every storage form against the FP64 oracle (mu 0.05: mu 0.1 is
      unstable, see C1).
  C4  the differentiation form at 128^3 is tests/test_gpu_adjoint.py
      (test_chain_finite_differences_128).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def close(a, b, rtol, atol):
    """(count of elements outside the tolerance, L2-relative difference);
    equal values (infinities included) and NaN on both sides agree, a NaN
    on one side does not."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    with np.errstate(invalid="ignore"):
        ok = (a == b) | (np.isnan(a) & np.isnan(b)) | (np.abs(a - b) <= rtol * np.abs(b) + atol)
        fin = np.isfinite(a) & np.isfinite(b)
        rel = float(np.linalg.norm((a - b)[fin]) / max(np.linalg.norm(b[fin]), 1e-30))
    return int((~ok).sum()), rel


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def hp_keys(pos):
    """The pyramid's traversal key of (x, y, z) rows: sum_b (x_b + 2 z_b + 4 y_b) 8^b."""
    key = np.zeros(len(pos), np.uint64)
    for b in range(21):
        x = (pos[:, 0] >> b) & 1
        y = (pos[:, 1] >> b) & 1
        z = (pos[:, 2] >> b) & 1
        key |= ((x + 2 * z + 4 * y).astype(np.uint64) << np.uint64(3 * b))
    return key


def test_config1_reference_16bit_mode(tsf, orc):
    from tsf import synth
    vol, _ = synth.config1()
    shape = vol.shape
    SN16 = tsf.STORAGE_SN16
    d_vol = tsf.DeviceArray.from_host(vol)
    d_b = tsf.DeviceArray(shape, np.float32)
    tsf.blur_gaussian(d_vol, d_b, 0.5)
    b = orc.blur(vol, 0.5, "f32")
    assert np.array_equal(d_b.download(), b)
    F = tsf.Field(shape, SN16)
    tsf.create_vector_field(d_b, F, 0.2, 1)
    Fo = orc.vector_field_sn16(b, 0.2, 1)
    assert np.array_equal(F.download(), Fo)
    V = tsf.Field(shape, SN16)
    tsf.gvf_euler(F, V, 0.1, 250, fuse=2)
    Vo = orc.gvf_euler_sn16(Fo, 0.1, 250)
    got = V.download()
    print(f"C1 16-bit GVF: identical {np.mean(got == Vo):.6f}, saturated {np.mean(np.abs(Vo[:3]) == 1.0):.4f}")
    assert np.array_equal(got, Vo)
    t, r = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    tsf.tdf_circle(V, 1.0, 5.0, 0.5, t, r)
    tsf.tdf_round_unorm16(t)
    to, ro = orc.tdf_circle(Vo, 1.0, 5.0, 0.5, "f32")
    to = orc.unorm16_round(to)
    gt, gr = t.download(), r.download()
    # the FP32 bar, then at most one UNORM16 code of rounding
    nbad, e = close(gt, to, 1e-4, 1.6e-5)
    print(f"C1 16-bit TDF: exact {np.mean(gt == to):.6f}, radius agree {np.mean(gr == ro):.6f}")
    assert nbad == 0, e
    assert np.array_equal(gr, ro)


def _c2_pipeline(tsf, vol, levels=3, mu=0.15, cycles=10, threshold=0.5):
    # gvf-iterations 10: the AAA-Vessels-CT preset (parameters/centerline-gpu/AAA-Vessels-CT:7)
    # dark lumen in bright walls: the reference's "black" mode (sign -1)
    shape = vol.shape
    F = tsf.Field(shape)
    tsf.create_vector_field(tsf.DeviceArray.from_host(vol), F, 0.2, -1)
    V = tsf.Field(shape)
    norms = tsf.gvf_fmg(F, V, mu, cycles, levels=levels, norms=True)
    t, r = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    tsf.tdf_circle(V, 1.0, 5.0, 0.5, t, r)
    m = tsf.DeviceArray(shape, np.uint8)
    tsf.candidates(t, m, threshold)
    cnt = tsf.hp_compact(m)
    pos = tsf.DeviceArray((max(cnt, 1), 3), np.int32)
    tsf.hp_compact(m, pos)
    return F, V, norms, t, r, m, cnt, pos


def test_config2_airway_fmg_64(tsf, orc):
    from tsf import synth
    vol, _ = synth.airway((64, 64, 64), 4, 202)
    F, V, norms, t, r, m, cnt, pos = _c2_pipeline(tsf, vol)
    Fh = F.download()
    assert np.array_equal(Fh, orc.vector_field(vol, 0.2, -1, "f32"))
    Vo = orc.gvf_fmg(Fh, 0.15, 10, 1, 2, 2, 3, "f32")
    nbad, e = close(V.download(), Vo, 1e-4, 1e-6)
    print(f"C2 64^3 FMG: L2 rel {e:.2e}; residual norms {norms[:, -1]}")
    assert nbad == 0, e
    assert np.all(norms[:, 1:] <= norms[:, :-1] * 1.0001)
    # TDF and compaction on the device's own FMG field: bit-exact
    Vg = V.download()
    to, ro = orc.tdf_circle(Vg, 1.0, 5.0, 0.5, "f32")
    nbad, e = close(t.download(), to, 1e-4, 1e-7)
    assert nbad == 0, e
    assert np.array_equal(r.download(), ro)
    mh = m.download()
    assert np.array_equal(mh, orc.candidates(t.download(), 0.5, "f32"))
    po = orc.hp_compact(mh)
    print(f"C2 64^3: {cnt} candidates")
    assert cnt == len(po) and cnt > 0
    assert np.array_equal(pos.download()[:cnt], po)


def test_config2_airway_256_properties(tsf):
    from tsf import synth
    vol, _ = synth.config2()
    F, V, norms, t, r, m, cnt, pos = _c2_pipeline(tsf, vol)
    # FMG: the residual of every component shrinks from cycle to cycle
    print(f"C2 256^3 FMG residual norms: {norms}")
    assert np.all(np.isfinite(norms)) and np.all(norms[:, 1:] <= norms[:, :-1] * 1.0001)
    th, rh = t.download(), r.download()
    assert np.all(np.isfinite(th)) and th.min() >= 0.0
    assert set(np.unique(rh)) <= set(np.arange(0.0, 5.01, 0.5).astype(np.float32).tolist())
    # compaction: every set voxel exactly once, in the pyramid's key order
    mh = m.download()
    assert cnt == int(mh.sum()) and cnt > 0
    p = pos.download()[:cnt]
    assert np.all(mh[p[:, 2], p[:, 1], p[:, 0]] == 1)
    k = hp_keys(p.astype(np.int64))
    assert np.all(k[1:] > k[:-1])
    print(f"C2 256^3: {cnt} candidates")
