"""The precision forms at the scale of the benchmark's workloads, against FP64.

C3 (512x512x400 CT-like vessels, GVF mu 0.05 x 300) and C5 (512^3 thin-vessel
MRA-like volume, GVF mu 0.05 x 400): the circle-fitting forward of
runCircleFittingMethod (tube-segmentation.cpp:212-939) -- blur 0.5, vector
field, TDF r = 0.5..3 on the vector field, GVF, TDF r = 2.5..4 on the GVF
field, combine -- on the GPU in each storage form and in the oracle's FP64
restatement, on the densest 192^3 block of each volume (the GVF then runs on
that block; both sides see the same input).  Errors are L2-relative:
  field        |V - V64| / |V64| over every voxel;
  eigenvalues  of the normalised Hessian, over the responding voxels
               (FP64 TDF > 0.1);
  TDF          over the responding voxels whose radius is unchanged (the
               first-maximum radius search is discontinuous), and the same
               without the basis-ambiguous voxels, whose two smallest
               |eigenvalues| are within 1e-3 of the largest (there the tube
               direction e1 is ill-conditioned: the FP64 and FP32 results
               already sit on either side of the switch for any
               perturbation of the field; their share is reported).
Also reported, on the combined TDF: the largest |TDF - TDF64| over the voxels
with TDF64 >= tdf-high (0.5, the parameter's default) and the number of
voxels whose candidate mask (TDF >= tdf-high) differs from FP64's.
Forms: F32 (tsf_gvf_euler, four steps per launch, FP32 TDF); the precision
form (tsf_gvf_euler_sr + tsf_tdf_circle_mixed: the 16-bit iterate with
stochastic rounding, a float tail of 32 iterations, the final iterate F32;
the TDF's Hessian from F32, its samples from 16-bit texels; the small
branch's vector field F32) with the GVF's f in SNORM16 and the iterate F16
(sn16, the benchmark's form), with f and iterate F16, and with both BF16;
and the plain F16 form (every field F16, round-to-nearest stores,
tsf_gvf_euler two steps per launch, tsf_tdf_circle), which stalls short of
the fixed point and is reported, not held to the bar.
The 1e-3 bar is the sn16 form's, on every responding voxel with unchanged
radius; jumps (voxels off by more than 1 %: eigenvector flips at nearly
equal eigenvalues or at the tie points of the Householder / QL sign
choices) are counted.  Each form's figures are printed, and one JSON line
per form and volume.

The full 512^3 C5 forward of the benchmark (the sn16 form) also runs once,
with property checks (finite, TDF >= 0, radii on the grid, w > 0, the
candidates compacted in key order).
"""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU = 0.05
TAIL = 32
F16_SCALE = 2.0 ** 14
TDF_HIGH = 0.5  # tdf-high's default (parameters/parameters:14)


def densest_block(vol, n, thr=0.7, step=32):
    """Origin of the n^3 block with the most voxels above thr, among origins
    on a grid `step` apart (n a multiple of step)."""
    Z, Y, X = (s // step for s in vol.shape)
    b = (vol[:Z * step, :Y * step, :X * step] > thr).reshape(Z, step, Y, step, X, step).sum(axis=(1, 3, 5))
    c = np.pad(b.cumsum(0).cumsum(1).cumsum(2), ((1, 0), (1, 0), (1, 0)))
    k = n // step
    s = (c[k:, k:, k:] - c[:-k, k:, k:] - c[k:, :-k, k:] - c[k:, k:, :-k]
         + c[:-k, :-k, k:] + c[:-k, k:, :-k] + c[k:, :-k, :-k] - c[:-k, :-k, :-k])
    z, y, x = np.unravel_index(int(np.argmax(s)), s.shape)
    return z * step, y * step, x * step


@pytest.fixture(scope="module")
def c5_volume():
    from tsf import synth
    vol, _ = synth.config5()
    return vol


@pytest.fixture(scope="module")
def c3_volume():
    from tsf import synth
    vol, _ = synth.config3()
    return vol


def crop(vol, n=192):
    z, y, x = densest_block(vol, n)
    return np.ascontiguousarray(vol[z:z + n, y:y + n, x:x + n])


def forward_fp64(orc, vol, iters):
    b = orc.blur(vol, 0.5, "f64")
    F = orc.vector_field(b, 0.2, 1, "f64")
    ts, rs = orc.tdf_circle(F, 0.5, 3.0, 0.5, "f64")
    V = orc.gvf_euler(F, MU, iters, "f64")
    tl, rl = orc.tdf_circle(V, 2.5, 4.0, 0.5, "f64")
    lam, _ = orc.hessian_eigen(V, 1, False, "f64")
    t, r = orc.combine(ts, rs, tl, rl, "f64")
    return V, lam, (ts, rs), (tl, rl), (t, r)


def forward_gpu(tsf, vol, iters, form):
    shape = vol.shape
    d_vol = tsf.DeviceArray.from_host(vol)
    d_b = tsf.DeviceArray(shape, np.float32)
    tsf.blur_gaussian(d_vol, d_b, 0.5)
    F32 = tsf.Field(shape)
    tsf.create_vector_field(d_b, F32, 0.2, 1)
    ts, rs = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    tl, rl = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    V = tsf.Field(shape)
    if form == "plain_f16":  # every field F16, round to nearest (tsf_gvf_euler, two steps per launch)
        F16 = tsf.Field(shape, tsf.STORAGE_F16, F16_SCALE)
        tsf.create_vector_field(d_b, F16, 0.2, 1)
        tsf.tdf_circle(F16, 0.5, 3.0, 0.5, ts, rs)
        V = tsf.Field(shape, tsf.STORAGE_F16, F16_SCALE)
        tsf.gvf_euler(F16, V, MU, iters, fuse=2)
        tsf.tdf_circle(V, 2.5, 4.0, 0.5, tl, rl)
    elif form == "f32":
        tsf.tdf_circle(F32, 0.5, 3.0, 0.5, ts, rs)
        tsf.gvf_euler(F32, V, MU, iters, fuse=4)
        tsf.tdf_circle(V, 2.5, 4.0, 0.5, tl, rl)
    else:
        S = tsf.STORAGE_BF16 if form == "bf16" else tsf.STORAGE_F16
        scale = 1.0 if form == "bf16" else F16_SCALE
        tsf.tdf_circle_mixed(F32, 0.5, 3.0, 0.5, ts, rs, sample_storage=S, sample_scale=scale)
        F16 = tsf.Field(shape, tsf.STORAGE_SN16) if form == "sn16" else tsf.Field(shape, S, scale)
        tsf.create_vector_field(d_b, F16, 0.2, 1)
        tsf.gvf_euler_sr(F16, V, MU, iters, seed=7, f32_tail=TAIL)
        tsf.tdf_circle_mixed(V, 2.5, 4.0, 0.5, tl, rl, sample_storage=S, sample_scale=scale)
    lam = tsf.DeviceArray((3,) + shape, np.float32)
    tsf.hessian_eigen(V, 1, lam)
    out = (V.download(), lam.download(), (ts.download(), rs.download()), (tl.download(), rl.download()))
    tsf.combine(ts, rs, tl, rl)
    return out + ((tl.download(), rl.download()),)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tdf_errors(t, r, t64, r64, cond):
    """(L2 over the responding voxels with unchanged radius, the same on the
    well-conditioned ones, radius changes, responding voxels, jumps -- voxels
    off by more than 1 % --, L2 without the jumps)."""
    resp = t64 > 0.1
    same = resp & (r == r64)
    good = same & cond
    jump = same & (np.abs(t.astype(np.float64) - t64) > 1e-2 * t64)
    rest = same & ~jump
    return (rel(t[same], t64[same]), rel(t[good], t64[good]), int((resp & (r != r64)).sum()), int(resp.sum()),
            int(jump.sum()), rel(t[rest], t64[rest]))


def precision_table(tsf, orc, vol, iters, label):
    V64, lam64, s64, l64, c64 = forward_fp64(orc, vol, iters)
    a = np.sort(np.abs(lam64), axis=0)
    cond = (a[1] - a[0]) >= 1e-3 * np.maximum(a[2], 1e-30)
    resp = c64[0] > 0.1
    res = {}
    for form in ("f32", "sn16", "f16", "bf16", "plain_f16"):
        V, lam, s, l, c = forward_gpu(tsf, vol, iters, form)
        ev = rel(V[:3], V64[:3])
        el = rel(lam[:, resp], lam64[:, resp])
        es = tdf_errors(s[0], s[1], s64[0], s64[1], cond)
        elg = tdf_errors(l[0], l[1], l64[0], l64[1], cond)
        ec = tdf_errors(c[0], c[1], c64[0], c64[1], cond)
        high = c64[0] >= TDF_HIGH
        mx = float(np.abs(c[0][high].astype(np.float64) - c64[0][high]).max()) if high.any() else 0.0
        flips = int(((c[0] >= TDF_HIGH) != high).sum())
        res[form] = (ev, el, es, elg, ec)
        print("PRECISION " + json.dumps({"volume": label, "form": form,
                                         "tail": TAIL if form in ("sn16", "f16", "bf16") else None,
                                         "field": ev, "eigenvalues": el,
                                         "tdf_small": es[0], "tdf_large": elg[0], "tdf_combined": ec[0],
                                         "tdf_combined_well_conditioned": ec[1], "radius_changed": ec[2],
                                         "responding": ec[3], "jumps_over_1pct": ec[4],
                                         "tdf_combined_without_jumps": ec[5],
                                         "basis_ambiguous": int((resp & ~cond).sum()),
                                         "tdf_max_abs_at_tdf_high": mx, "candidate_mask_flips": flips}))
        print(f"{label} {form:4s}: field {ev:.2e} eigenvalues {el:.2e} | TDF small {es[0]:.2e} ({es[1]:.2e} "
              f"well-cond.) large {elg[0]:.2e} ({elg[1]:.2e}) combined {ec[0]:.2e} ({ec[1]:.2e}); "
              f"radius changed {ec[2]} of {ec[3]} responding; basis-ambiguous {int((resp & ~cond).sum())}; "
              f"jumps > 1 %: {ec[4]}, L2 without them {ec[5]:.2e}; max |dT| at T >= {TDF_HIGH} {mx:.2e}, "
              f"candidate-mask flips {flips}")
    return res


@pytest.mark.timeout(900)
def test_c5_forward_512_properties(tsf, c5_volume):
    from test_gpu_configs import hp_keys
    vol = c5_volume
    shape = vol.shape
    d_vol = tsf.DeviceArray.from_host(vol)
    d_b = tsf.DeviceArray(shape, np.float32)
    tsf.blur_gaussian(d_vol, d_b, 0.5)
    del d_vol
    F32 = tsf.Field(shape)
    tsf.create_vector_field(d_b, F32, 0.2, 1)
    ts, rs = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    tsf.tdf_circle_mixed(F32, 0.5, 3.0, 0.5, ts, rs)
    del F32
    Fsn = tsf.Field(shape, tsf.STORAGE_SN16)
    tsf.create_vector_field(d_b, Fsn, 0.2, 1)
    V = tsf.Field(shape)
    tsf.gvf_euler_sr(Fsn, V, MU, 400, seed=7, f32_tail=TAIL)
    del Fsn
    tl, rl = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    tsf.tdf_circle_mixed(V, 2.5, 4.0, 0.5, tl, rl)
    tsf.combine(ts, rs, tl, rl)
    t, r = tl.download(), rl.download()
    assert np.all(np.isfinite(t)) and t.min() >= 0.0
    assert set(np.unique(r)) <= set(np.arange(0.0, 4.01, 0.5).astype(np.float32).tolist())
    w = V.planes[3].download()
    assert np.all(np.isfinite(w)) and w.min() > 0.0
    m = tsf.DeviceArray(shape, np.uint8)
    tsf.candidates(tl, m, 0.5)
    cnt = tsf.hp_compact(m)
    pos = tsf.DeviceArray((max(cnt, 1), 3), np.int32)
    assert tsf.hp_compact(m, pos) == cnt
    p = pos.download()[:cnt].astype(np.int64)
    mh = m.download()
    assert cnt == int(mh.sum()) > 0
    assert np.all(mh[p[:, 2], p[:, 1], p[:, 0]] == 1)
    k = hp_keys(p)
    assert np.all(k[1:] > k[:-1])
    print(f"C5 512^3: {cnt} candidates, TDF max {t.max():.3f}, responding {(t > 0.1).sum()}")


def check_bars(res):
    ev, el, es, elg, ec = res["f32"]
    assert ev < 1e-5 and ec[0] < 5e-4 and ec[5] < 2e-5
    ev, el, es, elg, ec = res["sn16"]
    assert ev < 1e-4 and ec[0] < 1e-3 and ec[2] < 5e-4 * ec[3]
    ev, el, es, elg, ec = res["f16"]
    assert ev < 4e-4 and ec[0] < 3e-3
    ev, el, es, elg, ec = res["bf16"]
    assert ev < 3e-3 and ec[0] < 3e-2
    # round to nearest stalls the iteration (measured: field 1.5e-3, TDF
    # 1e-2 at C3): the plain 16-bit form does not meet the 1e-3 bar
    ev, el, es, elg, ec = res["plain_f16"]
    assert ev < 5e-3 and ec[0] < 5e-2


@pytest.mark.timeout(900)
def test_c5_precision_vs_fp64(tsf, orc, c5_volume):
    check_bars(precision_table(tsf, orc, crop(c5_volume), 400, "C5"))


@pytest.mark.timeout(900)
def test_c3_precision_vs_fp64(tsf, orc, c3_volume):
    check_bars(precision_table(tsf, orc, crop(c3_volume), 300, "C3"))
