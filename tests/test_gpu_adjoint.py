"""GPU adjoints against the FP64 oracle gradient and against central finite
differences of the GPU forward.

Tolerances: each float adjoint stage against the FP64 oracle adjoint on the
same float inputs, L2-relative <= 1e-5 (blur, vector field, GVF) and 1e-3
(TDF, whose sampling positions depend on the float eigenvectors); scalar
gradients d/dmu, d/dsigma within 1e-4 per stage, 1e-3 through the chain;
gradients against central differences (h = 1e-3) within 1e-3 relative.

The TDF is piecewise smooth: the first-maximum radius search switches
radius where two circle responses cross.  Voxels whose radius differs
between the compared forwards are left out of the loss (gT = 0 there), so
that both sides differentiate the same smooth branch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def rng(s):
    return np.random.default_rng(s)


def f32(x):
    return float(np.float32(x))


def test_blur_adjoint(tsf, orc):
    g = rng(1)
    sh = (20, 22, 24)
    I = g.standard_normal(sh).astype(np.float32)
    dB = g.standard_normal(sh).astype(np.float32)
    for sigma in (0.5, 0.9, 1.7):
        dI = tsf.DeviceArray(sh, np.float32)
        ds = tsf.blur_adjoint(tsf.DeviceArray.from_host(dB), tsf.DeviceArray.from_host(I), dI, sigma)
        odI, ods = orc.blur_adjoint(dB.astype(np.float64), I.astype(np.float64), f32(sigma))
        print(f"blur sigma {sigma}: dI rel {rel(dI.download(), odI):.2e}, dsigma {ds:.8g} vs {ods:.8g}")
        assert rel(dI.download(), odI) < 1e-5
        assert abs(ds - ods) <= 1e-4 * abs(ods)


@pytest.mark.parametrize("sign", [1, -1])
def test_vector_field_adjoint(tsf, orc, sign):
    g = rng(2)
    sh = (20, 22, 24)
    B = (g.standard_normal(sh) * 0.1).astype(np.float32)
    dF = g.standard_normal((3,) + sh).astype(np.float32)
    f = tsf.Field.from_host(dF)
    out = tsf.DeviceArray(sh, np.float32)
    tsf.vector_field_adjoint(tsf.DeviceArray.from_host(B), f, 0.2, sign, out)
    exp = orc.vector_field_adjoint(B.astype(np.float64), dF.astype(np.float64), 0.2, sign)
    assert rel(out.download(), exp) < 1e-5


@pytest.mark.parametrize("iters,every,sh", [(10, 3, (18, 20, 22)), (12, 12, (18, 20, 22)), (7, 50, (18, 20, 22)),
                                           (1, 1, (18, 20, 22)), (5, 2, (70, 19, 130)), (4, 4, (5, 9, 4))])
def test_gvf_adjoint_checkpointed(tsf, orc, iters, every, sh):
    # shapes (Z, Y, X): one tile; several x/y tiles and z chunks; the smallest
    g = rng(3)
    F = np.zeros((4,) + sh, np.float32)
    F[:3] = g.uniform(-1, 1, (3,) + sh) / np.sqrt(3)
    F[3] = 1
    dV = g.standard_normal((3,) + sh).astype(np.float32)
    mu = 0.07
    f = tsf.Field.from_host(F)
    V = tsf.Field(sh)
    k = min(every, iters)
    cps = tsf.DeviceArray((tsf.gvf_checkpoint_count(iters, k), 3) + sh, np.float32)
    tsf.gvf_euler_checkpointed(f, V, mu, iters, k, cps)
    # the checkpointed forward is the plain forward
    V2 = tsf.Field(sh)
    tsf.gvf_euler(f, V2, mu, iters)
    assert np.array_equal(V.download(), V2.download())
    dF = tsf.DeviceArray((3,) + sh, np.float32)
    dmu = tsf.gvf_euler_adjoint(f, mu, iters, k, cps, tsf.DeviceArray.from_host(dV), dF)
    odF, odmu = orc.gvf_adjoint(F.astype(np.float64), f32(mu), iters, dV.astype(np.float64))
    print(f"GVF adjoint {iters}/{k}: dF rel {rel(dF.download(), odF):.2e}, dmu {dmu:.8g} vs {odmu:.8g}")
    assert rel(dF.download(), odF) < 1e-5
    assert abs(dmu - odmu) <= 1e-4 * abs(odmu)


@pytest.mark.parametrize("sh", [(18, 20, 22), (70, 19, 130), (9, 6, 5)])
def test_gvf_adjoint_recompute_bit_exact(tsf, sh):
    # every state stored (checkpoint every iteration, nothing recomputed,
    # every reverse step a single launch) against segments of 2 and 11
    # recomputed by four-step launches and single steps and reversed two
    # steps per launch (the last pair of segment 0 ends at v^0 = f): the
    # recomputed states are the stored ones and a two-step launch is two
    # single ones bit for bit, so dF is identical; dmu sums the same terms
    # in double over other blocks, so it agrees to double rounding
    g = rng(5)
    F = np.zeros((4,) + sh, np.float32)
    F[:3] = g.uniform(-1, 1, (3,) + sh) / np.sqrt(3)
    F[3] = 1
    dV = tsf.DeviceArray.from_host(g.standard_normal((3,) + sh).astype(np.float32))
    f = tsf.Field.from_host(F)
    iters, mu = 23, 0.07
    out = []
    for k in (1, 2, 11):
        V = tsf.Field(sh)
        cps = tsf.DeviceArray((tsf.gvf_checkpoint_count(iters, k), 3) + sh, np.float32)
        tsf.gvf_euler_checkpointed(f, V, mu, iters, k, cps)
        dF = tsf.DeviceArray((3,) + sh, np.float32)
        dmu = tsf.gvf_euler_adjoint(f, mu, iters, k, cps, dV, dF)
        out.append((dF.download(), dmu))
    for k, (dF, dmu) in zip((2, 11), out[1:]):
        print(f"k {k}: dF identical {np.mean(dF == out[0][0]):.6f}, dmu {dmu!r} vs {out[0][1]!r}")
        assert np.array_equal(dF, out[0][0])
        assert abs(dmu - out[0][1]) <= 1e-12 * abs(out[0][1])


def test_gvf_finish_adjoint(tsf):
    g = rng(7)
    sh = (6, 7, 8)
    V = g.standard_normal((3,) + sh).astype(np.float32)
    V[:, 0, 0, 0] = 0  # |v| = 0 there: w = 1, no gradient
    W = np.sqrt((V.astype(np.float64) ** 2).sum(0))
    d = g.standard_normal((4,) + sh).astype(np.float32)
    v = tsf.Field.from_host(np.concatenate([V, W[None].astype(np.float32)]))
    dv = tsf.Field.from_host(d)
    tsf.gvf_finish_adjoint(v, dv)
    got = dv.download()
    Wn = np.where(W > 0, W, 1.0)
    exp = d[:3] + np.where(W > 0, d[3] / Wn, 0.0) * V
    assert rel(got[:3], exp) < 1e-6


@pytest.mark.parametrize("rmin,rmax,rstep", [(1.0, 3.0, 0.5), (1.0, 5.0, 1.0)])
def test_tdf_adjoint(tsf, orc, rmin, rmax, rstep):
    from tsf import synth
    vol, _ = synth.tube_volume((24, 26, 28), 5, 1.5, 3.5, 3, noise=0.05)
    V = orc.gvf_euler(orc.vector_field(orc.blur(vol, 0.8, "f32"), 0.2, 1, "f32"), 0.05, 30, "f32")
    fld = tsf.Field.from_host(V)
    T = tsf.DeviceArray(vol.shape, np.float32)
    R = tsf.DeviceArray(vol.shape, np.float32)
    tsf.tdf_circle(fld, rmin, rmax, rstep, T, R)
    t, r = T.download(), R.download()
    _, r64 = orc.tdf_circle(V.astype(np.float64), rmin, rmax, rstep, "f64")
    same = r64 == r
    print(f"TDF radius agreement float/double: {same.mean():.5f}")
    assert same.mean() > 0.99
    g = rng(4)
    gT = np.where((t > 0.02) & same, g.uniform(0.5, 1.5, t.shape), 0.0).astype(np.float32)
    d = tsf.Field(V.shape[1:])
    for p in d.planes:
        p.zero()
    tsf.tdf_adjoint(fld, rmin, rmax, rstep, tsf.DeviceArray.from_host(gT), d)
    got = d.download()
    odV, odW = orc.tdf_adjoint(V.astype(np.float64), rmin, rmax, rstep, gT.astype(np.float64))
    e = rel(got[:3], odV)
    ew = rel(got[3], odW) if rmax >= 4 else 0.0
    print(f"TDF adjoint vs FP64: dV {e:.2e}, dW {ew:.2e}")
    assert e < 1e-3
    assert ew < 1e-3
    # r* read from the forward's radius output instead of a repeated search:
    # the same contributions (only the order of the float atomics differs),
    # also with gradient on voxels without response (T = 0, no r*)
    for gg in (gT, np.ones_like(gT)):
        outs = []
        for rad in (None, R):
            d2 = tsf.Field(V.shape[1:])
            for p in d2.planes:
                p.zero()
            tsf.tdf_adjoint(fld, rmin, rmax, rstep, tsf.DeviceArray.from_host(gg), d2, radius=rad)
            outs.append(d2.download())
        er = rel(outs[1], outs[0])
        print(f"TDF adjoint with the forward's radius vs repeated search: {er:.2e}")
        assert er < 1e-6


def test_tdf_adjoint_gradient_spanning_six_decades(tsf, orc):
    """dL/dT spread over six decades inside every brick (10^-6u, u uniform).
    The scatter tile's fixed point is scaled by the brick's largest |dL/dT|;
    its quantum must stay far below the float rounding of the values it sums,
    also in cells fed only by the brick's small gradients.  Per voxel, the
    error against FP64 is bounded by 1e-3 of the sum of the magnitudes of
    the contributions it adds up (the oracle's |contribution| scatter; sums
    that cancel are judged against what they summed), in every decade of
    that magnitude.  Measured: 2e-5 to 4e-4 across ten decades, the float
    contributions' own level; with a quantum of 2^-31 of the brick's largest
    gradient the error grew to the whole value in the smallest decades."""
    from tsf import synth
    vol, _ = synth.tube_volume((24, 26, 28), 5, 1.5, 3.5, 3, noise=0.05)
    V = orc.gvf_euler(orc.vector_field(orc.blur(vol, 0.8, "f32"), 0.2, 1, "f32"), 0.05, 30, "f32")
    fld = tsf.Field.from_host(V)
    rmin, rmax, rstep = 1.0, 3.5, 0.5
    T = tsf.DeviceArray(vol.shape, np.float32)
    R = tsf.DeviceArray(vol.shape, np.float32)
    tsf.tdf_circle(fld, rmin, rmax, rstep, T, R)
    t, r = T.download(), R.download()
    _, r64 = orc.tdf_circle(V.astype(np.float64), rmin, rmax, rstep, "f64")
    g = rng(12)
    gT = np.where((t > 0.02) & (r64 == r), 10.0 ** (-6.0 * g.random(t.shape)), 0.0).astype(np.float32)
    d = tsf.Field(V.shape[1:])
    for p in d.planes:
        p.zero()
    tsf.tdf_adjoint(fld, rmin, rmax, rstep, tsf.DeviceArray.from_host(gT), d, radius=R)
    got = d.download()[:3].astype(np.float64)
    V64, g64 = V.astype(np.float64), gT.astype(np.float64)
    odV, _ = orc.tdf_adjoint(V64, rmin, rmax, rstep, g64)
    A, _ = orc.tdf_adjoint_magnitude(V64, rmin, rmax, rstep, g64)
    q = np.abs(got - odV)[A > 0] / A[A > 0]
    a = A[A > 0]
    for k in range(0, 10):
        sel = (a >= 10.0 ** (-k - 1)) & (a < 10.0 ** -k)
        if sel.sum():
            print(f"scale in [1e-{k + 1}, 1e-{k}): {int(sel.sum())} values, error / scale: max {q[sel].max():.2e}, "
                  f"99.9 % {np.quantile(q[sel], 0.999):.2e}")
    assert q.max() < 1e-3


def test_tdf_adjoint_linear_in_gradient_scale(tsf, orc):
    # the scatter tile accumulates in fixed point scaled by each brick's
    # largest |dL/dT|: the gradient must still be linear in dL/dT, for a
    # scale that moves the fixed-point scale (2^20: the same integers) and
    # one below the tile's range (2^-110: the float atomics)
    from tsf import synth
    vol, _ = synth.tube_volume((24, 26, 28), 5, 1.5, 3.5, 3, noise=0.05)
    V = orc.gvf_euler(orc.vector_field(orc.blur(vol, 0.8, "f32"), 0.2, 1, "f32"), 0.05, 30, "f32")
    fld = tsf.Field.from_host(V)
    rmin, rmax, rstep = 1.0, 3.5, 0.5
    T = tsf.DeviceArray(vol.shape, np.float32)
    R = tsf.DeviceArray(vol.shape, np.float32)
    tsf.tdf_circle(fld, rmin, rmax, rstep, T, R)
    g = rng(9)
    gT = np.where(T.download() > 0, g.uniform(-1.5, 1.5, vol.shape), 0.0).astype(np.float32)
    out = {}
    for scale in (1.0, 2.0 ** 20, 2.0 ** -110):
        d = tsf.Field(V.shape[1:])
        for p in d.planes:
            p.zero()
        tsf.tdf_adjoint(fld, rmin, rmax, rstep, tsf.DeviceArray.from_host((gT * scale).astype(np.float32)), d,
                        radius=R)
        out[scale] = d.download().astype(np.float64) / scale
    for scale in (2.0 ** 20, 2.0 ** -110):
        e = rel(out[scale], out[1.0])
        print(f"TDF adjoint at dL/dT x {scale:g}, rescaled: {e:.2e}")
        assert np.all(np.isfinite(out[scale]))
        assert e < 1e-6


def _forward(tsf, I, **kw):
    from tsf.chain import Ctx, TubeFilterFunction
    ctx = Ctx()
    T, R = TubeFilterFunction.forward(ctx, tsf.DeviceArray.from_host(I), **kw)
    return ctx, T.download(), R.download()


def _backward(tsf, ctx, gT):
    from tsf.chain import TubeFilterFunction
    dI, dmu, dsig = TubeFilterFunction.backward(ctx, tsf.DeviceArray.from_host(gT.astype(np.float32)))
    return dI.download(), dmu, dsig


def test_chain_gradient_vs_fp64_oracle(tsf, orc):
    from tsf import synth
    vol, _ = synth.tube_volume((32, 32, 32), 6, 1.0, 3.0, 401, noise=0.1)
    mu, sigma, iters = 0.06, 0.8, 40
    kw = dict(mu=mu, sigma=sigma, iterations=iters, rmin=1.0, rmax=5.0, rstep=1.0, checkpoint_every=10)
    ctx, t, r = _forward(tsf, vol, **kw)
    _, r64 = orc.chain_forward(vol.astype(np.float64), f32(sigma), 0.2, 1, f32(mu), iters, 1.0, 5.0, 1.0)
    same = r64 == r
    print(f"chain radius agreement float/double: {same.mean():.5f}")
    assert same.mean() > 0.99
    g = rng(5)
    gT = np.where((t > 0.05) & same, g.uniform(0.5, 1.5, t.shape), 0.0)
    dI, dmu, dsig = _backward(tsf, ctx, gT)
    _, odI, odmu, odsig = orc.chain_gradient(vol.astype(np.float64), f32(sigma), 0.2, 1, f32(mu), iters,
                                             1.0, 5.0, 1.0, gT.astype(np.float32).astype(np.float64))
    e = rel(dI, odI)
    print(f"chain vs FP64: dI rel {e:.2e}, dmu {dmu:.8g} vs {odmu:.8g}, dsigma {dsig:.8g} vs {odsig:.8g}")
    assert e < 1e-3
    assert abs(dmu - odmu) <= 1e-3 * abs(odmu)
    assert abs(dsig - odsig) <= 1e-3 * abs(odsig)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("seed", [401, 402, 403, 404])
def test_chain_finite_differences_128(tsf, orc, seed):
    """Differentiation form at 128^3 on the four seeded volumes of the
    benchmark's C4 (tube tree r ~ U(1, 5), noise 0.1): 500 GVF iterations
    checkpointed every 50, TDF radii 1..5.  mu is 0.05: mu = 0.1 breaks the
    explicit scheme's stability bound 12 mu + |f|^2 <= 2 where |f| = 1, and
    the field diverges (in FP64 as well).

    The gradient of the float GPU chain (dL/dmu, dL/dsigma, and dL/dvolume
    along a signed random subset of 64 voxels) against central differences of
    the FP64 forward with h = 1e-4.  Voxels whose radius switches, or whose
    response has a kink inside [-h, h] (a second difference above
    (h / 1e-3)^2 * 1e-4, i.e. a curvature beyond 100), are left out of the
    loss, and so are voxels whose Hessian eigenvalues lie within 1e-2 of each
    other (the eigenvector derivative ~ 1/gap amplifies the float / double
    difference of the field there).  With h = 1e-3 a kink inside the
    interval can pass the detector: FP64 analytic against FP64 differences
    then disagree by up to 8e-3 on seed 403, and agree to 3e-7 at 1e-4."""
    from tsf import synth
    vol, _ = synth.config4(seeds=(seed,))[0]
    mu, sigma, h = 0.05, 0.8, 1e-4
    kw = dict(iterations=500, rmin=1.0, rmax=5.0, rstep=1.0, checkpoint_every=50)
    ctx, t0, r0 = _forward(tsf, vol, mu=mu, sigma=sigma, **kw)
    g = rng(6)
    gT = np.where(t0 > 0.05, g.uniform(0.5, 1.5, t0.shape), 0.0)
    V64 = vol.astype(np.float64)

    def fwd64(I, m, sg):
        return orc.chain_forward(I, sg, 0.2, 1, m, 500, 1.0, 5.0, 1.0, "f64")

    t64, r64 = fwd64(V64, float(f32(mu)), float(f32(sigma)))
    smooth = (gT != 0) & (r64 == r0)
    kthr = 1e-4 * (h / 1e-3) ** 2

    def kinks(name, tp, tm, rp, rm):
        flip = (rp != r64) | (rm != r64)
        kink = np.abs(tp - 2.0 * t64 + tm) > kthr
        print(f"{name}: radius switches {int((flip & (gT != 0)).sum())}, "
              f"other kinks {int((kink & ~flip & (gT != 0)).sum())}")
        return ~flip & ~kink

    # volume direction: 64 voxels of large |dL/dI|, signed as the gradient
    dI0, _, _ = _backward(tsf, ctx, gT)
    nz = np.flatnonzero(np.abs(dI0) >= np.median(np.abs(dI0[dI0 != 0])))
    sub = g.choice(nz, 64, replace=False)
    D = np.zeros(vol.size)
    D[sub] = np.sign(dI0.ravel()[sub])
    D = D.reshape(vol.shape)
    diffs = {}
    m64, s64 = float(f32(mu)), float(f32(sigma))
    for name, a, b, dh in (("mu", (V64, m64 + h, s64), (V64, m64 - h, s64), 2 * h),
                           ("sigma", (V64, m64, s64 + h), (V64, m64, s64 - h), 2 * h),
                           ("volume", (V64 + h * D, m64, s64), (V64 - h * D, m64, s64), None)):
        tp, rp = fwd64(*a)
        tm, rm = fwd64(*b)
        smooth &= kinks(name, tp, tm, rp, rm)
        diffs[name] = (tp - tm, dh)
    lam = tsf.DeviceArray((3,) + vol.shape, np.float32)
    tsf.hessian_eigen(ctx.saved["V"], True, lam)
    L = lam.download()
    gap = np.minimum(np.abs(L[1] - L[0]), np.minimum(np.abs(L[1] - L[2]), np.abs(L[2] - L[0])))
    print(f"eigenvalue gap < 1e-2: {int(((gap < 1e-2) & (gT != 0)).sum())} voxels")
    smooth &= gap >= 1e-2
    print(f"voxels in loss: {int(smooth.sum())} of {int((gT != 0).sum())}")
    assert smooth.sum() > 0.8 * (gT != 0).sum()
    gTm = np.where(smooth, gT, 0.0).astype(np.float32).astype(np.float64)
    dI, dmu, dsig = _backward(tsf, ctx, gTm)  # the saved forward serves any number of backwards
    an = {"mu": dmu, "sigma": dsig, "volume": float((dI.astype(np.float64) * (2 * h * D)).sum())}
    for name, (dt, dh) in diffs.items():
        dL = float((gTm * dt).sum())
        fd = dL if dh is None else dL / dh
        print(f"FD {name}: analytic {an[name]:.8g} vs central difference {fd:.8g} "
              f"(rel {abs(an[name] - fd) / abs(fd):.2e})")
        assert abs(an[name] - fd) <= 1e-3 * abs(fd)
