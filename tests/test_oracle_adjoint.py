"""Pins the oracle's FP64 adjoints (the reference gradient of the
differentiation form) against central finite differences of the oracle's own
FP64 forward functions.

Chain: B = blur_sigma(I), F = VF(B), V = GVF_N(F; mu), W = |V|, T = TDF(V, W),
L = sum gT * T.  Discrete choices (mask size, Fmax branch, eigenvalue order,
radius r*) are held fixed; the checks use perturbations far below the
switching points.  Parameters that the forward takes in float (mu in
gvf_euler) are checked through the chain functions, which take them in
double.
"""
import numpy as np
import pytest

H = 1e-6


def rng(s=0):
    return np.random.default_rng(s)


def fd(f, x, h=H):
    return (f(x + h) - f(x - h)) / (2 * h)


def test_blur_adjoint_fd(orc):
    g = rng(1)
    sh = (10, 11, 12)
    I, dB, s = g.standard_normal(sh), g.standard_normal(sh), 0.9
    dI, ds = orc.blur_adjoint(dB, I, s)
    L = lambda I_, s_: float((dB * orc.blur_sigma(I_, s_)).sum())
    for p in [(0, 0, 0), (5, 5, 5), (9, 10, 11), (0, 5, 11)]:
        E = np.zeros(sh)
        E[p] = 1
        assert np.isclose(fd(lambda t: L(I + t * E, s), 0.0), dI[p], rtol=1e-6, atol=1e-9)
    assert np.isclose(fd(lambda t: L(I, s + t), 0.0), ds, rtol=1e-6)


@pytest.mark.parametrize("sign", [1, -1])
def test_vector_field_adjoint_fd(orc, sign):
    g = rng(2)
    sh = (10, 11, 12)
    B = g.standard_normal(sh) * 0.1  # both Fmax branches occur
    dF = g.standard_normal((3,) + sh)
    dB = orc.vector_field_adjoint(B, dF, 0.2, sign)
    L = lambda B_: float((dF * orc.vector_field(B_, 0.2, sign, "f64")[:3]).sum())
    for p in [(0, 0, 0), (5, 5, 5), (9, 10, 11), (3, 0, 7)]:
        E = np.zeros(sh)
        E[p] = 1
        assert np.isclose(fd(lambda t: L(B + t * E), 0.0), dB[p], rtol=1e-6, atol=1e-9)


def test_gvf_adjoint_fd(orc):
    g = rng(3)
    sh = (10, 11, 12)
    F = np.zeros((4,) + sh)
    F[:3] = g.uniform(-1, 1, (3,) + sh) / np.sqrt(3)
    F[3] = 1
    dV = g.standard_normal((3,) + sh)
    mu, it = float(np.float32(0.07)), 6
    dF, dmu = orc.gvf_adjoint(F, mu, it, dV)
    L = lambda F_: float((dV * orc.gvf_euler(F_, mu, it, "f64")[:3]).sum())
    # boundary voxels (mirror transpose), interior, the mirror sources 2 and N-3
    for c, p in [(0, (0, 0, 0)), (1, (5, 5, 5)), (2, (9, 10, 11)), (0, (2, 2, 2)), (1, (1, 5, 7)), (2, (7, 8, 9))]:
        E = np.zeros((4,) + sh)
        E[(c,) + p] = 1
        assert np.isclose(fd(lambda t: L(F + t * E), 0.0), dF[(c,) + p], rtol=1e-6, atol=1e-9)
    # mu enters gvf_euler as float: use a step the float grid resolves
    Lm = lambda m_: float((dV * orc.gvf_euler(F, m_, it, "f64")[:3]).sum())
    assert np.isclose(fd(Lm, mu, 1e-4), dmu, rtol=1e-4)


def _gvf_field(orc):
    from tsf import synth
    vol, _ = synth.tube_volume((16, 18, 20), 5, 1.5, 3.5, 3, noise=0.05)
    F = orc.vector_field(orc.blur_sigma(vol.astype(np.float64), 0.8), 0.2, 1, "f64")
    return orc.gvf_euler(F, 0.05, 30, "f64")


@pytest.mark.parametrize("rmin,rmax,rstep", [(1.0, 3.0, 0.5), (1.0, 5.0, 1.0)])
def test_tdf_adjoint_fd(orc, rmin, rmax, rstep):
    g = rng(4)
    V = _gvf_field(orc)
    t, r = orc.tdf_circle(V, rmin, rmax, rstep, "f64")
    gT = np.zeros(t.shape)
    for i in np.argwhere(t > 0.05)[:40]:
        gT[tuple(i)] = g.standard_normal()
    dV, dW = orc.tdf_adjoint(V, rmin, rmax, rstep, gT)
    L = lambda V_: float((gT * orc.tdf_circle(V_, rmin, rmax, rstep, "f64")[0]).sum())
    checked = 0
    for p in np.argwhere(np.abs(dV).sum(0) > 1e-3)[:6]:
        for c in range(3):
            E = np.zeros(V.shape)
            E[(c,) + tuple(p)] = 1
            assert np.isclose(fd(lambda s: L(V + s * E), 0.0), dV[(c,) + tuple(p)], rtol=1e-5, atol=1e-9)
            checked += 1
    if rmax >= 4:  # normalised Hessian: the w plane carries gradient too
        for p in np.argwhere(np.abs(dW) > 1e-4)[:4]:
            E = np.zeros(V.shape)
            E[(3,) + tuple(p)] = 1
            assert np.isclose(fd(lambda s: L(V + s * E), 0.0), dW[tuple(p)], rtol=1e-5, atol=1e-9)
            checked += 1
    assert checked >= 18


def test_chain_gradient_fd(orc):
    """dL/dI, dL/dmu, dL/dsigma of the whole chain (FP64) by central differences."""
    from tsf import synth
    vol, _ = synth.tube_volume((16, 16, 18), 4, 1.5, 3.0, 11, noise=0.05)
    I = vol.astype(np.float64)
    args = dict(fmax=0.2, sign=1, mu=0.06, iters=12, rmin=1.0, rmax=4.0, rstep=1.0)
    sigma = 0.8
    t, r = orc.chain_forward(I, sigma, **args)
    g = rng(5)
    gT = np.zeros(t.shape)
    for i in np.argwhere(t > 0.05)[:30]:
        gT[tuple(i)] = g.uniform(0.5, 1.5)
    t2, dI, dmu, dsig = orc.chain_gradient(I, sigma, gT=gT, **args)
    assert np.array_equal(t, t2)

    def L(I_=I, s_=sigma, m_=args["mu"]):
        a = dict(args)
        a["mu"] = m_
        return float((gT * orc.chain_forward(I_, s_, **a)[0]).sum())

    for p in np.argwhere(np.abs(dI) > 1e-3 * np.abs(dI).max())[:: max(1, int(np.sum(np.abs(dI) > 0)) // 8)][:8]:
        E = np.zeros(I.shape)
        E[tuple(p)] = 1
        assert np.isclose(fd(lambda s: L(I_=I + s * E), 0.0), dI[tuple(p)], rtol=1e-4, atol=1e-9)
    assert np.isclose(fd(lambda s: L(m_=s), args["mu"]), dmu, rtol=1e-4)
    assert np.isclose(fd(lambda s: L(s_=s), sigma), dsig, rtol=1e-4)


def test_tdf_adjoint_magnitude_bounds_the_adjoint(orc):
    """The |contribution| scatter bounds the adjoint voxel by voxel (triangle
    inequality), is linear in |dL/dT|, and exceeds |dV| where contributions
    of both signs meet."""
    from tsf import synth
    vol, _ = synth.tube_volume((14, 15, 16), 3, 1.5, 3.0, 5, noise=0.05)
    V = orc.gvf_euler(orc.vector_field(orc.blur(vol, 0.8, "f64"), 0.2, 1, "f64"), 0.05, 20, "f64")
    gT = np.random.default_rng(13).uniform(-1.5, 1.5, vol.shape)
    dV, dW = orc.tdf_adjoint(V, 1.0, 4.0, 0.5, gT)
    A, AW = orc.tdf_adjoint_magnitude(V, 1.0, 4.0, 0.5, gT)
    assert np.all(A >= np.abs(dV) * (1 - 1e-12)) and np.all(AW >= np.abs(dW) * (1 - 1e-12))
    A2, _ = orc.tdf_adjoint_magnitude(V, 1.0, 4.0, 0.5, 2.0 * gT)
    assert np.allclose(A2, 2.0 * A, rtol=1e-12, atol=0)
    assert (A > 1.01 * np.abs(dV)).mean() > 0.5
