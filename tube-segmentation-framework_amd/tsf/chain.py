"""Differentiable tube-detection chain on the device, autograd.Function-style.

    B = blur_sigma(I) -> F = createVectorField(B) -> V = GVF_N(F; mu)
      -> W = |V| -> T = circleFittingTDF(V, W)

`TubeFilterFunction.forward(ctx, I, mu, sigma, ...)` runs the chain (the GVF
with checkpoints every `checkpoint_every` iterations) and saves what the
backward needs in `ctx`; `backward(ctx, gT)` returns (dL/dI, dL/dmu,
dL/dsigma) for L = sum gT * T.  Every step is a call into libtsf_hip.so; the
same class exists in C++ (host/tsf_autograd.hpp).
"""
import numpy as np

from . import _capi as C


class Ctx:
    pass


class TubeFilterFunction:
    @staticmethod
    def forward(ctx, I, mu, sigma, iterations, rmin, rmax, rstep, fmax=0.2, sign=1, checkpoint_every=50,
                stream=None):
        """I: DeviceArray (Z, Y, X) float32.  Returns (T, R) DeviceArrays."""
        shape = I.shape
        B = C.DeviceArray(shape, np.float32)
        C.blur_gaussian(I, B, sigma, stream)
        F = C.Field(shape)
        C.create_vector_field(B, F, fmax, sign, stream)
        V = C.Field(shape)
        k = max(1, min(checkpoint_every, iterations))
        ncp = C.gvf_checkpoint_count(iterations, k)
        cps = C.DeviceArray((ncp, 3) + tuple(shape), np.float32)
        C.gvf_euler_checkpointed(F, V, mu, iterations, k, cps, stream)
        T = C.DeviceArray(shape, np.float32)
        R = C.DeviceArray(shape, np.float32)
        C.tdf_circle(V, rmin, rmax, rstep, T, R, None, stream)
        ctx.saved = dict(I=I, B=B, F=F, V=V, R=R, cps=cps, k=k, mu=mu, sigma=sigma, iterations=iterations, rmin=rmin,
                         rmax=rmax, rstep=rstep, fmax=fmax, sign=sign, shape=shape)
        return T, R

    @staticmethod
    def backward(ctx, gT, stream=None):
        """gT: DeviceArray of dL/dT.  Returns (dI DeviceArray, dmu, dsigma)."""
        s = ctx.saved
        shape = tuple(s["shape"])
        # dV: x, y, z, w planes in one allocation -- a field for the TDF and
        # finish adjoints, the 3-plane input of the GVF adjoint
        dV = C.DeviceArray((4,) + shape, np.float32)
        dV.zero(stream)
        dVf = C.Field.over(dV)
        ws = C.DeviceArray((max(C.gvf_euler_adjoint_workspace_size(shape, s["k"]),
                                C.tdf_adjoint_workspace_size(shape)),), np.uint8)
        C.tdf_adjoint(s["V"], s["rmin"], s["rmax"], s["rstep"], gT, dVf, stream, ws, radius=s["R"])
        C.gvf_finish_adjoint(s["V"], dVf, stream)
        dF = C.DeviceArray((3,) + shape, np.float32)
        dmu = C.gvf_euler_adjoint(s["F"], s["mu"], s["iterations"], s["k"], s["cps"], dV, dF, ws, stream)
        dB = C.DeviceArray(shape, np.float32)
        C.vector_field_adjoint(s["B"], C.Field.over(dF), s["fmax"], s["sign"], dB, ws, stream)
        dI = C.DeviceArray(shape, np.float32)
        dsigma = C.blur_adjoint(dB, s["I"], dI, s["sigma"], True, stream)
        # device memory in use with every buffer of forward and backward live
        free, total = C.mem_info()
        ctx.backward_used_bytes = total - free
        return dI, dmu, dsigma


def fit_mu_sigma(I, target, mu, sigma, steps, iterations, rmin, rmax, rstep, checkpoint_every=40, lr=0.05, fmax=0.2,
                 sign=1, stream=None, log=None):
    """Gradient descent (Adam) on (mu, sigma) minimising mean (T - target)^2,
    T the chain's TDF response of volume I (DeviceArray).  The step size of
    each parameter is lr times its starting value; mu is kept inside the
    explicit scheme's stability bound 12 mu + max|f|^2 <= 2 (|f| <= 1, so
    mu <= 1/12) and sigma inside the blur mask's support (0.25 .. 2.5).
    Returns one record per step: loss (before the update), mu, sigma and the
    gradients, and the step's wall time in ms."""
    import time
    shape = I.shape
    grad = C.DeviceArray(shape, np.float32)
    mws = C.DeviceArray((C.lib().tsf_mse_workspace_size(),), np.uint8)
    theta = np.array([mu, sigma], np.float64)
    scale = lr * np.abs(theta)
    lo, hi = np.array([1e-4, 0.25]), np.array([1.0 / 12.0 - 1e-4, 2.5])
    m = np.zeros(2)
    v = np.zeros(2)
    b1, b2, eps = 0.9, 0.999, 1e-12
    hist = []
    for k in range(1, steps + 1):
        t0 = time.time()
        ctx = Ctx()
        T, _ = TubeFilterFunction.forward(ctx, I, float(theta[0]), float(theta[1]), iterations, rmin, rmax, rstep,
                                          fmax, sign, checkpoint_every, stream)
        loss = C.mse(T, target, grad, mws, stream)
        _, dmu, dsigma = TubeFilterFunction.backward(ctx, grad, stream)
        g = np.array([dmu, dsigma])
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh, vh = m / (1 - b1 ** k), v / (1 - b2 ** k)
        rec = dict(step=k, loss=loss, mu=float(theta[0]), sigma=float(theta[1]), dmu=float(dmu),
                   dsigma=float(dsigma))
        theta = np.clip(theta - scale * mh / (np.sqrt(vh) + eps), lo, hi)
        rec["ms"] = (time.time() - t0) * 1e3
        rec["hbm_bytes"] = getattr(ctx, "backward_used_bytes", None)
        hist.append(rec)
        if log:
            log(f"fit step {k}: loss {loss:.6e} mu {rec['mu']:.5f} sigma {rec['sigma']:.4f} "
                f"dmu {dmu:.4e} dsigma {dsigma:.4e} ({rec['ms']:.0f} ms)")
        del ctx, T
    return hist, float(theta[0]), float(theta[1])
