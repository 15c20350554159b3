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
        ctx.saved = dict(I=I, B=B, F=F, V=V, cps=cps, k=k, mu=mu, sigma=sigma, iterations=iterations, rmin=rmin,
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
        C.tdf_adjoint(s["V"], s["rmin"], s["rmax"], s["rstep"], gT, dVf, stream)
        C.gvf_finish_adjoint(s["V"], dVf, stream)
        dF = C.DeviceArray((3,) + shape, np.float32)
        ws = C.DeviceArray((C.gvf_euler_adjoint_workspace_size(shape, s["k"]),), np.uint8)
        dmu = C.gvf_euler_adjoint(s["F"], s["mu"], s["iterations"], s["k"], s["cps"], dV, dF, ws, stream)
        dB = C.DeviceArray(shape, np.float32)
        C.vector_field_adjoint(s["B"], C.Field.over(dF), s["fmax"], s["sign"], dB, ws, stream)
        dI = C.DeviceArray(shape, np.float32)
        dsigma = C.blur_adjoint(dB, s["I"], dI, s["sigma"], True, stream)
        return dI, dmu, dsigma
