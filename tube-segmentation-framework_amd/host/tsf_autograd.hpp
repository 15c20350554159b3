// Differentiable tube-detection chain with a torch.autograd.Function-style
// interface (static forward(ctx, ...) saving what backward needs in ctx;
// static backward(ctx, grad_outputs) returning the input gradients).
//
//   B = blur_sigma(I) -> F = createVectorField(B) -> V = GVF_N(F; mu)
//     -> W = |V| -> (T, R) = circleFittingTDF(V, W)
//
// backward(ctx, dL/dT) returns dL/dI, dL/dmu and dL/dsigma for
// L = sum_p dL/dT(p) T(p), with the discrete choices of the forward held
// fixed (blur mask size, the Fmax branch, eigenvalue order, radius R).  The
// GVF runs in F32 with a checkpoint every `checkpoint_every` iterations; the
// backward recomputes one segment at a time, so device memory is
// (iterations / k + k + 3) three-plane states besides the volumes.
// No reference counterpart: the reference computes no gradients.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "tsf_dropin.hpp"

namespace tsf {

struct TubeFilterOptions {
    int iterations = 250;
    float radiusMin = 1.0f, radiusMax = 5.0f, radiusStep = 1.0f;
    float Fmax = 0.2f;        // createVectorField's vector maximum
    int sign = 1;             // -1 for dark tubes ("mode black")
    int checkpoint_every = 50;
};

// The saved state of one forward (as torch's AutogradContext: tensors in
// save_for_backward, scalars in saved_data).
class AutogradContext {
   public:
    void save_for_backward(std::vector<Image3D> v) { saved = std::move(v); }
    const std::vector<Image3D>& get_saved_variables() const { return saved; }
    std::map<std::string, double> saved_data;
    Buffer checkpoints;  // GVF states v^0, v^k, ...
    Buffer workspace;    // backward scratch, sized at forward
    TubeFilterOptions options;

   private:
    std::vector<Image3D> saved;
};

struct TubeFilterGrads {
    Image3D volume;  // dL/dI, 1 channel F32
    double mu = 0.0;
    double sigma = 0.0;
};

class TubeFilterFunction {
   public:
    // I: 1-channel F32 volume.  Returns {TDF, radius} (1-channel F32).
    static std::vector<Image3D> forward(AutogradContext* ctx, OpenCL& ocl, const Image3D& I, float mu, float sigma,
                                        const TubeFilterOptions& options);
    // grad_outputs[0] = dL/dT (1-channel F32); a second entry (dL/dR) is
    // ignored, R being piecewise constant.  The context can serve several
    // backwards.
    static TubeFilterGrads backward(AutogradContext* ctx, OpenCL& ocl, const std::vector<Image3D>& grad_outputs);
    // Bytes of device memory a forward + backward holds at this size.
    static size_t memoryBytes(int3 size, const TubeFilterOptions& options);
};

}  // namespace tsf
