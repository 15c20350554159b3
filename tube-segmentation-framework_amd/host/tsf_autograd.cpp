// TubeFilterFunction: forward and backward as sequences of tsf_* calls on
// the context's stream.  Device buffers of the backward are laid out so no
// copies are needed between stages: dV is one 4-plane allocation (x, y, z,
// w contiguous) that the TDF adjoint fills as a field and the GVF adjoint
// reads as its 3-plane input; dF is one 3-plane allocation read back as a
// field by the vector-field adjoint.
#include "tsf_autograd.hpp"

#include <algorithm>

namespace tsf {

static size_t voxels(int3 s) { return (size_t)s.x * (size_t)s.y * (size_t)s.z; }

static tsf_field planes_field(void* base, size_t n, bool with_w) {
    float* p = static_cast<float*>(base);
    return tsf_field{p, p + n, p + 2 * n, with_w ? p + 3 * n : nullptr, TSF_STORAGE_F32, 1.0f};
}

static int clamp_every(const TubeFilterOptions& o) { return std::max(1, std::min(o.checkpoint_every, o.iterations)); }

static size_t backward_workspace(int3 size, int k) {
    const tsf_dims d{size.x, size.y, size.z};
    size_t w = tsf_gvf_euler_adjoint_workspace_size(d, k);
    w = std::max(w, tsf_vector_field_adjoint_workspace_size(d));
    w = std::max(w, tsf_blur_adjoint_workspace_size());
    w = std::max(w, tsf_tdf_adjoint_workspace_size(d));
    return w;
}

size_t TubeFilterFunction::memoryBytes(int3 size, const TubeFilterOptions& o) {
    const size_t N = voxels(size), k = (size_t)clamp_every(o);
    const size_t C = (size_t)tsf_gvf_checkpoint_count(o.iterations, (int)k);
    // I, B, F(4), V(4), T, R, checkpoints, backward workspace, dV(4), dF(3), dB, dI
    return N * 4 * (1 + 1 + 4 + 4 + 1 + 1 + 4 + 3 + 1 + 1) + C * 3 * N * 4 + backward_workspace(size, (int)k);
}

std::vector<Image3D> TubeFilterFunction::forward(AutogradContext* ctx, OpenCL& ocl, const Image3D& I, float mu,
                                                 float sigma, const TubeFilterOptions& o) {
    if (!ctx) throw Error(TSF_ERR_INVALID_ARGUMENT, "TubeFilterFunction::forward: NULL context");
    if (!I.valid() || I.channels() != 1 || I.storage() != TSF_STORAGE_F32)
        throw Error(TSF_ERR_INVALID_ARGUMENT, "TubeFilterFunction::forward: I must be a 1-channel F32 volume");
    if (o.iterations < 1) throw Error(TSF_ERR_INVALID_ARGUMENT, "TubeFilterFunction::forward: iterations < 1");
    const int3 size = I.size();
    const tsf_dims d = I.dims();
    const size_t N = voxels(size);
    const int k = clamp_every(o);
    Image3D B(size, 1), F(size, 4), V(size, 4), T(size, 1), R(size, 1);
    check(tsf_blur_gaussian((const float*)I.plane(0), (float*)B.plane(0), d, sigma, ocl.queue));
    const tsf_field f = F.field();
    check(tsf_create_vector_field((const float*)B.plane(0), &f, d, o.Fmax, o.sign, ocl.queue));
    const int64_t C = tsf_gvf_checkpoint_count(o.iterations, k);
    ctx->checkpoints = Buffer((size_t)C * 3 * N * sizeof(float));
    const tsf_field v = V.field();
    check(tsf_gvf_euler_checkpointed(&f, &v, d, mu, o.iterations, k, (float*)ctx->checkpoints.get(),
                                     ctx->checkpoints.size(), ocl.queue));
    check(tsf_tdf_circle(&v, d, o.radiusMin, o.radiusMax, o.radiusStep, (float*)T.plane(0), (float*)R.plane(0),
                         nullptr, 0, ocl.queue));
    ctx->workspace = Buffer(backward_workspace(size, k));
    ctx->options = o;
    ctx->options.checkpoint_every = k;
    ctx->saved_data["mu"] = mu;
    ctx->saved_data["sigma"] = sigma;
    ctx->save_for_backward({I, B, F, V, R});
    return {T, R};
}

TubeFilterGrads TubeFilterFunction::backward(AutogradContext* ctx, OpenCL& ocl,
                                             const std::vector<Image3D>& grad_outputs) {
    if (!ctx || ctx->get_saved_variables().size() != 5)
        throw Error(TSF_ERR_INVALID_ARGUMENT, "TubeFilterFunction::backward: context holds no forward");
    const auto& sv = ctx->get_saved_variables();
    const Image3D &I = sv[0], &B = sv[1], &F = sv[2], &V = sv[3], &R = sv[4];
    if (grad_outputs.empty() || !grad_outputs[0].valid() || grad_outputs[0].channels() != 1 ||
        grad_outputs[0].storage() != TSF_STORAGE_F32 || grad_outputs[0].size().x != I.size().x ||
        grad_outputs[0].size().y != I.size().y || grad_outputs[0].size().z != I.size().z)
        throw Error(TSF_ERR_INVALID_ARGUMENT, "TubeFilterFunction::backward: dL/dT must match the volume");
    const TubeFilterOptions& o = ctx->options;
    const int3 size = I.size();
    const tsf_dims d = I.dims();
    const size_t N = voxels(size);
    const float mu = (float)ctx->saved_data.at("mu"), sigma = (float)ctx->saved_data.at("sigma");
    void* ws = ctx->workspace.get();
    const size_t wsb = ctx->workspace.size();

    Buffer dV(4 * N * sizeof(float));
    check(tsf_memset(dV.get(), 0, dV.size(), ocl.queue));
    const tsf_field dv = planes_field(dV.get(), N, true);
    const tsf_field v = V.field();
    check(tsf_tdf_adjoint(&v, d, o.radiusMin, o.radiusMax, o.radiusStep, (const float*)grad_outputs[0].plane(0),
                          (const float*)R.plane(0), &dv, ws, wsb, ocl.queue));
    check(tsf_gvf_finish_adjoint(&v, &dv, d, ocl.queue));
    Buffer dF(3 * N * sizeof(float));
    const tsf_field f = F.field();
    TubeFilterGrads g;
    check(tsf_gvf_euler_adjoint(&f, d, mu, o.iterations, o.checkpoint_every, (const float*)ctx->checkpoints.get(),
                                (const float*)dV.get(), (float*)dF.get(), &g.mu, ws, wsb, ocl.queue));
    Buffer dB(N * sizeof(float));
    const tsf_field df = planes_field(dF.get(), N, false);
    check(tsf_vector_field_adjoint((const float*)B.plane(0), &df, d, o.Fmax, o.sign, (float*)dB.get(), ws, wsb,
                                   ocl.queue));
    g.volume = Image3D(size, 1);
    check(tsf_blur_adjoint((const float*)dB.get(), (const float*)I.plane(0), (float*)g.volume.plane(0), d, sigma,
                           &g.sigma, ws, wsb, ocl.queue));
    return g;
}

}  // namespace tsf
