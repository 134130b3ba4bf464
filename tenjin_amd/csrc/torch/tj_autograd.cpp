// libtorch autograd over the tangent-linear form of the solid path
// (tj_tangent_*, include/tenjin.h): the probe-trace calibration loss
//   L(theta) = sum_n (u_n - u*_n)^2,
// u_n the probe component after loop body n of a solid level whose materials
// carry theta = [Gamma0, s, yield stress] per material, as a
// torch::autograd::Function of theta.  The forward evaluates the loss and its
// gradient in one tangent run (one direction per parameter); the backward
// scales the stored gradient by the incoming one.  The reference has no
// differentiated form: this serves gradient-based calibration of the EOS and
// yield parameters against a measured trace.
#include <torch/cuda.h>
#include <torch/extension.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tenjin.h"

namespace {

struct Handles {
    tj_ctx* c = nullptr;
    tj_tangent* t = nullptr;
    ~Handles() {
        if (t) tj_tangent_destroy(t);
        if (c) tj_destroy(c);
    }
};

void check(int rc, const tj_ctx* c, const char* what) {
    if (rc != 0)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " +
                                 (c ? tj_last_error(c) : ""));
}

tj_mf view(const torch::Tensor& t, int ncomp) {
    tj_mf v{};
    v.data = t.data_ptr<double>();
    v.jstride = t.size(3);
    v.kstride = t.size(3) * t.size(2);
    v.nstride = t.size(3) * t.size(2) * t.size(1);
    v.ncomp = ncomp;
    return v;
}

// theta -> the level's materials: Gamma0 enters as adiabaticIndex = Gamma0 + 1
// (GruneisenGamma = adiabaticIndex - 1, equationofstate.cpp:47-75).
std::pair<double, std::vector<double>> evaluate(const torch::Tensor& theta, const torch::Tensor& prims,
                                                const torch::Tensor& target, const torch::Tensor& desc_bytes,
                                                int64_t nsteps, const std::vector<int64_t>& probe, int64_t chunk,
                                                int64_t shift) {
    TORCH_CHECK(desc_bytes.dtype() == torch::kUInt8 && desc_bytes.numel() == (int64_t)sizeof(tj_level_desc),
                "desc: the bytes of a tj_level_desc");
    tj_level_desc d;
    std::memcpy(&d, desc_bytes.contiguous().data_ptr<uint8_t>(), sizeof d);
    const torch::Tensor th = theta.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    TORCH_CHECK(th.numel() == 3 * d.nmat, "theta: [Gamma0, s, Y] per material (", 3 * d.nmat, " values)");
    const double* tp = th.data_ptr<double>();
    for (int m = 0; m < d.nmat; ++m) {
        d.mat[m].gamma = tp[3 * m] + 1.0;
        d.mat[m].s = tp[3 * m + 1];
        d.mat[m].yield_stress = tp[3 * m + 2];
    }
    TORCH_CHECK(prims.is_cuda() && prims.dtype() == torch::kFloat64 && prims.dim() == 4 && prims.is_contiguous(),
                "prims: contiguous float64 [ncomp][nz][ny][nx] on the GPU");
    TORCH_CHECK(probe.size() == 4, "probe: (component, i, j, k)");
    const torch::Tensor tg = target.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    TORCH_CHECK(tg.numel() == nsteps, "target: one value per step");

    Handles h;
    check(tj_create(&h.c, prims.get_device()), nullptr, "tj_create");
    check(tj_bind_level(h.c, &d), h.c, "tj_bind_level");
    const int ncomp = tj_ncomp(h.c);
    TORCH_CHECK(prims.size(0) == ncomp, "prims: ", ncomp, " components for this level");
    std::vector<int> mat, par;
    for (int m = 0; m < d.nmat; ++m)
        for (int p = TJ_TP_GAMMA0; p <= TJ_TP_YIELD; ++p) {
            mat.push_back(m);
            par.push_back(p);
        }
    const int ndir = (int)mat.size();
    check(tj_tangent_create(h.c, ndir, mat.data(), par.data(), (int)chunk, &h.t), h.c, "tj_tangent_create");
    torch::Tensor S = torch::empty_like(prims), T = torch::empty_like(prims);
    torch::cuda::synchronize();  // prims and the new buffers are torch's; the library runs on its own stream
    tj_mf vP = view(prims, ncomp), vS = view(S, ncomp), vT = view(T, ncomp);
    std::vector<double> trace((size_t)nsteps * (1 + ndir));
    int done = 0;
    check(tj_tangent_run(h.t, &vP, (int)shift, &vS, &vT, INFINITY, (int)nsteps, (int)probe[0], (int)probe[1],
                         (int)probe[2], (int)probe[3], trace.data(), &done),
          h.c, "tj_tangent_run");
    TORCH_CHECK(done == nsteps, "tangent run stopped after ", done, " of ", nsteps, " steps");
    const double* us = tg.data_ptr<double>();
    double loss = 0.0;
    std::vector<double> grad(ndir, 0.0);
    for (int64_t q = 0; q < nsteps; ++q) {
        const double* row = trace.data() + q * (1 + ndir);
        const double r = row[0] - us[q];
        loss += r * r;
        for (int t = 0; t < ndir; ++t) grad[t] += 2.0 * r * row[1 + t];
    }
    return {loss, grad};
}

class ProbeLoss : public torch::autograd::Function<ProbeLoss> {
   public:
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor theta, torch::Tensor prims,
                                 torch::Tensor target, torch::Tensor desc, int64_t nsteps, std::vector<int64_t> probe,
                                 int64_t chunk, int64_t shift) {
        auto res = evaluate(theta, prims, target, desc, nsteps, probe, chunk, shift);
        // the gradient takes theta's device, dtype and shape, the loss theta's device and dtype, so
        // autograd accepts them for a theta of any shape (3 * nmat values) on the CPU or the GPU
        ctx->saved_data["grad"] =
            torch::tensor(res.second, torch::kFloat64).to(theta.device(), theta.scalar_type()).view(theta.sizes());
        return torch::tensor(res.first, torch::kFloat64).to(theta.device(), theta.scalar_type());
    }
    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out) {
        const torch::Tensor g = ctx->saved_data["grad"].toTensor();
        return {grad_out[0].to(g.device(), g.scalar_type()) * g, torch::Tensor(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "torch::autograd::Function of the probe-trace loss over the tangent-linear form (libtenjin)";
    m.def(
        "probe_loss",
        [](torch::Tensor theta, torch::Tensor prims, torch::Tensor target, torch::Tensor desc, int64_t nsteps,
           std::vector<int64_t> probe, int64_t chunk, int64_t shift) {
            return ProbeLoss::apply(theta, prims, target, desc, nsteps, probe, chunk, shift);
        },
        "L(theta) = sum_n (u_n - u*_n)^2 with its gradient by the tangent-linear form");
}
