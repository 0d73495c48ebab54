// pbf_torch.cpp -- the differentiable rollout as a torch::autograd::Function
// (libtorch C++, built as the Python module pbf_amd.pbf_torch).
//
// rollout(adjoint, ctx, x0, v0, coef) -> (xT, vT) runs pbf_adjoint_forward on
// ROCm tensors; its backward runs pbf_adjoint_backward and returns dL/dx0,
// dL/dv0 and dL/dcoef (coef: per scene {viscosityFactor, vorticityFactor,
// g0, g1, g2}, include/pbf/pbf_adjoint.h).  Positions and velocities are
// (n, 3) float32 on the GPU, in particle-identity order.  The library runs
// on its own stream: every call waits for the current torch stream first and
// returns after the library's work has finished.
#include <c10/hip/HIPStream.h>
#include <torch/extension.h>

#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/pbf/pbf_adjoint.h"

namespace {

void check(int rc, pbf_ctx* ctx) {
    if (rc != PBF_OK) throw std::runtime_error("pbf error " + std::to_string(rc) + ": " + pbf_last_error(ctx));
}

// Each forward of an adjoint gets a generation; a backward must belong to the
// latest forward of its adjoint (the tape holds one rollout).
std::mutex g_mu;
std::unordered_map<int64_t, int64_t> g_generation;

torch::Tensor as_float4(const torch::Tensor& t, int64_t n, const char* what) {
    TORCH_CHECK(t.is_cuda(), what, " must be a GPU tensor");
    TORCH_CHECK(t.dim() == 2 && t.size(0) == n && (t.size(1) == 3 || t.size(1) == 4), what, " must be (", n,
                ", 3) or (", n, ", 4)");
    auto o = torch::zeros({n, 4}, t.options().dtype(torch::kFloat32));
    o.narrow(1, 0, 3).copy_(t.narrow(1, 0, 3));
    return o;
}

class PbfRollout : public torch::autograd::Function<PbfRollout> {
public:
    static torch::autograd::variable_list forward(torch::autograd::AutogradContext* actx, int64_t adj, int64_t ctxp,
                                                  torch::Tensor x0, torch::Tensor v0, torch::Tensor coef) {
        auto* a = reinterpret_cast<pbf_adjoint*>(adj);
        auto* ctx = reinterpret_cast<pbf_ctx*>(ctxp);
        TORCH_CHECK(a && ctx, "null adjoint or context");
        const int S = pbf_scene_count(ctx);
        TORCH_CHECK(S > 0, "bad context");
        TORCH_CHECK(coef.dim() == 2 && coef.size(0) == S && coef.size(1) == 5, "coef must be (", S, ", 5)");
        // the rows must be the context's particles: the library reads and
        // writes exactly that many float4 rows of every buffer it is given
        pbf_dev_bufs db{};
        check(pbf_device_buffers(ctx, &db), ctx);
        const int64_t n = (int64_t)db.n;
        TORCH_CHECK(x0.dim() >= 1 && x0.size(0) == n, "x0 must have ", n, " rows (the context's particles), got ",
                    x0.sizes());
        TORCH_CHECK(v0.dim() >= 1 && v0.size(0) == n, "v0 must have ", n, " rows (the context's particles), got ",
                    v0.sizes());
        auto ch = coef.detach().to(torch::kCPU, torch::kFloat32).contiguous();
        const float* cp = ch.data_ptr<float>();
        for (int s = 0; s < S; ++s) check(pbf_set_scene_coefficients(ctx, (uint32_t)s, cp[5 * s], cp[5 * s + 1], cp + 5 * s + 2), ctx);
        torch::NoGradGuard ng;
        auto X = as_float4(x0.detach(), n, "x0");
        auto V = as_float4(v0.detach(), n, "v0");
        auto XT = torch::empty_like(X), VT = torch::empty_like(V);
        c10::hip::getCurrentHIPStream().synchronize();
        check(pbf_adjoint_forward(a, X.data_ptr<float>(), V.data_ptr<float>(), XT.data_ptr<float>(),
                                  VT.data_ptr<float>()),
              ctx);
        int64_t gen;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            gen = ++g_generation[adj];
        }
        actx->saved_data["adj"] = adj;
        actx->saved_data["ctx"] = ctxp;
        actx->saved_data["gen"] = gen;
        actx->saved_data["n"] = n;
        actx->saved_data["coef_device"] = coef.device().str();
        return {XT.narrow(1, 0, 3).contiguous(), VT.narrow(1, 0, 3).contiguous()};
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* actx,
                                                   torch::autograd::variable_list grads) {
        const int64_t adj = actx->saved_data["adj"].toInt();
        auto* a = reinterpret_cast<pbf_adjoint*>(adj);
        auto* ctx = reinterpret_cast<pbf_ctx*>(actx->saved_data["ctx"].toInt());
        const int64_t n = actx->saved_data["n"].toInt();
        {
            std::lock_guard<std::mutex> lk(g_mu);
            TORCH_CHECK(g_generation[adj] == actx->saved_data["gen"].toInt(),
                        "this rollout's tape was replaced by a later forward of the same adjoint");
        }
        const int S = pbf_scene_count(ctx);
        const torch::Tensor& gx = grads[0];
        const torch::Tensor& gv = grads[1];
        auto opts = (gx.defined() ? gx : gv).options().dtype(torch::kFloat32);
        auto GX = gx.defined() ? as_float4(gx, n, "dL/dxT") : torch::Tensor();
        auto GV = gv.defined() ? as_float4(gv, n, "dL/dvT") : torch::Tensor();
        auto gx0 = torch::empty({n, 4}, opts), gv0 = torch::empty({n, 4}, opts);
        auto gc = torch::empty({S, 5}, opts);
        c10::hip::getCurrentHIPStream().synchronize();
        check(pbf_adjoint_backward(a, GX.defined() ? GX.data_ptr<float>() : nullptr,
                                   GV.defined() ? GV.data_ptr<float>() : nullptr, gx0.data_ptr<float>(),
                                   gv0.data_ptr<float>(), gc.data_ptr<float>()),
              ctx);
        const torch::Device cdev(actx->saved_data["coef_device"].toStringRef());
        return {torch::Tensor(), torch::Tensor(), gx0.narrow(1, 0, 3).contiguous(), gv0.narrow(1, 0, 3).contiguous(),
                gc.to(cdev)};
    }
};

std::vector<torch::Tensor> rollout(int64_t adjoint, int64_t ctx, torch::Tensor x0, torch::Tensor v0,
                                   torch::Tensor coef) {
    auto r = PbfRollout::apply(adjoint, ctx, x0, v0, coef);
    return {r[0], r[1]};
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "PBF rollout with a hand-written HIP adjoint (torch::autograd::Function)";
    m.def("rollout", &rollout, "Differentiable rollout: (adjoint, ctx, x0, v0, coef) -> (xT, vT)",
          pybind11::arg("adjoint"), pybind11::arg("ctx"), pybind11::arg("x0"), pybind11::arg("v0"),
          pybind11::arg("coef"));
}
