// A torch::autograd::Function over a static load path of an lpm context: the test harness of
// the adjoint. forward(Kn, Tv, bdamage) puts the bond parameters on the device, starts from a
// checkpointed state, runs SolverStatic::solveProblemStep (solver_static.h:97-149) for each
// load step, checkpoints every converged state and returns the objective of each step,
//   G_k = w_pin . Pin_k + w_x . x_k.
// backward(dL/dG) restores each step's checkpoint and runs one adjoint solve there
// (lpm_adjoint_linear), accumulating dL/dKn, dL/dTv and dL/dbdamage.
//
// The derivative is the one at a frozen broken-bond set and frozen plastic bond lengths: exact
// for elastic paths (each step's equilibrium depends on the parameters only through that
// step), the frozen-state derivative for path-dependent materials. A bond that is broken at a
// step contributes nothing there (its geometry is zero).
//
// Built as the Python module lpm_autograd next to liblpm_mi355x.so; it calls the library only
// through its C ABI.
#include <torch/extension.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "lpm_mi355x.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void chk(lpm_ctx *c, int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + lpm_last_error(c));
}

Tensor host_f64(const Tensor &t, int64_t n, const char *what) {
    if (!t.defined()) return t;
    if (t.numel() != n) throw std::invalid_argument(std::string(what) + ": expected " + std::to_string(n) + " entries");
    return t.detach().to(torch::kCPU, torch::kFloat64).contiguous();
}

void upload(lpm_ctx *c, const char *field, const Tensor &t) {
    chk(c, lpm_upload(c, field, t.data_ptr<double>(), t.numel()), field);
}

struct LoadPath : public torch::autograd::Function<LoadPath> {
    static Tensor forward(AutogradContext *ac, Tensor Kn, Tensor Tv, Tensor bdamage, int64_t handle,
                          std::vector<int64_t> steps, int64_t base, Tensor w_pin, Tensor w_x) {
        lpm_ctx *c = reinterpret_cast<lpm_ctx *>(handle);
        const int64_t n = lpm_size(c, "n"), dim = lpm_size(c, "dim"), nb = lpm_size(c, "nbonds");
        if (base < 0 || base + (int64_t)steps.size() >= LPM_MAX_CHECKPOINTS)
            throw std::invalid_argument("checkpoint slots out of range");
        Kn = host_f64(Kn, nb, "Kn");
        Tv = host_f64(Tv, nb, "Tv");
        bdamage = host_f64(bdamage, nb, "bdamage");
        w_pin = host_f64(w_pin, dim * n, "w_pin");
        w_x = host_f64(w_x, dim * n, "w_x");
        chk(c, lpm_checkpoint_restore(c, (int32_t)base), "restore the starting state");
        upload(c, "Kn", Kn);
        upload(c, "Tv", Tv);
        upload(c, "bdamage", bdamage);
        chk(c, lpm_update_geometry(c), "updateGeometry");
        chk(c, lpm_update_force_state(c), "updateForceState");
        const int64_t K = (int64_t)steps.size();
        Tensor G = torch::zeros({K}, torch::kFloat64);
        std::vector<double> pin(3 * n), x(3 * n);
        for (int64_t k = 0; k < K; ++k) {
            int32_t ok = 0;
            chk(c, lpm_solve_step_static(c, reinterpret_cast<const lpm_load_step *>(steps[k]), -1, &ok), "solveProblemStep");
            if (!ok) throw std::runtime_error("load step " + std::to_string(k) + " did not converge");
            chk(c, lpm_checkpoint_save(c, (int32_t)(base + 1 + k)), "checkpoint");
            chk(c, lpm_download(c, "Pin", pin.data(), 3 * n), "Pin");
            chk(c, lpm_download(c, "xyz", x.data(), 3 * n), "xyz");
            double g = 0.0;
            const double *wp = w_pin.defined() ? w_pin.data_ptr<double>() : nullptr;
            const double *wx = w_x.defined() ? w_x.data_ptr<double>() : nullptr;
            for (int64_t i = 0; i < n; ++i)
                for (int64_t a = 0; a < dim; ++a) {
                    if (wp) g += wp[dim * i + a] * pin[3 * i + a];
                    if (wx) g += wx[dim * i + a] * x[3 * i + a];
                }
            G[k] = g;
        }
        ac->save_for_backward({Kn, Tv, bdamage});
        ac->saved_data["handle"] = handle;
        ac->saved_data["steps"] = K;
        ac->saved_data["base"] = base;
        ac->saved_data["w_pin"] = w_pin;
        ac->saved_data["w_x"] = w_x;
        return G;
    }

    static tensor_list backward(AutogradContext *ac, tensor_list grad_out) {
        lpm_ctx *c = reinterpret_cast<lpm_ctx *>(ac->saved_data["handle"].toInt());
        const int64_t K = ac->saved_data["steps"].toInt(), base = ac->saved_data["base"].toInt();
        const Tensor w_pin = ac->saved_data["w_pin"].toTensor(), w_x = ac->saved_data["w_x"].toTensor();
        const auto saved = ac->get_saved_variables();
        const int64_t nb = lpm_size(c, "nbonds");
        const Tensor gG = grad_out[0].detach().to(torch::kCPU, torch::kFloat64).contiguous();
        Tensor gKn = torch::zeros({nb}, torch::kFloat64), gTv = torch::zeros({nb}, torch::kFloat64),
               gBd = torch::zeros({nb}, torch::kFloat64);
        std::vector<double> a(nb), b(nb), d(nb);
        // the parameters of this forward pass (another pass may have changed the context's)
        upload(c, "Kn", saved[0]);
        upload(c, "Tv", saved[1]);
        for (int64_t k = 0; k < K; ++k) {
            const double s = gG[k].item<double>();
            if (s == 0.0) continue;
            chk(c, lpm_checkpoint_restore(c, (int32_t)(base + 1 + k)), "restore");
            double G = 0.0, rr = 0.0;
            int32_t it = 0;
            chk(c,
                lpm_adjoint_linear(c, w_pin.defined() ? w_pin.data_ptr<double>() : nullptr,
                                   w_x.defined() ? w_x.data_ptr<double>() : nullptr, a.data(), b.data(), d.data(), &G, &it,
                                   &rr),
                "adjoint");
            double *pk = gKn.data_ptr<double>(), *pt = gTv.data_ptr<double>(), *pd = gBd.data_ptr<double>();
            for (int64_t e = 0; e < nb; ++e) {
                pk[e] += s * a[e];
                pt[e] += s * b[e];
                pd[e] += s * d[e];
            }
        }
        // leave the context at the end of the path, as forward did
        chk(c, lpm_checkpoint_restore(c, (int32_t)(base + K)), "restore");
        return {gKn, gTv, gBd, Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

// handle: the lpm_ctx* as an integer; steps: lpm_load_step* as integers; base: the checkpoint
// slot holding the starting state (steps use base+1 .. base+len(steps)). w_pin, w_x: DoF-order
// weights (dim*n), either may be None.
Tensor load_path(Tensor Kn, Tensor Tv, Tensor bdamage, int64_t handle, std::vector<int64_t> steps, int64_t base,
                 c10::optional<Tensor> w_pin, c10::optional<Tensor> w_x) {
    if (!handle) throw std::invalid_argument("null context");
    if (steps.empty()) throw std::invalid_argument("no load steps");
    return LoadPath::apply(Kn, Tv, bdamage, handle, steps, base, w_pin.value_or(Tensor()), w_x.value_or(Tensor()));
}

}  // namespace

PYBIND11_MODULE(lpm_autograd, m) {
    m.doc() = "torch.autograd.Function over a static load path of an lpm context (adjoint test harness)";
    m.def("load_path", &load_path, "G[k] = w_pin.Pin_k + w_x.x_k over the load steps; differentiable in Kn, Tv, bdamage",
          pybind11::arg("Kn"), pybind11::arg("Tv"), pybind11::arg("bdamage"), pybind11::arg("handle"),
          pybind11::arg("steps"), pybind11::arg("base"), pybind11::arg("w_pin") = pybind11::none(),
          pybind11::arg("w_x") = pybind11::none());
}
