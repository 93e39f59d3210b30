// torch_ext.cpp - torch::autograd::Function wrappers of the batched force
// and adjoint entry points (include/mdx_batch.h), for force matching:
//
//   F = LJForces(eps_pairs, sigma_pairs)        -> dL/d(eps, sigma) in backward
//   F = EAMForces(frho, rhor, z2r knot values)  -> dL/d(knots) in backward
//
// Parameters are CPU float64 tensors; forces are a float64 device tensor
// [N, 3] in the batch's atom order.  Backward hands dL/dF (a device tensor)
// to the hand-written adjoint kernels, which recompute pair distances from
// the positions (no per-pair state is kept).  Pair parameters are listed
// for type pairs (a, b) with a <= b in row order: (1,1), (1,2), ..,
// (1,nt), (2,2), ...; the cutoffs are fixed.
//
// Ordering: the caller passes PyTorch's current stream (an integer handle);
// the batch queues its kernels on that stream (mdx_batch_set_stream), so
// they run after the kernels that produced their inputs and before those
// that consume their outputs, with no device-wide synchronisation.
#include <torch/extension.h>

#include <cstring>
#include <vector>

#include "../../include/mdx.h"
#include "../../include/mdx_batch.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

mdx_batch *handle(int64_t h) { return reinterpret_cast<mdx_batch *>(h); }

// the batch's work goes to the caller's stream
mdx_batch *on_stream(int64_t h, int64_t stream) {
  mdx_batch *b = handle(h);
  TORCH_CHECK(mdx_batch_set_stream(b, 1, reinterpret_cast<void *>(stream)) == MDX_OK, "libmdx: ",
              mdx_last_error());
  return b;
}

void ok(int rc) { TORCH_CHECK(rc == MDX_OK, "libmdx: ", mdx_last_error()); }

std::vector<double> host_vector(const at::Tensor &t) {
  at::Tensor c = t.detach().to(at::kCPU, at::kDouble).contiguous();
  return std::vector<double>(c.data_ptr<double>(), c.data_ptr<double>() + c.numel());
}

// pair list (a <= b) -> symmetric (nt+1)^2 table
std::vector<double> table(const std::vector<double> &pairs, int nt) {
  const int nt1 = nt + 1;
  TORCH_CHECK((int)pairs.size() == nt * (nt + 1) / 2, "expected ", nt * (nt + 1) / 2, " pair values");
  std::vector<double> t(nt1 * nt1, 0.0);
  int k = 0;
  for (int a = 1; a <= nt; a++)
    for (int b = a; b <= nt; b++, k++) t[a * nt1 + b] = t[b * nt1 + a] = pairs[k];
  return t;
}

at::Tensor pairs_of(const std::vector<double> &tab, int nt) {
  const int nt1 = nt + 1;
  at::Tensor out = torch::zeros({nt * (nt + 1) / 2}, torch::dtype(torch::kDouble));
  double *o = out.data_ptr<double>();
  int k = 0;
  for (int a = 1; a <= nt; a++)
    for (int b = a; b <= nt; b++, k++) o[k] = tab[a * nt1 + b];
  return out;
}

at::Tensor device_forces(int64_t n, int64_t device) {
  return torch::empty({n, 3}, torch::dtype(torch::kDouble).device(torch::kCUDA, device));
}

struct LJForces : public torch::autograd::Function<LJForces> {
  static at::Tensor forward(AutogradContext *ctx, const at::Tensor &eps, const at::Tensor &sigma,
                            const at::Tensor &cut, int64_t h, int64_t ntypes, int64_t n, int64_t device,
                            int64_t stream) {
    const int nt = (int)ntypes;
    auto e = table(host_vector(eps), nt), s = table(host_vector(sigma), nt), c = table(host_vector(cut), nt);
    mdx_batch *b = on_stream(h, stream);
    ok(mdx_batch_lj_init(b, e.data(), s.data(), c.data()));
    at::Tensor f = device_forces(n, device);
    ok(mdx_batch_lj_forces(b, f.data_ptr<double>(), nullptr));
    ctx->save_for_backward({eps, sigma, cut});
    ctx->saved_data["h"] = h;
    ctx->saved_data["nt"] = ntypes;
    ctx->saved_data["dev"] = device;
    ctx->saved_data["stream"] = stream;
    return f;
  }

  static variable_list backward(AutogradContext *ctx, variable_list grad) {
    auto saved = ctx->get_saved_variables();
    const int64_t h = ctx->saved_data["h"].toInt();
    const int nt = (int)ctx->saved_data["nt"].toInt();
    auto e = table(host_vector(saved[0]), nt), s = table(host_vector(saved[1]), nt),
         c = table(host_vector(saved[2]), nt);
    mdx_batch *b = on_stream(h, ctx->saved_data["stream"].toInt());
    ok(mdx_batch_lj_init(b, e.data(), s.data(), c.data()));
    at::Tensor gf = grad[0].to(torch::kDouble).contiguous();
    TORCH_CHECK(gf.is_cuda(), "dL/dF must be on the batch's device");
    const int nt1 = nt + 1;
    std::vector<double> ge(nt1 * nt1), gs(nt1 * nt1);
    ok(mdx_batch_lj_vjp(b, gf.data_ptr<double>(), ge.data(), gs.data()));
    return {pairs_of(ge, nt), pairs_of(gs, nt), at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(),
            at::Tensor(), at::Tensor()};
  }
};

struct EAMForces : public torch::autograd::Function<EAMForces> {
  static at::Tensor forward(AutogradContext *ctx, const at::Tensor &frho, const at::Tensor &rhor,
                            const at::Tensor &z2r, int64_t h, int64_t n, int64_t device, double drho, double dr,
                            double cut, int64_t stream) {
    auto f1 = host_vector(frho), r1 = host_vector(rhor), z1 = host_vector(z2r);
    TORCH_CHECK(r1.size() == z1.size(), "rhor and z2r need the same number of knots");
    mdx_batch *b = on_stream(h, stream);
    ok(mdx_batch_eam_init(b, (int)f1.size(), drho, (int)r1.size(), dr, cut, f1.data(), r1.data(), z1.data()));
    at::Tensor f = device_forces(n, device);
    ok(mdx_batch_eam_forces(b, f.data_ptr<double>(), nullptr, nullptr, nullptr));
    ctx->save_for_backward({frho, rhor, z2r});
    ctx->saved_data["h"] = h;
    ctx->saved_data["dev"] = device;
    ctx->saved_data["stream"] = stream;
    ctx->saved_data["drho"] = drho;
    ctx->saved_data["dr"] = dr;
    ctx->saved_data["cut"] = cut;
    return f;
  }

  static variable_list backward(AutogradContext *ctx, variable_list grad) {
    auto saved = ctx->get_saved_variables();
    const int64_t h = ctx->saved_data["h"].toInt();
    auto f1 = host_vector(saved[0]), r1 = host_vector(saved[1]), z1 = host_vector(saved[2]);
    mdx_batch *b = on_stream(h, ctx->saved_data["stream"].toInt());
    // unchanged tables keep the forward's rho/fp on the device
    ok(mdx_batch_eam_init(b, (int)f1.size(), ctx->saved_data["drho"].toDouble(), (int)r1.size(),
                          ctx->saved_data["dr"].toDouble(), ctx->saved_data["cut"].toDouble(), f1.data(), r1.data(),
                          z1.data()));
    at::Tensor gf = grad[0].to(torch::kDouble).contiguous();
    TORCH_CHECK(gf.is_cuda(), "dL/dF must be on the batch's device");
    at::Tensor gfr = torch::zeros({(int64_t)f1.size()}, torch::dtype(torch::kDouble));
    at::Tensor grr = torch::zeros({(int64_t)r1.size()}, torch::dtype(torch::kDouble));
    at::Tensor gzr = torch::zeros({(int64_t)z1.size()}, torch::dtype(torch::kDouble));
    ok(mdx_batch_eam_vjp(b, gf.data_ptr<double>(), gfr.data_ptr<double>(), grr.data_ptr<double>(),
                         gzr.data_ptr<double>()));
    return {gfr, grr, gzr, at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(),
            at::Tensor()};
  }
};

at::Tensor lj_forces(const at::Tensor &eps, const at::Tensor &sigma, const at::Tensor &cut, int64_t h,
                     int64_t ntypes, int64_t n, int64_t device, int64_t stream) {
  return LJForces::apply(eps, sigma, cut, h, ntypes, n, device, stream);
}

at::Tensor eam_forces(const at::Tensor &frho, const at::Tensor &rhor, const at::Tensor &z2r, int64_t h, int64_t n,
                      int64_t device, double drho, double dr, double cut, int64_t stream) {
  return EAMForces::apply(frho, rhor, z2r, h, n, device, drho, dr, cut, stream);
}

}  // namespace

PYBIND11_MODULE(_mdx_torch, m) {
  m.doc() = "autograd Functions over libmdx's batched forces and force-matching adjoints";
  m.def("lj_forces", &lj_forces, "lj/cut forces of every replica; differentiable in (eps, sigma) pair values");
  m.def("eam_forces", &eam_forces, "eam forces of every replica; differentiable in the knot values");
  m.def("version", []() { return mdx_version(); });
}
