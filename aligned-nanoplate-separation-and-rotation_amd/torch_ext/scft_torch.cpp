// libtorch wrapper of the plate energy: a torch::autograd::Function whose
// forward moves the two plates of one system of a libscft_amd context to the
// geometry (D, theta), optionally relaxes the fields to the saddle point, and
// returns Re H; its backward returns the analytic plate derivatives of H
// (scft_grad_plates) in place of finite-difference sweeps over separate runs
// (the reference sweeps d and theta by running code/simulate.cpp once per
// geometry).
//
// Geometry convention (the one of the separation / angle sweeps): plate
// centres at fractional x = 0.5 -+ D / (2 Lx), y = z = 0.5; both normals
// u = (sin theta, 0, cos theta), i.e. rotated by theta about y.  D is in b
// units, theta in radians.
//
// At a converged saddle point (max_iter > 0, tol small) dH/dw = 0, so the
// explicit derivative is the total one: force dH/dD and torque dH/dtheta.
// With max_iter = 0 the fields stay fixed and the derivative is exact for H
// at those fields.
//
// The context is passed as its integer handle (scft_ctx*).  This module only
// calls the C ABI; import torch before loading it, so that the process holds
// one HIP runtime.
#include <torch/extension.h>

#include <cmath>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "scft.h"

namespace {

void check(int rc, const char* what) {
  if (rc != SCFT_OK) throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                                              "): " + scft_last_error());
}

void set_pair(scft_ctx* c, int sys, double D, double th, double Lx) {
  const double d = D / Lx;
  const double cen[6] = {0.5 - 0.5 * d, 0.5, 0.5, 0.5 + 0.5 * d, 0.5, 0.5};
  const double u[6] = {std::sin(th), 0.0, std::cos(th), std::sin(th), 0.0, std::cos(th)};
  check(scft_set_plates(c, sys, cen, u), "scft_set_plates");
}

// systems of the context: the first index scft_get_info rejects
int count_systems(scft_ctx* c) {
  int n = 0;
  scft_sysinfo t;
  while (scft_get_info(c, n, &t) == SCFT_OK) n++;
  return n;
}

// Re H of system `sys` after relaxing all systems of the context in chunks
// of `chunk` iterations until Re H of `sys` changes by less than tol
// (relative) over one chunk, at most max_iter iterations.  The density is
// recomputed first, so H and Qd belong to the current plates and fields.
double relax(scft_ctx* c, int sys, int64_t max_iter, double tol, int64_t chunk, int64_t* done) {
  check(scft_calc_poly_density(c), "scft_calc_poly_density");
  std::vector<double> h(2 * (size_t)count_systems(c));
  check(scft_calc_H(c, h.data()), "scft_calc_H");
  double prev = h[2 * sys];
  *done = 0;
  while (*done < max_iter) {
    const int64_t k = std::min(chunk, max_iter - *done);
    check(scft_iterate(c, (int)k), "scft_iterate");
    *done += k;
    check(scft_calc_H(c, h.data()), "scft_calc_H");
    const double cur = h[2 * sys];
    const double change = std::fabs(cur - prev) / std::fabs(cur);
    prev = cur;
    if (change < tol) break;
  }
  return prev;
}

}  // namespace

class PlateEnergy : public torch::autograd::Function<PlateEnergy> {
 public:
  static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor geom, int64_t handle,
                               int64_t sys, double Lx, int64_t max_iter, double tol) {
    TORCH_CHECK(geom.numel() == 2, "geom must hold (D, theta)");
    TORCH_CHECK(handle != 0, "null context handle");
    scft_ctx* c = reinterpret_cast<scft_ctx*>(handle);
    auto g = geom.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    const double D = g.data_ptr<double>()[0], th = g.data_ptr<double>()[1];
    set_pair(c, (int)sys, D, th, Lx);
    int64_t done = 0;
    const double H = relax(c, (int)sys, max_iter, tol, 100, &done);
    // analytic derivatives at this state: centres (b units) and rotation
    // about y; the symmetric pair moves its centres by -+ dD / 2
    const double axis[3] = {0.0, 1.0, 0.0};
    double out[16];
    check(scft_grad_plates(c, (int)sys, axis, out), "scft_grad_plates");
    const double dHdD = 0.5 * (out[8 + 0] - out[0]);
    const double dHdth = out[6] + out[8 + 6];
    ctx->saved_data["dHdD"] = dHdD;
    ctx->saved_data["dHdth"] = dHdth;
    ctx->saved_data["iterations"] = done;
    ctx->saved_data["dtype"] = (int64_t)geom.scalar_type();
    return torch::tensor(H, torch::dtype(torch::kFloat64));
  }

  static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::variable_list grad) {
    const double gD = ctx->saved_data["dHdD"].toDouble();
    const double gT = ctx->saved_data["dHdth"].toDouble();
    auto dt = static_cast<c10::ScalarType>(ctx->saved_data["dtype"].toInt());
    auto go = grad[0].to(torch::kFloat64);
    auto gg = torch::stack({go * gD, go * gT}).to(dt);
    return {gg, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

torch::Tensor plate_energy(torch::Tensor geom, int64_t handle, int64_t sys, double Lx, int64_t max_iter,
                           double tol) {
  return PlateEnergy::apply(geom, handle, sys, Lx, max_iter, tol);
}

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "libtorch autograd wrapper of the plate energy of libscft_amd";
  m.def("plate_energy", &plate_energy, "Re H at plate geometry (D, theta); backward: dH/dD, dH/dtheta",
        pybind11::arg("geom"), pybind11::arg("handle"), pybind11::arg("sys"), pybind11::arg("Lx"),
        pybind11::arg("max_iter") = 0, pybind11::arg("tol") = 0.0);
}
