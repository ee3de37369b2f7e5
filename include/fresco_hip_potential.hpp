/*
 * fresco_hip_potential.hpp -- a potential of libfresco_hip.so behind the
 * reference's own C++ potential interface, fresco::BasePotential
 * (source/fresco/base_potential.hpp:15-43), so that the reference's
 * optimizers (fresco::LBFGS, source/lbfgs.cpp; fresco::MODIFIED_FIRE,
 * source/fresco/modified_fire.hpp) drive the GPU potential unchanged through
 * their get_energy_gradient call (optimizer.hpp:208-215).
 *
 * This header belongs on the reference's side: it is compiled together with
 * the reference's sources (its include path must reach
 * "fresco/base_potential.hpp") and links libfresco_hip.so and the HIP
 * runtime.  Host vectors are copied to and from device buffers owned by the
 * adapter around every call (O(ndof) bytes each way).
 *
 * Errors follow the reference's convention: FRESCO_E_INVALID becomes
 * std::invalid_argument (ValueError through Cython's `except +`), any other
 * failure std::runtime_error.
 */
#ifndef FRESCO_HIP_POTENTIAL_HPP
#define FRESCO_HIP_POTENTIAL_HPP

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "fresco/base_potential.hpp"
#include "fresco_hip.h"

namespace fresco {

class HipPotential : public BasePotential {
 public:
  /* h: a potential made by one of the fresco_*_create entry points.  With
   * owns = true the adapter destroys it with itself. */
  explicit HipPotential(fresco_pot* h, bool owns = true) : h_(h), owns_(owns) {
    int64_t n = 0;
    check(fresco_pot_ndof(h_, &n));
    n_ = (size_t)n;
    hip(hipMalloc((void**)&x_, n_ * sizeof(double)), "hipMalloc");
    hip(hipMalloc((void**)&g_, n_ * sizeof(double)), "hipMalloc");
  }
  HipPotential(const HipPotential&) = delete;
  HipPotential& operator=(const HipPotential&) = delete;
  ~HipPotential() override {
    (void)hipFree(x_);
    (void)hipFree(g_);
    if (owns_) (void)fresco_pot_destroy(h_);
  }

  size_t get_ndim() override { return n_; }

  /* BasePotential::get_energy (base_potential.hpp:28-31) */
  double get_energy(const std::vector<double>& x) override {
    upload(x);
    double e = 0.0;
    check(fresco_pot_energy(h_, x_, &e, nullptr));
    return e;
  }

  /* BasePotential::get_energy_gradient (base_potential.hpp:38-43) */
  double get_energy_gradient(const std::vector<double>& x, std::vector<double>& grad) override {
    if (grad.size() != n_) throw std::invalid_argument("the gradient has the wrong size");
    upload(x);
    double e = 0.0;
    check(fresco_pot_energy_gradient(h_, x_, g_, &e, nullptr));
    hip(hipMemcpy(grad.data(), g_, n_ * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    return e;
  }

  fresco_pot* handle() const { return h_; }

 private:
  void upload(const std::vector<double>& x) {
    if (x.size() != n_) throw std::invalid_argument("the coordinates have the wrong size");
    hip(hipMemcpy(x_, x.data(), n_ * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  }
  static void check(int rc) {
    if (rc == FRESCO_OK) return;
    if (rc == FRESCO_E_INVALID) throw std::invalid_argument(fresco_last_error());
    throw std::runtime_error(fresco_last_error());
  }
  static void hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }

  fresco_pot* h_;
  bool owns_;
  size_t n_ = 0;
  double* x_ = nullptr;
  double* g_ = nullptr;
};

}  // namespace fresco

#endif /* FRESCO_HIP_POTENTIAL_HPP */
