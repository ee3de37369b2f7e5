// Composite and host-callback potentials.
//   CombinedPotential: sum of potentials over the same coordinates
//     (source/fresco/combine_potentials.hpp:16-59, add_energy_gradient of
//     base_potential.hpp:50-58) with device-side gradient accumulation.
//   CallbackPotential: a potential implemented by the caller on the host
//     (the PythonPotential route of source/fresco/python_potential_wrapper.hpp
//     :43-184 and fresco/potentials/_pythonpotential.pyx:58-65).
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "potential.hpp"

namespace fresco_hip {

namespace {

__global__ void k_add(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += x[i];
}

// e = ((0 + e[0]) + e[1]) + ...: the host loop's order
__global__ void k_sum_energies(const double* __restrict__ e, int n, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double acc = 0.;
    for (int i = 0; i < n; ++i) acc += e[i];
    *out = acc;
  }
}

__global__ void k_zero(double* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = 0.0;
}

unsigned eb(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256)); }

class Combined final : public Potential {
 public:
  explicit Combined(std::vector<std::shared_ptr<Potential>> pots) : pots_(std::move(pots)) {
    if (pots_.empty()) throw InvalidArgument("CombinedPotential needs at least one potential");
    n_ = pots_[0]->ndof();
    for (auto& p : pots_)
      if (p->ndof() != n_) throw InvalidArgument("the gradient has the wrong size");
    tmp_.allocate(n_);
    ebuf_.allocate(pots_.size());
  }
  int64_t ndof() const override { return n_; }
  // recordable when every term is; the terms' energies are summed on the
  // device in the order of energy_gradient
  bool capture_ready() const override {
    for (auto& p : pots_)
      if (!p->capture_ready()) return false;
    return true;
  }
  uint64_t graph_epoch() const override {
    uint64_t e = 0;
    for (auto& p : pots_) e = e * 1000003ull + p->graph_epoch();
    return e;
  }
  bool energy_gradient_async(const double* x, double* g, double* e_dev, hipStream_t s) override {
    if (!capture_ready()) return false;
    hipLaunchKernelGGL(k_zero, dim3(eb(n_)), dim3(256), 0, s, g, n_);
    for (size_t i = 0; i < pots_.size(); ++i) {
      if (!pots_[i]->energy_gradient_async(x, tmp_.get(), ebuf_.get() + i, s))
        throw RuntimeFailure("a recordable potential refused its device-side evaluation");
      hipLaunchKernelGGL(k_add, dim3(eb(n_)), dim3(256), 0, s, g, tmp_.get(), n_);
    }
    hipLaunchKernelGGL(k_sum_energies, dim3(1), dim3(64), 0, s, ebuf_.get(), (int)pots_.size(), e_dev);
    FRESCO_HIP_CHECK(hipGetLastError());
    return true;
  }
  double energy(const double* x, hipStream_t s) override {
    double e = 0.;
    for (auto& p : pots_) e += p->energy(x, s);
    return e;
  }
  double energy_gradient(const double* x, double* g, hipStream_t s) override {
    double e = 0.;
    hipLaunchKernelGGL(k_zero, dim3(eb(n_)), dim3(256), 0, s, g, n_);
    for (auto& p : pots_) {
      e += p->energy_gradient(x, tmp_.get(), s);
      hipLaunchKernelGGL(k_add, dim3(eb(n_)), dim3(256), 0, s, g, tmp_.get(), n_);
    }
    FRESCO_HIP_CHECK(hipGetLastError());
    return e;
  }

 private:
  std::vector<std::shared_ptr<Potential>> pots_;
  int64_t n_;
  DeviceBuffer<double> tmp_, ebuf_;
};

class Callback final : public Potential {
 public:
  Callback(int64_t n, HostEgCallback cb, void* user) : n_(n), cb_(cb), user_(user) {
    if (n <= 0) throw InvalidArgument("ndof must be positive");
    if (!cb) throw InvalidArgument("null callback");
    FRESCO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&xh_), n * sizeof(double)));
    FRESCO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&gh_), n * sizeof(double)));
  }
  ~Callback() override {
    if (xh_) (void)hipHostFree(xh_);
    if (gh_) (void)hipHostFree(gh_);
  }
  int64_t ndof() const override { return n_; }
  double energy(const double* x, hipStream_t s) override {
    pull(x, s);
    double e = 0.;
    if (cb_(xh_, nullptr, &e, user_) != 0) throw RuntimeFailure("potential callback failed");
    return e;
  }
  double energy_gradient(const double* x, double* g, hipStream_t s) override {
    pull(x, s);
    double e = 0.;
    if (cb_(xh_, gh_, &e, user_) != 0) throw RuntimeFailure("potential callback failed");
    FRESCO_HIP_CHECK(hipMemcpyAsync(g, gh_, n_ * sizeof(double), hipMemcpyHostToDevice, s));
    FRESCO_HIP_CHECK(hipStreamSynchronize(s));
    return e;
  }

 private:
  void pull(const double* x, hipStream_t s) {
    FRESCO_HIP_CHECK(hipMemcpyAsync(xh_, x, n_ * sizeof(double), hipMemcpyDeviceToHost, s));
    FRESCO_HIP_CHECK(hipStreamSynchronize(s));
  }
  int64_t n_;
  HostEgCallback cb_;
  void* user_;
  double* xh_ = nullptr;
  double* gh_ = nullptr;
};

}  // namespace

std::shared_ptr<Potential> make_combined(std::vector<std::shared_ptr<Potential>> pots) {
  return std::make_shared<Combined>(std::move(pots));
}

std::shared_ptr<Potential> make_callback(int64_t n, HostEgCallback cb, void* user) {
  return std::make_shared<Callback>(n, cb, user);
}

}  // namespace fresco_hip
