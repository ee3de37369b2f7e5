// Device-resident potentials: the C++ counterpart of fresco::BasePotential
// (source/fresco/base_potential.hpp:15-156) with x / grad in device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <thread>
#include <vector>

#include "common.hpp"
#include "prof.hpp"

namespace fresco_hip {

class Potential {
 public:
  virtual ~Potential() = default;
  virtual int64_t ndof() const = 0;
  virtual int64_t nmodes() const { return 0; }
  // base_potential.hpp:28-31
  virtual double energy(const double* x, hipStream_t s) = 0;
  // base_potential.hpp:38-43; grad (device, ndof) is overwritten
  virtual double energy_gradient(const double* x, double* grad, hipStream_t s) = 0;
  // energy_gradient that returns once the energy has reached the host; the
  // rest of the gradient pass may still be running on s (stream-ordered
  // callers).  Potentials without that form synchronise fully.
  virtual double energy_gradient_stream(const double* x, double* grad, hipStream_t s) {
    return energy_gradient(x, grad, s);
  }
  // energy_gradient without the wait: enqueues the evaluation on s and
  // leaves the energy in *e_dev (device).  Returns false (enqueuing nothing)
  // when the potential has no such path.  What it enqueues is kernels only
  // (no memory fills, host waits or allocations), so it may be recorded
  // into a HIP graph once capture_ready() holds; graph_epoch() changes
  // whenever later calls would enqueue other work or other buffers.
  virtual bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) {
    (void)x;
    (void)grad;
    (void)e_dev;
    (void)s;
    return false;
  }
  virtual bool capture_ready() const { return false; }
  virtual uint64_t graph_epoch() const { return 0; }
  // rho_k of the last evaluation, complex interleaved float64 (device)
  virtual void get_rho(double* rho, hipStream_t s) {
    (void)rho;
    (void)s;
    throw RuntimeFailure("this potential does not expose rho");
  }
  // d phi / d S0_k and d phi / d V_k at the last evaluation's rho (device,
  // nmodes each, the potential's own mode order)
  virtual void target_grads(double* dS0, double* dV, hipStream_t s) {
    (void)dS0;
    (void)dV;
    (void)s;
    throw RuntimeFailure("this potential has no S0 / V targets");
  }
  // energy at x, keeping what gradient_kept needs (the split evaluation of
  // a differentiable loss: forward here, backward later); *token names the
  // kept evaluation.  False (nothing done) when the potential has no split
  virtual bool energy_keep(const double* x, double* e, uint64_t* token, hipStream_t s) {
    (void)x;
    (void)e;
    (void)token;
    (void)s;
    return false;
  }
  // the gradient (device, ndof) of the kept evaluation named token; false
  // when that is not what the potential keeps (another evaluation or a
  // target update came in between).  grad null: only that check
  virtual bool gradient_kept(uint64_t token, double* grad, hipStream_t s) {
    (void)token;
    (void)grad;
    (void)s;
    return false;
  }
  // replace the targets S0 and weights V (device, nmodes each; V null: the
  // weights stay)
  virtual void set_targets(const double* Sk, const double* V, hipStream_t s) {
    (void)Sk;
    (void)V;
    (void)s;
    throw RuntimeFailure("this potential has no S0 / V targets");
  }
  // energy of every pattern of a batched potential (host, n entries) at the
  // last evaluation
  virtual void pattern_energies(double* host, int n, hipStream_t s) {
    (void)host;
    (void)n;
    (void)s;
    throw RuntimeFailure("not a batched potential");
  }
  // spreader configuration in use: mode 0 register tiles / owner tiles,
  // 1 halo tiles; parts = point-range slices per bin (-1 when not applicable)
  // where spread_config's choice came from: 0 fixed, 1 tuning cache or
  // table, 2 timed on the first evaluation, 3 pinned by the environment
  virtual int spread_source() const { return 0; }
  virtual void spread_config(int* mode, int* parts) const {
    *mode = -1;
    *parts = -1;
  }
  virtual void plan_info(int64_t* nf, int* w, double* beta, int* tile) const {
    for (int a = 0; a < 3; ++a) {
      nf[a] = 0;
      tile[a] = 0;
    }
    *w = 0;
    *beta = 0;
  }
};

// Pinned host scalar used to return energies from the device.
class HostScalar {
 public:
  HostScalar() { FRESCO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p_), sizeof(double))); }
  ~HostScalar() {
    if (p_) (void)hipHostFree(p_);
  }
  HostScalar(const HostScalar&) = delete;
  HostScalar& operator=(const HostScalar&) = delete;
  double fetch(const double* dev, hipStream_t s) {
    FRESCO_HIP_CHECK(hipMemcpyAsync(p_, dev, sizeof(double), hipMemcpyDeviceToHost, s));
    FRESCO_HIP_CHECK(hipStreamSynchronize(s));
    if (prof_enabled()) prof_collect();
    return *p_;
  }
  // the scalar as kernels address it (a kernel writes it directly: no copy)
  double* dev() {
    if (!d_) FRESCO_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_), p_, 0));
    return d_;
  }
  // after a kernel enqueued on s wrote it through dev()
  double wait(hipStream_t s) {
    FRESCO_HIP_CHECK(hipStreamSynchronize(s));
    if (prof_enabled()) prof_collect();
    return *p_;
  }
  // before enqueuing a kernel that writes it: a marker the kernel's value
  // replaces (a NaN payload no arithmetic produces)
  void arm() { *reinterpret_cast<volatile uint64_t*>(p_) = kMarker; }
  // the value, as soon as the kernel (enqueued on s after arm()) has written
  // it: the host polls the pinned word instead of waiting for the whole
  // stream; if the stream drains without the value (a failed launch), the
  // stream is synchronised, which reports the error.  Stage-profiler events
  // completed so far are collected (fresco_prof_get collects the rest).
  double wait_value(hipStream_t s) {
    const volatile uint64_t* w = reinterpret_cast<const volatile uint64_t*>(p_);
    for (unsigned it = 1;; ++it) {
      const uint64_t a = *w;
      // two equal reads: the value is whole (not read while being written)
      if (a != kMarker && a == *w) break;
      if ((it & 255u) == 0u) {
        if (hipStreamQuery(s) != hipErrorNotReady) {
          FRESCO_HIP_CHECK(hipStreamSynchronize(s));
          if (*w == kMarker) throw RuntimeFailure("the evaluation finished without delivering its energy");
          break;
        }
        if (it >= 4096u) std::this_thread::yield();
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    if (prof_enabled()) prof_collect();
    return *p_;
  }


 private:
  static constexpr uint64_t kMarker = 0x7ff4f3e5c0de0000ull;
  double* p_ = nullptr;
  double* d_ = nullptr;
};

// Per-mode derivatives of phi = sum V D^2, D = (|rho|^2/norm - S0)[/S0 where
// S0 != 0]:  dphi/dV = D^2;  dphi/dS0 = -2 V D |rho|^2/(norm S0^2) where
// S0 != 0, else -2 V D.  norm is *norm_dev when given, else norm.
// rho: complex interleaved float64 [m].
void launch_target_grads(const double* rho, const double* Sk, const double* V, int64_t m, double norm,
                         const double* norm_dev, double* dS0, double* dV, hipStream_t s);

// Host callback: x (host, ndof) -> *energy and, if grad != NULL, grad (host).
// Returns 0 on success.
typedef int (*HostEgCallback)(const double* x, double* grad, double* energy, void* user);

std::shared_ptr<Potential> make_combined(std::vector<std::shared_ptr<Potential>> pots);
std::shared_ptr<Potential> make_callback(int64_t n, HostEgCallback cb, void* user);

std::unique_ptr<Potential> make_nuwu(const double* radii, int64_t n, int d, double K, const double* Sk,
                                     const double* V, int64_t m, const double* L, double eps, int precision,
                                     double sigma);

// shape[d]: grid, first entry slowest (row-major, the FFTW plan of uwu.hpp:48)
std::unique_ptr<Potential> make_uwu(const int64_t* shape, int d, const double* Sk, const double* V, int64_t m,
                                    int precision, int exact);
// shape[d]: mode grid, first entry fastest (the FINUFFT n_modes of uwnu.hpp:44-82)
std::unique_ptr<Potential> make_uwnu(const int64_t* shape, int d, const double* K, int64_t nkd, const double* Sk,
                                     const double* V, int64_t nk, const double* L, double eps, int precision,
                                     double sigma);
// method 0: type-3 NUFFT (eps, sigma); 1: exact sums with float64 phases
std::unique_ptr<Potential> make_nuwnu(const double* radii, int64_t n, int d, const double* K, int64_t nkd,
                                      const double* Sk, const double* V, int64_t m, const double* L, double eps,
                                      int precision, int method, double sigma);

// kind 0 inverse power (p1 = exponent a), 1 WCA (p1 = sigma); method 1
// cartesian, 2 periodic (box[d]); cells != 0: cell lists (periodic only)
std::unique_ptr<Potential> make_pair(int kind, double p1, double eps, const double* radii, int64_t n, int d,
                                     const double* box, int method, int cells, double ncellx_scale);

// npat independent 1D / 2D NUwU patterns: npts[p] points (radii concatenated),
// nmodes[p] = (2K_p+1)^d modes (Sk, V concatenated), one box L
std::unique_ptr<Potential> make_nuwu_batch(int npat, const int64_t* npts, const double* radii, int d,
                                           const int64_t* nmodes, const double* Sk, const double* V, const double* L,
                                           double eps, int precision, double sigma);

}  // namespace fresco_hip
