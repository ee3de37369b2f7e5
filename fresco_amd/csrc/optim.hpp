// Device-resident LBFGS and modified FIRE (see optim.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <iostream>
#include <memory>
#include <vector>

#include "common.hpp"
#include "potential.hpp"

namespace fresco_hip {

// FP64 vector operations on n-element device vectors with fixed-order
// (reproducible) reductions.
struct VecOps {
  int64_t n = 0;
  DeviceBuffer<double> part;  // [3][1024] block partials
  DeviceBuffer<double> scal;  // device scalars
  double* host = nullptr;     // pinned host mirror of scal
  void init(int64_t n_);
  ~VecOps();
  // out[q] = add[q] + mul * (a_q . b_q), q < nd   (device result)
  void dots_dev(int nd, const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
                const double* b2, double mul, const double* add, double* out, hipStream_t s);
  void fetch(int cnt, double* dst, hipStream_t s);
  double dot(const double* a, const double* b, hipStream_t s);
  void dots3(const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
             const double* b2, double* r, hipStream_t s);
  void axpbz(double* y, double a, const double* x, double b, const double* z, hipStream_t s);  // y = a x + b z
  void axpy_dev(double* y, const double* x, const double* coef, double sign, hipStream_t s);   // y += sign c x
  void copy(double* dst, const double* src, hipStream_t s);
  void zero(double* y, hipStream_t s);
};

// optimizer.hpp:55-241
class GradientOptimizer {
 public:
  GradientOptimizer(std::shared_ptr<Potential> pot, const double* x0_host, int64_t n, double tol);
  virtual ~GradientOptimizer() = default;
  virtual void one_iteration() = 0;
  virtual void reset(const double* x0_host) = 0;
  virtual void initialize_func_gradient();
  virtual void set_func_gradient(double f, const double* grad_host);
  void run(int niter);
  void run();
  bool stop_criterion_satisfied();

  void set_tol(double v) { tol_ = v; }
  virtual void set_maxstep(double v) { maxstep_ = v; }
  void set_max_iter(int v) { maxiter_ = v; }
  void set_iprint(int v) { iprint_ = v; }
  void set_verbosity(int v) { verbosity_ = v; }
  void set_stream(hipStream_t s) { stream_ = s; }

  double get_f() const { return f_; }
  double get_rms() const { return rms_; }
  int get_nfev() const { return nfev_; }
  int get_niter() const { return iter_number_; }
  int get_maxiter() const { return maxiter_; }
  bool success() { return stop_criterion_satisfied(); }
  virtual double get_H0() const { return 0.0; }
  int64_t size() const { return n_; }
  void get_x(double* host);
  void get_g(double* host);
  double* x_device() { return x_.get(); }
  double* g_device() { return g_.get(); }
  // HIP graph replay of whole iterations (where the optimizer supports it)
  virtual void set_use_graph(bool on) { (void)on; }
  virtual void graph_stats(int64_t* captures, int64_t* replays) const {
    *captures = 0;
    *replays = 0;
  }

 protected:
  void compute_func_gradient(double* x, double& f, double* g);

  std::shared_ptr<Potential> pot_;
  int64_t n_;
  double tol_;
  double maxstep_ = 0.1;
  int maxiter_ = 1000;
  int iprint_ = -1;
  int verbosity_ = 0;
  int iter_number_ = 0;
  int nfev_ = 0;
  DeviceBuffer<double> x_, g_;
  double f_ = 0.;
  double rms_ = 1e10;
  double gg_ = 0.;  // g . g at x
  bool func_initialized_ = false;
  hipStream_t stream_ = nullptr;
  VecOps ops_;
};

// lbfgs.cpp:9-257.  Scalars (H0, rho, k, step factor, energies) live on the
// device so that one iteration is a fixed kernel sequence, replayed as a HIP
// graph when the potential allows it (see optim.hip).
class LBFGS final : public GradientOptimizer {
 public:
  LBFGS(std::shared_ptr<Potential> pot, const double* x0, int64_t n, double tol, int M);
  ~LBFGS() override;
  LBFGS(const LBFGS&) = delete;
  LBFGS& operator=(const LBFGS&) = delete;
  void one_iteration() override;
  void reset(const double* x0_host) override;
  void initialize_func_gradient() override;
  void set_func_gradient(double f, const double* grad_host) override;
  void set_H0(double H0);
  void set_maxstep(double v) override;
  void set_max_f_rise(double v);
  void set_use_relative_f(int v);
  double get_H0() const override { return H0_; }
  void set_use_graph(bool on) override { use_graph_ = on; }
  // 0: vector-free recursion from VL_MIN_N unknowns on (M <= 16), else the
  // slot-by-slot two-loop recursion; 1: always slot by slot; 2: always
  // vector-free (M <= 16).  Only before the first iteration.
  void set_form(int form);
  bool vector_free() const { return vl_; }
  static constexpr int64_t VL_MIN_N = 1 << 16;
  void graph_stats(int64_t* captures, int64_t* replays) const override {
    *captures = captures_;
    *replays = replays_;
  }

 private:
  void set_scalar(int idx, double v);
  void read_scalars();
  void enqueue_iteration(hipStream_t s);
  void enqueue_trial(bool dev_factor, double host_factor, hipStream_t s);
  void enqueue_eval(hipStream_t s);
  void enqueue_commit(int force, hipStream_t s);
  void enqueue_gdots(hipStream_t s);
  bool run_graph();

  int M_;
  double max_f_rise_ = 1e-4;
  bool use_relative_f_ = false;
  double H0_ = 0.1;
  int k_ = 0;
  double inv_sqrt_size_ = 1.0;
  DeviceBuffer<double> xold_, gold_, step_, s_, y_, alpha_, dsc_;
  bool vl_ = false;                             // vector-free recursion
  DeviceBuffer<double> gram_, delta_, part_vl_;  // its Gram matrix, coefficients, partial sums
  double* hsc_ = nullptr;  // pinned mirror of the device scalars
  bool use_graph_ = true;
  hipStream_t cap_ = nullptr;
  hipGraphExec_t gexec_ = nullptr;
  uint64_t gepoch_ = 0;
  int64_t captures_ = 0, replays_ = 0;
};

// modified_fire.hpp:44-230
class ModifiedFire final : public GradientOptimizer {
 public:
  ModifiedFire(std::shared_ptr<Potential> pot, const double* x0, int64_t n, double dtstart, double dtmax,
               double maxstep, int64_t Nmin, double finc, double fdec, double fa, double astart, double tol,
               bool stepback);
  ~ModifiedFire() override;
  ModifiedFire(const ModifiedFire&) = delete;
  ModifiedFire& operator=(const ModifiedFire&) = delete;
  void one_iteration() override;
  void reset(const double* x0_host) override;
  void initialize_func_gradient() override;
  void set_func_gradient(double f, const double* grad_host) override;
  void set_use_graph(bool on) override { use_graph_ = on; }
  void graph_stats(int64_t* captures, int64_t* replays) const override {
    *captures = captures_;
    *replays = replays_;
  }

 private:
  // iterations on device scalars, recorded as HIP graphs (one per buffer
  // assignment) when the potential can be recorded
  void enqueue_iteration(hipStream_t s);
  void graph_iteration();
  void upload_scalars();

  double dtstart_, dt_, dtmax_, fmaxstep_, Nmin_, finc_, fdec_, fa_, astart_, a_;
  double fold_ = 0, ifnorm_ = 0, vnorm_ = 0;
  size_t fire_iter_number_ = 0;
  bool stepback_;
  DeviceBuffer<double> v_, dx_, xold_, gold_, dsc_;
  double* hsc_ = nullptr;  // pinned mirror of the device scalars
  bool dirty_ = true;      // host scalars newer than the device ones
  bool use_graph_ = true;
  hipStream_t cap_ = nullptr;
  hipGraphExec_t gexec_[2] = {nullptr, nullptr};
  const double* gx_[2] = {nullptr, nullptr};  // x buffer each graph starts from
  uint64_t gepoch_ = 0;
  int64_t captures_ = 0, replays_ = 0;
};

}  // namespace fresco_hip
