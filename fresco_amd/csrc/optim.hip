// Device-resident optimizers driving a Potential: LBFGS with backtracking line
// search (source/lbfgs.cpp:9-257, source/fresco/lbfgs.hpp) and modified FIRE
// (source/fresco/modified_fire.hpp:44-230, source/modified_fire.cpp:43-118),
// with the shared driver of source/fresco/optimizer.hpp:55-241.
//
// All vectors (x, g, step, LBFGS memory, FIRE velocity) live in device memory
// and are updated by FP64 kernels.  Dot products are two-stage fixed-order
// reductions (bitwise reproducible).  The two-loop recursion keeps its
// alpha/beta scalars on the device, so the only host synchronisations per
// iteration are the ones the control flow needs (the uphill test, step norm,
// the energy of each line-search trial, rms, ys / yy).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>

#include "optim.hpp"

namespace fresco_hip {

namespace {

constexpr int RB = 256;      // reduction block
constexpr int RMAXB = 1024;  // maximum partial blocks

__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int h = RB / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// up to three dot products in one pass: (a0.b0, a1.b1, a2.b2)
__global__ void __launch_bounds__(RB)
    k_dots(const double* __restrict__ a0, const double* __restrict__ b0, const double* __restrict__ a1,
           const double* __restrict__ b1, const double* __restrict__ a2, const double* __restrict__ b2, int64_t n,
           int nd, double* __restrict__ part) {
  __shared__ double sh[RB];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    s0 += a0[i] * b0[i];
    if (nd > 1) s1 += a1[i] * b1[i];
    if (nd > 2) s2 += a2[i] * b2[i];
  }
  s0 = block_sum(s0, sh);
  if (nd > 1) s1 = block_sum(s1, sh);
  if (nd > 2) s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[RMAXB + blockIdx.x] = s1;
    part[2 * RMAXB + blockIdx.x] = s2;
  }
}

// out[q] = add[q] (if given) + mul * sum_b part[q][b]
__global__ void __launch_bounds__(RB)
    k_dots_final(const double* __restrict__ part, int nb, int nd, double mul, const double* __restrict__ add,
                 double* __restrict__ out) {
  __shared__ double sh[RB];
  for (int q = 0; q < nd; ++q) {
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += RB) s += part[q * RMAXB + b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[q] = (add ? add[q] : 0.0) + mul * s;
  }
}

// y += sign * coef[0] * x     (coef on the device)
__global__ void k_axpy_dev(double* __restrict__ y, const double* __restrict__ x, const double* __restrict__ coef,
                           double sign, int64_t n) {
  const double c = sign * coef[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += c * x[i];
}

// y = a * x + b * z
__global__ void k_axpbz(double* __restrict__ y, double a, const double* __restrict__ x, double b,
                        const double* __restrict__ z, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + (z ? b * z[i] : 0.0);
}

// y[k] = g - gold ; s[k] = x - xold ; partial sums of ys and yy
__global__ void __launch_bounds__(RB)
    k_lbfgs_memory(const double* __restrict__ x, const double* __restrict__ xold, const double* __restrict__ g,
                   const double* __restrict__ gold, double* __restrict__ s, double* __restrict__ y, int64_t n,
                   double* __restrict__ part) {
  __shared__ double sh[RB];
  double ys = 0, yy = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    double yv = g[i] - gold[i];
    double sv = x[i] - xold[i];
    y[i] = yv;
    s[i] = sv;
    ys += yv * sv;
    yy += yv * yv;
  }
  ys = block_sum(ys, sh);
  yy = block_sum(yy, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ys;
    part[RMAXB + blockIdx.x] = yy;
  }
}

// FIRE: v = (1-a) v - a g ifnorm vnorm ; v -= dt g ; dx = dt v ; partial |dx|^2
__global__ void __launch_bounds__(RB)
    k_fire_euler(double* __restrict__ v, const double* __restrict__ g, double* __restrict__ dx, double a,
                 double gscale, double dt, int64_t n, double* __restrict__ part) {
  __shared__ double sh[RB];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    double vi = (1.0 - a) * v[i] - a * g[i] * gscale;  // modified_fire.hpp:172-174
    vi -= dt * g[i];                                   // modified_fire.hpp:141-144
    double d = dt * vi;
    v[i] = vi;
    dx[i] = d;
    s += d * d;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

inline unsigned nblocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(RMAXB, (n + RB - 1) / RB)); }
inline unsigned eblocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256)); }

}  // namespace

// ---------------------------------------------------------------------------
// VecOps
// ---------------------------------------------------------------------------

void VecOps::init(int64_t n_) {
  n = n_;
  part.allocate(3 * RMAXB);
  scal.allocate(8);
  FRESCO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host), 8 * sizeof(double)));
}

VecOps::~VecOps() {
  if (host) (void)hipHostFree(host);
}

void VecOps::dots_dev(int nd, const double* a0, const double* b0, const double* a1, const double* b1,
                      const double* a2, const double* b2, double mul, const double* add, double* out,
                      hipStream_t s) {
  unsigned nb = nblocks(n);
  hipLaunchKernelGGL(k_dots, dim3(nb), dim3(RB), 0, s, a0, b0, a1, b1, a2, b2, n, nd, part.get());
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(RB), 0, s, part.get(), (int)nb, nd, mul, add, out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

void VecOps::fetch(int cnt, double* dst, hipStream_t s) {
  FRESCO_HIP_CHECK(hipMemcpyAsync(host, scal.get(), cnt * sizeof(double), hipMemcpyDeviceToHost, s));
  FRESCO_HIP_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < cnt; ++i) dst[i] = host[i];
}

double VecOps::dot(const double* a, const double* b, hipStream_t s) {
  dots_dev(1, a, b, nullptr, nullptr, nullptr, nullptr, 1.0, nullptr, scal.get(), s);
  double r;
  fetch(1, &r, s);
  return r;
}

void VecOps::dots3(const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
                   const double* b2, double* r, hipStream_t s) {
  dots_dev(3, a0, b0, a1, b1, a2, b2, 1.0, nullptr, scal.get(), s);
  fetch(3, r, s);
}

void VecOps::axpbz(double* y, double a, const double* x, double b, const double* z, hipStream_t s) {
  hipLaunchKernelGGL(k_axpbz, dim3(eblocks(n)), dim3(256), 0, s, y, a, x, b, z, n);
  FRESCO_HIP_CHECK(hipGetLastError());
}

void VecOps::axpy_dev(double* y, const double* x, const double* coef, double sign, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_dev, dim3(eblocks(n)), dim3(256), 0, s, y, x, coef, sign, n);
  FRESCO_HIP_CHECK(hipGetLastError());
}

void VecOps::copy(double* dst, const double* src, hipStream_t s) {
  FRESCO_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void VecOps::zero(double* y, hipStream_t s) { FRESCO_HIP_CHECK(hipMemsetAsync(y, 0, n * sizeof(double), s)); }

// ---------------------------------------------------------------------------
// GradientOptimizer (optimizer.hpp:55-241)
// ---------------------------------------------------------------------------

GradientOptimizer::GradientOptimizer(std::shared_ptr<Potential> pot, const double* x0_host, int64_t n, double tol)
    : pot_(std::move(pot)), n_(n), tol_(tol) {
  if (!pot_) throw InvalidArgument("null potential");
  if (pot_->ndof() != n)
    throw InvalidArgument("x0 has " + std::to_string(n) + " entries, the potential expects " +
                          std::to_string(pot_->ndof()));
  x_.allocate(n);
  g_.allocate(n);
  x_.upload(x0_host, n, stream_);
  ops_.init(n);
}

void GradientOptimizer::compute_func_gradient(double* x, double& f, double* g) {
  nfev_ += 1;
  f = pot_->energy_gradient(x, g, stream_);
}

void GradientOptimizer::initialize_func_gradient() {
  compute_func_gradient(x_.get(), f_, g_.get());
  double gg = ops_.dot(g_.get(), g_.get(), stream_);
  rms_ = std::sqrt(gg / (double)n_);
  func_initialized_ = true;
}

void GradientOptimizer::set_func_gradient(double f, const double* grad_host) {
  if (iter_number_ > 0) std::cout << "warning: setting f and grad after the first iteration.  this is dangerous.\n";
  f_ = f;
  g_.upload(grad_host, n_, stream_);
  double gg = ops_.dot(g_.get(), g_.get(), stream_);
  rms_ = std::sqrt(gg / (double)n_);
  func_initialized_ = true;
}

bool GradientOptimizer::stop_criterion_satisfied() {
  if (!func_initialized_) initialize_func_gradient();
  return rms_ <= tol_;
}

void GradientOptimizer::run(int niter) {
  if (!func_initialized_) initialize_func_gradient();
  for (int i = 0; i < niter; ++i) {
    if (stop_criterion_satisfied()) break;
    one_iteration();
  }
}

void GradientOptimizer::run() { run(maxiter_ - iter_number_); }

void GradientOptimizer::get_x(double* host) {
  FRESCO_HIP_CHECK(hipMemcpyAsync(host, x_.get(), n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
  FRESCO_HIP_CHECK(hipStreamSynchronize(stream_));
}

void GradientOptimizer::get_g(double* host) {
  FRESCO_HIP_CHECK(hipMemcpyAsync(host, g_.get(), n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
  FRESCO_HIP_CHECK(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------
// LBFGS (lbfgs.cpp)
// ---------------------------------------------------------------------------

LBFGS::LBFGS(std::shared_ptr<Potential> pot, const double* x0, int64_t n, double tol, int M)
    : GradientOptimizer(std::move(pot), x0, n, tol), M_(M), rho_(M, 0.0) {
  if (M < 1) throw InvalidArgument("LBFGS memory M must be >= 1");
  inv_sqrt_size_ = 1.0 / std::sqrt((double)n);
  xold_.allocate(n);
  gold_.allocate(n);
  step_.allocate(n);
  s_.allocate((size_t)n * M);
  y_.allocate((size_t)n * M);
  alpha_.allocate(M);
  coef_.allocate(1);
}

void LBFGS::one_iteration() {
  if (!func_initialized_) initialize_func_gradient();
  ops_.copy(xold_.get(), x_.get(), stream_);
  ops_.copy(gold_.get(), g_.get(), stream_);
  compute_lbfgs_step();
  double stepnorm = backtracking_linesearch();
  update_memory();
  if (iprint_ > 0 && iter_number_ % iprint_ == 0) {
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10) << "lbgs: " << iter_number_ << " E "
              << f_ << " rms " << rms_ << " nfev " << nfev_ << " step norm " << stepnorm << std::endl;
  }
  iter_number_ += 1;
}

void LBFGS::update_memory() {
  const int kl = k_ % M_;
  double* s = s_.get() + (size_t)kl * n_;
  double* y = y_.get() + (size_t)kl * n_;
  unsigned nb = nblocks(n_);
  hipLaunchKernelGGL(k_lbfgs_memory, dim3(nb), dim3(RB), 0, stream_, x_.get(), xold_.get(), g_.get(), gold_.get(), s,
                     y, n_, ops_.part.get());
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(RB), 0, stream_, ops_.part.get(), (int)nb, 2, 1.0, nullptr,
                     ops_.scal.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  double r[2];
  ops_.fetch(2, r, stream_);
  double ys = r[0], yy = r[1];
  if (ys == 0.) {
    if (verbosity_ > 0) std::cout << "warning: resetting YS to 1." << std::endl;
    ys = 1.;
  }
  rho_[kl] = 1. / ys;
  if (yy == 0.) {
    if (verbosity_ > 0) std::cout << "warning: resetting YY to 1." << std::endl;
    yy = 1.;
  }
  H0_ = ys / yy;
  k_ += 1;
}

void LBFGS::compute_lbfgs_step() {
  if (k_ == 0) {
    double gnorm = std::sqrt(ops_.dot(g_.get(), g_.get(), stream_));
    if (gnorm > 1.) gnorm = 1. / gnorm;
    ops_.axpbz(step_.get(), -gnorm * H0_, g_.get(), 0.0, nullptr, stream_);
    return;
  }
  ops_.copy(step_.get(), g_.get(), stream_);
  const int jmin = std::max(0, k_ - M_);
  const int jmax = k_;
  for (int j = jmax - 1; j >= jmin; --j) {
    const int i = j % M_;
    const double* s = s_.get() + (size_t)i * n_;
    const double* y = y_.get() + (size_t)i * n_;
    // alpha_i = rho_i s_i.step ; step -= alpha_i y_i
    ops_.dots_dev(1, s, step_.get(), nullptr, nullptr, nullptr, nullptr, rho_[i], nullptr, alpha_.get() + i,
                  stream_);
    ops_.axpy_dev(step_.get(), y, alpha_.get() + i, -1.0, stream_);
  }
  ops_.axpbz(step_.get(), -H0_, step_.get(), 0.0, nullptr, stream_);
  for (int j = jmin; j < jmax; ++j) {
    const int i = j % M_;
    const double* s = s_.get() + (size_t)i * n_;
    const double* y = y_.get() + (size_t)i * n_;
    // beta = -rho_i y_i.step ; step -= s_i (alpha_i - beta)
    ops_.dots_dev(1, y, step_.get(), nullptr, nullptr, nullptr, nullptr, rho_[i], alpha_.get() + i, coef_.get(),
                  stream_);
    ops_.axpy_dev(step_.get(), s, coef_.get(), -1.0, stream_);
  }
}

double LBFGS::backtracking_linesearch() {
  double dotprod = ops_.dot(step_.get(), g_.get(), stream_);
  if (dotprod > 0.) {
    if (verbosity_ > 1) std::cout << "warning: step direction was uphill.  inverting" << std::endl;
    ops_.axpbz(step_.get(), -1.0, step_.get(), 0.0, nullptr, stream_);
  }
  double factor = 1.;
  double stepnorm = std::sqrt(ops_.dot(step_.get(), step_.get(), stream_));
  if (factor * stepnorm > maxstep_) factor = maxstep_ / stepnorm;
  double fnew = f_;
  int nred;
  const int nred_max = 10;
  for (nred = 0; nred < nred_max; ++nred) {
    ops_.axpbz(x_.get(), 1.0, xold_.get(), factor, step_.get(), stream_);
    compute_func_gradient(x_.get(), fnew, g_.get());
    double df = fnew - f_;
    if (use_relative_f_) {
      double absf = 1e-100;
      if (f_ != 0) absf = std::abs(f_);
      df /= absf;
    }
    if (df < max_f_rise_) break;
    factor *= 0.5;
    if (verbosity_ > 2)
      std::cout << "energy increased by " << df << " to " << fnew << " from " << f_ << " reducing step norm to "
                << factor * stepnorm << " H0 " << H0_ << std::endl;
  }
  if (nred >= nred_max && verbosity_ > 0) std::cout << "warning: the line search backtracked too many times" << std::endl;
  f_ = fnew;
  rms_ = std::sqrt(ops_.dot(g_.get(), g_.get(), stream_)) * inv_sqrt_size_;
  return stepnorm * factor;
}

void LBFGS::reset(const double* x0_host) {
  k_ = 0;
  iter_number_ = 0;
  nfev_ = 0;
  x_.upload(x0_host, n_, stream_);
  initialize_func_gradient();
}

// ---------------------------------------------------------------------------
// MODIFIED_FIRE (modified_fire.hpp / modified_fire.cpp)
// ---------------------------------------------------------------------------

ModifiedFire::ModifiedFire(std::shared_ptr<Potential> pot, const double* x0, int64_t n, double dtstart, double dtmax,
                           double maxstep, int64_t Nmin, double finc, double fdec, double fa, double astart,
                           double tol, bool stepback)
    : GradientOptimizer(std::move(pot), x0, n, tol),
      dtstart_(dtstart),
      dt_(dtstart),
      dtmax_(dtmax),
      fmaxstep_(maxstep),
      Nmin_((double)Nmin),
      finc_(finc),
      fdec_(fdec),
      fa_(fa),
      astart_(astart),
      a_(astart),
      stepback_(stepback) {
  v_.allocate(n);
  dx_.allocate(n);
  xold_.allocate(n);
  gold_.allocate(n);
  ops_.zero(v_.get(), stream_);
  ops_.copy(xold_.get(), x_.get(), stream_);
  ops_.zero(gold_.get(), stream_);
}

void ModifiedFire::initialize_func_gradient() {
  nfev_ += 1;
  f_ = pot_->energy_gradient(x_.get(), g_.get(), stream_);
  fold_ = f_;
  ops_.axpbz(v_.get(), -dt_, g_.get(), 0.0, nullptr, stream_);
  double r[3];
  ops_.dots3(g_.get(), g_.get(), v_.get(), v_.get(), g_.get(), g_.get(), r, stream_);
  ifnorm_ = 1. / std::sqrt(r[0]);
  vnorm_ = std::sqrt(r[1]);
  rms_ = 1. / (ifnorm_ * std::sqrt((double)n_));
  func_initialized_ = true;
}

void ModifiedFire::set_func_gradient(double f, const double* grad_host) {
  if (iter_number_ > 0) std::cout << "warning: setting f and grad after the first iteration.  this is dangerous.\n";
  f_ = f;
  fold_ = f_;
  g_.upload(grad_host, n_, stream_);
  ops_.axpbz(v_.get(), -dt_, g_.get(), 0.0, nullptr, stream_);
  double r[3];
  ops_.dots3(g_.get(), g_.get(), v_.get(), v_.get(), g_.get(), g_.get(), r, stream_);
  ifnorm_ = 1. / std::sqrt(r[0]);
  vnorm_ = std::sqrt(r[1]);
  rms_ = 1. / (ifnorm_ * std::sqrt((double)n_));
  func_initialized_ = true;
}

void ModifiedFire::one_iteration() {
  if (!func_initialized_) initialize_func_gradient();
  nfev_ += 1;
  iter_number_ += 1;
  fire_iter_number_ += 1;
  fold_ = f_;
  ops_.copy(xold_.get(), x_.get(), stream_);
  // velocity mixing + forward Euler (modified_fire.hpp:136-158, 172-177)
  unsigned nb = nblocks(n_);
  hipLaunchKernelGGL(k_fire_euler, dim3(nb), dim3(RB), 0, stream_, v_.get(), g_.get(), dx_.get(), a_,
                     ifnorm_ * vnorm_, dt_, n_, ops_.part.get());
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(RB), 0, stream_, ops_.part.get(), (int)nb, 1, 1.0, nullptr,
                     ops_.scal.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  ops_.copy(gold_.get(), g_.get(), stream_);
  double nd;
  ops_.fetch(1, &nd, stream_);
  double normdx = std::sqrt(nd);
  double factor = 1;
  if (normdx > fmaxstep_) factor = fmaxstep_ / normdx;
  ops_.axpbz(x_.get(), 1.0, x_.get(), factor, dx_.get(), stream_);
  f_ = pot_->energy_gradient(x_.get(), g_.get(), stream_);
  double r[3];
  ops_.dots3(v_.get(), g_.get(), g_.get(), g_.get(), v_.get(), v_.get(), r, stream_);
  double P = -r[0];
  if (P > 0) {
    if ((double)fire_iter_number_ > Nmin_) {
      dt_ = std::min(dt_ * finc_, dtmax_);
      a_ *= fa_;
    }
    ifnorm_ = 1. / std::sqrt(r[1]);
    vnorm_ = std::sqrt(r[2]);
    rms_ = 1. / (ifnorm_ * std::sqrt((double)n_));
  } else {
    dt_ *= fdec_;
    a_ = astart_;
    fire_iter_number_ = 0;
    ops_.zero(v_.get(), stream_);
    if (stepback_) {
      f_ = fold_;
      ops_.copy(x_.get(), xold_.get(), stream_);
      ops_.copy(g_.get(), gold_.get(), stream_);
    }
    if (verbosity_ > 1) std::cout << "warning: step direction was uphill.  inverting\n";
  }
  if (iprint_ > 0 && iter_number_ % iprint_ == 0) {
    std::cout << "fire: " << iter_number_ << " fire_iter_number " << fire_iter_number_ << " dt " << dt_ << " a " << a_
              << " P " << P << " vnorm " << vnorm_ << " E " << f_ << " rms " << rms_ << " nfev " << nfev_ << "\n";
  }
}

void ModifiedFire::reset(const double* x0_host) {
  // modified_fire.hpp:84-112
  if (!func_initialized_) initialize_func_gradient();
  iter_number_ = 0;
  x_.upload(x0_host, n_, stream_);
  f_ = pot_->energy_gradient(x_.get(), g_.get(), stream_);
  nfev_ = 1;
  fire_iter_number_ = 0;
  dt_ = dtstart_;
  a_ = astart_;
  fold_ = f_;
  ops_.copy(xold_.get(), x_.get(), stream_);
  ops_.copy(gold_.get(), g_.get(), stream_);
  ops_.axpbz(v_.get(), -dt_, g_.get(), 0.0, nullptr, stream_);
  double r[3];
  ops_.dots3(g_.get(), g_.get(), v_.get(), v_.get(), g_.get(), g_.get(), r, stream_);
  ifnorm_ = 1. / std::sqrt(r[0]);
  vnorm_ = std::sqrt(r[1]);
  rms_ = 1. / (ifnorm_ * std::sqrt((double)n_));
}

}  // namespace fresco_hip
