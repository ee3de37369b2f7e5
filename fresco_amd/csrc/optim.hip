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
#include <utility>

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

__device__ __forceinline__ double wave_sum_xor(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
  return x;
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

// y = a * x + b * z   (b = *bdev when bdev is given: a device scalar)
__global__ void k_axpbz(double* __restrict__ y, double a, const double* __restrict__ x, double b,
                        const double* __restrict__ z, int64_t n, const double* __restrict__ bdev = nullptr) {
  if (bdev) b = *bdev;
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

// FIRE device scalars (the recorded iteration keeps its state here)
enum : int {
  FS_F = 0,   // energy at the current x
  FS_FOLD,    // energy at the previous x
  FS_DT,      // time step
  FS_A,       // mixing
  FS_IFNORM,  // 1 / |g|
  FS_VNORM,   // |v|
  FS_NITER,   // iterations since the last uphill step
  FS_P,       // power -v . g of the last step
  FS_ND,      // |dx|^2
  FS_FACTOR,  // step scale (maxstep)
  FS_UPHILL,  // 1 when the last step went uphill (P <= 0)
  FS_E,       // energy of the last evaluation
  FS_R0,      // v . g, g . g, v . v at the new x
  FS_R1,
  FS_R2,
  FS_RMS,
  FS_COUNT
};

// FIRE: v = (1-a) v - a g ifnorm vnorm ; v -= dt g ; dx = dt v ; partial |dx|^2
// (a, ifnorm vnorm and dt from dsc when it is given)
__global__ void __launch_bounds__(RB)
    k_fire_euler(double* __restrict__ v, const double* __restrict__ g, double* __restrict__ dx, double a,
                 double gscale, double dt, int64_t n, double* __restrict__ part,
                 const double* __restrict__ dsc = nullptr) {
  __shared__ double sh[RB];
  if (dsc) {
    a = dsc[FS_A];
    gscale = dsc[FS_IFNORM] * dsc[FS_VNORM];
    dt = dsc[FS_DT];
  }
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

// modified_fire.hpp:146-158: step scale from |dx| (FS_ND); the iteration's
// bookkeeping (previous energy, iteration count)
__global__ void k_fire_factor(double* __restrict__ dsc, double fmaxstep) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    dsc[FS_FOLD] = dsc[FS_F];
    dsc[FS_NITER] += 1.0;
    const double normdx = sqrt(dsc[FS_ND]);
    double factor = 1;
    if (normdx > fmaxstep) factor = fmaxstep / normdx;
    dsc[FS_FACTOR] = factor;
  }
}

// modified_fire.hpp:178-216 after the evaluation: FS_R0..2 = v.g, g.g, v.v
__global__ void k_fire_update(double* __restrict__ dsc, double Nmin, double finc, double dtmax, double fa,
                              double fdec, double astart, double sqrtn, int stepback) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double P = -dsc[FS_R0];
    dsc[FS_P] = P;
    dsc[FS_F] = dsc[FS_E];
    if (P > 0) {
      if (dsc[FS_NITER] > Nmin) {
        const double dtn = dsc[FS_DT] * finc;
        dsc[FS_DT] = dtmax < dtn ? dtmax : dtn;  // std::min order
        dsc[FS_A] *= fa;
      }
      const double ifn = 1. / sqrt(dsc[FS_R1]);
      dsc[FS_IFNORM] = ifn;
      dsc[FS_VNORM] = sqrt(dsc[FS_R2]);
      dsc[FS_RMS] = 1. / (ifn * sqrtn);
      dsc[FS_UPHILL] = 0.0;
    } else {
      dsc[FS_DT] *= fdec;
      dsc[FS_A] = astart;
      dsc[FS_NITER] = 0.0;
      dsc[FS_UPHILL] = 1.0;
      if (stepback) dsc[FS_F] = dsc[FS_FOLD];
    }
  }
}

// v = 0 after an uphill step
__global__ void k_fire_zero_v(double* __restrict__ v, const double* __restrict__ dsc, int64_t n) {
  if (dsc[FS_UPHILL] == 0.0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.0;
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
  if (prof_enabled()) prof_collect();
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
  gg_ = gg;
  rms_ = std::sqrt(gg / (double)n_);
  func_initialized_ = true;
}

void GradientOptimizer::set_func_gradient(double f, const double* grad_host) {
  if (iter_number_ > 0) std::cout << "warning: setting f and grad after the first iteration.  this is dangerous.\n";
  f_ = f;
  g_.upload(grad_host, n_, stream_);
  double gg = ops_.dot(g_.get(), g_.get(), stream_);
  gg_ = gg;
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
//
// One iteration is a fixed sequence of kernels whose scalars (H0, rho, the
// memory count k, the step factor, the energies) live in dsc_ on the device:
//   begin       xold = x, gold = g, step = g (or the k = 0 first step)
//   2 x M slots the two-loop recursion; slot kernels of memory entries not
//               yet filled (k < M) return at once
//   line        uphill test, step norm and factor of the first trial
//   trial       x = xold + factor step, then the potential's evaluation
//   memory      s, y of slot k % M and the sums ys, yy, g.g; committed (rho,
//               H0, f, k += 1) only when the trial meets the energy-rise
//               criterion
// The same sequence runs launch by launch or, once the potential can be
// recorded, as one replayed HIP graph; either way the host reads the scalars
// back once.  A rejected first trial continues the backtracking from the host
// (lbfgs.cpp:175-215) and then commits the memory unconditionally.
// ---------------------------------------------------------------------------

namespace {

enum : int {
  LS_F = 0,     // energy at the accepted x
  LS_FNEW,      // energy of the last trial
  LS_H0,        // lbfgs.cpp:100
  LS_FACTOR,    // signed factor of the first trial (negative: step inverted)
  LS_STEPNORM,  // |step|
  LS_DOT,       // step . g before the inversion
  LS_GG,        // g . g at the current x
  LS_YS,        // raw y . s and y . y of the last memory update
  LS_YY,
  LS_K,           // memory updates so far
  LS_ACCEPT,      // 1 when the last committed trial was accepted
  LS_COEF,        // coefficient of the current two-loop slot
  LS_MAXSTEP,     // parameters
  LS_MAXRISE,
  LS_REL,
  LS_RHO = 16  // rho[M]
};

// memory slot i (0 <= i < M) of two-loop slot jj in loop 0 (backwards) or 1
// (forwards), -1 when that entry is not filled yet (lbfgs.cpp:128-160)
__device__ __forceinline__ int lb_slot(const double* dsc, int M, int loop, int jj) {
  const int k = (int)dsc[LS_K];
  const int cnt = k < M ? k : M;
  if (jj >= cnt) return -1;
  const int j = loop == 0 ? k - 1 - jj : (k - cnt) + jj;
  return j % M;
}

__global__ void k_lb_set(double* __restrict__ dsc, int i0, double v0, int i1, double v1, int i2, double v2, int i3,
                         double v3) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (i0 >= 0) dsc[i0] = v0;
    if (i1 >= 0) dsc[i1] = v1;
    if (i2 >= 0) dsc[i2] = v2;
    if (i3 >= 0) dsc[i3] = v3;
  }
}

// lbfgs.cpp:110-127 and the copies of lbfgs.cpp:40-41
__global__ void k_lb_begin(const double* __restrict__ x, const double* __restrict__ g, double* __restrict__ xold,
                           double* __restrict__ gold, double* __restrict__ step, const double* __restrict__ dsc,
                           int64_t n) {
  const bool first = dsc[LS_K] == 0.0;
  double pref = 1.0;
  if (first) {
    double gn = sqrt(dsc[LS_GG]);
    if (gn > 1.) gn = 1. / gn;
    pref = -gn * dsc[LS_H0];
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double gi = g[i];
    xold[i] = x[i];
    gold[i] = gi;
    step[i] = first ? pref * gi : gi;
  }
}

// partial sums of v_i . step for the memory entry of a two-loop slot
__global__ void __launch_bounds__(RB)
    k_lb_dot(const double* __restrict__ vecs, const double* __restrict__ step, int64_t n,
             const double* __restrict__ dsc, int M, int loop, int jj, double* __restrict__ part) {
  const int i = lb_slot(dsc, M, loop, jj);
  if (i < 0) return;
  __shared__ double sh[RB];
  const double* v = vecs + (int64_t)i * n;
  double acc = 0;
  for (int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x; t < n; t += (int64_t)gridDim.x * RB) acc += v[t] * step[t];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// loop 0: alpha_i = rho_i (s_i . step)            (lbfgs.cpp:133-138)
// loop 1: coef = alpha_i - beta, beta = -rho_i (y_i . step)  (lbfgs.cpp:149-155)
__global__ void __launch_bounds__(RB)
    k_lb_coef(const double* __restrict__ part, int nb, double* __restrict__ dsc, double* __restrict__ alpha, int M,
              int loop, int jj) {
  const int i = lb_slot(dsc, M, loop, jj);
  if (i < 0) return;
  __shared__ double sh[RB];
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += RB) s += part[b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    const double r = dsc[LS_RHO + i] * s;
    if (loop == 0) {
      alpha[i] = r;
      dsc[LS_COEF] = r;
    } else {
      dsc[LS_COEF] = alpha[i] + r;
    }
  }
}

// step -= coef v_i   (lbfgs.cpp:139-142, 157-160)
__global__ void k_lb_axpy(double* __restrict__ step, const double* __restrict__ vecs, int64_t n,
                          const double* __restrict__ dsc, int M, int loop, int jj) {
  const int i = lb_slot(dsc, M, loop, jj);
  if (i < 0) return;
  const double c = dsc[LS_COEF];
  const double* v = vecs + (int64_t)i * n;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    step[t] -= c * v[t];
}

// step *= -H0 once the memory holds an entry (lbfgs.cpp:144-147)
__global__ void k_lb_scale(double* __restrict__ step, const double* __restrict__ dsc, int64_t n) {
  if (dsc[LS_K] == 0.0) return;
  const double c = -dsc[LS_H0];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    step[t] *= c;
}

// uphill inversion and the maxstep cap (lbfgs.cpp:166-186) from the sums
// (step . g, step . step)
__global__ void __launch_bounds__(RB) k_lb_line(const double* __restrict__ part, int nb, double* __restrict__ dsc) {
  __shared__ double sh[RB];
  double s0 = 0, s1 = 0;
  for (int b = threadIdx.x; b < nb; b += RB) {
    s0 += part[b];
    s1 += part[RMAXB + b];
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    const double sn = sqrt(s1);
    double factor = 1.;
    if (factor * sn > dsc[LS_MAXSTEP]) factor = dsc[LS_MAXSTEP] / sn;
    dsc[LS_DOT] = s0;
    dsc[LS_STEPNORM] = sn;
    dsc[LS_FACTOR] = s0 > 0. ? -factor : factor;
  }
}

// x = xold + factor step (lbfgs.cpp:193)
__global__ void k_lb_trial(double* __restrict__ x, const double* __restrict__ xold, const double* __restrict__ step,
                           const double* __restrict__ dsc, int64_t n) {
  const double f = dsc[LS_FACTOR];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    x[t] = xold[t] + f * step[t];
}

// y_k = g - gold, s_k = x - xold in slot k % M; sums ys, yy, g.g
// (lbfgs.cpp:66-84)
__global__ void __launch_bounds__(RB)
    k_lb_memory(const double* __restrict__ x, const double* __restrict__ xold, const double* __restrict__ g,
                const double* __restrict__ gold, double* __restrict__ sbase, double* __restrict__ ybase, int64_t n,
                const double* __restrict__ dsc, int M, double* __restrict__ part) {
  __shared__ double sh[RB];
  const int kl = (int)dsc[LS_K] % M;
  double* s = sbase + (int64_t)kl * n;
  double* y = ybase + (int64_t)kl * n;
  double ys = 0, yy = 0, gg = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    const double gi = g[i];
    const double yv = gi - gold[i];
    const double sv = x[i] - xold[i];
    y[i] = yv;
    s[i] = sv;
    ys += yv * sv;
    yy += yv * yv;
    gg += gi * gi;
  }
  ys = block_sum(ys, sh);
  yy = block_sum(yy, sh);
  gg = block_sum(gg, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ys;
    part[RMAXB + blockIdx.x] = yy;
    part[2 * RMAXB + blockIdx.x] = gg;
  }
}

// energy-rise test of the trial (lbfgs.cpp:195-204); on acceptance (or
// force) rho, H0, f and k as in lbfgs.cpp:86-104.  Returns the decision.
__device__ bool lb_commit_scalars(double* __restrict__ dsc, int M, int force, double ys, double yy, double gg) {
  const double f = dsc[LS_F], fnew = dsc[LS_FNEW];
  double df = fnew - f;
  if (dsc[LS_REL] != 0.0) {
    double absf = 1e-100;
    if (f != 0) absf = fabs(f);
    df /= absf;
  }
  const bool accept = force != 0 || df < dsc[LS_MAXRISE];
  dsc[LS_GG] = gg;
  dsc[LS_YS] = ys;
  dsc[LS_YY] = yy;
  dsc[LS_ACCEPT] = accept ? 1.0 : 0.0;
  if (accept) {
    const int k = (int)dsc[LS_K];
    const double ys1 = ys == 0. ? 1. : ys;
    const double yy1 = yy == 0. ? 1. : yy;
    dsc[LS_RHO + k % M] = 1. / ys1;
    dsc[LS_H0] = ys1 / yy1;
    dsc[LS_F] = fnew;
    dsc[LS_K] = (double)(k + 1);
  }
  return accept;
}

__global__ void __launch_bounds__(RB)
    k_lb_commit(const double* __restrict__ part, int nb, double* __restrict__ dsc, int M, int force) {
  __shared__ double sh[RB];
  double ys = 0, yy = 0, gg = 0;
  for (int b = threadIdx.x; b < nb; b += RB) {
    ys += part[b];
    yy += part[RMAXB + b];
    gg += part[2 * RMAXB + b];
  }
  ys = block_sum(ys, sh);
  yy = block_sum(yy, sh);
  gg = block_sum(gg, sh);
  if (threadIdx.x == 0) lb_commit_scalars(dsc, M, force, ys, yy, gg);
}

// ---------------------------------------------------------------------------
// Small problems (n <= SMALL_T * 16): begin, the slot-by-slot two-loop
// recursion (lbfgs.cpp:110-160), the uphill test / maxstep factor
// (lbfgs.cpp:166-186) and the first trial x = xold + factor step in ONE
// workgroup, the step held in registers (PER values per thread).  Same
// arithmetic as the multi-kernel recursion except the reduction tree of the
// dot products; one launch instead of ~6M.
// ---------------------------------------------------------------------------

constexpr int SMALL_T = 1024;
constexpr int SMALL_MAXM = 64;

// sum over the workgroup (fixed order); every thread gets the result
__device__ __forceinline__ double small_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wave_sum_xor(v);
  __syncthreads();  // red may still be read by the previous call
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < SMALL_T / 64; ++w) t += red[w];
  return t;
}

template <int PER>
__global__ void __launch_bounds__(SMALL_T)
    k_lb_small(double* __restrict__ x, const double* __restrict__ g, double* __restrict__ xold,
               double* __restrict__ gold, double* __restrict__ step, const double* __restrict__ sb,
               const double* __restrict__ yb, int64_t n, double* __restrict__ dsc, int M) {
  __shared__ double red[SMALL_T / 64];
  __shared__ double alpha[SMALL_MAXM];
  const int tid = threadIdx.x;
  const int k = (int)dsc[LS_K];
  const bool first = k == 0;
  double pref = 1.0;
  if (first) {
    double gn = sqrt(dsc[LS_GG]);
    if (gn > 1.) gn = 1. / gn;
    pref = -gn * dsc[LS_H0];
  }
  double st[PER], gv[PER], xo[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int64_t t = tid + (int64_t)q * SMALL_T;
    gv[q] = 0.0;
    xo[q] = 0.0;
    if (t < n) {
      gv[q] = g[t];
      xo[q] = x[t];
      xold[t] = xo[q];
      gold[t] = gv[q];
    }
    st[q] = first ? pref * gv[q] : gv[q];
  }
  const int cnt = k < M ? k : M;
  for (int j = k - 1; j >= k - cnt; --j) {  // lbfgs.cpp:133-143
    const int i = j % M;
    const double* sv = sb + (int64_t)i * n;
    const double* yv = yb + (int64_t)i * n;
    double d = 0.0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t t = tid + (int64_t)q * SMALL_T;
      if (t < n) d += sv[t] * st[q];
    }
    const double a = dsc[LS_RHO + i] * small_sum(d, red);
    if (tid == 0) alpha[i] = a;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t t = tid + (int64_t)q * SMALL_T;
      if (t < n) st[q] -= a * yv[t];
    }
  }
  if (!first) {  // lbfgs.cpp:144-147
    const double h = -dsc[LS_H0];
#pragma unroll
    for (int q = 0; q < PER; ++q) st[q] *= h;
  }
  __syncthreads();  // alpha (LDS) written by thread 0
  for (int j = k - cnt; j < k; ++j) {  // lbfgs.cpp:149-160
    const int i = j % M;
    const double* sv = sb + (int64_t)i * n;
    const double* yv = yb + (int64_t)i * n;
    double d = 0.0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t t = tid + (int64_t)q * SMALL_T;
      if (t < n) d += yv[t] * st[q];
    }
    const double c = alpha[i] + dsc[LS_RHO + i] * small_sum(d, red);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t t = tid + (int64_t)q * SMALL_T;
      if (t < n) st[q] -= c * sv[t];
    }
  }
  double d0 = 0.0, d1 = 0.0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    d0 += st[q] * gv[q];
    d1 += st[q] * st[q];
  }
  d0 = small_sum(d0, red);
  d1 = small_sum(d1, red);
  const double sn = sqrt(d1);
  double factor = 1.;
  if (factor * sn > dsc[LS_MAXSTEP]) factor = dsc[LS_MAXSTEP] / sn;
  const double f = d0 > 0. ? -factor : factor;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int64_t t = tid + (int64_t)q * SMALL_T;
    if (t < n) {
      step[t] = st[q];
      x[t] = xo[q] + f * st[q];
    }
  }
  __syncthreads();  // every thread has read dsc before thread 0 writes it
  if (tid == 0) {
    dsc[LS_DOT] = d0;
    dsc[LS_STEPNORM] = sn;
    dsc[LS_FACTOR] = f;
  }
}

// ---------------------------------------------------------------------------
// Vector-free two-loop recursion (Chen, Wang & Zhou, "Large-scale L-BFGS
// using MapReduce", NIPS 2014).  The step is a combination of the basis
// {s_q, y_q (filled slots), g}; the two-loop recursion of lbfgs.cpp:128-160
// runs on its 2M+1 coefficients with the basis' dot products (one thread),
// and the step is formed in one pass over the basis.  The dot products of a
// new pair with every slot and of every slot with g come out of the memory
// pass that writes the pair, so an iteration reads each memory vector twice
// (vs about five times in the slot-by-slot recursion).
// ---------------------------------------------------------------------------

constexpr int VL_MAXM = 16;

// Gram storage: ss[p*M+q] = s_p.s_q, sy[p*M+q] = s_p.y_q, yy[p*M+q] = y_p.y_q,
// sg[q] = s_q.g, yg[q] = y_q.g (g the gradient at the committed x)
struct VlGram {
  double* ss;
  double* sy;
  double* yy;
  double* sg;
  double* yg;
};

// slots holding a pair after k memory updates
__device__ __forceinline__ unsigned vl_mask(int k, int M) {
  const int c = k < M ? k : M;
  return c >= 32 ? 0xffffffffu : ((1u << c) - 1u);
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

// 64-lane sums of V values, then the 4 wavefronts in LDS;
// part[v * RMAXB + blockIdx.x] = block sum of v
template <int V>
__device__ __forceinline__ void vl_block_store(double (&v)[V], double* __restrict__ part) {
  __shared__ double sh[RB / 64][V];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const double x = wave_sum(v[i]);
    if (lane == 0) sh[wv][i] = x;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < V; i += RB) {
    double t = 0;
#pragma unroll
    for (int w = 0; w < RB / 64; ++w) t += sh[w][i];
    part[(int64_t)i * RMAXB + blockIdx.x] = t;
  }
}

// tot[v] = sum of the nb block partials of value v (one workgroup per value)
__global__ void __launch_bounds__(RB) k_vl_totals(const double* __restrict__ part, int nb, double* __restrict__ tot) {
  __shared__ double sh[RB];
  const int v = blockIdx.x;
  double x = 0;
  for (int b = threadIdx.x; b < nb; b += RB) x += part[(int64_t)v * RMAXB + b];
  x = block_sum(x, sh);
  if (threadIdx.x == 0) tot[v] = x;
}

// coefficients delta = [s_0..s_15, y_0..y_15, g] of the step: one wavefront,
// lane l owns basis vector l (s_l for l < M, y_{l-M} for l < 2M, g for 2M);
// the Gram entries are staged in LDS first
__global__ void __launch_bounds__(64)
    k_vl_coef(const double* __restrict__ dsc, VlGram G, double* __restrict__ delta, int M) {
  __shared__ double ss[VL_MAXM * VL_MAXM], sy[VL_MAXM * VL_MAXM], yy[VL_MAXM * VL_MAXM];
  __shared__ double sg[VL_MAXM], yg[VL_MAXM], rho[VL_MAXM], alpha[VL_MAXM];
  const int l = threadIdx.x;
  for (int i = l; i < M * M; i += 64) {
    ss[i] = G.ss[i];
    sy[i] = G.sy[i];
    yy[i] = G.yy[i];
  }
  if (l < M) {
    sg[l] = G.sg[l];
    yg[l] = G.yg[l];
    rho[l] = dsc[LS_RHO + l];
  }
  __syncthreads();
  const int k = (int)dsc[LS_K];
  const int cnt = k < M ? k : M;
  const unsigned mask = vl_mask(k, M);
  const bool is_s = l < M, is_y = l >= M && l < 2 * M, is_g = l == 2 * M;
  const int q = is_s ? l : l - M;
  const bool on = is_g || ((is_s || is_y) && ((mask >> q) & 1u));
  double dl = 0.0;
  if (k == 0) {
    // lbfgs.cpp:110-125
    double gn = sqrt(dsc[LS_GG]);
    if (gn > 1.) gn = 1. / gn;
    if (is_g) dl = -gn * dsc[LS_H0];
  } else {
    if (is_g) dl = 1.0;
    for (int j = k - 1; j >= k - cnt; --j) {  // lbfgs.cpp:133-143
      const int i = j % M;
      double term = 0.0;
      if (on) term = dl * (is_g ? sg[i] : (is_s ? ss[i * M + q] : sy[i * M + q]));
      const double a = rho[i] * wave_sum(term);
      if (l == 0) alpha[i] = a;
      if (is_y && q == i) dl -= a;
    }
    dl *= -dsc[LS_H0];  // lbfgs.cpp:144-147
    __syncthreads();
    for (int j = k - cnt; j < k; ++j) {  // lbfgs.cpp:149-160
      const int i = j % M;
      double term = 0.0;
      if (on) term = dl * (is_g ? yg[i] : (is_s ? sy[q * M + i] : yy[i * M + q]));
      const double t = wave_sum(term);
      if (is_s && q == i) dl -= alpha[i] + rho[i] * t;
    }
  }
  if (is_s) delta[q] = dl;
  if (is_y) delta[VL_MAXM + q] = dl;
  if (is_g) delta[2 * VL_MAXM] = dl;
}

// step = sum of delta times the basis; xold = x, gold = g; partial sums of
// step.g and step.step (k_lb_line finishes them).  Loads of slots not filled
// yet are issued anyway (always in bounds) and their values discarded, so all
// loads of an element can be in flight together.
template <int M>
__global__ void __launch_bounds__(RB)
    k_vl_step(const double* __restrict__ x, const double* __restrict__ g, double* __restrict__ xold,
              double* __restrict__ gold, double* __restrict__ step, const double* __restrict__ sb,
              const double* __restrict__ yb, int64_t n, const double* __restrict__ delta,
              const double* __restrict__ dsc, double* __restrict__ part) {
  __shared__ double sh[RB];
  const unsigned mask = vl_mask((int)dsc[LS_K], M);
  double ds[M], dy[M];
#pragma unroll
  for (int q = 0; q < M; ++q) {
    const bool on = (mask >> q) & 1u;
    ds[q] = on ? delta[q] : 0.0;
    dy[q] = on ? delta[VL_MAXM + q] : 0.0;
  }
  const double dg = delta[2 * VL_MAXM];
  double a0 = 0, a1 = 0;
  for (int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x; t < n; t += (int64_t)gridDim.x * RB) {
    const double gt = g[t];
    const double xt = x[t];
    double sv[M], yv[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      sv[q] = sb[(int64_t)q * n + t];
      yv[q] = yb[(int64_t)q * n + t];
    }
    double st = dg * gt;
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const bool on = (mask >> q) & 1u;
      st += on ? ds[q] * sv[q] + dy[q] * yv[q] : 0.0;
    }
    step[t] = st;
    xold[t] = xt;
    gold[t] = gt;
    a0 += st * gt;
    a1 += st * st;
  }
  a0 = block_sum(a0, sh);
  a1 = block_sum(a1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a0;
    part[RMAXB + blockIdx.x] = a1;
  }
}

// memory pass: s, y of slot k % M; ys, yy, g.g; and for every slot filled
// after this update, its dot products with g and with the new pair
template <int M>
__global__ void __launch_bounds__(RB)
    k_vl_memory(const double* __restrict__ x, const double* __restrict__ xold, const double* __restrict__ g,
                const double* __restrict__ gold, double* __restrict__ sb, double* __restrict__ yb, int64_t n,
                const double* __restrict__ dsc, double* __restrict__ part) {
  constexpr int V = 6 * M + 3;
  const int k = (int)dsc[LS_K];
  const int kl = k % M;
  const unsigned mask = vl_mask(k + 1, M);
  double acc[V];
#pragma unroll
  for (int i = 0; i < V; ++i) acc[i] = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x; t < n; t += (int64_t)gridDim.x * RB) {
    const double gi = g[t];
    const double snew = x[t] - xold[t];
    const double ynew = gi - gold[t];
    double sv[M], yv[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      sv[q] = sb[(int64_t)q * n + t];
      yv[q] = yb[(int64_t)q * n + t];
    }
    acc[0] += ynew * snew;
    acc[1] += ynew * ynew;
    acc[2] += gi * gi;
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const bool on = (mask >> q) & 1u;
      const double sq = q == kl ? snew : (on ? sv[q] : 0.0);
      const double yq = q == kl ? ynew : (on ? yv[q] : 0.0);
      acc[3 + 6 * q] += sq * gi;
      acc[4 + 6 * q] += yq * gi;
      acc[5 + 6 * q] += snew * sq;
      acc[6 + 6 * q] += snew * yq;
      acc[7 + 6 * q] += ynew * sq;
      acc[8 + 6 * q] += ynew * yq;
    }
    sb[(int64_t)kl * n + t] = snew;
    yb[(int64_t)kl * n + t] = ynew;
  }
  vl_block_store<V>(acc, part);
}

// commit of the memory pass from the totals: the scalars as k_lb_commit, and
// on acceptance the Gram rows of the new pair and every slot's products with g
__global__ void k_vl_commit(const double* __restrict__ tot, double* __restrict__ dsc, VlGram G, int M, int force) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int k = (int)dsc[LS_K];
  const int kl = k % M;
  const unsigned mask = vl_mask(k + 1, M);
  if (!lb_commit_scalars(dsc, M, force, tot[0], tot[1], tot[2])) return;
  for (int q = 0; q < M; ++q) {
    if (!((mask >> q) & 1u)) continue;
    const double* a = tot + 3 + 6 * q;
    G.sg[q] = a[0];
    G.yg[q] = a[1];
    G.ss[kl * M + q] = a[2];
    G.ss[q * M + kl] = a[2];
    G.sy[kl * M + q] = a[3];  // s_new . y_q
    G.sy[q * M + kl] = a[4];  // s_q . y_new
    G.yy[kl * M + q] = a[5];
    G.yy[q * M + kl] = a[5];
  }
}

// s_q.g and y_q.g of the filled slots (after the gradient is replaced)
template <int M>
__global__ void __launch_bounds__(RB)
    k_vl_gdots(const double* __restrict__ g, const double* __restrict__ sb, const double* __restrict__ yb, int64_t n,
               double* __restrict__ part) {
  constexpr int V = 2 * M;
  double acc[V];
#pragma unroll
  for (int i = 0; i < V; ++i) acc[i] = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x; t < n; t += (int64_t)gridDim.x * RB) {
    const double gi = g[t];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      acc[2 * q] += sb[(int64_t)q * n + t] * gi;
      acc[2 * q + 1] += yb[(int64_t)q * n + t] * gi;
    }
  }
  vl_block_store<V>(acc, part);
}

__global__ void k_vl_gdots_set(const double* __restrict__ tot, const double* __restrict__ dsc, VlGram G, int M) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const unsigned mask = vl_mask((int)dsc[LS_K], M);
  for (int q = 0; q < M; ++q) {
    if (!((mask >> q) & 1u)) continue;
    G.sg[q] = tot[2 * q];
    G.yg[q] = tot[2 * q + 1];
  }
}

}  // namespace

LBFGS::LBFGS(std::shared_ptr<Potential> pot, const double* x0, int64_t n, double tol, int M)
    : GradientOptimizer(std::move(pot), x0, n, tol), M_(M) {
  if (M < 1) throw InvalidArgument("LBFGS memory M must be >= 1");
  inv_sqrt_size_ = 1.0 / std::sqrt((double)n);
  xold_.allocate(n);
  gold_.allocate(n);
  step_.allocate(n);
  s_.allocate((size_t)n * M);
  y_.allocate((size_t)n * M);
  alpha_.allocate(M);
  dsc_.allocate(LS_RHO + (size_t)M);
  FRESCO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hsc_), LS_RHO * sizeof(double)));
  // rho_ starts at 0 (lbfgs.cpp:21) and is only read for filled entries
  std::vector<double> init(LS_RHO + (size_t)M, 0.0);
  init[LS_H0] = H0_;
  init[LS_MAXSTEP] = maxstep_;
  init[LS_MAXRISE] = max_f_rise_;
  init[LS_REL] = use_relative_f_ ? 1.0 : 0.0;
  dsc_.upload(init.data(), init.size(), stream_);
  set_form(0);
}

void LBFGS::set_form(int form) {
  if (form < 0 || form > 2) throw InvalidArgument("LBFGS form must be 0 (auto), 1 (two-loop) or 2 (vector-free)");
  if (iter_number_ > 0 || k_ > 0) throw InvalidArgument("the LBFGS form can only be chosen before the first iteration");
  if (form == 2 && M_ > VL_MAXM)
    throw InvalidArgument("the vector-free form takes M <= " + std::to_string(VL_MAXM));
  vl_ = form == 2 || (form == 0 && n_ >= VL_MIN_N && M_ <= VL_MAXM);
  if (vl_ && !gram_.get()) {
    gram_.allocate(3 * (size_t)M_ * M_ + 2 * (size_t)M_);
    delta_.allocate(2 * VL_MAXM + 1 + 6 * VL_MAXM + 3);  // coefficients, then the pass totals
    part_vl_.allocate((size_t)(6 * VL_MAXM + 3) * RMAXB);
  }
  if (gexec_) {
    FRESCO_HIP_CHECK(hipGraphExecDestroy(gexec_));
    gexec_ = nullptr;
  }
}

LBFGS::~LBFGS() {
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (cap_) (void)hipStreamDestroy(cap_);
  if (hsc_) (void)hipHostFree(hsc_);
}

void LBFGS::set_H0(double H0) {
  if (iter_number_ > 0) std::cout << "warning: setting H0 after the first iteration.\n";
  H0_ = H0;
  set_scalar(LS_H0, H0);
}

void LBFGS::set_maxstep(double v) {
  maxstep_ = v;
  set_scalar(LS_MAXSTEP, v);
}

void LBFGS::set_max_f_rise(double v) {
  max_f_rise_ = v;
  set_scalar(LS_MAXRISE, v);
}

void LBFGS::set_use_relative_f(int v) {
  use_relative_f_ = v != 0;
  set_scalar(LS_REL, use_relative_f_ ? 1.0 : 0.0);
}

void LBFGS::set_scalar(int idx, double v) {
  hipLaunchKernelGGL(k_lb_set, dim3(1), dim3(64), 0, stream_, dsc_.get(), idx, v, -1, 0.0, -1, 0.0, -1, 0.0);
  FRESCO_HIP_CHECK(hipGetLastError());
}

void LBFGS::initialize_func_gradient() {
  GradientOptimizer::initialize_func_gradient();
  hipLaunchKernelGGL(k_lb_set, dim3(1), dim3(64), 0, stream_, dsc_.get(), (int)LS_F, f_, (int)LS_GG, gg_, -1, 0.0, -1,
                     0.0);
  FRESCO_HIP_CHECK(hipGetLastError());
}

void LBFGS::set_func_gradient(double f, const double* grad_host) {
  GradientOptimizer::set_func_gradient(f, grad_host);
  hipLaunchKernelGGL(k_lb_set, dim3(1), dim3(64), 0, stream_, dsc_.get(), (int)LS_F, f_, (int)LS_GG, gg_, -1, 0.0, -1,
                     0.0);
  FRESCO_HIP_CHECK(hipGetLastError());
  // the vector-free recursion keeps the memory's products with g
  if (vl_ && k_ > 0) enqueue_gdots(stream_);
}

void LBFGS::read_scalars() {
  FRESCO_HIP_CHECK(hipMemcpyAsync(hsc_, dsc_.get(), LS_RHO * sizeof(double), hipMemcpyDeviceToHost, stream_));
  FRESCO_HIP_CHECK(hipStreamSynchronize(stream_));
  if (prof_enabled()) prof_collect();
}

// trial x = xold + factor step and its evaluation; the energy lands in
// dsc_[LS_FNEW].  dev_factor: the first trial's factor, LS_FACTOR on the
// device; otherwise host_factor (the host's backtracking).
void LBFGS::enqueue_trial(bool dev_factor, double host_factor, hipStream_t s) {
  const unsigned eb = eblocks(n_);
  if (dev_factor)
    hipLaunchKernelGGL(k_lb_trial, dim3(eb), dim3(256), 0, s, x_.get(), xold_.get(), step_.get(), dsc_.get(), n_);
  else
    hipLaunchKernelGGL(k_axpbz, dim3(eb), dim3(256), 0, s, x_.get(), 1.0, xold_.get(), host_factor, step_.get(), n_);
  FRESCO_HIP_CHECK(hipGetLastError());
  enqueue_eval(s);
}

// the potential's evaluation at x_ (energy to dsc_[LS_FNEW])
void LBFGS::enqueue_eval(hipStream_t s) {
  nfev_ += 1;
  if (!pot_->energy_gradient_async(x_.get(), g_.get(), dsc_.get() + LS_FNEW, s)) {
    const double f = pot_->energy_gradient(x_.get(), g_.get(), s);
    hipLaunchKernelGGL(k_lb_set, dim3(1), dim3(64), 0, s, dsc_.get(), (int)LS_FNEW, f, -1, 0.0, -1, 0.0, -1, 0.0);
    FRESCO_HIP_CHECK(hipGetLastError());
  }
}

namespace {

VlGram vl_gram(double* base, int M) {
  VlGram G;
  G.ss = base;
  G.sy = base + (size_t)M * M;
  G.yy = base + 2 * (size_t)M * M;
  G.sg = base + 3 * (size_t)M * M;
  G.yg = G.sg + M;
  return G;
}

// launches F<M> for the memory length M (1..16) of this optimizer
#define FRESCO_VL_DISPATCH(M, CALL) \
  switch (M) {                      \
    case 1: CALL(1); break;         \
    case 2: CALL(2); break;         \
    case 3: CALL(3); break;         \
    case 4: CALL(4); break;         \
    case 5: CALL(5); break;         \
    case 6: CALL(6); break;         \
    case 7: CALL(7); break;         \
    case 8: CALL(8); break;         \
    case 9: CALL(9); break;         \
    case 10: CALL(10); break;       \
    case 11: CALL(11); break;       \
    case 12: CALL(12); break;       \
    case 13: CALL(13); break;       \
    case 14: CALL(14); break;       \
    case 15: CALL(15); break;       \
    default: CALL(16); break;       \
  }

}  // namespace

void LBFGS::enqueue_commit(int force, hipStream_t s) {
  const unsigned nb = nblocks(n_);
  if (vl_) {
    const VlGram G = vl_gram(gram_.get(), M_);
    double* part = part_vl_.get();
    double* tot = delta_.get() + 2 * VL_MAXM + 1;
#define FRESCO_VL_MEM(MM)                                                                                     \
  hipLaunchKernelGGL(k_vl_memory<MM>, dim3(nb), dim3(RB), 0, s, x_.get(), xold_.get(), g_.get(), gold_.get(), \
                     s_.get(), y_.get(), n_, dsc_.get(), part)
    FRESCO_VL_DISPATCH(M_, FRESCO_VL_MEM)
#undef FRESCO_VL_MEM
    hipLaunchKernelGGL(k_vl_totals, dim3(6 * M_ + 3), dim3(RB), 0, s, part, (int)nb, tot);
    hipLaunchKernelGGL(k_vl_commit, dim3(1), dim3(64), 0, s, tot, dsc_.get(), G, M_, force);
    FRESCO_HIP_CHECK(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_lb_memory, dim3(nb), dim3(RB), 0, s, x_.get(), xold_.get(), g_.get(), gold_.get(), s_.get(),
                     y_.get(), n_, dsc_.get(), M_, ops_.part.get());
  hipLaunchKernelGGL(k_lb_commit, dim3(1), dim3(RB), 0, s, ops_.part.get(), (int)nb, dsc_.get(), M_, force);
  FRESCO_HIP_CHECK(hipGetLastError());
}

void LBFGS::enqueue_gdots(hipStream_t s) {
  const unsigned nb = nblocks(n_);
  const VlGram G = vl_gram(gram_.get(), M_);
  double* part = part_vl_.get();
  double* tot = delta_.get() + 2 * VL_MAXM + 1;
#define FRESCO_VL_GD(MM) \
  hipLaunchKernelGGL(k_vl_gdots<MM>, dim3(nb), dim3(RB), 0, s, g_.get(), s_.get(), y_.get(), n_, part)
  FRESCO_VL_DISPATCH(M_, FRESCO_VL_GD)
#undef FRESCO_VL_GD
  hipLaunchKernelGGL(k_vl_totals, dim3(2 * M_), dim3(RB), 0, s, part, (int)nb, tot);
  hipLaunchKernelGGL(k_vl_gdots_set, dim3(1), dim3(64), 0, s, tot, dsc_.get(), G, M_);
  FRESCO_HIP_CHECK(hipGetLastError());
}

// direction, first trial and the conditional memory update (lbfgs.cpp:31-44)
void LBFGS::enqueue_iteration(hipStream_t s) {
  const unsigned eb = eblocks(n_);
  const unsigned nb = nblocks(n_);
  double* part = ops_.part.get();
  double* dsc = dsc_.get();
  if (vl_) {
    const VlGram G = vl_gram(gram_.get(), M_);
    hipLaunchKernelGGL(k_vl_coef, dim3(1), dim3(64), 0, s, dsc, G, delta_.get(), M_);
#define FRESCO_VL_STEP(MM)                                                                                   \
  hipLaunchKernelGGL(k_vl_step<MM>, dim3(nb), dim3(RB), 0, s, x_.get(), g_.get(), xold_.get(), gold_.get(), \
                     step_.get(), s_.get(), y_.get(), n_, delta_.get(), dsc, part)
    FRESCO_VL_DISPATCH(M_, FRESCO_VL_STEP)
#undef FRESCO_VL_STEP
    hipLaunchKernelGGL(k_lb_line, dim3(1), dim3(RB), 0, s, part, (int)nb, dsc);
    FRESCO_HIP_CHECK(hipGetLastError());
    enqueue_trial(true, 0.0, s);
    enqueue_commit(0, s);
    return;
  }
  if (n_ <= 16 * SMALL_T && M_ <= SMALL_MAXM) {
    const int per = (int)((n_ + SMALL_T - 1) / SMALL_T);
#define FRESCO_LB_SMALL(P)                                                                                        \
  hipLaunchKernelGGL(k_lb_small<P>, dim3(1), dim3(SMALL_T), 0, s, x_.get(), g_.get(), xold_.get(), gold_.get(), \
                     step_.get(), s_.get(), y_.get(), n_, dsc, M_)
    if (per <= 1) FRESCO_LB_SMALL(1);
    else if (per <= 2) FRESCO_LB_SMALL(2);
    else if (per <= 4) FRESCO_LB_SMALL(4);
    else if (per <= 8) FRESCO_LB_SMALL(8);
    else FRESCO_LB_SMALL(16);
#undef FRESCO_LB_SMALL
    FRESCO_HIP_CHECK(hipGetLastError());
    enqueue_eval(s);  // the kernel above wrote the trial x
    enqueue_commit(0, s);
    return;
  }
  hipLaunchKernelGGL(k_lb_begin, dim3(eb), dim3(256), 0, s, x_.get(), g_.get(), xold_.get(), gold_.get(), step_.get(),
                     dsc, n_);
  for (int jj = 0; jj < M_; ++jj) {
    hipLaunchKernelGGL(k_lb_dot, dim3(nb), dim3(RB), 0, s, s_.get(), step_.get(), n_, dsc, M_, 0, jj, part);
    hipLaunchKernelGGL(k_lb_coef, dim3(1), dim3(RB), 0, s, part, (int)nb, dsc, alpha_.get(), M_, 0, jj);
    hipLaunchKernelGGL(k_lb_axpy, dim3(eb), dim3(256), 0, s, step_.get(), y_.get(), n_, dsc, M_, 0, jj);
  }
  hipLaunchKernelGGL(k_lb_scale, dim3(eb), dim3(256), 0, s, step_.get(), dsc, n_);
  for (int jj = 0; jj < M_; ++jj) {
    hipLaunchKernelGGL(k_lb_dot, dim3(nb), dim3(RB), 0, s, y_.get(), step_.get(), n_, dsc, M_, 1, jj, part);
    hipLaunchKernelGGL(k_lb_coef, dim3(1), dim3(RB), 0, s, part, (int)nb, dsc, alpha_.get(), M_, 1, jj);
    hipLaunchKernelGGL(k_lb_axpy, dim3(eb), dim3(256), 0, s, step_.get(), s_.get(), n_, dsc, M_, 1, jj);
  }
  hipLaunchKernelGGL(k_dots, dim3(nb), dim3(RB), 0, s, step_.get(), g_.get(), step_.get(), step_.get(), nullptr,
                     nullptr, n_, 2, part);
  hipLaunchKernelGGL(k_lb_line, dim3(1), dim3(RB), 0, s, part, (int)nb, dsc);
  FRESCO_HIP_CHECK(hipGetLastError());
  enqueue_trial(true, 0.0, s);
  enqueue_commit(0, s);
}

bool LBFGS::run_graph() {
  const uint64_t ep = pot_->graph_epoch();
  if (!gexec_ || ep != gepoch_) {
    if (gexec_) {
      FRESCO_HIP_CHECK(hipGraphExecDestroy(gexec_));
      gexec_ = nullptr;
    }
    if (!cap_) FRESCO_HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
    // the captured evaluation must not count as one
    const int nfev = nfev_;
    hipGraph_t g = nullptr;
    FRESCO_HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_iteration(cap_);
    } catch (...) {
      (void)hipStreamEndCapture(cap_, &g);
      if (g) (void)hipGraphDestroy(g);
      nfev_ = nfev;
      throw;
    }
    nfev_ = nfev;
    FRESCO_HIP_CHECK(hipStreamEndCapture(cap_, &g));
    const hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FRESCO_HIP_CHECK(e);
    gepoch_ = ep;
    ++captures_;
  }
  FRESCO_HIP_CHECK(hipGraphLaunch(gexec_, stream_));
  nfev_ += 1;
  ++replays_;
  return true;
}

void LBFGS::one_iteration() {
  if (!func_initialized_) initialize_func_gradient();
  // lbfgs.cpp:31-62
  const bool graph = use_graph_ && pot_->capture_ready() && !prof_enabled();
  if (graph)
    run_graph();
  else
    enqueue_iteration(stream_);
  read_scalars();
  const double f0 = hsc_[LS_F];
  if (hsc_[LS_DOT] > 0. && verbosity_ > 1) std::cout << "warning: step direction was uphill.  inverting" << std::endl;
  const double stepnorm = hsc_[LS_STEPNORM];
  double factor = std::fabs(hsc_[LS_FACTOR]);
  double fnew = hsc_[LS_FNEW];
  if (hsc_[LS_ACCEPT] == 0.0) {
    // backtracking continues from the second trial (lbfgs.cpp:188-215)
    const double sign = hsc_[LS_FACTOR] < 0 ? -1.0 : 1.0;
    const int nred_max = 10;
    int nred = 0;
    for (;;) {
      double df = fnew - f0;
      if (use_relative_f_) {
        double absf = 1e-100;
        if (f0 != 0) absf = std::abs(f0);
        df /= absf;
      }
      if (df < max_f_rise_) break;
      factor *= 0.5;
      if (verbosity_ > 2)
        std::cout << "energy increased by " << df << " to " << fnew << " from " << f0 << " reducing step norm to "
                  << factor * stepnorm << " H0 " << hsc_[LS_H0] << std::endl;
      if (++nred >= nred_max) break;
      enqueue_trial(false, sign * factor, stream_);
      read_scalars();
      fnew = hsc_[LS_FNEW];
    }
    if (nred >= nred_max && verbosity_ > 0) std::cout << "warning: the line search backtracked too many times" << std::endl;
    enqueue_commit(1, stream_);
    read_scalars();
  }
  if (verbosity_ > 0 && hsc_[LS_YS] == 0.) std::cout << "warning: resetting YS to 1." << std::endl;
  if (verbosity_ > 0 && hsc_[LS_YY] == 0.) std::cout << "warning: resetting YY to 1." << std::endl;
  f_ = hsc_[LS_F];
  gg_ = hsc_[LS_GG];
  rms_ = std::sqrt(gg_) * inv_sqrt_size_;
  H0_ = hsc_[LS_H0];
  k_ = (int)hsc_[LS_K];
  if (iprint_ > 0 && iter_number_ % iprint_ == 0) {
    std::cout << std::setprecision(std::numeric_limits<double>::max_digits10) << "lbgs: " << iter_number_ << " E "
              << f_ << " rms " << rms_ << " nfev " << nfev_ << " step norm " << stepnorm * factor << std::endl;
  }
  iter_number_ += 1;
}

void LBFGS::reset(const double* x0_host) {
  k_ = 0;
  iter_number_ = 0;
  nfev_ = 0;
  set_scalar(LS_K, 0.0);
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
  dsc_.allocate(FS_COUNT);
  FRESCO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hsc_), FS_COUNT * sizeof(double)));
  ops_.zero(v_.get(), stream_);
  ops_.copy(xold_.get(), x_.get(), stream_);
  ops_.zero(gold_.get(), stream_);
}

ModifiedFire::~ModifiedFire() {
  for (int q = 0; q < 2; ++q)
    if (gexec_[q]) (void)hipGraphExecDestroy(gexec_[q]);
  if (cap_) (void)hipStreamDestroy(cap_);
  if (hsc_) (void)hipHostFree(hsc_);
}

// The iteration below as a fixed kernel sequence on the device scalars: the
// new x goes to the other position buffer and its gradient to the other
// gradient buffer (the current ones become x_old, g_old), so one recorded
// graph per assignment of the two buffer pairs.
void ModifiedFire::enqueue_iteration(hipStream_t s) {
  double* dsc = dsc_.get();
  const unsigned nb = nblocks(n_);
  hipLaunchKernelGGL(k_fire_euler, dim3(nb), dim3(RB), 0, s, v_.get(), g_.get(), dx_.get(), 0.0, 0.0, 0.0, n_,
                     ops_.part.get(), dsc);
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(RB), 0, s, ops_.part.get(), (int)nb, 1, 1.0, nullptr, dsc + FS_ND);
  hipLaunchKernelGGL(k_fire_factor, dim3(1), dim3(64), 0, s, dsc, fmaxstep_);
  hipLaunchKernelGGL(k_axpbz, dim3(eblocks(n_)), dim3(256), 0, s, xold_.get(), 1.0, x_.get(), 0.0, dx_.get(), n_,
                     dsc + FS_FACTOR);
  FRESCO_HIP_CHECK(hipGetLastError());
  if (!pot_->energy_gradient_async(xold_.get(), gold_.get(), dsc + FS_E, s))
    throw RuntimeFailure("potential has no device-side evaluation");
  hipLaunchKernelGGL(k_dots, dim3(nb), dim3(RB), 0, s, v_.get(), gold_.get(), gold_.get(), gold_.get(), v_.get(),
                     v_.get(), n_, 3, ops_.part.get());
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(RB), 0, s, ops_.part.get(), (int)nb, 3, 1.0, nullptr, dsc + FS_R0);
  hipLaunchKernelGGL(k_fire_update, dim3(1), dim3(64), 0, s, dsc, Nmin_, finc_, dtmax_, fa_, fdec_, astart_,
                     std::sqrt((double)n_), stepback_ ? 1 : 0);
  hipLaunchKernelGGL(k_fire_zero_v, dim3(eblocks(n_)), dim3(256), 0, s, v_.get(), dsc, n_);
  FRESCO_HIP_CHECK(hipGetLastError());
}

void ModifiedFire::upload_scalars() {
  hsc_[FS_F] = f_;
  hsc_[FS_FOLD] = fold_;
  hsc_[FS_DT] = dt_;
  hsc_[FS_A] = a_;
  hsc_[FS_IFNORM] = ifnorm_;
  hsc_[FS_VNORM] = vnorm_;
  hsc_[FS_NITER] = (double)fire_iter_number_;
  hsc_[FS_RMS] = rms_;
  FRESCO_HIP_CHECK(hipMemcpyAsync(dsc_.get(), hsc_, FS_COUNT * sizeof(double), hipMemcpyHostToDevice, stream_));
  dirty_ = false;
}

void ModifiedFire::graph_iteration() {
  if (dirty_) upload_scalars();
  const uint64_t ep = pot_->graph_epoch();
  if (ep != gepoch_) {
    for (int q = 0; q < 2; ++q) {
      if (gexec_[q]) FRESCO_HIP_CHECK(hipGraphExecDestroy(gexec_[q]));
      gexec_[q] = nullptr;
      gx_[q] = nullptr;
    }
    gepoch_ = ep;
  }
  int q = gx_[0] == x_.get() ? 0 : (gx_[1] == x_.get() ? 1 : -1);
  if (q < 0 || !gexec_[q]) {
    q = gexec_[0] ? 1 : 0;
    if (!cap_) FRESCO_HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    FRESCO_HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_iteration(cap_);
    } catch (...) {
      (void)hipStreamEndCapture(cap_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    FRESCO_HIP_CHECK(hipStreamEndCapture(cap_, &g));
    const hipError_t e = hipGraphInstantiate(&gexec_[q], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FRESCO_HIP_CHECK(e);
    gx_[q] = x_.get();
    ++captures_;
  }
  FRESCO_HIP_CHECK(hipGraphLaunch(gexec_[q], stream_));
  ++replays_;
  FRESCO_HIP_CHECK(hipMemcpyAsync(hsc_, dsc_.get(), FS_COUNT * sizeof(double), hipMemcpyDeviceToHost, stream_));
  FRESCO_HIP_CHECK(hipStreamSynchronize(stream_));
  f_ = hsc_[FS_F];
  fold_ = hsc_[FS_FOLD];
  dt_ = hsc_[FS_DT];
  a_ = hsc_[FS_A];
  ifnorm_ = hsc_[FS_IFNORM];
  vnorm_ = hsc_[FS_VNORM];
  fire_iter_number_ = (size_t)hsc_[FS_NITER];
  rms_ = hsc_[FS_RMS];
  const bool uphill = hsc_[FS_UPHILL] != 0.0;
  if (!(uphill && stepback_)) {
    // the new x, g were written to the other pair of buffers
    std::swap(x_, xold_);
    std::swap(g_, gold_);
  }
  if (uphill && verbosity_ > 1) std::cout << "warning: step direction was uphill.  inverting\n";
  if (iprint_ > 0 && iter_number_ % iprint_ == 0) {
    std::cout << "fire: " << iter_number_ << " fire_iter_number " << fire_iter_number_ << " dt " << dt_ << " a " << a_
              << " P " << hsc_[FS_P] << " vnorm " << vnorm_ << " E " << f_ << " rms " << rms_ << " nfev " << nfev_
              << "\n";
  }
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
  dirty_ = true;
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
  dirty_ = true;
}

void ModifiedFire::one_iteration() {
  if (!func_initialized_) initialize_func_gradient();
  nfev_ += 1;
  iter_number_ += 1;
  if (use_graph_ && pot_->capture_ready() && !prof_enabled()) {
    graph_iteration();
    return;
  }
  dirty_ = true;
  fire_iter_number_ += 1;
  fold_ = f_;
  // x_old = x, g_old = g (modified_fire.hpp:146, 168) by exchanging buffer
  // ownership instead of copying: the step below writes the new positions
  // into the other buffer, the evaluation the new gradient
  std::swap(x_, xold_);
  // velocity mixing + forward Euler (modified_fire.hpp:136-158, 172-177)
  unsigned nb = nblocks(n_);
  hipLaunchKernelGGL(k_fire_euler, dim3(nb), dim3(RB), 0, stream_, v_.get(), g_.get(), dx_.get(), a_,
                     ifnorm_ * vnorm_, dt_, n_, ops_.part.get());
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(RB), 0, stream_, ops_.part.get(), (int)nb, 1, 1.0, nullptr,
                     ops_.scal.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  std::swap(g_, gold_);
  double nd;
  ops_.fetch(1, &nd, stream_);
  double normdx = std::sqrt(nd);
  double factor = 1;
  if (normdx > fmaxstep_) factor = fmaxstep_ / normdx;
  ops_.axpbz(x_.get(), 1.0, xold_.get(), factor, dx_.get(), stream_);
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
      // x = x_old, g = g_old (modified_fire.hpp:207-211): back to the old buffers
      f_ = fold_;
      std::swap(x_, xold_);
      std::swap(g_, gold_);
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
  dirty_ = true;
}

}  // namespace fresco_hip
