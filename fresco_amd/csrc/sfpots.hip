// UwU, UwNU and NUwNU structure-factor potentials.
//
//   UwU   source/fresco/uwu.hpp:17-195   grid intensities, FFT-only
//   UwNU  source/fresco/uwnu.hpp:17-159  grid values as NUFFT modes, k points
//   NUwNU source/fresco/nuwnu.hpp:17-291 non-uniform points and k vectors
//
// All share the per-mode loss of the reference: s = |rho|^2 / norm,
// D = (s - S0)[/S0 where S0 != 0], phi = sum V D^2, with the potential's own
// pullback constant.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "nufft.hpp"
#include "potential.hpp"

namespace fresco_hip {

namespace {

constexpr int TB = 256;

__device__ __forceinline__ double block_reduce(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int h = TB / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  return sh[0];
}

// the loss of one mode: returns V D^2 and sets g = V D [/S0]
__device__ __forceinline__ double mode_loss(double a2, double norm, double S0, double V, double& g) {
  double D = a2 / norm - S0;
  if (S0 != 0) D /= S0;
  g = V * D;
  if (S0 != 0) g /= S0;
  return V * D * D;
}

// ---------------------------------------------------------------------------
// UwU
// ---------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(TB) k_uwu_load(const double* __restrict__ x, int64_t m, Cplx<T>* __restrict__ c,
                                                 double* __restrict__ part) {
  __shared__ double sh[TB];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < m; i += (int64_t)gridDim.x * TB) {
    const double v = x[i];
    c[i] = Cplx<T>{(T)v, T(0)};
    s += v;
  }
  s = block_reduce(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// uwu.hpp:166-182: f = (2/rhotot) V D rho / S0 (reference) or the exact
// derivative's (4/rhotot) V D rho / S0 with the sum-rule term reduced in cpart
template <typename T, int MODE>
__global__ void __launch_bounds__(TB)
    k_uwu_loss(const Cplx<T>* __restrict__ rho, const double* __restrict__ Sk, const double* __restrict__ V,
               int64_t m, const double* __restrict__ rhotot, Cplx<T>* __restrict__ f, double fscale,
               double* __restrict__ epart, double* __restrict__ cpart) {
  __shared__ double sh[TB];
  const double R = rhotot[0];
  double e = 0, q = 0;
  for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < m; i += (int64_t)gridDim.x * TB) {
    const double re = (double)rho[i].re, im = (double)rho[i].im;
    const double a2 = re * re + im * im;
    double g;
    e += mode_loss(a2, R, Sk[i], V[i], g);
    if (MODE > 0) {
      const double fs = fscale / R * g;
      f[i] = Cplx<T>{(T)(fs * re), (T)(fs * im)};
      if (MODE == 2) q += 2.0 * g * a2 / (R * R);
    }
  }
  e = block_reduce(e, sh);
  if (threadIdx.x == 0) epart[blockIdx.x] = e;
  if (MODE == 2) {
    q = block_reduce(q, sh);
    if (threadIdx.x == 0) cpart[blockIdx.x] = q;
  }
}

template <typename T>
__global__ void k_real_part(const Cplx<T>* __restrict__ c, int64_t m, double scale, const double* __restrict__ sub,
                            double* __restrict__ grad) {
  const double off = sub ? sub[0] : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    grad[i] = scale * (double)c[i].re - off;
}

template <typename T>
__global__ void k_copy_c(const Cplx<T>* __restrict__ a, int64_t m, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    out[2 * i] = (double)a[i].re;
    out[2 * i + 1] = (double)a[i].im;
  }
}

unsigned rblocks(int64_t m) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (m + TB - 1) / TB)); }

template <typename T>
class UwU final : public Potential {
 public:
  UwU(const int64_t* shape, int d, const double* Sk, const double* V, int64_t m, int exact) : d_(d), exact_(exact) {
    if (d < 1 || d > 3) throw InvalidArgument("Invalid dimension: " + std::to_string(d));
    M_ = 1;
    for (int a = 0; a < d; ++a) {
      if (shape[a] <= 0) throw InvalidArgument("grid shape must be positive");
      M_ *= shape[a];
    }
    if (m != M_) throw InvalidArgument("Sk / V size must equal the grid size");
    // row-major grid (FFTW plan over N[0..d-1], last fastest, uwu.hpp:48)
    int len[3];
    for (int a = 0; a < d; ++a) len[a] = (int)shape[d - 1 - a];
    const bool dbl = sizeof(T) == 8;
    fwd_.create(FftPlan::C2C_FWD, dbl, d, len);
    inv_.create(FftPlan::C2C_INV, dbl, d, len);
    c_.allocate(M_);
    rho_.allocate(M_);
    f_.allocate(M_);
    Sk_.allocate(M_);
    V_.allocate(M_);
    Sk_.upload(Sk, M_);
    V_.upload(V, M_);
    nb_ = rblocks(M_);
    epart_.allocate(nb_);
    cpart_.allocate(nb_);
    tpart_.allocate(nb_);
    rhotot_.allocate(1);
    e_.allocate(1);
    cterm_.allocate(1);
  }
  int64_t ndof() const override { return M_; }
  int64_t nmodes() const override { return M_; }

  double energy(const double* x, hipStream_t s) override {
    forward(x, s);
    hipLaunchKernelGGL((k_uwu_loss<T, 0>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), M_,
                       rhotot_.get(), f_.get(), 0.0, epart_.get(), cpart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  double energy_gradient(const double* x, double* grad, hipStream_t s) override {
    enqueue_eg(x, grad, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  // energy_gradient without the wait (kernels only: recordable)
  bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) override {
    enqueue_eg(x, grad, e_dev, s);
    return true;
  }
  bool capture_ready() const override { return true; }

  void enqueue_eg(const double* x, double* grad, double* e_out, hipStream_t s) {
    forward(x, s);
    if (exact_) {
      hipLaunchKernelGGL((k_uwu_loss<T, 2>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), M_,
                         rhotot_.get(), f_.get(), 4.0, epart_.get(), cpart_.get());
      reduce_partials(cpart_.get(), nb_, cterm_.get(), s);
    } else {
      hipLaunchKernelGGL((k_uwu_loss<T, 1>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), M_,
                         rhotot_.get(), f_.get(), 2.0, epart_.get(), cpart_.get());
    }
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_out, s);
    fwd_.exec(f_.get(), c_.get(), s);  // FFTW sign -1 (uwu.hpp:49,187)
    hipLaunchKernelGGL(k_real_part<T>, dim3(rblocks(M_)), dim3(TB), 0, s, c_.get(), M_, 1.0,
                       exact_ ? cterm_.get() : nullptr, grad);
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  void get_rho(double* rho, hipStream_t s) override {
    hipLaunchKernelGGL(k_copy_c<T>, dim3(rblocks(M_)), dim3(TB), 0, s, rho_.get(), M_, rho);
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  void target_grads(double* dS0, double* dV, hipStream_t s) override {
    if (!rho_tmp_.get()) rho_tmp_.allocate(2 * (size_t)M_);
    get_rho(rho_tmp_.get(), s);
    launch_target_grads(rho_tmp_.get(), Sk_.get(), V_.get(), M_, 0.0, rhotot_.get(), dS0, dV, s);
  }

  void set_targets(const double* Sk, const double* V, hipStream_t s) override {
    FRESCO_HIP_CHECK(hipMemcpyAsync(Sk_.get(), Sk, M_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (V) FRESCO_HIP_CHECK(hipMemcpyAsync(V_.get(), V, M_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  }

 private:
  void forward(const double* x, hipStream_t s) {
    // rhotot = sum x (uwu.hpp:128-133, accumulated deterministically)
    hipLaunchKernelGGL(k_uwu_load<T>, dim3(nb_), dim3(TB), 0, s, x, M_, c_.get(), tpart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(tpart_.get(), nb_, rhotot_.get(), s);
    inv_.exec(c_.get(), rho_.get(), s);  // FFTW sign +1 (uwu.hpp:48,134)
  }

  int d_;
  int exact_;
  int64_t M_ = 0;
  FftPlan fwd_, inv_;
  DeviceBuffer<Cplx<T>> c_, rho_, f_;
  DeviceBuffer<double> Sk_, V_, epart_, cpart_, tpart_, rhotot_, e_, cterm_, rho_tmp_;
  int nb_ = 1;
  HostScalar host_;
};

// ---------------------------------------------------------------------------
// UwNU
// ---------------------------------------------------------------------------

struct ModeGridDev {
  int d;
  int64_t n[3];   // grid shape, first dimension fastest
  int64_t nf[3];  // fine grid
  const double* ipsi[3];
};

// place the modes c_m / Psi(m) on the zeroed complex fine grid (m mod nf)
template <typename T>
__global__ void k_uwnu_pad(const double* __restrict__ x, ModeGridDev g, int64_t m, Cplx<T>* __restrict__ fine) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = j;
    int64_t fi = 0, mul = 1;
    double w = x[j];
    for (int a = 0; a < g.d; ++a) {
      const int64_t i = r % g.n[a];
      r /= g.n[a];
      const int64_t mm = i - g.n[a] / 2;
      w *= g.ipsi[a][mm < 0 ? -mm : mm];
      fi += mul * (mm < 0 ? mm + g.nf[a] : mm);
      mul *= g.nf[a];
    }
    fine[fi] = Cplx<T>{(T)w, T(0)};
  }
}

// grad_m = Re(B[m mod nf] / Psi(m)) / (2 pi)  (uwnu.hpp:151-156)
template <typename T>
__global__ void k_uwnu_extract(const Cplx<T>* __restrict__ fine, ModeGridDev g, int64_t m, double* __restrict__ grad) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = j;
    int64_t fi = 0, mul = 1;
    double w = 1.0;
    for (int a = 0; a < g.d; ++a) {
      const int64_t i = r % g.n[a];
      r /= g.n[a];
      const int64_t mm = i - g.n[a] / 2;
      w *= g.ipsi[a][mm < 0 ? -mm : mm];
      fi += mul * (mm < 0 ? mm + g.nf[a] : mm);
      mul *= g.nf[a];
    }
    grad[j] = (double)fine[fi].re * w * (1.0 / (2.0 * M_PI));
  }
}

// loss over the k points (rhotot = 1, uwnu.hpp:120-148): f = 4 V D rho / S0
template <typename T, int MODE>
__global__ void __launch_bounds__(TB)
    k_pts_loss(const double* __restrict__ rho, const double* __restrict__ Sk, const double* __restrict__ V,
               int64_t nk, double norm, double fre, double fim, Cplx<T>* __restrict__ f, double* __restrict__ epart) {
  __shared__ double sh[TB];
  double e = 0;
  for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < nk; i += (int64_t)gridDim.x * TB) {
    const double re = rho[2 * i], im = rho[2 * i + 1];
    double g;
    e += mode_loss(re * re + im * im, norm, Sk[i], V[i], g);
    if (MODE > 0) {
      // f = (fre + i fim) g rho
      const double ar = fre * g, ai = fim * g;
      f[i] = Cplx<T>{(T)(ar * re - ai * im), (T)(ar * im + ai * re)};
    }
  }
  e = block_reduce(e, sh);
  if (threadIdx.x == 0) epart[blockIdx.x] = e;
}

template <typename T>
class UwNU final : public Potential {
 public:
  UwNU(const int64_t* shape, int d, const double* K, int64_t nkd, const double* Sk, const double* V, int64_t nk,
       const double* L, double eps, double sigma)
      : d_(d), nk_(nk) {
    if (d < 1 || d > 3) throw InvalidArgument("Invalid dimension: " + std::to_string(d));
    if (nk <= 0) throw InvalidArgument("need at least one k vector");
    if (nkd / d < nk) throw InvalidArgument("K has fewer than len(Sk) vectors");
    M_ = 1;
    for (int a = 0; a < d; ++a) {
      if (shape[a] <= 0) throw InvalidArgument("grid shape must be positive");
      if (!(L[a] > 0)) throw InvalidArgument("box lengths must be positive");
      n_[a] = shape[a];
      M_ *= shape[a];
    }
    if (sigma <= 0) sigma = 2.0;
    es_ = es_params(eps, sigma);
    int nf[3] = {1, 1, 1};
    for (int a = 0; a < d; ++a) nf[a] = next_smooth_even((int)std::ceil(sigma * n_[a]), es_.w);
    geom_ = make_geom(d, nf, es_.w, (int)sizeof(T));
    // k scaled by 2 pi / L (uwnu.hpp:48,58,73): the NUFFT points
    std::vector<double> kp((size_t)nk * d);
    for (int64_t j = 0; j < nk; ++j)
      for (int a = 0; a < d; ++a) kp[(size_t)j * d + a] = K[(size_t)j * d + a] * 2.0 * M_PI / L[a];
    DeviceBuffer<double> kd(kp.size());
    kd.upload(kp.data(), kp.size());
    pts_.init(geom_, es_, nk);
    spreader_.init(geom_, es_);
    CoordMap cm;
    cm.wrap = 0;
    pts_.set(kd.get(), cm, nullptr, nullptr);
    FRESCO_HIP_CHECK(hipDeviceSynchronize());
    spreader_.plan_complex_slices(pts_);
    fine_.allocate(geom_.ncells());
    fine2_.allocate(geom_.ncells());
    const bool dbl = sizeof(T) == 8;
    fwd_.create(FftPlan::C2C_FWD, dbl, d, geom_.nf);
    inv_.create(FftPlan::C2C_INV, dbl, d, geom_.nf);
    for (int a = 0; a < 3; ++a) {
      std::vector<double> ps = a < d ? es_deconv(es_, geom_.nf[a], (int)(n_[a] / 2 + 1)) : std::vector<double>(1, 1.0);
      for (auto& v : ps) v = 1.0 / v;
      ipsi_[a].allocate(ps.size());
      ipsi_[a].upload(ps.data(), ps.size());
      mg_.ipsi[a] = ipsi_[a].get();
      mg_.n[a] = a < d ? n_[a] : 1;
      mg_.nf[a] = geom_.nf[a];
    }
    mg_.d = d;
    rho_.allocate(2 * (size_t)nk);
    f_.allocate(nk);
    Sk_.allocate(nk);
    V_.allocate(nk);
    Sk_.upload(Sk, nk);
    V_.upload(V, nk);
    nb_ = rblocks(nk);
    epart_.allocate(nb_);
    e_.allocate(1);
  }
  int64_t ndof() const override { return M_; }
  int64_t nmodes() const override { return nk_; }

  double energy(const double* x, hipStream_t s) override {
    forward(x, s);
    hipLaunchKernelGGL((k_pts_loss<T, 0>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), nk_, 1.0, 0.0,
                       0.0, f_.get(), epart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  double energy_gradient(const double* x, double* grad, hipStream_t s) override {
    enqueue_eg(x, grad, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  // energy_gradient without the wait (kernels only: recordable)
  bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) override {
    enqueue_eg(x, grad, e_dev, s);
    return true;
  }
  bool capture_ready() const override { return true; }

  void enqueue_eg(const double* x, double* grad, double* e_out, hipStream_t s) {
    forward(x, s);
    hipLaunchKernelGGL((k_pts_loss<T, 1>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), nk_, 1.0, 4.0,
                       0.0, f_.get(), epart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_out, s);
    // type-1 (-1) of f at the k points (uwnu.hpp:151)
    spreader_.spread_complex(pts_, f_.get(), fine_.get(), s);
    fwd_.exec(fine_.get(), fine2_.get(), s);
    hipLaunchKernelGGL(k_uwnu_extract<T>, dim3(rblocks(M_)), dim3(TB), 0, s, fine2_.get(), mg_, M_, grad);
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  void get_rho(double* rho, hipStream_t s) override {
    FRESCO_HIP_CHECK(hipMemcpyAsync(rho, rho_.get(), 2 * nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  }

  void target_grads(double* dS0, double* dV, hipStream_t s) override {
    launch_target_grads(rho_.get(), Sk_.get(), V_.get(), nk_, 1.0, nullptr, dS0, dV, s);
  }

  void set_targets(const double* Sk, const double* V, hipStream_t s) override {
    FRESCO_HIP_CHECK(hipMemcpyAsync(Sk_.get(), Sk, nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (V) FRESCO_HIP_CHECK(hipMemcpyAsync(V_.get(), V, nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  }

  void plan_info(int64_t* nf, int* w, double* beta, int* tile) const override {
    for (int a = 0; a < 3; ++a) {
      nf[a] = geom_.nf[a];
      tile[a] = geom_.tile[a];
    }
    *w = es_.w;
    *beta = es_.beta;
  }

 private:
  void forward(const double* x, hipStream_t s) {
    // type-2 (+1): the grid values are the modes (uwnu.hpp:99,129)
    fill_zero(fine_.get(), geom_.ncells(), s);
    hipLaunchKernelGGL(k_uwnu_pad<T>, dim3(rblocks(M_)), dim3(TB), 0, s, x, mg_, M_, fine_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    inv_.exec(fine_.get(), fine2_.get(), s);
    interpolate_complex(pts_, fine2_.get(), rho_.get(), s);
  }

  int d_;
  int64_t nk_;
  int64_t n_[3] = {1, 1, 1};
  int64_t M_ = 1;
  EsParams es_;
  GridGeom geom_;
  SortedPoints<T> pts_;
  Spreader<T> spreader_;
  DeviceBuffer<Cplx<T>> fine_, fine2_, f_;  // fine2_: out-of-place FFT result
  FftPlan fwd_, inv_;
  DeviceBuffer<double> ipsi_[3];
  ModeGridDev mg_{};
  DeviceBuffer<double> rho_, Sk_, V_, epart_, e_;
  int nb_ = 1;
  HostScalar host_;
};

// ---------------------------------------------------------------------------
// NUwNU: exact direct sums with float64 phases (type-3 of nuwnu.hpp:77-78)
// ---------------------------------------------------------------------------

struct T3Dev {
  int d;
  double L[3];
};

__device__ __forceinline__ double wrap_coord(double p, double L) {
  // (p - round(p/L) + L) * 2 pi / L   (nuwnu.hpp:134,193)
  return (p - round(p / L) + L) * 2.0 * M_PI / L;
}

// reduced phase in [-pi, pi]
__device__ __forceinline__ double reduce_phase(double th) {
  return th - 2.0 * M_PI * rint(th * (1.0 / (2.0 * M_PI)));
}

template <typename T>
__device__ __forceinline__ void sincos_t(double ph, T& sn, T& cs);
template <>
__device__ __forceinline__ void sincos_t<double>(double ph, double& sn, double& cs) {
  sincos(ph, &sn, &cs);
}
template <>
__device__ __forceinline__ void sincos_t<float>(double ph, float& sn, float& cs) {
  __sincosf((float)ph, &sn, &cs);
}

// rho_q = sum_j c_j exp(+i k_q . x_j): one thread per target, points staged in LDS
template <typename T>
__global__ void __launch_bounds__(TB)
    k_t3_forward(const double* __restrict__ p, const double* __restrict__ c, int64_t n, const double* __restrict__ kv,
                 int64_t nk, T3Dev g, double* __restrict__ rho) {
  __shared__ double sx[TB * 3];
  __shared__ double sc[TB];
  const int64_t q = (int64_t)blockIdx.x * TB + threadIdx.x;
  double k[3] = {0, 0, 0};
  if (q < nk)
    for (int a = 0; a < g.d; ++a) k[a] = kv[q * g.d + a];
  double ar = 0, ai = 0;
  for (int64_t base = 0; base < n; base += TB) {
    __syncthreads();
    const int64_t j = base + threadIdx.x;
    if (j < n) {
      for (int a = 0; a < g.d; ++a) sx[threadIdx.x * 3 + a] = wrap_coord(p[j * g.d + a], g.L[a]);
      sc[threadIdx.x] = c[j];
    }
    __syncthreads();
    const int cnt = (int)min<int64_t>(TB, n - base);
    T pr = T(0), pi = T(0);
    for (int t = 0; t < cnt; ++t) {
      double th = k[0] * sx[t * 3];
      if (g.d > 1) th += k[1] * sx[t * 3 + 1];
      if (g.d > 2) th += k[2] * sx[t * 3 + 2];
      T sn, cs;
      sincos_t<T>(reduce_phase(th), sn, cs);
      const T cw = (T)sc[t];
      pr += cw * cs;
      pi += cw * sn;
    }
    ar += (double)pr;
    ai += (double)pi;
  }
  if (q < nk) {
    rho[2 * q] = ar;
    rho[2 * q + 1] = ai;
  }
}

// grad[d j + a] = Re(sum_q f_{a,q} exp(-i k_q . x_j)), f_a = k_a f, no conj(c) (nuwnu.hpp:265-287)
template <typename T>
__global__ void __launch_bounds__(TB)
    k_t3_adjoint(const double* __restrict__ p, int64_t n, const double* __restrict__ kv, const Cplx<T>* __restrict__ f,
                 int64_t nk, T3Dev g, double* __restrict__ grad) {
  __shared__ double sk[TB * 3];
  __shared__ double sf[TB * 2];
  const int64_t j = (int64_t)blockIdx.x * TB + threadIdx.x;
  double x[3] = {0, 0, 0};
  if (j < n)
    for (int a = 0; a < g.d; ++a) x[a] = wrap_coord(p[j * g.d + a], g.L[a]);
  double acc[3] = {0, 0, 0};
  for (int64_t base = 0; base < nk; base += TB) {
    __syncthreads();
    const int64_t q = base + threadIdx.x;
    if (q < nk) {
      for (int a = 0; a < g.d; ++a) sk[threadIdx.x * 3 + a] = kv[q * g.d + a];
      sf[2 * threadIdx.x] = (double)f[q].re;
      sf[2 * threadIdx.x + 1] = (double)f[q].im;
    }
    __syncthreads();
    const int cnt = (int)min<int64_t>(TB, nk - base);
    T s0 = T(0), s1 = T(0), s2 = T(0);
    for (int t = 0; t < cnt; ++t) {
      double th = sk[t * 3] * x[0];
      if (g.d > 1) th += sk[t * 3 + 1] * x[1];
      if (g.d > 2) th += sk[t * 3 + 2] * x[2];
      T sn, cs;
      sincos_t<T>(reduce_phase(th), sn, cs);
      // Re(f e^{-i th}) = fr cos + fi sin
      const T re = (T)sf[2 * t] * cs + (T)sf[2 * t + 1] * sn;
      s0 += (T)sk[t * 3] * re;
      if (g.d > 1) s1 += (T)sk[t * 3 + 1] * re;
      if (g.d > 2) s2 += (T)sk[t * 3 + 2] * re;
    }
    acc[0] += (double)s0;
    acc[1] += (double)s1;
    acc[2] += (double)s2;
  }
  if (j < n)
    for (int a = 0; a < g.d; ++a) grad[j * g.d + a] = acc[a];
}

template <typename T>
class NUwNUDirect final : public Potential {
 public:
  // the k vectors set the mode count, Nk = len(K) / d (nuwnu.hpp:40)
  NUwNUDirect(const double* radii, int64_t n, int d, const double* K, int64_t nkd, const double* Sk, const double* V,
              int64_t m, const double* L)
      : N_(n), d_(d), nk_(d >= 1 && d <= 3 ? nkd / d : 0) {
    if (d < 1 || d > 3) throw InvalidArgument("Invalid dimension: " + std::to_string(d));
    if (n <= 0 || nk_ <= 0) throw InvalidArgument("need points and k vectors");
    if (m < nk_) throw InvalidArgument("Sk / V hold fewer than len(K) / d entries");
    const int64_t nk = nk_;
    for (int a = 0; a < d; ++a) {
      if (!(L[a] > 0)) throw InvalidArgument("box lengths must be positive");
      g_.L[a] = L[a];
    }
    g_.d = d;
    // weights (nuwnu.hpp:81-105)
    std::vector<double> c(n, 1.0);
    double sum = 0;
    for (int64_t i = 0; i < n; ++i) {
      for (int a = 0; a < d; ++a) {
        if (d == 2) c[i] = c[i] * M_PI * radii[i] * radii[i];
        else if (d == 3) c[i] = c[i] * M_PI * radii[i] * radii[i] * radii[i] * 4.0 / 3.0;
      }
      sum += c[i];
    }
    for (auto& v : c) v *= (double)n / sum;
    c_.allocate(n);
    c_.upload(c.data(), n);
    k_.allocate((size_t)nk * d);
    k_.upload(K, (size_t)nk * d);
    Sk_.allocate(nk);
    V_.allocate(nk);
    Sk_.upload(Sk, nk);
    V_.upload(V, nk);
    rho_.allocate(2 * (size_t)nk);
    f_.allocate(nk);
    nb_ = rblocks(nk);
    epart_.allocate(nb_);
    e_.allocate(1);
  }
  int64_t ndof() const override { return N_ * d_; }
  int64_t nmodes() const override { return nk_; }

  double energy(const double* x, hipStream_t s) override {
    forward(x, s);
    hipLaunchKernelGGL((k_pts_loss<T, 0>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), nk_,
                       (double)N_, 0.0, 0.0, f_.get(), epart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  double energy_gradient(const double* x, double* grad, hipStream_t s) override {
    enqueue_eg(x, grad, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  // energy_gradient without the wait (kernels only: recordable)
  bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) override {
    enqueue_eg(x, grad, e_dev, s);
    return true;
  }
  bool capture_ready() const override { return true; }

  void enqueue_eg(const double* x, double* grad, double* e_out, hipStream_t s) {
    forward(x, s);
    // factor = (-4i/N) V D rho / S0 (nuwnu.hpp:228,247)
    hipLaunchKernelGGL((k_pts_loss<T, 1>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), nk_,
                       (double)N_, 0.0, -4.0 / (double)N_, f_.get(), epart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_out, s);
    hipLaunchKernelGGL(k_t3_adjoint<T>, dim3((unsigned)((N_ + TB - 1) / TB)), dim3(TB), 0, s, x, N_, k_.get(),
                       f_.get(), nk_, g_, grad);
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  void get_rho(double* rho, hipStream_t s) override {
    FRESCO_HIP_CHECK(hipMemcpyAsync(rho, rho_.get(), 2 * nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  }

  void target_grads(double* dS0, double* dV, hipStream_t s) override {
    launch_target_grads(rho_.get(), Sk_.get(), V_.get(), nk_, (double)N_, nullptr, dS0, dV, s);
  }

  void set_targets(const double* Sk, const double* V, hipStream_t s) override {
    FRESCO_HIP_CHECK(hipMemcpyAsync(Sk_.get(), Sk, nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (V) FRESCO_HIP_CHECK(hipMemcpyAsync(V_.get(), V, nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  }

 private:
  void forward(const double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_t3_forward<T>, dim3((unsigned)((nk_ + TB - 1) / TB)), dim3(TB), 0, s, x, c_.get(), N_,
                       k_.get(), nk_, g_, rho_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  int64_t N_;
  int d_;
  int64_t nk_;
  T3Dev g_{};
  DeviceBuffer<double> c_, k_, Sk_, V_, rho_, epart_, e_;
  DeviceBuffer<Cplx<T>> f_;
  int nb_ = 1;
  HostScalar host_;
};


// ---------------------------------------------------------------------------
// NUwNU by a type-3 NUFFT (nuwnu.hpp:77-78: FINUFFT type 3, +1 for rho, -1
// for the gradient).  With x = C + xt (C = 2 pi: the +L-shifted coordinates
// lie in [pi, 3 pi]) and k = D + kt (D the centre of the k box):
//   rho(k) = e^{i k.C} sum_j [c_j e^{i D.xt_j}] e^{i kt.xt_j}
// The bracket is spread with the ES kernel onto an outer periodic grid of
// spacing h (xt_j at cell xt_j / h + nf1/2), whose Fourier series at
// s = kt h is an inner type-2 NUFFT (the outer grid's cells are its modes);
// deconvolving by the kernel's continuous transform gives the sum:
//   sum_j c'_j e^{i kt.xt_j} = prod_a 2 / (w phihat(kt_a h_a w / 2)) B(kt h)
// Sizes per dimension (X = pi, S = half width of the k box):
//   nf1 >= 2 sigma X S / pi + w + 1,  h = 2 X / (nf1 - w),  nf2 = sigma nf1.
// The gradient grad_{j,a} = Re(sum_q k_{q,a} f_q e^{-i k_q.x_j}) runs the
// chain backwards for each component a: conj(deconvolution) -> inner type 1
// (spread at s, forward FFT, deconvolve) -> interpolation at xt_j ->
// times e^{-i D.xt_j}.
// ---------------------------------------------------------------------------

struct T3Map {
  int d;
  double L[3];
  double D[3];
};

// c'_j = c_j e^{i D.xt_j}, xt = (p - round(p/L)) 2 pi / L
template <typename T>
__global__ void k_t3_strength(const double* __restrict__ p, const double* __restrict__ c, int64_t n, T3Map m,
                              Cplx<T>* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double ph = 0.0;
    for (int a = 0; a < m.d; ++a) {
      const double v = p[j * m.d + a];
      ph += m.D[a] * (v - round(v / m.L[a])) * (2.0 * M_PI / m.L[a]);
    }
    double sn, cs;
    sincos(ph, &sn, &cs);
    out[j] = Cplx<T>{(T)(c[j] * cs), (T)(c[j] * sn)};
  }
}

// complex modes b (outer grid, index l, m = l - n/2, first dimension fastest)
// -> inner fine grid at m mod nf, times 1/Psi(m)
template <typename T>
__global__ void k_t3_pad(const Cplx<T>* __restrict__ b, ModeGridDev g, int64_t cnt, Cplx<T>* __restrict__ fine) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = j, fi = 0, mul = 1;
    double wgt = 1.0;
    for (int a = 0; a < g.d; ++a) {
      const int64_t i = r % g.n[a];
      r /= g.n[a];
      const int64_t mm = i - g.n[a] / 2;
      wgt *= g.ipsi[a][mm < 0 ? -mm : mm];
      fi += mul * (mm < 0 ? mm + g.nf[a] : mm);
      mul *= g.nf[a];
    }
    const Cplx<T> v = b[j];
    fine[fi] = Cplx<T>{(T)(wgt * (double)v.re), (T)(wgt * (double)v.im)};
  }
}

// inner fine grid -> outer grid b_l = fine[m mod nf] / Psi(m)
template <typename T>
__global__ void k_t3_extract(const Cplx<T>* __restrict__ fine, ModeGridDev g, int64_t cnt, Cplx<T>* __restrict__ b) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = j, fi = 0, mul = 1;
    double wgt = 1.0;
    for (int a = 0; a < g.d; ++a) {
      const int64_t i = r % g.n[a];
      r /= g.n[a];
      const int64_t mm = i - g.n[a] / 2;
      wgt *= g.ipsi[a][mm < 0 ? -mm : mm];
      fi += mul * (mm < 0 ? mm + g.nf[a] : mm);
      mul *= g.nf[a];
    }
    const Cplx<T> v = fine[fi];
    b[j] = Cplx<T>{(T)(wgt * (double)v.re), (T)(wgt * (double)v.im)};
  }
}

// rho_q = corr_q B_q (complex, float64 interleaved)
__global__ void k_t3_post(const double* __restrict__ B, const double* __restrict__ corr, int64_t nk,
                          double* __restrict__ rho) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nk; q += (int64_t)gridDim.x * blockDim.x) {
    const double br = B[2 * q], bi = B[2 * q + 1], cr = corr[2 * q], ci = corr[2 * q + 1];
    rho[2 * q] = br * cr - bi * ci;
    rho[2 * q + 1] = br * ci + bi * cr;
  }
}

// adjoint strengths u_q = k_{q,a} f_q conj(corr_q)
template <typename T>
__global__ void k_t3_adj_strength(const Cplx<T>* __restrict__ f, const double* __restrict__ corr,
                                  const double* __restrict__ kv, int d, int a, int64_t nk, Cplx<T>* __restrict__ u) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nk; q += (int64_t)gridDim.x * blockDim.x) {
    const double fr = (double)f[q].re, fi = (double)f[q].im, cr = corr[2 * q], ci = -corr[2 * q + 1];
    const double ka = kv[q * d + a];
    u[q] = Cplx<T>{(T)(ka * (fr * cr - fi * ci)), (T)(ka * (fr * ci + fi * cr))};
  }
}

// grad[j d + a] = Re(e^{-i D.xt_j} H_j)
__global__ void k_t3_grad(const double* __restrict__ p, const double* __restrict__ H, int64_t n, T3Map m, int a,
                          double* __restrict__ grad) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double ph = 0.0;
    for (int b = 0; b < m.d; ++b) {
      const double v = p[j * m.d + b];
      ph += m.D[b] * (v - round(v / m.L[b])) * (2.0 * M_PI / m.L[b]);
    }
    double sn, cs;
    sincos(ph, &sn, &cs);
    // Re((cs - i sn)(hr + i hi)) = cs hr + sn hi
    grad[j * m.d + a] = cs * H[2 * j] + sn * H[2 * j + 1];
  }
}

template <typename T>
class NUwNUType3 final : public Potential {
 public:
  NUwNUType3(const double* radii, int64_t n, int d, const double* K, int64_t nkd, const double* Sk, const double* V,
             int64_t m, const double* L, double eps, double sigma)
      : N_(n), d_(d), nk_(d >= 1 && d <= 3 ? nkd / d : 0) {
    if (d < 1 || d > 3) throw InvalidArgument("Invalid dimension: " + std::to_string(d));
    if (n <= 0 || nk_ <= 0) throw InvalidArgument("need points and k vectors");
    if (m < nk_) throw InvalidArgument("Sk / V hold fewer than len(K) / d entries");
    if (sigma <= 0) sigma = 2.0;
    es_ = es_params(eps, sigma);
    const int w = es_.w;
    tm_.d = d;
    for (int a = 0; a < 3; ++a) {
      tm_.L[a] = 1.0;
      tm_.D[a] = 0.0;
    }
    // k box per dimension
    double S[3] = {0, 0, 0};
    for (int a = 0; a < d; ++a) {
      if (!(L[a] > 0)) throw InvalidArgument("box lengths must be positive");
      tm_.L[a] = L[a];
      double lo = K[a], hi = K[a];
      for (int64_t q = 0; q < nk_; ++q) {
        lo = std::min(lo, K[q * d + a]);
        hi = std::max(hi, K[q * d + a]);
      }
      tm_.D[a] = 0.5 * (lo + hi);
      S[a] = 0.5 * (hi - lo);
    }
    // outer and inner grids
    const double X = M_PI;
    int nf1[3] = {1, 1, 1}, nf2[3] = {1, 1, 1};
    for (int a = 0; a < d; ++a) {
      nf1[a] = next_smooth_even((int)std::ceil(2.0 * sigma * X * S[a] / M_PI) + w + 1, w);
      h_[a] = 2.0 * X / (nf1[a] - w);
      nf2[a] = next_smooth_even((int)std::ceil(sigma * nf1[a]), w);
    }
    g1_ = make_geom(d, nf1, w, (int)sizeof(T));
    g2_ = make_geom(d, nf2, w, (int)sizeof(T));
    src_.init(g1_, es_, n);
    sp1_.init(g1_, es_);
    tgt_.init(g2_, es_, nk_);
    sp2_.init(g2_, es_);
    // sources: cell xt/h + nf1/2 on the outer grid
    for (int a = 0; a < 3; ++a) {
      const double La = a < d ? L[a] : 1.0;
      const double ha = a < d ? h_[a] : 1.0;
      const double n1 = a < d ? nf1[a] : 1.0;
      cmap_.L[a] = La;
      cmap_.shift[a] = La * ha * n1 / (4.0 * M_PI);
      cmap_.scale[a] = 4.0 * M_PI * M_PI / (La * ha * n1);
    }
    cmap_.wrap = 1;
    // targets s = (k - D) h on the inner grid, fixed: sorted once
    std::vector<double> s((size_t)nk_ * d), u;
    std::vector<double> corr(2 * (size_t)nk_);
    std::vector<double> fac((size_t)nk_, 1.0);
    for (int a = 0; a < d; ++a) {
      u.resize(nk_);
      for (int64_t q = 0; q < nk_; ++q) {
        const double kt = K[q * d + a] - tm_.D[a];
        s[(size_t)q * d + a] = kt * h_[a];
        u[q] = kt * h_[a] * w / 2.0;
      }
      const std::vector<double> ph = es_ft(es_, u);
      for (int64_t q = 0; q < nk_; ++q) fac[q] *= 2.0 / (w * ph[q]);
    }
    for (int64_t q = 0; q < nk_; ++q) {
      double ang = 0.0;
      for (int a = 0; a < d; ++a) ang += K[q * d + a] * 2.0 * M_PI;  // e^{i k.C}, C = 2 pi
      corr[2 * q] = fac[q] * std::cos(ang);
      corr[2 * q + 1] = fac[q] * std::sin(ang);
    }
    DeviceBuffer<double> sd(s.size());
    sd.upload(s.data(), s.size());
    CoordMap tmap;
    tmap.wrap = 0;
    tgt_.set(sd.get(), tmap, nullptr, nullptr);
    FRESCO_HIP_CHECK(hipDeviceSynchronize());
    sp2_.plan_complex_slices(tgt_);
    corr_.allocate(corr.size());
    corr_.upload(corr.data(), corr.size());
    k_.allocate((size_t)nk_ * d);
    k_.upload(K, (size_t)nk_ * d);
    // inner deconvolution over the outer grid's modes
    int64_t M1 = 1;
    for (int a = 0; a < 3; ++a) {
      std::vector<double> ps = a < d ? es_deconv(es_, nf2[a], nf1[a] / 2 + 1) : std::vector<double>(1, 1.0);
      for (auto& v : ps) v = 1.0 / v;
      ipsi_[a].allocate(ps.size());
      ipsi_[a].upload(ps.data(), ps.size());
      mg_.ipsi[a] = ipsi_[a].get();
      mg_.n[a] = a < d ? nf1[a] : 1;
      mg_.nf[a] = a < d ? nf2[a] : 1;
      M1 *= mg_.n[a];
    }
    mg_.d = d;
    M1_ = M1;
    const bool dbl = sizeof(T) == 8;
    fwd_.create(FftPlan::C2C_FWD, dbl, d, g2_.nf);
    inv_.create(FftPlan::C2C_INV, dbl, d, g2_.nf);
    outer_.allocate(g1_.ncells());
    fine_.allocate(g2_.ncells());
    fine2_.allocate(g2_.ncells());
    str_.allocate(n);
    fk_.allocate(nk_);
    uk_.allocate(nk_);
    B_.allocate(2 * (size_t)nk_);
    H_.allocate(2 * (size_t)n);
    // weights (nuwnu.hpp:81-105)
    std::vector<double> c(n, 1.0);
    double sum = 0;
    for (int64_t i = 0; i < n; ++i) {
      for (int a = 0; a < d; ++a) {
        if (d == 2) c[i] = c[i] * M_PI * radii[i] * radii[i];
        else if (d == 3) c[i] = c[i] * M_PI * radii[i] * radii[i] * radii[i] * 4.0 / 3.0;
      }
      sum += c[i];
    }
    for (auto& v : c) v *= (double)n / sum;
    c_.allocate(n);
    c_.upload(c.data(), n);
    Sk_.allocate(nk_);
    V_.allocate(nk_);
    Sk_.upload(Sk, nk_);
    V_.upload(V, nk_);
    rho_.allocate(2 * (size_t)nk_);
    nb_ = rblocks(nk_);
    epart_.allocate(nb_);
    e_.allocate(1);
  }
  int64_t ndof() const override { return N_ * d_; }
  int64_t nmodes() const override { return nk_; }

  double energy(const double* x, hipStream_t s) override {
    forward(x, s);
    hipLaunchKernelGGL((k_pts_loss<T, 0>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), nk_,
                       (double)N_, 0.0, 0.0, fk_.get(), epart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  double energy_gradient(const double* x, double* grad, hipStream_t s) override {
    enqueue_eg(x, grad, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  // energy_gradient without the wait (kernels only: recordable)
  bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) override {
    enqueue_eg(x, grad, e_dev, s);
    return true;
  }
  bool capture_ready() const override { return true; }

  void enqueue_eg(const double* x, double* grad, double* e_out, hipStream_t s) {
    forward(x, s);
    // f = (-4i/N) V D rho / S0 (nuwnu.hpp:228,247)
    hipLaunchKernelGGL((k_pts_loss<T, 1>), dim3(nb_), dim3(TB), 0, s, rho_.get(), Sk_.get(), V_.get(), nk_,
                       (double)N_, 0.0, -4.0 / (double)N_, fk_.get(), epart_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    reduce_partials(epart_.get(), nb_, e_out, s);
    for (int a = 0; a < d_; ++a) {
      hipLaunchKernelGGL(k_t3_adj_strength<T>, dim3(rblocks(nk_)), dim3(TB), 0, s, fk_.get(), corr_.get(), k_.get(),
                         d_, a, nk_, uk_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
      // inner type 1 (-1) at s, then the outer grid's modes
      sp2_.spread_complex(tgt_, uk_.get(), fine_.get(), s);
      fwd_.exec(fine_.get(), fine2_.get(), s);
      hipLaunchKernelGGL(k_t3_extract<T>, dim3(rblocks(M1_)), dim3(TB), 0, s, fine2_.get(), mg_, M1_, outer_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
      interpolate_complex(src_, outer_.get(), H_.get(), s);
      hipLaunchKernelGGL(k_t3_grad, dim3(rblocks(N_)), dim3(TB), 0, s, x, H_.get(), N_, tm_, a, grad);
      FRESCO_HIP_CHECK(hipGetLastError());
    }
  }

  void get_rho(double* rho, hipStream_t s) override {
    FRESCO_HIP_CHECK(hipMemcpyAsync(rho, rho_.get(), 2 * nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  }

  void target_grads(double* dS0, double* dV, hipStream_t s) override {
    launch_target_grads(rho_.get(), Sk_.get(), V_.get(), nk_, (double)N_, nullptr, dS0, dV, s);
  }

  void set_targets(const double* Sk, const double* V, hipStream_t s) override {
    FRESCO_HIP_CHECK(hipMemcpyAsync(Sk_.get(), Sk, nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (V) FRESCO_HIP_CHECK(hipMemcpyAsync(V_.get(), V, nk_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  }

  void plan_info(int64_t* nf, int* w, double* beta, int* tile) const override {
    // the inner (upsampled) grid; the outer spreading grid is nf / sigma
    for (int a = 0; a < 3; ++a) {
      nf[a] = g2_.nf[a];
      tile[a] = g2_.tile[a];
    }
    *w = es_.w;
    *beta = es_.beta;
  }

 private:
  void forward(const double* x, hipStream_t s) {
    src_.set(x, cmap_, nullptr, s);
    hipLaunchKernelGGL(k_t3_strength<T>, dim3(rblocks(N_)), dim3(TB), 0, s, x, c_.get(), N_, tm_, str_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    sp1_.spread_complex(src_, str_.get(), outer_.get(), s);
    fill_zero(fine_.get(), g2_.ncells(), s);
    hipLaunchKernelGGL(k_t3_pad<T>, dim3(rblocks(M1_)), dim3(TB), 0, s, outer_.get(), mg_, M1_, fine_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    inv_.exec(fine_.get(), fine2_.get(), s);
    interpolate_complex(tgt_, fine2_.get(), B_.get(), s);
    hipLaunchKernelGGL(k_t3_post, dim3(rblocks(nk_)), dim3(TB), 0, s, B_.get(), corr_.get(), nk_, rho_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  int64_t N_;
  int d_;
  int64_t nk_;
  EsParams es_;
  T3Map tm_{};
  double h_[3] = {1, 1, 1};
  GridGeom g1_, g2_;
  SortedPoints<T> src_, tgt_;
  Spreader<T> sp1_, sp2_;
  CoordMap cmap_;
  ModeGridDev mg_{};
  int64_t M1_ = 1;
  DeviceBuffer<double> ipsi_[3];
  FftPlan fwd_, inv_;
  DeviceBuffer<Cplx<T>> outer_, fine_, fine2_, str_, fk_, uk_;
  DeviceBuffer<double> corr_, k_, B_, H_, c_, Sk_, V_, rho_, epart_, e_;
  int nb_ = 1;
  HostScalar host_;
};

}  // namespace

std::unique_ptr<Potential> make_uwu(const int64_t* shape, int d, const double* Sk, const double* V, int64_t m,
                                    int precision, int exact) {
  if (precision == 1) return std::make_unique<UwU<float>>(shape, d, Sk, V, m, exact);
  return std::make_unique<UwU<double>>(shape, d, Sk, V, m, exact);
}

std::unique_ptr<Potential> make_uwnu(const int64_t* shape, int d, const double* K, int64_t nkd, const double* Sk,
                                     const double* V, int64_t nk, const double* L, double eps, int precision,
                                     double sigma) {
  if (precision == 1) return std::make_unique<UwNU<float>>(shape, d, K, nkd, Sk, V, nk, L, eps, sigma);
  return std::make_unique<UwNU<double>>(shape, d, K, nkd, Sk, V, nk, L, eps, sigma);
}

std::unique_ptr<Potential> make_nuwnu(const double* radii, int64_t n, int d, const double* K, int64_t nkd,
                                      const double* Sk, const double* V, int64_t m, const double* L, double eps,
                                      int precision, int method, double sigma) {
  if (method == 1) {
    if (precision == 1) return std::make_unique<NUwNUDirect<float>>(radii, n, d, K, nkd, Sk, V, m, L);
    return std::make_unique<NUwNUDirect<double>>(radii, n, d, K, nkd, Sk, V, m, L);
  }
  if (method != 0) throw InvalidArgument("unknown NUwNU method " + std::to_string(method));
  if (precision == 1) return std::make_unique<NUwNUType3<float>>(radii, n, d, K, nkd, Sk, V, m, L, eps, sigma);
  return std::make_unique<NUwNUType3<double>>(radii, n, d, K, nkd, Sk, V, m, L, eps, sigma);
}

}  // namespace fresco_hip
