// NUwU: structure-factor loss of non-uniform points on a uniform k grid.
//
// Reference: source/fresco/nuwu.hpp (constructor 35-52, mode grid 54-79,
// weights 125-149, energy 152-208, energy+gradient 210-317).
//
// Per evaluation (E+grad):
//   bin/sort points -> spread (real fine grid, weights c real) -> R2C FFT ->
//   fused deconvolve + loss + pullback over the active (+-k) mode pairs ->
//   C2R FFT -> interpolation with the kernel derivative (all d gradient
//   components from one grid).
// The FP64 "plain" variant keeps the reference's structure of d separate
// type-2 transforms (d C2R FFTs + d value interpolations).
//
// Because c is real, rho_{-k} = conj(rho_k): the loss visits each pair
// {k, -k} once, from the half spectrum of the real-to-complex FFT.  Modes whose
// weight V is 0 contribute exactly 0 to the energy and the gradient
// (nuwu.hpp:271-272) and are skipped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <string>

#include "banddft.hpp"
#include "nuwu_loss.hpp"
#include "nufft.hpp"
#include "potential.hpp"

namespace fresco_hip {

namespace {

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
    k_nuwu_loss(const Cplx<T>* __restrict__ S, PairDev pd, int npairs, double norm, double gscale,
                Cplx<T>* __restrict__ X0, Cplx<T>* __restrict__ X1, Cplx<T>* __restrict__ X2, int d,
                double* __restrict__ partial, unsigned* __restrict__ ticket, double* __restrict__ e_out,
                double* __restrict__ e_host) {
  __shared__ double red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double e = 0.0;
  if (i < npairs) {
    double xr, xi;
    const PairVals pv = pair_vals(pd, i);
    const bool self = pv.self;
    e = pair_loss(pv, S[pd.h[i]], norm, gscale, MODE > 0, xr, xi);
    if (MODE > 0) {
      const int h = pd.h[i], hn = pd.hneg[i];
      if (MODE == 1) {
        // k = 0 (the only self-paired mode) moves no point: leaving it out
        // keeps its constant, of size N, off the grid whose derivative is
        // interpolated (the constant's derivative is not exactly zero on a
        // finite kernel)
        if (!self) {
          X0[h] = Cplx<T>{(T)xr, (T)xi};
          if (hn >= 0) X0[hn] = Cplx<T>{(T)xr, (T)(-xi)};
        }
      } else {
        // f_a = -i k_a factor  =>  X_a = i k_a X
        Cplx<T>* Xs[3] = {X0, X1, X2};
        for (int a = 0; a < d; ++a) {
          const double ka = (double)pd.kv[3 * i + a];
          const double yr = -ka * xi, yi = ka * xr;
          Xs[a][h] = Cplx<T>{(T)yr, (T)yi};
          if (hn >= 0) Xs[a][hn] = Cplx<T>{(T)yr, (T)(-yi)};
        }
      }
    }
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int hh = 128; hh >= 1; hh >>= 1) {
    if ((int)threadIdx.x < hh) red[threadIdx.x] += red[threadIdx.x + hh];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
  // the last workgroup sums the partials in their fixed order (the energy of
  // the evaluation, nuwu.hpp:270-271), with no extra launch
  if (last_block_arrived(ticket, gridDim.x)) block_sum_partials(partial, gridDim.x, e_out, e_host);
}

struct RhoDev {
  int d, kp, n;
  int nf[3];
  int hx;  // nf[0]/2 + 1
  int cb;  // >= 0: S in the compact band layout of BandDft with band cb
  const double* psi0;
  const double* psi1;
  const double* psi2;
};

template <typename T>
__global__ void k_nuwu_rho(const Cplx<T>* __restrict__ S, RhoDev r, int64_t m, double* __restrict__ rho) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int k[3] = {0, 0, 0};
  k[0] = (int)(i % r.n) - r.kp;
  if (r.d >= 2) k[1] = (int)((i / r.n) % r.n) - r.kp;
  if (r.d >= 3) k[2] = (int)(i / ((int64_t)r.n * r.n)) - r.kp;
  bool rep = k[0] > 0 || (k[0] == 0 && (k[1] > 0 || (k[1] == 0 && k[2] >= 0)));
  int q[3];
  for (int a = 0; a < 3; ++a) q[a] = rep ? k[a] : -k[a];
  int hy = q[1] < 0 ? q[1] + r.nf[1] : q[1];
  int hz = q[2] < 0 ? q[2] + r.nf[2] : q[2];
  int64_t h = q[0] + (int64_t)r.hx * (hy + (int64_t)r.nf[1] * hz);
  if (r.cb >= 0) h = ((int64_t)(q[2] + r.cb) * (2 * r.cb + 1) + (q[1] + r.cb)) * (r.cb + 1) + q[0];
  double ps = r.psi0[abs(k[0])];
  if (r.d >= 2) ps *= r.psi1[abs(k[1])];
  if (r.d >= 3) ps *= r.psi2[abs(k[2])];
  Cplx<T> sv = S[h];
  double re = (double)sv.re / ps, im = (double)sv.im / ps;
  // rep: rho = conj(S)/Psi;  mirrored: rho_k = conj(rho_-k) = S_-k / Psi
  rho[2 * i] = re;
  rho[2 * i + 1] = rep ? -im : im;
}

// pair tables <- full-grid targets (device), for an unchanged active set
__global__ void k_gather_targets(const double* __restrict__ Sk, const double* __restrict__ V, PairDev pd, int npairs,
                                 double* __restrict__ vk, double* __restrict__ vm, double* __restrict__ sk,
                                 double* __restrict__ sm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  vk[i] = V[pd.ik[i]];
  vm[i] = V[pd.imk[i]];
  sk[i] = Sk[pd.ik[i]];
  sm[i] = Sk[pd.imk[i]];
}

// number of modes with V != 0 outside the active pair set
__global__ void k_count_inactive(const double* __restrict__ V, const unsigned char* __restrict__ active, int64_t m,
                                 unsigned int* __restrict__ count) {
  unsigned int c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    c += (V[i] != 0.0 && !active[i]) ? 1u : 0u;
  if (c) atomicAdd(count, c);
}

int kvec_size(int64_t m, int d) {
  // calculate_Kvec (nuwu.hpp:54-79)
  if (d == 1) return (int)((m - 1) / 2);
  if (d == 2) return (int)((std::sqrt((double)m) - 1.0) / 2.0);
  if (d == 3) return (int)((std::cbrt((double)m) - 1.0) / 2.0);
  throw RuntimeFailure("Invalid dimension: " + std::to_string(d));
}

template <typename T>
class NUwU final : public Potential {
 public:
  // force_band >= 0: fine grid for modes |k_a| <= force_band regardless of V
  NUwU(const double* radii, int64_t n, int d, double K, const double* Sk, const double* V, int64_t m,
       const double* L, double eps, int precision, double sigma, int force_band = -1)
      : N_(n), d_(d), plain_(precision == 0) {
    (void)K;  // stored but unused by the arithmetic (nuwu.hpp:23,39)
    if (d < 1 || d > 3) throw InvalidArgument("Invalid dimension: " + std::to_string(d));
    if (n <= 0) throw InvalidArgument("radii must be non-empty");
    if (m <= 0) throw InvalidArgument("Sk / V must be non-empty");
    kp_ = kvec_size(m, d);
    nm_ = 2 * kp_ + 1;
    int64_t mm = 1;
    for (int a = 0; a < d; ++a) mm *= nm_;
    if (mm != m)
      throw InvalidArgument("V.size() = " + std::to_string(m) + " is not (2K+1)^d; nearest grid is " +
                            std::to_string(nm_) + "^" + std::to_string(d));
    M_ = m;
    for (int a = 0; a < d; ++a) {
      if (!(L[a] > 0)) throw InvalidArgument("box lengths must be positive");
      L_[a] = L[a];
    }
    if (sigma <= 0) sigma = 2.0;
    sigma_ = sigma;
    eps_ = eps;
    precision_ = precision;
    // the adjoint path interpolates the kernel derivative: one extra width
    // (two in 1D) keeps its gradient at the accuracy the eps rule gives values
    es_ = es_params(eps, sigma, plain_ ? 0 : (d == 1 ? 2 : 1));

    // weights c (nuwu.hpp:125-149)
    std::vector<double> c(n, 1.0);
    double sum = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      for (int a = 0; a < d; ++a) {
        if (d == 2) c[i] = c[i] * M_PI * radii[i] * radii[i];
        else if (d == 3) c[i] = c[i] * M_PI * radii[i] * radii[i] * radii[i] * 4.0 / 3.0;
      }
      sum += c[i];
    }
    std::vector<T> ct(n);
    for (int64_t i = 0; i < n; ++i) ct[i] = (T)(c[i] * ((double)n / sum));
    c_.allocate(n);
    c_.upload(ct.data(), n);

    radii_host_.assign(radii, radii + n);
    K_ = K;
    Sk_full_.allocate(M_);
    V_full_.allocate(M_);
    Sk_full_.upload(Sk, M_);
    V_full_.upload(V, M_);
    setup_grid(force_band >= 0 ? force_band : active_band(V));
    build_pairs(Sk, V);

    // coordinate maps: energy path (nuwu.hpp:160) has +L, gradient path (219) not
    for (int a = 0; a < 3; ++a) {
      double La = a < d ? L_[a] : 1.0;
      map_e_.L[a] = map_g_.L[a] = La;
      map_e_.shift[a] = La;
      map_g_.shift[a] = 0.0;
      map_e_.scale[a] = map_g_.scale[a] = 2.0 * M_PI / La;
    }
  }

  int64_t ndof() const override { return N_ * d_; }
  int64_t nmodes() const override { return M_; }

  double energy(const double* x, hipStream_t s) override {
    last_x_ = x;
    kept_ = false;
    {
      ProfScope p(ST_SORT, s);
      pts_.set(x, map_e_, c_.get(), s);
    }
    spreader_.spread(pts_, grid_.get(), s);  // times ST_SPREAD (+ ST_MERGE) itself
    if (fz_) {
      {
        ProfScope p(ST_FFT_FWD, s);
        bdft_.forward_planes(grid_.get(), s);
      }
      ProfScope p(ST_LOSS, s);
      launch_zloss(0, s, e_dev_.get(), host_.dev());
      return host_.wait(s);
    }
    {
      ProfScope p(ST_FFT_FWD, s);
      forward_spectrum(s);
    }
    {
      ProfScope p(ST_LOSS, s);
      launch_loss<0>(s, e_dev_.get(), host_.dev());
    }
    return host_.wait(s);
  }

  double energy_gradient(const double* x, double* grad, hipStream_t s) override {
    enqueue_eg(x, grad, e_dev_.get(), host_.dev(), s);
    return host_.wait(s);
  }

  // the energy reaches the host before the interpolation (and, in the fused
  // band form, most of the inverse transform) has run
  double energy_gradient_stream(const double* x, double* grad, hipStream_t s) override {
    host_.arm();
    enqueue_eg(x, grad, e_dev_.get(), host_.dev(), s);
    return host_.wait_value(s);
  }

  bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) override {
    enqueue_eg(x, grad, e_dev, nullptr, s);
    return true;
  }

  // the gradient pass split at the pullback spectrum: spreading, transform
  // and loss now, the inverse transform and interpolation in gradient_kept
  // (the sorted points and the pullback spectrum stay until the next
  // evaluation or target update)
  bool energy_keep(const double* x, double* e, uint64_t* token, hipStream_t s) override {
    enqueue_front(x, e_dev_.get(), host_.dev(), s);
    *e = host_.wait(s);
    kept_ = true;
    *token = ++keep_serial_;
    return true;
  }
  bool gradient_kept(uint64_t token, double* grad, hipStream_t s) override {
    if (!kept_ || token != keep_serial_) return false;
    if (grad) enqueue_back(grad, s);
    return true;
  }
  // the spreader's first call may time its candidates (host waits)
  bool capture_ready() const override { return spreader_.tuned; }
  uint64_t graph_epoch() const override { return epoch_; }

  void get_rho(double* rho, hipStream_t s) override {
    if (band_ < kp_) {
      // the fine grid covers only the active band: evaluate the full mode
      // grid from the last coordinates
      if (!last_x_) throw RuntimeFailure("get_rho needs a previous evaluation");
      if (!full_) {
        const std::vector<double> zero(M_, 0.0);
        full_ = std::make_unique<NUwU<T>>(radii_host_.data(), N_, d_, K_, zero.data(), zero.data(), M_, L_, eps_,
                                          precision_, sigma_, kp_);
      }
      full_->energy(last_x_, s);
      full_->get_rho(rho, s);
      return;
    }
    RhoDev r;
    r.d = d_;
    r.kp = kp_;
    r.n = nm_;
    for (int a = 0; a < 3; ++a) r.nf[a] = geom_.nf[a];
    r.hx = hx_;
    r.cb = bd_ ? bdft_.band() : -1;
    r.psi0 = psi_[0].get();
    r.psi1 = psi_[1].get();
    r.psi2 = psi_[2].get();
    hipLaunchKernelGGL(k_nuwu_rho<T>, dim3(blocks_for(M_, 256)), dim3(256), 0, s, S_.get(), r, M_, rho);
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  void plan_info(int64_t* nf, int* w, double* beta, int* tile) const override {
    for (int a = 0; a < 3; ++a) {
      nf[a] = geom_.nf[a];
      tile[a] = geom_.tile[a];
    }
    *w = es_.w;
    *beta = es_.beta;
  }

  int spread_source() const override { return spreader_.halo_ok ? spreader_.tune_source : 0; }
  void spread_config(int* mode, int* parts) const override {
    *mode = spreader_.mode;
    *parts = spreader_.mode == Spreader<T>::REG ? spreader_.parts : spreader_.halo_parts;
  }

  void target_grads(double* dS0, double* dV, hipStream_t s) override {
    if (!rho_tmp_.get()) rho_tmp_.allocate(2 * (size_t)M_);
    get_rho(rho_tmp_.get(), s);
    launch_target_grads(rho_tmp_.get(), Sk_full_.get(), V_full_.get(), M_, (double)N_, nullptr, dS0, dV, s);
  }

  void set_targets(const double* Sk, const double* V, hipStream_t s) override {
    kept_ = false;
    FRESCO_HIP_CHECK(hipMemcpyAsync(Sk_full_.get(), Sk, M_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (!V) {
      // weights unchanged: the active modes stay, no host check
      regather_targets(s);
      return;
    }
    FRESCO_HIP_CHECK(hipMemcpyAsync(V_full_.get(), V, M_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    // modes switched on outside the active pair set need new pair tables
    fill_zero(count_.get(), 1, s);
    hipLaunchKernelGGL(k_count_inactive, dim3(blocks_for(std::min<int64_t>(M_, 1 << 20), 256)), dim3(256), 0, s,
                       V_full_.get(), active_.get(), M_, count_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    unsigned int cnt = 0;
    FRESCO_HIP_CHECK(hipMemcpyAsync(&cnt, count_.get(), sizeof(cnt), hipMemcpyDeviceToHost, s));
    FRESCO_HIP_CHECK(hipStreamSynchronize(s));
    if (cnt) {
      std::vector<double> hs(M_), hv(M_);
      FRESCO_HIP_CHECK(hipMemcpy(hs.data(), Sk_full_.get(), M_ * sizeof(double), hipMemcpyDeviceToHost));
      FRESCO_HIP_CHECK(hipMemcpy(hv.data(), V_full_.get(), M_ * sizeof(double), hipMemcpyDeviceToHost));
      const int nb = active_band(hv.data());
      if (nb > band_) setup_grid(nb);
      build_pairs(hs.data(), hv.data());
      return;
    }
    regather_targets(s);
  }

 private:
  // the pair tables' targets and weights from the full arrays
  void regather_targets(hipStream_t s) {
    if (npairs_ > 0) {
      hipLaunchKernelGGL(k_gather_targets, dim3((npairs_ + 255) / 256), dim3(256), 0, s, Sk_full_.get(),
                         V_full_.get(), pd_, npairs_, p_vk_.get(), p_vm_.get(), p_sk_.get(), p_sm_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
    }
  }

  // e_host: device-visible pinned host memory that also receives the energy
  // (the host-synchronous entry points), or null (device-only: graphs)
  void enqueue_eg(const double* x, double* grad, double* e_out, double* e_host, hipStream_t s) {
    if (fz_) {
      // the tiles' energies are summed by the inverse transform's first
      // workgroup (the launch after the loss: no fence, no extra launch)
      enqueue_front(x, e_out, e_host, s, true);
      kept_ = false;
      enqueue_back(grad, s, e_out, e_host);
      return;
    }
    enqueue_front(x, e_out, e_host, s);
    kept_ = false;
    enqueue_back(grad, s);
  }

  // binning, spreading, forward transform, loss and pullback spectrum
  // defer (fused band form): the energy is left as per-tile partials for
  // enqueue_back to sum
  void enqueue_front(const double* x, double* e_out, double* e_host, hipStream_t s, bool defer = false) {
    last_x_ = x;
    {
      ProfScope p(ST_SORT, s);
      pts_.set(x, map_g_, c_.get(), s);
    }
    spreader_.spread(pts_, grid_.get(), s);  // times ST_SPREAD (+ ST_MERGE) itself
    if (fz_) {
      // x and y passes per plane, then the z pass, the loss and the inverse
      // z pass of the pullback per tile of band columns
      {
        ProfScope p(ST_FFT_FWD, s);
        bdft_.forward_planes(grid_.get(), s);
      }
      ProfScope p(ST_LOSS, s);
      launch_zloss(1, s, e_out, e_host, defer);
      return;
    }
    if (bd_) {
      // the band transform's last pass also clears the pullback spectrum
      ProfScope p(ST_FFT_FWD, s);
      bdft_.forward(grid_.get(), S_.get(), s, X_[0].get());
    } else {
      {
        ProfScope p(ST_FFT_FWD, s);
        forward_spectrum(s);
      }
      ProfScope p(ST_ZERO, s);
      for (int a = 0; a < nX_; ++a) fill_zero(X_[a].get(), nspec_, s);
    }
    {
      ProfScope p(ST_LOSS, s);
      if (plain_)
        launch_loss<2>(s, e_out, e_host);
      else
        launch_loss<1>(s, e_out, e_host);
    }
  }

  // inverse transform(s) of the pullback spectrum and the interpolation of
  // the gradient at the sorted points
  // e_out / e_host not null: the fused band form's deferred energy sum
  void enqueue_back(double* grad, hipStream_t s, double* e_out = nullptr, double* e_host = nullptr) {
    if (plain_) {
      for (int a = 0; a < d_; ++a) {
        {
          ProfScope p(ST_FFT_INV, s);
          c2r_.exec(X_[a].get(), grid_.get(), s);
        }
        ProfScope p(ST_INTERP, s);
        interpolate(pts_, grid_.get(), false, nullptr, grad, d_, a, s);
      }
    } else {
      {
        ProfScope p(ST_FFT_INV, s);
        if (fz_)
          bdft_.inverse_planes(grid_.get(), s, e_out ? partial_.get() : nullptr, e_out, e_host);
        else if (bd_)
          bdft_.inverse(X_[0].get(), grid_.get(), s);
        else
          c2r_.exec(X_[0].get(), grid_.get(), s);
      }
      ProfScope p(ST_INTERP, s);
      interpolate(pts_, grid_.get(), true, dscale_, grad, d_, 0, s);
    }
  }

  void forward_spectrum(hipStream_t s) {
    if (bd_)
      bdft_.forward(grid_.get(), S_.get(), s);
    else
      r2c_.exec(grid_.get(), S_.get(), s);
  }

  // largest |k_a| over modes whose pair {k, -k} carries weight; the FP64
  // plain path keeps the reference's full mode grid
  int active_band(const double* V) const {
    if (plain_ || !V) return kp_;
    int b = 0;
    const int n = nm_;
    for (int64_t i = 0; i < M_; ++i) {
      if (V[i] == 0.0) continue;
      int k[3] = {0, 0, 0};
      k[0] = (int)(i % n) - kp_;
      if (d_ >= 2) k[1] = (int)((i / n) % n) - kp_;
      if (d_ >= 3) k[2] = (int)(i / ((int64_t)n * n)) - kp_;
      for (int a = 0; a < d_; ++a) b = std::max(b, std::abs(k[a]));
    }
    return b;
  }

  // fine grid sized for the modes |k_a| <= band (sigma (2 band + 1) cells, the
  // NUFFT's accuracy for those modes is that of the full grid)
  void setup_grid(int band) {
    ++epoch_;
    band_ = band;
    const int nmb = 2 * band + 1;
    int nf[3] = {1, 1, 1};
    for (int a = 0; a < d_; ++a) nf[a] = next_smooth_even((int)std::ceil(sigma_ * nmb), es_.w);
    geom_ = make_geom(d_, nf, es_.w, (int)sizeof(T));
    pts_.init(geom_, es_, N_);
    spreader_.init(geom_, es_);
    grid_.allocate(geom_.ncells());
    hx_ = geom_.nf[0] / 2 + 1;
    // the adjoint path's 3D spectrum comes from pruned DFTs of the band when
    // the band is small against the grid (FRESCO_BANDDFT=0: always the FFTs)
    const char* env = std::getenv("FRESCO_BANDDFT");
    bd_ = !plain_ && d_ == 3 && !(env && env[0] == '0') && BandDft<T>::suits(geom_.nf, band);
    if (bd_) {
      bdft_.init(geom_.nf, band);
      nspec_ = bdft_.size();
      fz_ = bdft_.fused();
    } else {
      fz_ = false;
      nspec_ = (int64_t)hx_ * geom_.nf[1] * geom_.nf[2];
    }
    S_.allocate(nspec_);
    nX_ = plain_ ? d_ : 1;
    for (int a = 0; a < nX_; ++a) X_[a].allocate(nspec_);
    const bool dbl = sizeof(T) == 8;
    if (!bd_) {
      r2c_.create(FftPlan::R2C, dbl, d_, geom_.nf);
      c2r_.create(FftPlan::C2R, dbl, d_, geom_.nf);
    }
    for (int a = 0; a < 3; ++a) {
      std::vector<double> ps = a < d_ ? es_deconv(es_, geom_.nf[a], kp_) : std::vector<double>(kp_ + 1, 1.0);
      psi_host_[a] = ps;
      psi_[a].allocate(ps.size());
      psi_[a].upload(ps.data(), ps.size());
    }
    for (int a = 0; a < 3; ++a) dscale_[a] = a < d_ ? -(2.0 / es_.w) * (geom_.nf[a] / (2.0 * M_PI)) : 0.0;
  }

  template <int MODE>
  void launch_loss(hipStream_t s, double* e_out, double* e_host) {
    Cplx<T>* X0 = X_[0].get();
    Cplx<T>* X1 = nX_ > 1 ? X_[1].get() : nullptr;
    Cplx<T>* X2 = nX_ > 2 ? X_[2].get() : nullptr;
    hipLaunchKernelGGL((k_nuwu_loss<T, MODE>), dim3(nblk_), dim3(256), 0, s, S_.get(), pd_, npairs_,
                       (double)N_, 4.0 / (double)N_, X0, X1, X2, d_, partial_.get(), ticket_.get(), e_out, e_host);
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  void launch_zloss(int mode, hipStream_t s, double* e_out, double* e_host, bool defer = false) {
    bdft_.zloss(mode, pd_, bl_, (double)N_, 4.0 / (double)N_, S_.get(), partial_.get(), e_out, e_host, defer, s);
  }

  void build_pairs(const double* Sk, const double* V) {
    ++epoch_;
    std::vector<int> h, hn, ik, imk, kv;
    std::vector<double> vk, vm, sk, sm, ip;
    const int n = nm_;
    for (int64_t i = 0; i < M_; ++i) {
      int k[3] = {0, 0, 0};
      k[0] = (int)(i % n) - kp_;
      if (d_ >= 2) k[1] = (int)((i / n) % n) - kp_;
      if (d_ >= 3) k[2] = (int)(i / ((int64_t)n * n)) - kp_;
      bool rep = k[0] > 0 || (k[0] == 0 && (k[1] > 0 || (k[1] == 0 && k[2] >= 0)));
      if (!rep) continue;
      int64_t j = (int64_t)(-k[0] + kp_);
      if (d_ >= 2) j += (int64_t)n * (-k[1] + kp_);
      if (d_ >= 3) j += (int64_t)n * n * (-k[2] + kp_);
      if (V[i] == 0.0 && V[j] == 0.0) continue;
      int hy = k[1] < 0 ? k[1] + geom_.nf[1] : k[1];
      int hz = k[2] < 0 ? k[2] + geom_.nf[2] : k[2];
      h.push_back(bd_ ? (int)bdft_.index(k[0], k[1], k[2])
                      : (int)(k[0] + (int64_t)hx_ * (hy + (int64_t)geom_.nf[1] * hz)));
      if (k[0] == 0 && j != i) {
        int ny = -k[1] < 0 ? -k[1] + geom_.nf[1] : -k[1];
        int nz = -k[2] < 0 ? -k[2] + geom_.nf[2] : -k[2];
        hn.push_back(bd_ ? (int)bdft_.index(0, -k[1], -k[2]) : (int)((int64_t)hx_ * (ny + (int64_t)geom_.nf[1] * nz)));
      } else {
        hn.push_back(-1);
      }
      ik.push_back((int)i);
      imk.push_back((int)j);
      for (int a = 0; a < 3; ++a) kv.push_back(k[a]);
      vk.push_back(V[i]);
      vm.push_back(V[j]);
      sk.push_back(Sk[i]);
      sm.push_back(Sk[j]);
      double ps = psi_host_[0][std::abs(k[0])];
      if (d_ >= 2) ps *= psi_host_[1][std::abs(k[1])];
      if (d_ >= 3) ps *= psi_host_[2][std::abs(k[2])];
      ip.push_back(1.0 / ps);
    }
    npairs_ = (int)h.size();
    std::vector<int> loc, mir, pstart;
    if (fz_) {
      // fused band form: pairs grouped by the z-pass tile of their
      // representative's column (stable), with their places in the tile
      const int B = bdft_.band();
      std::vector<int> tile(npairs_);
      loc.resize(npairs_);
      mir.resize(npairs_);
      for (int q = 0; q < npairs_; ++q) {
        const int k0 = kv[3 * q], k1 = kv[3 * q + 1], k2 = kv[3 * q + 2];
        int t, sl;
        bdft_.column_slot(k1, k0, &t, &sl);
        tile[q] = t;
        loc[q] = (k2 + B) * 16 + sl;
        mir[q] = -1;
        if (hn[q] >= 0) {
          int t2, sl2;
          bdft_.column_slot(-k1, 0, &t2, &sl2);
          if (t2 != t) throw RuntimeFailure("band transform: a column and its mirror in different tiles");
          mir[q] = (-k2 + B) * 16 + sl2;
        }
      }
      std::vector<int> order(npairs_);
      for (int q = 0; q < npairs_; ++q) order[q] = q;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tile[a] < tile[b]; });
      auto perm_i = [&](std::vector<int>& v, int per) {
        std::vector<int> o((size_t)npairs_ * per);
        for (int q = 0; q < npairs_; ++q)
          for (int c = 0; c < per; ++c) o[(size_t)q * per + c] = v[(size_t)order[q] * per + c];
        v.swap(o);
      };
      auto perm_d = [&](std::vector<double>& v) {
        std::vector<double> o(npairs_);
        for (int q = 0; q < npairs_; ++q) o[q] = v[order[q]];
        v.swap(o);
      };
      perm_i(h, 1);
      perm_i(hn, 1);
      perm_i(ik, 1);
      perm_i(imk, 1);
      perm_i(kv, 3);
      perm_i(loc, 1);
      perm_i(mir, 1);
      perm_d(vk);
      perm_d(vm);
      perm_d(sk);
      perm_d(sm);
      perm_d(ip);
      pstart.assign(bdft_.ntiles() + 1, 0);
      for (int q = 0; q < npairs_; ++q) ++pstart[tile[order[q]] + 1];
      for (int t = 0; t < bdft_.ntiles(); ++t) pstart[t + 1] += pstart[t];
    }
    std::vector<unsigned char> act(M_, 0);
    for (int q = 0; q < npairs_; ++q) act[ik[q]] = act[imk[q]] = 1;
    active_.allocate(M_);
    active_.upload(act.data(), M_);
    if (!count_.get()) count_.allocate(1);
    const size_t np = std::max<size_t>(npairs_, 1);
    auto up_i = [&](DeviceBuffer<int>& b, std::vector<int>& v, size_t cnt) {
      v.resize(cnt, 0);
      b.allocate(cnt);
      b.upload(v.data(), cnt);
    };
    auto up_d = [&](DeviceBuffer<double>& b, std::vector<double>& v) {
      v.resize(np, 0.0);
      b.allocate(np);
      b.upload(v.data(), np);
    };
    up_i(p_h_, h, np);
    up_i(p_hn_, hn, np);
    up_i(p_ik_, ik, np);
    up_i(p_imk_, imk, np);
    up_i(p_kv_, kv, 3 * np);
    up_d(p_vk_, vk);
    up_d(p_vm_, vm);
    up_d(p_sk_, sk);
    up_d(p_sm_, sm);
    up_d(p_ip_, ip);
    pd_ = PairDev{p_h_.get(), p_hn_.get(), p_ik_.get(), p_imk_.get(), p_kv_.get(), p_vk_.get(),
                  p_vm_.get(), p_sk_.get(), p_sm_.get(), p_ip_.get()};
    if (fz_) {
      up_i(p_loc_, loc, np);
      up_i(p_mir_, mir, np);
      up_i(p_pstart_, pstart, pstart.size());
      bl_ = BandLossDev{p_pstart_.get(), p_loc_.get(), p_mir_.get()};
    }
    nblk_ = std::max(1, (npairs_ + 255) / 256);
    partial_.allocate(std::max(nblk_, fz_ ? bdft_.ntiles() : 1));
    if (!ticket_.get()) {
      ticket_.allocate(1);
      FRESCO_HIP_CHECK(hipMemset(ticket_.get(), 0, sizeof(unsigned)));
    }
    e_dev_.allocate(1);
  }

  int64_t N_;
  int d_;
  bool plain_;
  int kp_ = 0, nm_ = 1;
  int64_t M_ = 0;
  double L_[3] = {1, 1, 1};
  EsParams es_;
  GridGeom geom_;
  SortedPoints<T> pts_;
  Spreader<T> spreader_;
  DeviceBuffer<T> c_;
  DeviceBuffer<T> grid_;
  int hx_ = 1;
  int64_t nspec_ = 0;
  DeviceBuffer<Cplx<T>> S_;
  DeviceBuffer<Cplx<T>> X_[3];
  int nX_ = 1;
  FftPlan r2c_, c2r_;
  bool bd_ = false;  // spectrum through bdft_ (compact band layout)
  bool fz_ = false;  // bdft_'s fused form (transforms and loss in three launches)
  BandDft<T> bdft_;
  std::vector<double> psi_host_[3];
  DeviceBuffer<double> psi_[3];
  DeviceBuffer<int> p_h_, p_hn_, p_ik_, p_imk_, p_kv_;
  DeviceBuffer<double> p_vk_, p_vm_, p_sk_, p_sm_, p_ip_;
  PairDev pd_{};
  DeviceBuffer<int> p_loc_, p_mir_, p_pstart_;  // pair tiles of the fused band form
  BandLossDev bl_{};
  int npairs_ = 0, nblk_ = 1;
  DeviceBuffer<double> partial_, e_dev_;
  DeviceBuffer<unsigned> ticket_;  // last-workgroup ticket of the loss kernel
  DeviceBuffer<double> Sk_full_, V_full_, rho_tmp_;  // targets on the full mode grid
  DeviceBuffer<unsigned char> active_;               // mode belongs to an active pair
  DeviceBuffer<unsigned int> count_;
  HostScalar host_;
  uint64_t epoch_ = 0;  // bumped whenever buffers or launch shapes change
  CoordMap map_e_, map_g_;
  double dscale_[3] = {0, 0, 0};
  double sigma_ = 2.0, eps_ = 1e-6, K_ = 0.0;
  int precision_ = 0;
  int band_ = 0;                          // modes |k_a| <= band_ are on the fine grid
  std::vector<double> radii_host_;
  const double* last_x_ = nullptr;        // coordinates of the last evaluation
  bool kept_ = false;                     // energy_keep's state still current
  uint64_t keep_serial_ = 0;              // token of the last energy_keep
  std::unique_ptr<NUwU<T>> full_;         // full-mode-grid evaluator for get_rho
};

}  // namespace

std::unique_ptr<Potential> make_nuwu(const double* radii, int64_t n, int d, double K, const double* Sk,
                                     const double* V, int64_t m, const double* L, double eps, int precision,
                                     double sigma) {
  if (precision == 1) return std::make_unique<NUwU<float>>(radii, n, d, K, Sk, V, m, L, eps, precision, sigma);
  if (precision == 0 || precision == 2)
    return std::make_unique<NUwU<double>>(radii, n, d, K, Sk, V, m, L, eps, precision, sigma);
  throw InvalidArgument("unknown precision " + std::to_string(precision));
}

}  // namespace fresco_hip
