// Batched NUwU: many independent 1D / 2D patterns of uneven size evaluated
// together.  Every pattern keeps the loss of source/fresco/nuwu.hpp:152-317
// (its own N, mode grid, S0, V and weights); the batch shares one fine-grid
// size (the largest active band over the patterns, so each pattern is
// upsampled by at least sigma), one radix sort of all points by
// (pattern, bin), one spreading launch and one interpolation launch over all
// patterns' bins (pattern = blockIdx.y), and one batched FFT per direction.
// The loss runs over every pattern's active mode pairs, with each pattern's
// pairs padded to whole blocks so that block energies reduce per pattern in a
// fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "nufft.hpp"
#include "potential.hpp"

namespace fresco_hip {

namespace {

constexpr int BT = 256;

struct BPairDev {
  const int64_t* h;     // spectrum index of k (pattern offset included)
  const int64_t* hneg;  // index of -k when kx == 0 and k != 0, else -1
  const int* kv;        // k components [3 per pair]
  const unsigned char* self;  // k == -k
  const double* Vk;
  const double* Vm;
  const double* Sk;
  const double* Sm;
  const double* ipsi;   // 1 / Psi(k)
  const int* blk_pat;   // pattern of every block
  const int* blk_first; // first pair of every block
  const int* blk_count; // pairs in every block
  const double* npts;   // N of every pattern (the loss normalisation)
};

template <typename T, int MODE>
__global__ void __launch_bounds__(BT)
    k_batch_loss(const Cplx<T>* __restrict__ S, BPairDev pd, Cplx<T>* __restrict__ X, double* __restrict__ partial) {
  __shared__ double red[BT];
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  double e = 0.0;
  if (t < pd.blk_count[b]) {
    const int i = pd.blk_first[b] + t;
    const double norm = pd.npts[pd.blk_pat[b]];
    const Cplx<T> sv = S[pd.h[i]];
    const double ip = pd.ipsi[i];
    const double re = (double)sv.re * ip;  // rho_k = conj(S_k) / Psi(k)
    const double im = -(double)sv.im * ip;
    const double s = (re * re + im * im) / norm;  // nuwu.hpp:262-264
    const double Sk = pd.Sk[i], Vk = pd.Vk[i];
    double D = s - Sk;
    if (Sk != 0) D /= Sk;
    e = Vk * D * D;
    double gk = Vk * D;
    if (Sk != 0) gk /= Sk;
    double gm = gk;
    if (!pd.self[i]) {
      const double Sm = pd.Sm[i], Vm = pd.Vm[i];
      double Dm = s - Sm;
      if (Sm != 0) Dm /= Sm;
      e += Vm * Dm * Dm;
      gm = Vm * Dm;
      if (Sm != 0) gm /= Sm;
    }
    if (MODE == 1) {
      // H = (factor_k + conj(factor_-k)) / 2, factor = (4/N) g rho
      // (nuwu.hpp:272); C2R input conj(H) / Psi
      const double gs = 4.0 / norm;
      double hr, hi;
      if (pd.self[i]) {
        hr = gs * gk * re;
        hi = 0.0;
      } else {
        const double gg = 0.5 * gs * (gk + gm);
        hr = gg * re;
        hi = gg * im;
      }
      const double xr = hr * ip, xi = -hi * ip;
      // k = 0 moves no point: kept off the derivative-interpolated grid
      if (!pd.self[i]) {
        X[pd.h[i]] = Cplx<T>{(T)xr, (T)xi};
        const int64_t hn = pd.hneg[i];
        if (hn >= 0) X[hn] = Cplx<T>{(T)xr, (T)(-xi)};
      }
    }
  }
  red[t] = e;
  __syncthreads();
  for (int hh = BT / 2; hh >= 1; hh >>= 1) {
    if (t < hh) red[t] += red[t + hh];
    __syncthreads();
  }
  if (t == 0) partial[b] = red[0];
}

// energy of every pattern: its blocks' partials in order (compensated), and
// the total as the in-order sum over patterns
__global__ void k_pattern_sums(const double* __restrict__ partial, const int* __restrict__ pat_blk0,
                               int npat, double* __restrict__ epat, double* __restrict__ total) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npat) {
    double s = 0.0, c = 0.0;
    for (int b = pat_blk0[p]; b < pat_blk0[p + 1]; ++b) {
      const double y = partial[b] - c;
      const double t = s + y;
      c = (t - s) - y;
      s = t;
    }
    epat[p] = s;
  }
}

// total_host (device-visible pinned host memory, or null) receives it too
__global__ void k_total(const double* __restrict__ epat, int npat, double* __restrict__ total,
                        double* __restrict__ total_host) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0, c = 0.0;
    for (int p = 0; p < npat; ++p) {
      const double y = epat[p] - c;
      const double t = s + y;
      c = (t - s) - y;
      s = t;
    }
    total[0] = s;
    if (total_host) {
      total_host[0] = s;
      __threadfence_system();
    }
  }
}

int kvec_size_b(int64_t m, int d) {
  // calculate_Kvec (nuwu.hpp:54-79)
  if (d == 1) return (int)((m - 1) / 2);
  if (d == 2) return (int)((std::sqrt((double)m) - 1.0) / 2.0);
  throw InvalidArgument("batched plans are 1D / 2D");
}

template <typename T>
class BatchNUwU final : public Potential {
 public:
  BatchNUwU(int npat, const int64_t* npts, const double* radii, int d, const int64_t* nmodes, const double* Sk,
            const double* V, const double* L, double eps, double sigma)
      : P_(npat), d_(d) {
    if (d != 1 && d != 2) throw InvalidArgument("batched plans are 1D / 2D, got d = " + std::to_string(d));
    if (npat <= 0) throw InvalidArgument("need at least one pattern");
    if (sigma <= 0) sigma = 2.0;
    // the adjoint gradient's kernel-derivative interpolation: one extra width (two in 1D)
    es_ = es_params(eps, sigma, d == 1 ? 2 : 1);
    for (int a = 0; a < d; ++a)
      if (!(L[a] > 0)) throw InvalidArgument("box lengths must be positive");
    pt0_.assign(npat + 1, 0);
    mo0_.assign(npat + 1, 0);
    kp_.resize(npat);
    int band = 0, kpmax = 0;
    for (int p = 0; p < npat; ++p) {
      if (npts[p] <= 0) throw InvalidArgument("pattern " + std::to_string(p) + " has no points");
      pt0_[p + 1] = pt0_[p] + npts[p];
      mo0_[p + 1] = mo0_[p] + nmodes[p];
      kp_[p] = kvec_size_b(nmodes[p], d);
      int64_t mm = 1;
      for (int a = 0; a < d; ++a) mm *= 2 * kp_[p] + 1;
      if (mm != nmodes[p])
        throw InvalidArgument("pattern " + std::to_string(p) + ": V.size() = " + std::to_string(nmodes[p]) +
                              " is not (2K+1)^d");
      kpmax = std::max(kpmax, kp_[p]);
      const int n = 2 * kp_[p] + 1;
      for (int64_t i = 0; i < nmodes[p]; ++i) {
        if (V[mo0_[p] + i] == 0.0) continue;
        band = std::max(band, std::abs((int)(i % n) - kp_[p]));
        if (d == 2) band = std::max(band, std::abs((int)(i / n) - kp_[p]));
      }
    }
    Ntot_ = pt0_[npat];
    // one fine-grid size for the batch: the widest active band
    int nf[3] = {1, 1, 1};
    for (int a = 0; a < d; ++a) nf[a] = next_smooth_even((int)std::ceil(sigma * (2 * band + 1)), es_.w);
    geom_ = make_geom(d, nf, es_.w, (int)sizeof(T), npat);
    pts_.init(geom_, es_, Ntot_);
    pts_.set_segments(pt0_);
    spreader_.init(geom_, es_);
    G_ = geom_.ncells();
    grid_.allocate((size_t)G_ * npat);
    hx_ = geom_.nf[0] / 2 + 1;
    nspec_ = (int64_t)hx_ * geom_.nf[1];
    S_.allocate((size_t)nspec_ * npat);
    X_.allocate((size_t)nspec_ * npat);
    const bool dbl = sizeof(T) == 8;
    r2c_.create(FftPlan::R2C, dbl, d, geom_.nf, npat);
    c2r_.create(FftPlan::C2R, dbl, d, geom_.nf, npat);
    std::vector<double> psi[2];
    for (int a = 0; a < d; ++a) psi[a] = es_deconv(es_, geom_.nf[a], kpmax);

    // weights per pattern (nuwu.hpp:125-149), pattern index per point
    std::vector<T> ct(Ntot_);
    std::vector<uint32_t> pat(Ntot_);
    for (int p = 0; p < npat; ++p) {
      double sum = 0.0;
      std::vector<double> c(npts[p], 1.0);
      for (int64_t i = 0; i < npts[p]; ++i) {
        const double r = radii[pt0_[p] + i];
        for (int a = 0; a < d; ++a)
          if (d == 2) c[i] = c[i] * M_PI * r * r;
        sum += c[i];
      }
      for (int64_t i = 0; i < npts[p]; ++i) {
        ct[pt0_[p] + i] = (T)(c[i] * ((double)npts[p] / sum));
        pat[pt0_[p] + i] = (uint32_t)p;
      }
    }
    c_.allocate(Ntot_);
    c_.upload(ct.data(), Ntot_);
    pat_.allocate(Ntot_);
    pat_.upload(pat.data(), Ntot_);

    // active pairs per pattern, padded to whole blocks
    std::vector<int64_t> h, hn;
    std::vector<int> kv, bpat, bfirst, bcount, pblk0(npat + 1, 0);
    std::vector<unsigned char> self;
    std::vector<double> vk, vm, sk, sm, ip, np(npat);
    for (int p = 0; p < npat; ++p) {
      np[p] = (double)npts[p];
      const int n = 2 * kp_[p] + 1;
      const int kq = kp_[p];
      const int64_t first = (int64_t)h.size();
      for (int64_t i = 0; i < nmodes[p]; ++i) {
        int k[2] = {(int)(i % n) - kq, d == 2 ? (int)(i / n) - kq : 0};
        const bool rep = k[0] > 0 || (k[0] == 0 && k[1] >= 0);
        if (!rep) continue;
        int64_t j = (int64_t)(-k[0] + kq);
        if (d == 2) j += (int64_t)n * (-k[1] + kq);
        const double Vi = V[mo0_[p] + i], Vj = V[mo0_[p] + j];
        if (Vi == 0.0 && Vj == 0.0) continue;
        const int hy = k[1] < 0 ? k[1] + geom_.nf[1] : k[1];
        h.push_back((int64_t)p * nspec_ + k[0] + (int64_t)hx_ * hy);
        if (k[0] == 0 && j != i) {
          const int ny = -k[1] < 0 ? -k[1] + geom_.nf[1] : -k[1];
          hn.push_back((int64_t)p * nspec_ + (int64_t)hx_ * ny);
        } else {
          hn.push_back(-1);
        }
        self.push_back(j == i ? 1 : 0);
        kv.push_back(k[0]);
        kv.push_back(k[1]);
        kv.push_back(0);
        vk.push_back(Vi);
        vm.push_back(Vj);
        sk.push_back(Sk[mo0_[p] + i]);
        sm.push_back(Sk[mo0_[p] + j]);
        double ps = psi[0][std::abs(k[0])];
        if (d == 2) ps *= psi[1][std::abs(k[1])];
        ip.push_back(1.0 / ps);
      }
      const int64_t cnt = (int64_t)h.size() - first;
      pblk0[p] = (int)bpat.size();
      for (int64_t o = 0; o < cnt; o += BT) {
        bpat.push_back(p);
        bfirst.push_back((int)(first + o));
        bcount.push_back((int)std::min<int64_t>(BT, cnt - o));
      }
    }
    pblk0[npat] = (int)bpat.size();
    nblk_ = std::max(1, (int)bpat.size());
    if (bpat.empty()) {
      bpat.push_back(0);
      bfirst.push_back(0);
      bcount.push_back(0);
    }
    auto up64 = [](DeviceBuffer<int64_t>& b, std::vector<int64_t>& v) {
      if (v.empty()) v.push_back(0);
      b.allocate(v.size());
      b.upload(v.data(), v.size());
    };
    auto upi = [](DeviceBuffer<int>& b, std::vector<int>& v) {
      if (v.empty()) v.push_back(0);
      b.allocate(v.size());
      b.upload(v.data(), v.size());
    };
    auto upd = [](DeviceBuffer<double>& b, std::vector<double>& v) {
      if (v.empty()) v.push_back(0.0);
      b.allocate(v.size());
      b.upload(v.data(), v.size());
    };
    up64(h_, h);
    up64(hn_, hn);
    upi(kv_, kv);
    if (self.empty()) self.push_back(0);
    self_.allocate(self.size());
    self_.upload(self.data(), self.size());
    upd(vk_, vk);
    upd(vm_, vm);
    upd(sk_, sk);
    upd(sm_, sm);
    upd(ip_, ip);
    upi(bpat_, bpat);
    upi(bfirst_, bfirst);
    upi(bcount_, bcount);
    upi(pblk0_, pblk0);
    upd(np_, np);
    pd_ = BPairDev{h_.get(), hn_.get(), kv_.get(), self_.get(), vk_.get(), vm_.get(), sk_.get(), sm_.get(),
                   ip_.get(), bpat_.get(), bfirst_.get(), bcount_.get(), np_.get()};
    partial_.allocate(nblk_);
    epat_.allocate(npat);
    e_.allocate(1);

    for (int a = 0; a < 3; ++a) {
      const double La = a < d ? L[a] : 1.0;
      map_e_.L[a] = map_g_.L[a] = La;
      map_e_.shift[a] = La;  // energy path: + L (nuwu.hpp:160)
      map_g_.shift[a] = 0.0;
      map_e_.scale[a] = map_g_.scale[a] = 2.0 * M_PI / La;
      dscale_[a] = a < d ? -(2.0 / es_.w) * (geom_.nf[a] / (2.0 * M_PI)) : 0.0;
    }
  }

  int64_t ndof() const override { return Ntot_ * d_; }
  int64_t nmodes() const override { return mo0_[P_]; }

  double energy(const double* x, hipStream_t s) override {
    forward(x, map_e_, s);
    hipLaunchKernelGGL((k_batch_loss<T, 0>), dim3(nblk_), dim3(BT), 0, s, S_.get(), pd_, X_.get(), partial_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    finish_energy(s, e_.get());
    return host_.fetch(e_.get(), s);
  }

  double energy_gradient(const double* x, double* grad, hipStream_t s) override {
    enqueue_eg(x, grad, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }

  // the energy reaches the host before the inverse transforms and the
  // interpolation have run
  double energy_gradient_stream(const double* x, double* grad, hipStream_t s) override {
    host_.arm();
    enqueue_eg(x, grad, e_.get(), s, host_.dev());
    return host_.wait_value(s);
  }

  bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) override {
    enqueue_eg(x, grad, e_dev, s);
    return true;
  }
  bool capture_ready() const override { return spreader_.tuned; }

  void pattern_energies(double* host, int n, hipStream_t s) override {
    if (n != P_) throw InvalidArgument("expected " + std::to_string(P_) + " patterns");
    FRESCO_HIP_CHECK(hipMemcpyAsync(host, epat_.get(), P_ * sizeof(double), hipMemcpyDeviceToHost, s));
    FRESCO_HIP_CHECK(hipStreamSynchronize(s));
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
  void enqueue_eg(const double* x, double* grad, double* e_out, hipStream_t s, double* e_host = nullptr) {
    forward(x, map_g_, s);
    {
      ProfScope p(ST_ZERO, s);
      fill_zero(X_.get(), (size_t)nspec_ * P_, s);
    }
    {
      ProfScope p(ST_LOSS, s);
      hipLaunchKernelGGL((k_batch_loss<T, 1>), dim3(nblk_), dim3(BT), 0, s, S_.get(), pd_, X_.get(),
                         partial_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
      finish_energy(s, e_out, e_host);
    }
    {
      ProfScope p(ST_FFT_INV, s);
      c2r_.exec(X_.get(), grid_.get(), s);
    }
    {
      ProfScope p(ST_INTERP, s);
      interpolate(pts_, grid_.get(), true, dscale_, grad, d_, 0, s);
    }
  }

  void forward(const double* x, const CoordMap& m, hipStream_t s) {
    {
      ProfScope p(ST_SORT, s);
      pts_.set(x, m, c_.get(), s, pat_.get());
    }
    spreader_.spread(pts_, grid_.get(), s);
    ProfScope p(ST_FFT_FWD, s);
    r2c_.exec(grid_.get(), S_.get(), s);
  }

  void finish_energy(hipStream_t s, double* e_out, double* e_host = nullptr) {
    hipLaunchKernelGGL(k_pattern_sums, dim3((P_ + 127) / 128), dim3(128), 0, s, partial_.get(), pblk0_.get(), P_,
                       epat_.get(), e_out);
    hipLaunchKernelGGL(k_total, dim3(1), dim3(64), 0, s, epat_.get(), P_, e_out, e_host);
    FRESCO_HIP_CHECK(hipGetLastError());
  }

  int P_;
  int d_;
  EsParams es_;
  std::vector<int64_t> pt0_, mo0_;
  std::vector<int> kp_;
  int64_t Ntot_ = 0;
  GridGeom geom_;
  SortedPoints<T> pts_;
  Spreader<T> spreader_;
  int64_t G_ = 0;
  int hx_ = 1;
  int64_t nspec_ = 0;
  DeviceBuffer<T> grid_, c_;
  DeviceBuffer<uint32_t> pat_;
  DeviceBuffer<Cplx<T>> S_, X_;
  FftPlan r2c_, c2r_;
  DeviceBuffer<int64_t> h_, hn_;
  DeviceBuffer<int> kv_, bpat_, bfirst_, bcount_, pblk0_;
  DeviceBuffer<unsigned char> self_;
  DeviceBuffer<double> vk_, vm_, sk_, sm_, ip_, np_, partial_, epat_, e_;
  BPairDev pd_{};
  int nblk_ = 1;
  HostScalar host_;
  CoordMap map_e_, map_g_;
  double dscale_[3] = {0, 0, 0};
};

}  // namespace

std::unique_ptr<Potential> make_nuwu_batch(int npat, const int64_t* npts, const double* radii, int d,
                                           const int64_t* nmodes, const double* Sk, const double* V, const double* L,
                                           double eps, int precision, double sigma) {
  if (precision == 1) return std::make_unique<BatchNUwU<float>>(npat, npts, radii, d, nmodes, Sk, V, L, eps, sigma);
  return std::make_unique<BatchNUwU<double>>(npat, npts, radii, d, nmodes, Sk, V, L, eps, sigma);
}

}  // namespace fresco_hip
