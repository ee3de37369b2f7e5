// Pair potentials: inverse power (source/fresco/inversepower_potential.hpp:
// 18-124) and Weeks-Chandler-Andersen (source/fresco/wca.hpp:14-138), with
// cartesian or periodic minimum-image distances (distance.hpp:266-379), over
// all pairs (simple_pairwise_potential.hpp) or through cell lists
// (cell_list_potential.hpp:382-546, cell_lists.hpp:739-1044).
//
// GPU layout: one thread per point i accumulates E_i = 1/2 sum_j E_ij and its
// own gradient g_i = -sum_j gij (x_i - x_j) (gij = -(dV/dr)/r, the
// reference's convention), so no atomics and no second pass over pairs.
//   all pairs : the j loop streams 256-point tiles through LDS;
//   cell lists: points are binned into cells of side >= the interaction
//               cutoff, sorted by cell (the stable bin sort of sort.hip, so
//               the whole evaluation is kernels only and can be recorded
//               in a HIP graph), and each thread visits the
//               distinct neighbour cells (3^d stencil, wrapped, deduplicated
//               for boxes of fewer than 3 cells per side).
// Energies are summed per block in FP64 and reduced in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "nufft.hpp"
#include "potential.hpp"

namespace fresco_hip {

namespace {

constexpr int PT = 256;

struct PairInter {
  int kind;  // 0 inverse power, 1 WCA
  double pow, eps;
  double C6, C12, coff2;
};

// energy and gij (= -(dV/dr)/r) of one pair at squared distance r2
template <int KIND>
__device__ __forceinline__ double pair_eg(const PairInter& P, double r2, double rs, double& gij) {
  if (KIND == 0) {
    // inversepower_potential.hpp:44-58
    if (r2 >= rs * rs) {
      gij = 0.0;
      return 0.0;
    }
    const double r = sqrt(r2);
    const double f = pow(1.0 - r / rs, P.pow) * P.eps;
    gij = -f / ((r - rs) * r);
    return f / P.pow;
  } else {
    // wca.hpp:48-63
    const double ir2 = 1.0 / r2;
    const double ir6 = ir2 * ir2 * ir2;
    const double ir12 = ir6 * ir6;
    if (r2 < P.coff2) {
      gij = 4.0 * P.eps * (-6.0 * P.C6 * ir6 + 12.0 * P.C12 * ir12) * ir2;
      return 4.0 * P.eps * (-P.C6 * ir6 + P.C12 * ir12) + P.eps;
    }
    gij = 0.0;
    return 0.0;
  }
}

struct BoxDev {
  double box[4];
  double ibox[4];
};

template <int D, bool PER>
__device__ __forceinline__ void min_image(double* dr, const BoxDev& b) {
  if (PER) {
#pragma unroll
    for (int a = 0; a < D; ++a) dr[a] -= rint(dr[a] * b.ibox[a]) * b.box[a];  // round_fast, distance.hpp:32-36,179
  }
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int h = PT / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  return sh[0];
}

// all pairs, j tiles through LDS
template <int D, bool PER, int KIND, bool GRAD>
__global__ void __launch_bounds__(PT)
    k_pair_all(const double* __restrict__ x, const double* __restrict__ radii, int64_t n, BoxDev b, PairInter P,
               double* __restrict__ grad, double* __restrict__ epart) {
  __shared__ double sx[PT * D];
  __shared__ double sr[PT];
  __shared__ double red[PT];
  const int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
  double xi[D], gi[D];
  double ri = 0.0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    xi[a] = i < n ? x[i * D + a] : 0.0;
    gi[a] = 0.0;
  }
  if (i < n) ri = radii[i];
  double e = 0.0;
  for (int64_t base = 0; base < n; base += PT) {
    __syncthreads();
    const int64_t j = base + threadIdx.x;
    if (j < n) {
#pragma unroll
      for (int a = 0; a < D; ++a) sx[threadIdx.x * D + a] = x[j * D + a];
      sr[threadIdx.x] = radii[j];
    }
    __syncthreads();
    const int cnt = (int)min<int64_t>(PT, n - base);
    if (i < n) {
      for (int t = 0; t < cnt; ++t) {
        if (base + t == i) continue;
        double dr[D];
        double r2 = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) dr[a] = xi[a] - sx[t * D + a];
        min_image<D, PER>(dr, b);
#pragma unroll
        for (int a = 0; a < D; ++a) r2 += dr[a] * dr[a];
        double gij;
        e += pair_eg<KIND>(P, r2, ri + sr[t], gij);
        if (GRAD && gij != 0) {
#pragma unroll
          for (int a = 0; a < D; ++a) gi[a] -= gij * dr[a];
        }
      }
    }
  }
  if (GRAD && i < n) {
#pragma unroll
    for (int a = 0; a < D; ++a) grad[i * D + a] = gi[a];
  }
  const double tot = block_sum(0.5 * e, red);
  if (threadIdx.x == 0) epart[blockIdx.x] = tot;
}

struct CellDev {
  int nc[4];       // cells per dimension
  int noff[4];     // distinct neighbour offsets per dimension (1..3)
  int off[4][3];   // the offsets
};

// cell of each point (positions wrapped into the box); non-finite
// coordinates raise *bad and go to cell 0
template <int D>
__global__ void k_cell_keys(const double* __restrict__ x, int64_t n, BoxDev b, CellDev c, uint32_t* __restrict__ key,
                            uint32_t* __restrict__ idx, unsigned int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = 0, mul = 1;
  bool ok = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double v = x[i * D + a];
    ok = ok && isfinite(v);
    double u = v * b.ibox[a];
    u -= floor(u);
    int ci = isfinite(u) ? (int)(u * c.nc[a]) : 0;
    ci = ci < 0 ? 0 : (ci >= c.nc[a] ? c.nc[a] - 1 : ci);
    k += mul * (uint32_t)ci;
    mul *= (uint32_t)c.nc[a];
  }
  if (!ok) atomicOr(bad, 1u);
  key[i] = k;
  idx[i] = (uint32_t)i;
}

__global__ void k_cell_start(const uint32_t* __restrict__ keys, int64_t n, int ncells, int* __restrict__ start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < (uint32_t)c)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[c] = (int)lo;
}

template <int D>
__global__ void k_cell_gather(const double* __restrict__ x, const double* __restrict__ radii,
                              const uint32_t* __restrict__ idx, int64_t n, double* __restrict__ xs,
                              double* __restrict__ rs) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = idx[k];
#pragma unroll
  for (int a = 0; a < D; ++a) xs[k * D + a] = x[i * D + a];
  rs[k] = radii[i];
}

// one thread per sorted point; neighbour cells of its cell
template <int D, int KIND, bool GRAD>
__global__ void __launch_bounds__(PT)
    k_pair_cells(const double* __restrict__ xs, const double* __restrict__ rs, const uint32_t* __restrict__ key,
                 const uint32_t* __restrict__ idx, const int* __restrict__ start, int64_t n, BoxDev b, CellDev c,
                 PairInter P, const unsigned int* __restrict__ bad, double* __restrict__ grad,
                 double* __restrict__ epart) {
  __shared__ double red[PT];
  const int64_t k = (int64_t)blockIdx.x * PT + threadIdx.x;
  const bool nonfinite = *bad != 0;
  double e = 0.0;
  if (k < n && !nonfinite) {
    double xi[D], gi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      xi[a] = xs[k * D + a];
      gi[a] = 0.0;
    }
    const double ri = rs[k];
    int ci[D];
    uint32_t kk = key[k];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      ci[a] = (int)(kk % (uint32_t)c.nc[a]);
      kk /= (uint32_t)c.nc[a];
    }
    int o[4] = {0, 0, 0, 0};
    const int total = c.noff[0] * (D > 1 ? c.noff[1] : 1) * (D > 2 ? c.noff[2] : 1) * (D > 3 ? c.noff[3] : 1);
    for (int s = 0; s < total; ++s) {
      int r = s;
      uint32_t cell = 0, mul = 1;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        o[a] = r % c.noff[a];
        r /= c.noff[a];
        int cc = ci[a] + c.off[a][o[a]];
        if (cc < 0) cc += c.nc[a];
        if (cc >= c.nc[a]) cc -= c.nc[a];
        cell += mul * (uint32_t)cc;
        mul *= (uint32_t)c.nc[a];
      }
      const int j0 = start[cell], j1 = start[cell + 1];
      for (int j = j0; j < j1; ++j) {
        if (j == k) continue;
        double dr[D];
        double r2 = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) dr[a] = xi[a] - xs[(int64_t)j * D + a];
        min_image<D, true>(dr, b);
#pragma unroll
        for (int a = 0; a < D; ++a) r2 += dr[a] * dr[a];
        double gij;
        e += pair_eg<KIND>(P, r2, ri + rs[j], gij);
        if (GRAD && gij != 0) {
#pragma unroll
          for (int a = 0; a < D; ++a) gi[a] -= gij * dr[a];
        }
      }
    }
    if (GRAD) {
      const int64_t i = idx[k];
#pragma unroll
      for (int a = 0; a < D; ++a) grad[i * D + a] = gi[a];
    }
  } else if (k < n && nonfinite) {
    // non-finite coordinates: NaN energy and gradient (as the reference's
    // arithmetic gives, cell_lists tests getEnergy_nan / _inf)
    e = NAN;
    if (GRAD) {
      const int64_t i = idx[k];
#pragma unroll
      for (int a = 0; a < D; ++a) grad[i * D + a] = NAN;
    }
  }
  const double tot = block_sum(0.5 * e, red);
  if (threadIdx.x == 0) epart[blockIdx.x] = tot;
}

class PairPotential final : public Potential {
 public:
  PairPotential(int kind, double p1, double eps, const double* radii, int64_t n, int d, const double* box,
                int method, int cells, double ncellx_scale)
      : kind_(kind), n_(n), d_(d), periodic_(method == 2), cells_(cells != 0) {
    if (d < 1 || d > 4) throw InvalidArgument("pair potentials support 1 to 4 dimensions, got " + std::to_string(d));
    if (kind != 0 && kind != 1) throw InvalidArgument("unknown pair interaction");
    if (method != 1 && method != 2) throw InvalidArgument("method must be 1 (cartesian) or 2 (periodic)");
    if (cells_ && !periodic_) throw InvalidArgument("cell lists need periodic boundaries (as in the reference)");
    if (n <= 0) throw InvalidArgument("radii must be non-empty");
    (void)ncellx_scale;
    P_.kind = kind;
    if (kind == 0) {
      // InversePowerInteraction(a, eps)
      P_.pow = p1;
      P_.eps = eps;
    } else {
      // WCA_interaction(sig, eps), wca.hpp:23-30
      const double s2 = p1 * p1;
      P_.C6 = s2 * s2 * s2;
      P_.C12 = P_.C6 * P_.C6;
      P_.coff2 = std::pow(2.0, 1.0 / 3) * s2;
      P_.eps = eps;
    }
    for (int a = 0; a < 4; ++a) {
      b_.box[a] = 1.0;
      b_.ibox[a] = 1.0;
    }
    if (periodic_) {
      if (!box) throw InvalidArgument("periodic boundaries need a box");
      for (int a = 0; a < d; ++a) {
        if (!(box[a] > 0)) throw InvalidArgument("box lengths must be positive");
        b_.box[a] = box[a];
        b_.ibox[a] = 1.0 / box[a];
      }
    }
    double rmax = 0.0;
    for (int64_t i = 0; i < n; ++i) rmax = std::max(rmax, radii[i]);
    radii_.allocate(n);
    radii_.upload(radii, n);
    nblk_ = (int)((n + PT - 1) / PT);
    epart_.allocate(nblk_);
    e_.allocate(1);
    if (cells_) {
      // cell side >= the interaction range: 2 max(r) for the inverse power
      // (its rcut, inversepower_potential.hpp:115), and the WCA cutoff
      // 2^(1/6) sigma when that is larger, so every interacting pair is seen
      double rcut = 2.0 * rmax;
      if (kind == 1) rcut = std::max(rcut, std::sqrt(P_.coff2));
      if (!(rcut > 0)) throw InvalidArgument("interaction range must be positive for cell lists");
      // at most ~2 cells per point in total
      int64_t total = 1;
      for (int a = 0; a < d; ++a) {
        int nc = (int)std::floor(box[a] / rcut);
        c_.nc[a] = std::max(1, nc);
        total *= c_.nc[a];
      }
      while (total > std::max<int64_t>(2 * n, 1)) {
        int big = 0;
        for (int a = 1; a < d; ++a)
          if (c_.nc[a] > c_.nc[big]) big = a;
        total /= c_.nc[big];
        c_.nc[big] = std::max(1, c_.nc[big] / 2);
        total *= c_.nc[big];
      }
      ncells_ = (int)total;
      for (int a = 0; a < 4; ++a) {
        if (a >= d) {
          c_.nc[a] = 1;
          c_.noff[a] = 1;
          c_.off[a][0] = 0;
          continue;
        }
        // distinct offsets of {-1, 0, 1} modulo nc
        int cnt = 0;
        std::vector<int> seen;
        for (int o : {0, -1, 1}) {
          int m = ((o % c_.nc[a]) + c_.nc[a]) % c_.nc[a];
          if (std::find(seen.begin(), seen.end(), m) != seen.end()) continue;
          seen.push_back(m);
          c_.off[a][cnt++] = o;
        }
        c_.noff[a] = cnt;
      }
      key_.allocate(n);
      key_s_.allocate(n);
      idx_in_.allocate(n);
      idx_.allocate(n);
      start_.allocate((size_t)ncells_ + 1);
      xs_.allocate((size_t)n * d);
      rs_.allocate(n);
      bad_.allocate(1);
      sorter_.init(n, bits_for(ncells_));
    }
  }

  int64_t ndof() const override { return n_ * d_; }

  double energy(const double* x, hipStream_t s) override {
    run<false>(x, nullptr, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }
  double energy_gradient(const double* x, double* grad, hipStream_t s) override {
    run<true>(x, grad, e_.get(), s);
    return host_.fetch(e_.get(), s);
  }
  // kernels only (cell keys, bin sort, gather, pairs, fixed-order energy sum)
  bool energy_gradient_async(const double* x, double* grad, double* e_dev, hipStream_t s) override {
    run<true>(x, grad, e_dev, s);
    return true;
  }
  bool capture_ready() const override { return true; }
  uint64_t graph_epoch() const override { return 1; }

 private:
  static int bits_for(int ncells) {
    int b = 1;
    while ((1ll << b) < (long long)ncells + 1) ++b;
    return b;
  }

  // enqueues the evaluation; the energy goes to e_out (device)
  template <bool GRAD>
  void run(const double* x, double* grad, double* e_out, hipStream_t s) {
    switch (d_) {
      case 1: run_d<1, GRAD>(x, grad, e_out, s); break;
      case 2: run_d<2, GRAD>(x, grad, e_out, s); break;
      case 3: run_d<3, GRAD>(x, grad, e_out, s); break;
      default: run_d<4, GRAD>(x, grad, e_out, s); break;
    }
  }

  template <int D, bool GRAD>
  void run_d(const double* x, double* grad, double* e_out, hipStream_t s) {
    if (kind_ == 0)
      run_k<D, 0, GRAD>(x, grad, e_out, s);
    else
      run_k<D, 1, GRAD>(x, grad, e_out, s);
  }

  template <int D, int KIND, bool GRAD>
  void run_k(const double* x, double* grad, double* e_out, hipStream_t s) {
    if (cells_) {
      const unsigned nb = blocks_for(n_, PT);
      fill_zero(bad_.get(), 1, s);
      hipLaunchKernelGGL(k_cell_keys<D>, dim3(nb), dim3(PT), 0, s, x, n_, b_, c_, key_.get(), idx_in_.get(),
                         bad_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
      sorter_.sort(key_.get(), idx_in_.get(), key_s_.get(), idx_.get(), s);
      hipLaunchKernelGGL(k_cell_start, dim3(blocks_for((size_t)ncells_ + 1, PT)), dim3(PT), 0, s, key_s_.get(), n_,
                         ncells_, start_.get());
      hipLaunchKernelGGL(k_cell_gather<D>, dim3(nb), dim3(PT), 0, s, x, radii_.get(), idx_.get(), n_, xs_.get(),
                         rs_.get());
      hipLaunchKernelGGL((k_pair_cells<D, KIND, GRAD>), dim3(nb), dim3(PT), 0, s, xs_.get(), rs_.get(),
                         key_s_.get(), idx_.get(), start_.get(), n_, b_, c_, P_, bad_.get(), grad, epart_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
    } else if (periodic_) {
      hipLaunchKernelGGL((k_pair_all<D, true, KIND, GRAD>), dim3(nblk_), dim3(PT), 0, s, x, radii_.get(), n_, b_, P_,
                         grad, epart_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
    } else {
      hipLaunchKernelGGL((k_pair_all<D, false, KIND, GRAD>), dim3(nblk_), dim3(PT), 0, s, x, radii_.get(), n_, b_,
                         P_, grad, epart_.get());
      FRESCO_HIP_CHECK(hipGetLastError());
    }
    reduce_partials(epart_.get(), nblk_, e_out, s);
  }

  int kind_;
  int64_t n_;
  int d_;
  bool periodic_, cells_;
  PairInter P_{};
  BoxDev b_{};
  CellDev c_{};
  int ncells_ = 1;
  int nblk_ = 1;
  DeviceBuffer<double> radii_, epart_, e_, xs_, rs_;
  DeviceBuffer<uint32_t> key_, key_s_, idx_in_, idx_;
  DeviceBuffer<int> start_;
  DeviceBuffer<unsigned int> bad_;
  BinSort sorter_;
  HostScalar host_;
};

}  // namespace

std::unique_ptr<Potential> make_pair(int kind, double p1, double eps, const double* radii, int64_t n, int d,
                                     const double* box, int method, int cells, double ncellx_scale) {
  return std::make_unique<PairPotential>(kind, p1, eps, radii, n, d, box, method, cells, ncellx_scale);
}

}  // namespace fresco_hip
