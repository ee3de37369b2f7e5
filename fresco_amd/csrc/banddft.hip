// Pruned separable DFTs between a real 3D fine grid and its low band of
// modes (banddft.hpp).  Every pass is one launch of k_band_pass: for each
// outer index o (grid.x) it contracts the axis j of a [J x ITILE] slab of the
// input, staged in LDS with its [KTILE x J] twiddle rows, into ITILE x KTILE
// outputs; a thread keeps NPT accumulators.  LANE_K false: lane = inner index
// i, the thread's outputs are k = kg + 8 m; LANE_K true: lane = output index
// k, the thread's inputs are i = ig + 8 m (used where k has unit stride in
// the output, so the writes are coalesced).
#include "banddft.hpp"

#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

namespace fresco_hip {

namespace {

constexpr int BD_T = 256;
constexpr int BD_JMAX = 128;  // longest contracted axis (grids up to 128 cells per dimension)
constexpr int BD_NPT = 2;     // accumulators per thread (2: measured faster than 1 and 4)

template <typename T, bool IN_REAL, bool OUT_REAL, bool LANE_K, int NPT>
__global__ void __launch_bounds__(BD_T)
    k_band_pass(const void* __restrict__ in_, int64_t so, int64_t si, int64_t sj, int I, int J,
                const Cplx<T>* __restrict__ W, int K, void* __restrict__ out_, int64_t to, int64_t ti, int64_t tk,
                Cplx<T>* __restrict__ zero_out) {
  using IT = typename std::conditional<IN_REAL, T, Cplx<T>>::type;
  constexpr int ITILE = LANE_K ? 8 * NPT : 32;
  constexpr int KTILE = LANE_K ? 32 : 8 * NPT;
  static_assert(ITILE <= 32 && KTILE <= 32, "tiles of at most 32");
  extern __shared__ __align__(16) unsigned char smem[];
  const int JP = J | 1;  // odd twiddle row stride: lanes reading different rows hit different banks
  IT* slab = reinterpret_cast<IT*>(smem);  // [J][ITILE]
  Cplx<T>* ws = reinterpret_cast<Cplx<T>*>(smem + (((size_t)J * ITILE * sizeof(IT) + 15) & ~(size_t)15));
  const IT* in = static_cast<const IT*>(in_);
  const int o = blockIdx.x;
  const int i0 = blockIdx.y * ITILE, k0 = blockIdx.z * KTILE;
  const int tid = threadIdx.x;
  const int64_t ob = (int64_t)o * so;
  // staging: every load of a thread is issued before any LDS store (one
  // memory latency per workgroup instead of one per element); 32
  // consecutive threads walk the unit-stride index (j when sj == 1, else i)
  const int t32 = tid & 31, r8 = tid >> 5;
  constexpr int JB = BD_JMAX / 32;
  if (sj == 1) {
    IT v[ITILE / 8][JB];
#pragma unroll
    for (int a = 0; a < ITILE / 8; ++a) {
      const int il = r8 + 8 * a;
      const IT* src = in + ob + (int64_t)(i0 + il) * si;
#pragma unroll
      for (int b = 0; b < JB; ++b) {
        const int j = t32 + 32 * b;
        v[a][b] = (i0 + il < I && j < J) ? src[j] : IT{};
      }
    }
#pragma unroll
    for (int a = 0; a < ITILE / 8; ++a)
#pragma unroll
      for (int b = 0; b < JB; ++b)
        if (t32 + 32 * b < J) slab[(t32 + 32 * b) * ITILE + r8 + 8 * a] = v[a][b];
  } else {
    // lane t32 < ITILE takes inner index t32 (ITILE <= 32 here)
    IT v[JB * 4];  // j = r8 + 8 q covers BD_JMAX
    const bool ok = t32 < ITILE && i0 + t32 < I;
    const IT* src = in + ob + (int64_t)(i0 + t32) * si;
#pragma unroll
    for (int q = 0; q < JB * 4; ++q) {
      const int j = r8 + 8 * q;
      v[q] = (ok && j < J) ? src[(int64_t)j * sj] : IT{};
    }
#pragma unroll
    for (int q = 0; q < JB * 4; ++q)
      if (t32 < ITILE && r8 + 8 * q < J) slab[(r8 + 8 * q) * ITILE + t32] = v[q];
  }
  {
    Cplx<T> w[KTILE / 8][JB];
#pragma unroll
    for (int a = 0; a < KTILE / 8; ++a) {
      const int kl = r8 + 8 * a;
      const Cplx<T>* src = W + (int64_t)(k0 + kl) * J;
#pragma unroll
      for (int b = 0; b < JB; ++b) {
        const int j = t32 + 32 * b;
        w[a][b] = (k0 + kl < K && j < J) ? src[j] : Cplx<T>{};
      }
    }
#pragma unroll
    for (int a = 0; a < KTILE / 8; ++a)
#pragma unroll
      for (int b = 0; b < JB; ++b)
        if (t32 + 32 * b < J) ws[(r8 + 8 * a) * JP + t32 + 32 * b] = w[a][b];
  }
  __syncthreads();
  const int lane = tid & 31, grp = tid >> 5;
  T ar[NPT], ai[NPT];
#pragma unroll
  for (int m = 0; m < NPT; ++m) ar[m] = ai[m] = T(0);
  auto mac = [&](int m, const IT& v, const Cplx<T>& w) {
    if constexpr (IN_REAL) {
      ar[m] += v * w.re;
      ai[m] += v * w.im;
    } else if constexpr (OUT_REAL) {
      ar[m] += v.re * w.re - v.im * w.im;
    } else {
      ar[m] += v.re * w.re - v.im * w.im;
      ai[m] += v.re * w.im + v.im * w.re;
    }
  };
  // four j per step, the step's LDS reads issued before its multiply-adds
  int j = 0;
  for (; j + 4 <= J; j += 4) {
    IT v[4][NPT];
    Cplx<T> w[4][NPT];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int m = 0; m < NPT; ++m) {
        v[u][m] = LANE_K ? slab[(j + u) * ITILE + grp + 8 * m] : slab[(j + u) * ITILE + lane];
        w[u][m] = LANE_K ? ws[lane * JP + j + u] : ws[(grp + 8 * m) * JP + j + u];
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int m = 0; m < NPT; ++m) mac(m, v[u][m], w[u][m]);
  }
  for (; j < J; ++j) {
#pragma unroll
    for (int m = 0; m < NPT; ++m) {
      const IT v = LANE_K ? slab[j * ITILE + grp + 8 * m] : slab[j * ITILE + lane];
      const Cplx<T> w = LANE_K ? ws[lane * JP + j] : ws[(grp + 8 * m) * JP + j];
      mac(m, v, w);
    }
  }
#pragma unroll
  for (int m = 0; m < NPT; ++m) {
    const int i = i0 + (LANE_K ? grp + 8 * m : lane);
    const int k = k0 + (LANE_K ? lane : grp + 8 * m);
    if (i < I && k < K) {
      const int64_t at = (int64_t)o * to + (int64_t)i * ti + (int64_t)k * tk;
      if constexpr (OUT_REAL)
        static_cast<T*>(out_)[at] = ar[m];
      else
        static_cast<Cplx<T>*>(out_)[at] = Cplx<T>{ar[m], ai[m]};
      if (zero_out) zero_out[at] = Cplx<T>{T(0), T(0)};
    }
  }
}

template <typename T, bool IN_REAL, bool OUT_REAL, bool LANE_K, int NPT>
void launch_pass(const void* in, int64_t so, int64_t si, int64_t sj, int O, int I, int J, const Cplx<T>* W, int K,
                 void* out, int64_t to, int64_t ti, int64_t tk, hipStream_t s, Cplx<T>* zero_out = nullptr) {
  using IT = typename std::conditional<IN_REAL, T, Cplx<T>>::type;
  constexpr int ITILE = LANE_K ? 8 * NPT : 32;
  constexpr int KTILE = LANE_K ? 32 : 8 * NPT;
  const size_t lds = (((size_t)J * ITILE * sizeof(IT) + 15) & ~(size_t)15) + (size_t)KTILE * (J | 1) * sizeof(Cplx<T>);
  if (lds > 160 * 1024 || J > BD_JMAX)
    throw InvalidArgument("band transform pass: contracted axis over " + std::to_string(BD_JMAX) + " or LDS exceeded");
  const dim3 grd(O, (I + ITILE - 1) / ITILE, (K + KTILE - 1) / KTILE);
  hipLaunchKernelGGL((k_band_pass<T, IN_REAL, OUT_REAL, LANE_K, NPT>), grd, dim3(BD_T), lds, s, in, so, si, sj, I, J,
                     W, K, out, to, ti, tk, zero_out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

// W[k][j] = wt(j') exp(sign 2 pi i k' j' / n) with k' = k + koff, j' = j + joff
template <typename T>
void twiddles(DeviceBuffer<Cplx<T>>& buf, int K, int koff, int J, int joff, int n, int sign, bool weight_kx) {
  std::vector<Cplx<T>> h((size_t)K * J);
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < J; ++j) {
      const long long kk = k + koff, jj = j + joff;
      long long r = (kk * jj) % n;
      if (r < 0) r += n;
      const double ang = sign * 2.0 * M_PI * (double)r / (double)n;
      const double wt = weight_kx ? (jj == 0 ? 1.0 : 2.0) : 1.0;
      h[(size_t)k * J + j] = Cplx<T>{(T)(wt * std::cos(ang)), (T)(wt * std::sin(ang))};
    }
  buf.allocate(h.size());
  buf.upload(h.data(), h.size());
}

}  // namespace

template <typename T>
void BandDft<T>::init(const int* nf, int band) {
  if (!suits(nf, band)) throw InvalidArgument("band transform: band or grid out of range");
  for (int a = 0; a < 3; ++a) {
    nf_[a] = nf[a];
    if (nf[a] < 2 * band + 1) throw InvalidArgument("band transform: grid narrower than the band");
  }
  B_ = band;
  nb_ = band + 1;
  const int nk = 2 * band + 1;
  twiddles(wa_, nb_, 0, nf_[0], 0, nf_[0], -1, false);     // kx in [0, B]  x x
  twiddles(wb_, nk, -band, nf_[1], 0, nf_[1], -1, false);  // ky in [-B, B] x y
  twiddles(wc_, nk, -band, nf_[2], 0, nf_[2], -1, false);  // kz x z
  twiddles(wci_, nf_[2], 0, nk, -band, nf_[2], 1, false);  // z x kz
  twiddles(wbi_, nf_[1], 0, nk, -band, nf_[1], 1, false);  // y x ky
  twiddles(wai_, nf_[0], 0, nb_, 0, nf_[0], 1, true);      // x x kx, kx > 0 twice
  u1_.allocate((size_t)nf_[2] * nf_[1] * nb_);
  u2_.allocate((size_t)nf_[2] * nk * nb_);
}

template <typename T>
void BandDft<T>::forward(const T* grid, Cplx<T>* S, hipStream_t s, Cplx<T>* zero) {
  const int nx = nf_[0], ny = nf_[1], nz = nf_[2], nk = 2 * B_ + 1, nb = nb_;
  // x: grid[z][y][x] -> U1[z][y][kx]; lanes over kx (coalesced writes)
  launch_pass<T, true, false, true, BD_NPT>(grid, (int64_t)ny * nx, nx, 1, nz, ny, nx, wa_.get(), nb, u1_.get(),
                                       (int64_t)ny * nb, nb, 1, s);
  // y: U1[z][y][kx] -> U2[z][ky][kx]; lanes over kx
  launch_pass<T, false, false, false, BD_NPT>(u1_.get(), (int64_t)ny * nb, 1, nb, nz, nb, ny, wb_.get(), nk, u2_.get(),
                                         (int64_t)nk * nb, 1, nb, s);
  // z: U2[z][ky][kx] -> S[kz][ky][kx]
  launch_pass<T, false, false, false, BD_NPT>(u2_.get(), nb, 1, (int64_t)nk * nb, nk, nb, nz, wc_.get(), nk, S, nb, 1,
                                              (int64_t)nk * nb, s, zero);
}

template <typename T>
void BandDft<T>::inverse(const Cplx<T>* X, T* grid, hipStream_t s) {
  const int nx = nf_[0], ny = nf_[1], nz = nf_[2], nk = 2 * B_ + 1, nb = nb_;
  // z: X[kz][ky][kx] -> Y1[z][ky][kx] (in u2_)
  launch_pass<T, false, false, false, BD_NPT>(X, nb, 1, (int64_t)nk * nb, nk, nb, nk, wci_.get(), nz, u2_.get(), nb, 1,
                                         (int64_t)nk * nb, s);
  // y: Y1[z][ky][kx] -> Y2[z][y][kx] (in u1_)
  launch_pass<T, false, false, false, BD_NPT>(u2_.get(), (int64_t)nk * nb, 1, nb, nz, nb, nk, wbi_.get(), ny, u1_.get(),
                                         (int64_t)ny * nb, 1, nb, s);
  // x: Y2[z][y][kx] -> grid[z][y][x] (real part; lanes over x)
  launch_pass<T, false, true, true, BD_NPT>(u1_.get(), (int64_t)ny * nb, nb, 1, nz, ny, nb, wai_.get(), nx, grid,
                                       (int64_t)ny * nx, nx, 1, s);
}

template class BandDft<float>;
template class BandDft<double>;

}  // namespace fresco_hip
