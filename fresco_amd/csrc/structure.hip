// Structure factor of a point pattern on a full uniform mode grid, its
// real-space transform, and radial profiles of grids: the analysis helpers of
// fresco/utils/utils.py (export_structure 16-52, radial_profile 138-152).
//
// S(k) = |sum_j c_j e^{i k.x_j}|^2 / N for k_a = -nk/2 .. (nk-1)/2 with
// x = 2 pi p / L: points are sorted and spread with real weights onto an
// upsampled grid (the type-1 NUFFT of the NUwU path, all modes |k_a| <= nk/2),
// one R2C FFT, deconvolution.  Because the weights are real,
// |rho_k| = |rho_-k| and every mode is read from the half spectrum.
//
// g(r) = prod(L)/N * Re sum_k (S(k) - 1) e^{2 pi i k.r / nk}, centred
// (utils.py:40-46): the real part is the transform of the Hermitian part of
// S - 1 in FFT order, which one C2R FFT of its half computes.
//
// Radial profile: every cell's shell floor(|i - center|) (cells beyond
// R = shape[0]/2 share shell R+1), a stable sort of the cells by shell, and
// per-shell sums in fixed chunks of the sorted order: deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "nufft.hpp"
#include "structure.hpp"

namespace fresco_hip {

namespace {

struct SkDev {
  int d, nk, kp;
  int nf[3];
  int hx;
  const double* psi0;
  const double* psi1;
  const double* psi2;
  double invN;
};

template <typename T>
__global__ void k_to_weights(const double* __restrict__ c, int64_t n, T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)c[i];
}

// Sk[i], i = (ix * nk + iy) * nk + iz (first index slowest), k_a = i_a - nk/2
template <typename T>
__global__ void k_sk_extract(const Cplx<T>* __restrict__ S, SkDev r, int64_t m, double* __restrict__ Sk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int k[3] = {0, 0, 0};
  const int half = r.nk / 2;
  if (r.d == 1) {
    k[0] = (int)i - half;
  } else if (r.d == 2) {
    k[0] = (int)(i / r.nk) - half;
    k[1] = (int)(i % r.nk) - half;
  } else {
    k[0] = (int)(i / ((int64_t)r.nk * r.nk)) - half;
    k[1] = (int)((i / r.nk) % r.nk) - half;
    k[2] = (int)(i % r.nk) - half;
  }
  const bool rep = k[0] > 0 || (k[0] == 0 && (k[1] > 0 || (k[1] == 0 && k[2] >= 0)));
  int q[3];
  for (int a = 0; a < 3; ++a) q[a] = rep ? k[a] : -k[a];
  const int hy = q[1] < 0 ? q[1] + r.nf[1] : q[1];
  const int hz = q[2] < 0 ? q[2] + r.nf[2] : q[2];
  const int64_t h = q[0] + (int64_t)r.hx * (hy + (int64_t)r.nf[1] * hz);
  double ps = r.psi0[abs(k[0])];
  if (r.d >= 2) ps *= r.psi1[abs(k[1])];
  if (r.d >= 3) ps *= r.psi2[abs(k[2])];
  const Cplx<T> sv = S[h];
  const double re = (double)sv.re, im = (double)sv.im;
  Sk[i] = (re * re + im * im) / (ps * ps) * r.invN;
}

// FFT-order index j -> mode k as numpy's ifftshift places it
__device__ __forceinline__ int fft_mode(int j, int nk) { return j < (nk + 1) / 2 ? j : j - nk; }

// CMCL index of mode k and of its partner -k taken modulo nk
__device__ __forceinline__ int mode_index(int k, int nk) { return k + nk / 2; }
__device__ __forceinline__ int partner_index(int k, int nk) {
  int q = -k;
  if (q + nk / 2 >= nk) q -= nk;  // even nk: -(-nk/2) wraps to -nk/2
  return q + nk / 2;
}

// half spectrum of the Hermitian part of ifftshift(S) - 1 (x halved, lengths
// fastest first).  numpy's real(ifftn(A)) is the inverse transform of
// (A(k) + A(-k mod nk)) / 2; for odd nk that is A itself (S(k) = S(-k)
// exactly), for even nk it averages the Nyquist planes, whose -k partners
// are not -k.
__global__ void k_gr_fill(const double* __restrict__ Sk, int d, int nk, int hx2, int64_t mh,
                          Cplx<double>* __restrict__ H) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mh) return;
  int k[3] = {0, 0, 0};
  const int64_t rest = i / hx2;
  k[0] = fft_mode((int)(i % hx2), nk);
  if (d == 2) k[1] = fft_mode((int)rest, nk);
  if (d == 3) {
    k[1] = fft_mode((int)(rest % nk), nk);
    k[2] = fft_mode((int)(rest / nk), nk);
  }
  int64_t a = 0, b = 0;
  for (int c = 0; c < d; ++c) {
    a = a * nk + mode_index(k[c], nk);
    b = b * nk + partner_index(k[c], nk);
  }
  H[i] = Cplx<double>{0.5 * (Sk[a] + Sk[b]) - 1.0, 0.0};
}

// gr[(ix * nk + iy) * nk + iz] = scale * G[fftshift]; G has x fastest
__global__ void k_gr_shift(const double* __restrict__ G, int d, int nk, int64_t m, double scale,
                           double* __restrict__ gr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int half = nk / 2;
  auto sh = [&](int64_t v) { return (int)((v - half + nk) % nk); };
  int64_t src;
  if (d == 1) {
    src = sh(i);
  } else if (d == 2) {
    src = sh(i / nk) + (int64_t)nk * sh(i % nk);
  } else {
    src = sh(i / ((int64_t)nk * nk)) + (int64_t)nk * (sh((i / nk) % nk) + (int64_t)nk * sh(i % nk));
  }
  gr[i] = G[src] * scale;
}

// ---------------------------------------------------------------------------
// radial profile
// ---------------------------------------------------------------------------

constexpr int RP_CHUNK = 1024;

int bits_for(long long v) {
  int b = 0;
  while ((1LL << b) < v) ++b;
  return b;
}

struct RpShape {
  int d;
  int64_t s[3];
  double c[3];
  int R;
};

// shell of cell i (C order): floor of the float64 distance, as
// np.sqrt(np.sum((idx - center)**2, axis=0)).astype(int)
__global__ void k_rp_keys(RpShape sh, int64_t n, uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t idx[3] = {0, 0, 0};
  int64_t rem = i;
  for (int a = sh.d - 1; a >= 0; --a) {
    idx[a] = rem % sh.s[a];
    rem /= sh.s[a];
  }
  double acc = 0.0;
  for (int a = 0; a < sh.d; ++a) {
    const double t = (double)idx[a] - sh.c[a];
    acc += t * t;
  }
  int64_t r = (int64_t)sqrt(acc);
  // the floor of the exact root of acc (r^2 is exact for these magnitudes)
  if ((double)(r + 1) * (double)(r + 1) <= acc) r += 1;
  else if ((double)r * (double)r > acc) r -= 1;
  key[i] = (uint32_t)(r <= sh.R ? r : sh.R + 1);
  val[i] = (uint32_t)i;
}

__global__ void k_rp_clear(int64_t* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_rp_bounds(const uint32_t* __restrict__ keys, int64_t n, int64_t* __restrict__ start,
                            int64_t* __restrict__ end) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) start[k] = i;
  if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
}

// chunk prefix: shell b owns chunks [cpre[b], cpre[b+1])
__global__ void k_rp_chunks(const int64_t* __restrict__ start, const int64_t* __restrict__ end, int nb,
                            int64_t* __restrict__ cpre) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t t = 0;
  for (int b = 0; b < nb; ++b) {
    cpre[b] = t;
    t += (end[b] - start[b] + RP_CHUNK - 1) / RP_CHUNK;
  }
  cpre[nb] = t;
}

__global__ void k_rp_partial(const uint32_t* __restrict__ cells, const double* __restrict__ data,
                             const int64_t* __restrict__ start, const int64_t* __restrict__ end,
                             const int64_t* __restrict__ cpre, int nb, double* __restrict__ part) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cpre[nb]) return;
  int lo = 0, hi = nb - 1;  // largest b with cpre[b] <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (cpre[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int b = lo;
  const int64_t a = start[b] + (t - cpre[b]) * RP_CHUNK;
  const int64_t e = end[b] < a + RP_CHUNK ? end[b] : a + RP_CHUNK;
  double acc = 0.0;
  for (int64_t i = a; i < e; ++i) acc += data[cells[i]];
  part[t] = acc;
}

__global__ void k_rp_final(const int64_t* __restrict__ cpre, const double* __restrict__ part,
                           const int64_t* __restrict__ start, const int64_t* __restrict__ end, int nb,
                           double* __restrict__ sums, int64_t* __restrict__ counts) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double acc = 0.0;
  for (int64_t t = cpre[b]; t < cpre[b + 1]; ++t) acc += part[t];
  sums[b] = acc;
  counts[b] = end[b] - start[b];
}

template <typename T>
void structure_factor_t(const double* x, const double* c, int64_t n, int d, const double* L, int nk, double eps,
                        double* Sk, double* gr, hipStream_t s) {
  const int kp = nk / 2;
  const EsParams es = es_params(eps, 2.0);
  const int nmb = 2 * kp + 1;
  int nf[3] = {1, 1, 1};
  for (int a = 0; a < d; ++a) nf[a] = next_smooth_even((int)std::ceil(2.0 * nmb), es.w);
  const GridGeom geom = make_geom(d, nf, es.w, (int)sizeof(T));
  SortedPoints<T> pts;
  pts.init(geom, es, n);
  Spreader<T> spreader;
  spreader.init(geom, es);
  DeviceBuffer<T> grid, cw;
  grid.allocate(geom.ncells());
  cw.allocate(n);
  const int hx = geom.nf[0] / 2 + 1;
  const int64_t nspec = (int64_t)hx * geom.nf[1] * geom.nf[2];
  DeviceBuffer<Cplx<T>> S;
  S.allocate(nspec);
  FftPlan r2c;
  r2c.create(FftPlan::R2C, sizeof(T) == 8, d, geom.nf);
  DeviceBuffer<double> psi[3];
  for (int a = 0; a < 3; ++a) {
    const std::vector<double> ps = a < d ? es_deconv(es, geom.nf[a], kp) : std::vector<double>(kp + 1, 1.0);
    psi[a].allocate(ps.size());
    psi[a].upload(ps.data(), ps.size());
  }
  CoordMap map;
  map.wrap = 0;
  for (int a = 0; a < 3; ++a) map.scale[a] = a < d ? 2.0 * M_PI / L[a] : 0.0;

  hipLaunchKernelGGL(k_to_weights<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, c, n, cw.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  pts.set(x, map, cw.get(), s);
  spreader.spread(pts, grid.get(), s);
  r2c.exec(grid.get(), S.get(), s);
  int64_t m = 1;
  for (int a = 0; a < d; ++a) m *= nk;
  SkDev r;
  r.d = d;
  r.nk = nk;
  r.kp = kp;
  for (int a = 0; a < 3; ++a) r.nf[a] = geom.nf[a];
  r.hx = hx;
  r.psi0 = psi[0].get();
  r.psi1 = psi[1].get();
  r.psi2 = psi[2].get();
  r.invN = 1.0 / (double)n;
  hipLaunchKernelGGL(k_sk_extract<T>, dim3(blocks_for(m, 256)), dim3(256), 0, s, S.get(), r, m, Sk);
  FRESCO_HIP_CHECK(hipGetLastError());
  if (gr) {
    const int hx2 = nk / 2 + 1;
    int64_t mh = hx2;
    for (int a = 1; a < d; ++a) mh *= nk;
    DeviceBuffer<Cplx<double>> H;
    DeviceBuffer<double> G;
    H.allocate(mh);
    G.allocate(m);
    int len[3] = {nk, nk, nk};
    FftPlan c2r;
    c2r.create(FftPlan::C2R, true, d, len);
    hipLaunchKernelGGL(k_gr_fill, dim3(blocks_for(mh, 256)), dim3(256), 0, s, Sk, d, nk, hx2, mh, H.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    c2r.exec(H.get(), G.get(), s);
    double vol = 1.0;
    for (int a = 0; a < d; ++a) vol *= L[a];
    hipLaunchKernelGGL(k_gr_shift, dim3(blocks_for(m, 256)), dim3(256), 0, s, G.get(), d, nk, m, vol / (double)n,
                       gr);
    FRESCO_HIP_CHECK(hipGetLastError());
  }
  // the buffers above are freed on return: finish their work first
  FRESCO_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

void structure_factor(const double* x, const double* c, int64_t n, int d, const double* L, int64_t nk, double eps,
                      int precision, double* Sk, double* gr, hipStream_t s) {
  if (d < 1 || d > 3) throw InvalidArgument("Invalid dimension: " + std::to_string(d));
  if (n <= 0) throw InvalidArgument("no points");
  if (nk < 1 || nk > (1 << 20)) throw InvalidArgument("Nk out of range");
  if (!(eps > 0)) throw InvalidArgument("tol must be positive");
  for (int a = 0; a < d; ++a)
    if (!(L[a] > 0)) throw InvalidArgument("box lengths must be positive");
  if (!x || !c || !Sk) throw InvalidArgument("null pointer");
  if (precision == 1)
    structure_factor_t<float>(x, c, n, d, L, (int)nk, eps, Sk, gr, s);
  else
    structure_factor_t<double>(x, c, n, d, L, (int)nk, eps, Sk, gr, s);
}

namespace {

RpShape rp_shape(const double* data, int d, const int64_t* shape, const double* center, double* sums,
                 int64_t* counts, int64_t& n) {
  if (d < 1 || d > 3) throw InvalidArgument("radial profiles take 1 to 3 dimensions");
  if (!data || !sums || !counts) throw InvalidArgument("null pointer");
  RpShape sh;
  sh.d = d;
  n = 1;
  for (int a = 0; a < 3; ++a) {
    sh.s[a] = a < d ? shape[a] : 1;
    sh.c[a] = a < d ? center[a] : 0.0;
    if (sh.s[a] < 1) throw InvalidArgument("empty grid");
    n *= sh.s[a];
  }
  if (n >= (1LL << 31)) throw InvalidArgument("grid too large for the shell sort (2^31 cells)");
  sh.R = (int)(shape[0] / 2);
  return sh;
}

// Device buffers of one radial profile: n cells, nb shells.
struct RpBuffers {
  DeviceBuffer<uint32_t> key, val, ks, vs;
  DeviceBuffer<int64_t> start, end, cpre;
  DeviceBuffer<double> part;
  BinSort sorter;
};

int64_t rp_max_chunks(int64_t n, int nb) { return (n + RP_CHUNK - 1) / RP_CHUNK + nb; }

// shell keys, stable sort by shell, fixed-chunk partial sums, per-shell sums:
// every buffer is written before it is read (no fills)
void rp_run(const RpShape& sh, int64_t n, const double* data, double* sums, int64_t* counts, RpBuffers& b,
            hipStream_t s) {
  const int nb = sh.R + 2;
  const int64_t maxchunks = rp_max_chunks(n, nb);
  hipLaunchKernelGGL(k_rp_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, sh, n, b.key.get(), b.val.get());
  hipLaunchKernelGGL(k_rp_clear, dim3((nb + 255) / 256), dim3(256), 0, s, b.start.get(), nb);
  hipLaunchKernelGGL(k_rp_clear, dim3((nb + 255) / 256), dim3(256), 0, s, b.end.get(), nb);
  FRESCO_HIP_CHECK(hipGetLastError());
  b.sorter.sort(b.key.get(), b.val.get(), b.ks.get(), b.vs.get(), s);
  hipLaunchKernelGGL(k_rp_bounds, dim3(blocks_for(n, 256)), dim3(256), 0, s, b.ks.get(), n, b.start.get(),
                     b.end.get());
  hipLaunchKernelGGL(k_rp_chunks, dim3(1), dim3(64), 0, s, b.start.get(), b.end.get(), nb, b.cpre.get());
  hipLaunchKernelGGL(k_rp_partial, dim3(blocks_for(maxchunks, 256)), dim3(256), 0, s, b.vs.get(), data,
                     b.start.get(), b.end.get(), b.cpre.get(), nb, b.part.get());
  hipLaunchKernelGGL(k_rp_final, dim3((nb + 255) / 256), dim3(256), 0, s, b.cpre.get(), b.part.get(), b.start.get(),
                     b.end.get(), nb, sums, counts);
  FRESCO_HIP_CHECK(hipGetLastError());
  FRESCO_HIP_CHECK(hipStreamSynchronize(s));
}

// Scratch of the allocation-free radial profile: one array per GPU, held by
// the loaded code object until the process ends.  Its address on the current
// device comes from hipGetSymbolAddress (the name is only a host-side stand-in).
constexpr size_t RP_SCRATCH_BYTES = (size_t)20 << 20;
constexpr size_t rp_round(size_t b) { return (b + 255) & ~(size_t)255; }
// largest carve: 4 key/index arrays, shell bounds and chunk prefix, chunk
// partials, and the one-pass sort's count table (2^11 digits x 4096-key
// tiles) with its segment totals
static_assert(4 * rp_round(RP_NOALLOC_MAX_CELLS * 4) + 2 * rp_round(RP_NOALLOC_MAX_SHELLS * 8) +
                      rp_round((RP_NOALLOC_MAX_SHELLS + 1) * 8) +
                      rp_round(((RP_NOALLOC_MAX_CELLS + RP_CHUNK - 1) / RP_CHUNK + RP_NOALLOC_MAX_SHELLS) * 8) +
                      rp_round(2048 * ((RP_NOALLOC_MAX_CELLS + 4095) / 4096) * 4) +
                      rp_round((2048 * ((RP_NOALLOC_MAX_CELLS + 4095) / 4096) + 4095) / 4096 * 4) <=
                  RP_SCRATCH_BYTES,
              "radial profile scratch too small for its limits");
static __device__ __align__(256) unsigned char scratch[RP_SCRATCH_BYTES];

constexpr int RP_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[RP_MAX_DEVICES] = {};

void* rp_scratch_on_current_device() {
  int dev = 0;
  FRESCO_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= RP_MAX_DEVICES) throw RuntimeFailure("device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw HipFailure(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

void radial_profile(const double* data, int d, const int64_t* shape, const double* center, double* sums,
                    int64_t* counts, hipStream_t s) {
  int64_t n = 0;
  const RpShape sh = rp_shape(data, d, shape, center, sums, counts, n);
  const int nb = sh.R + 2;
  RpBuffers b;
  b.key.allocate(n);
  b.val.allocate(n);
  b.ks.allocate(n);
  b.vs.allocate(n);
  b.start.allocate(nb);
  b.end.allocate(nb);
  b.cpre.allocate(nb + 1);
  b.part.allocate(rp_max_chunks(n, nb));
  b.sorter.init(n, std::max(1, bits_for(nb)));
  rp_run(sh, n, data, sums, counts, b, s);
}

void radial_profile_noalloc(const double* data, int d, const int64_t* shape, const double* center, double* sums,
                            int64_t* counts, hipStream_t s) {
  int64_t n = 0;
  const RpShape sh = rp_shape(data, d, shape, center, sums, counts, n);
  const int nb = sh.R + 2;
  if (n > RP_NOALLOC_MAX_CELLS)
    throw InvalidArgument("radial_profile_noalloc: " + std::to_string(n) + " cells exceed its limit of " +
                          std::to_string(RP_NOALLOC_MAX_CELLS));
  if (nb > RP_NOALLOC_MAX_SHELLS)
    throw InvalidArgument("radial_profile_noalloc: shape[0] = " + std::to_string(shape[0]) + " gives " +
                          std::to_string(nb) + " shells, beyond its limit of " +
                          std::to_string(RP_NOALLOC_MAX_SHELLS) + " (shape[0] <= " +
                          std::to_string(2 * (RP_NOALLOC_MAX_SHELLS - 2) + 1) + ")");
  Carve ws(rp_scratch_on_current_device(), RP_SCRATCH_BYTES);
  RpBuffers b;
  ws.take(b.key, (size_t)n);
  ws.take(b.val, (size_t)n);
  ws.take(b.ks, (size_t)n);
  ws.take(b.vs, (size_t)n);
  ws.take(b.start, (size_t)nb);
  ws.take(b.end, (size_t)nb);
  ws.take(b.cpre, (size_t)nb + 1);
  ws.take(b.part, (size_t)rp_max_chunks(n, nb));
  b.sorter.init(n, std::max(1, bits_for(nb)), ws);
  rp_run(sh, n, data, sums, counts, b, s);
}

}  // namespace fresco_hip
