// Non-uniform FFT kernels for gfx950: point binning/sorting, owner-computes
// LDS-tiled spreading, wavefront-per-point interpolation, rocFFT plans.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <dlfcn.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <type_traits>

#include "nufft.hpp"
#include "prof.hpp"

namespace fresco_hip {

// ---------------------------------------------------------------------------
// ES kernel parameters and deconvolution factors (host, float64)
// ---------------------------------------------------------------------------

EsParams es_params(double eps, double sigma, int extra_w) {
  if (!(eps > 0)) throw InvalidArgument("eps must be > 0");
  if (!(sigma > 1.0)) throw InvalidArgument("upsampling factor sigma must be > 1");
  EsParams p;
  int w;
  double bw;
  if (sigma == 2.0) {
    w = (int)std::ceil(-std::log10(eps / 10.0)) + extra_w;
    w = std::max(2, w);
    bw = 2.30;
    if (w == 2) bw = 2.20;
    if (w == 3) bw = 2.26;
    if (w == 4) bw = 2.38;
  } else {
    w = (int)std::ceil(-std::log(eps) / (M_PI * std::sqrt(1.0 - 1.0 / sigma))) + extra_w;
    w = std::max(2, w);
    bw = 0.97 * M_PI * (1.0 - 1.0 / (2.0 * sigma));
  }
  if (w > 16) throw InvalidArgument("eps too small: kernel width would exceed 16");
  p.w = w;
  p.beta = bw * w;
  return p;
}

static void gauss_legendre(int q, std::vector<double>& x, std::vector<double>& wt) {
  x.resize(q);
  wt.resize(q);
  for (int i = 0; i < q; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (q + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 0; j < q; ++j) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * j + 1.0) * z * p1 - j * p2) / (j + 1.0);
      }
      double dp = q * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) {
        // recompute derivative at the converged node for the weight
        double a0 = 1.0, a1 = 0.0;
        for (int j = 0; j < q; ++j) {
          double a2 = a1;
          a1 = a0;
          a0 = ((2.0 * j + 1.0) * z * a1 - j * a2) / (j + 1.0);
        }
        dp = q * (z * a0 - a1) / (z * z - 1.0);
        x[i] = z;
        wt[i] = 2.0 / ((1.0 - z * z) * dp * dp);
        break;
      }
    }
  }
}

std::vector<double> es_deconv(const EsParams& p, int n, int kmax) {
  // Psi(k) = (w/2) * int_{-1}^{1} phi(z) cos(pi k w z / n) dz
  const int q = 4 * p.w + 64;
  std::vector<double> z, wt;
  gauss_legendre(q, z, wt);
  std::vector<double> out(kmax + 1);
  for (int k = 0; k <= kmax; ++k) {
    double s = 0.0;
    for (int i = 0; i < q; ++i) {
      double ph = std::exp(p.beta * (std::sqrt(1.0 - z[i] * z[i]) - 1.0));
      s += wt[i] * ph * std::cos(M_PI * k * p.w * z[i] / n);
    }
    out[k] = 0.5 * p.w * s;
  }
  return out;
}

std::vector<double> es_ft(const EsParams& p, const std::vector<double>& u) {
  const int q = 4 * p.w + 64;
  std::vector<double> z, wt;
  gauss_legendre(q, z, wt);
  std::vector<double> ph(q);
  for (int i = 0; i < q; ++i) ph[i] = wt[i] * std::exp(p.beta * (std::sqrt(1.0 - z[i] * z[i]) - 1.0));
  std::vector<double> out(u.size());
  for (size_t k = 0; k < u.size(); ++k) {
    double s = 0.0;
    for (int i = 0; i < q; ++i) s += ph[i] * std::cos(u[k] * z[i]);
    out[k] = s;
  }
  return out;
}

int next_smooth_even(int target, int w) {
  int n = std::max(target, 2 * w);
  if (n % 2) ++n;
  for (;; n += 2) {
    int m = n;
    for (int f : {2, 3, 5})
      while (m % f == 0) m /= f;
    if (m == 1) return n;
  }
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------

static int ceil_log2(long long v) {
  int b = 0;
  while ((1LL << b) < v) ++b;
  return std::max(b, 1);
}

int lds_row_stride(int n, int elem_bytes, int w) {
  // Row stride (elements) for the (ja, jb) lane pattern: consecutive kernel
  // rows of one half-wavefront should land on disjoint LDS banks.
  const int wpe = elem_bytes / 4;
  const int wmax = w <= 8 ? 8 : 16;
  const int target = (wmax * wpe) % 32;
  int words = n * wpe;
  if (target != 0) words += ((target - words % 32) + 32) % 32;
  return words / wpe;
}

size_t spread_lds_bytes(const GridGeom& g, int elem_bytes, int w) {
  int sx = g.d >= 2 ? lds_row_stride(g.tile[0], elem_bytes, w) : g.tile[0];
  return (size_t)elem_bytes * sx * (g.d >= 2 ? g.tile[1] : 1) * (g.d >= 3 ? g.tile[2] : 1);
}

size_t interp_lds_bytes(const GridGeom& g, int elem_bytes, int w) {
  int rx = g.tile[0] + w - 1;
  int ry = g.d >= 2 ? g.tile[1] + w - 1 : 1;
  int rz = g.d >= 3 ? g.tile[2] + w - 1 : 1;
  int srx = g.d >= 2 ? lds_row_stride(rx, elem_bytes, w) : rx;
  return (size_t)elem_bytes * srx * ry * rz;
}

GridGeom make_geom(int d, const int* nf, int w, int elem_bytes, int batch) {
  GridGeom g;
  g.d = d;
  if (batch < 1) throw InvalidArgument("batch must be positive");
  if (batch > 1 && d == 3) throw InvalidArgument("batched plans are 1D / 2D");
  g.batch = batch;
  for (int a = 0; a < 3; ++a) g.nf[a] = a < d ? nf[a] : 1;
  for (int a = 0; a < 3; ++a) g.tile[a] = 1;
  if (d == 3) {
    // 16 x 16 x 8: the register-tile spreader's layout; FP32 grids with a
    // kernel up to 8 wide use 16 x 8 x 8 bins for the halo spreader (23 KB
    // halo tiles: more wavefronts per CU, no point slicing needed, half the
    // halo bytes of 4-sliced 16 x 16 x 8 bins).  Not shrunk further.
    g.tile[0] = 16;
    g.tile[1] = (elem_bytes == 4 && w <= 8) ? 8 : 16;
    g.tile[2] = 8;
  } else if (d == 2) {
    g.tile[0] = g.tile[1] = 64;
  } else {
    g.tile[0] = 1024;
  }
  auto clamp = [&]() {
    for (int a = 0; a < d; ++a) g.tile[a] = std::max(1, std::min(g.tile[a], g.nf[a]));
  };
  auto ntiles = [&]() {
    long long t = 1;
    for (int a = 0; a < 3; ++a) t *= (g.nf[a] + g.tile[a] - 1) / g.tile[a];
    return t;
  };
  auto halve_largest = [&](int minsz) {
    int best = -1;
    for (int a = d - 1; a >= 0; --a)
      if (g.tile[a] > minsz && (best < 0 || g.tile[a] > g.tile[best])) best = a;
    if (best < 0) return false;
    g.tile[best] /= 2;
    return true;
  };
  clamp();
  if (d < 3) {
    // LDS budget: interpolation region <= 80 KB (two workgroups per CU),
    // spreading tile <= 40 KB; enough tiles to fill 256 CUs
    while (interp_lds_bytes(g, elem_bytes, w) > 80 * 1024 || spread_lds_bytes(g, elem_bytes, w) > 40 * 1024)
      if (!halve_largest(4)) break;
    // enough tiles per pattern that the largest pattern of a batch (whose
    // tiles take the longest) spreads over the GPU by itself
    const int minsz = d == 2 ? 16 : 64;
    while (ntiles() < 1024)
      if (!halve_largest(minsz)) break;
    clamp();
  }
  for (int a = 0; a < 3; ++a) g.ntile[a] = (g.nf[a] + g.tile[a] - 1) / g.tile[a];
  g.nbins = g.ntile[0] * g.ntile[1] * g.ntile[2];
  if ((long long)g.nbins * batch >= (1LL << 31)) throw InvalidArgument("too many bins");
  g.key_bits = ceil_log2(g.nbins * batch + 1);
  g.cell_bits = d == 3 ? 10 : (d == 2 ? 16 : 31);
  for (int a = 0; a < d; ++a)
    if (g.nf[a] > (1 << g.cell_bits) - 1)
      throw InvalidArgument("fine grid too large for the packed cell index (" + std::to_string(g.nf[a]) + ")");
  if (interp_lds_bytes(g, elem_bytes, w) > 160 * 1024) throw InvalidArgument("kernel too wide for LDS tiles");
  return g;
}

// ---------------------------------------------------------------------------
// point binning and sorting
// ---------------------------------------------------------------------------

struct MapDev {
  int d, wrap, w;
  double L[3], shift[3], scale[3];
  int nf[3], tile[3], ntile[3];
  int tshift[3];     // log2 tile[a] for power-of-two tiles, else -1
  double itile[3];   // 1 / tile[a]
  int cell_bits;
};

__device__ __forceinline__ void map_component(double v, const MapDev& m, int a, int& cell, double& off) {
  // no fused multiply-adds: every kernel that maps a point gets the same
  // cell and offset bit for bit (the direct binning counts in one kernel and
  // places in another, both mapping the point)
#pragma clang fp contract(off)
  double xs = m.wrap ? (v - round(v / m.L[a]) + m.shift[a]) * m.scale[a] : v * m.scale[a];
  double t = xs * (1.0 / (2.0 * M_PI));
  t -= floor(t);
  double gg = t * (double)m.nf[a];
  double i0 = floor(gg);
  double f = gg - i0;
  double sf = ceil(f - 0.5 * m.w);
  off = f - sf;  // in [w/2 - 1, w/2]
  // i0 in [0, nf] and sf in [-w/2, 1 - w/2]: one wrap step suffices; the
  // remainder only for grids narrower than the kernel
  int ic = (int)(isfinite(i0) ? i0 : 0.0) + (int)(isfinite(sf) ? sf : 0.0);
  if (ic >= m.nf[a])
    ic -= m.nf[a];
  else if (ic < 0)
    ic += m.nf[a];
  if (ic < 0 || ic >= m.nf[a]) {
    ic %= m.nf[a];
    if (ic < 0) ic += m.nf[a];
  }
  cell = ic;
}

// bin key of point j (b0 + tile of its first support cell), its record
// (fine-grid offsets and weight) and its packed first cell; the D
// coordinates are loaded together, the map is unrolled over them
template <typename T, int D>
__device__ __forceinline__ uint32_t point_record_d(const double* __restrict__ x, int64_t j, const MapDev& m,
                                                   uint32_t b0, const T* __restrict__ c, uint32_t& packed,
                                                   typename Vec4<T>::type& r) {
  double v[D];
#pragma unroll
  for (int a = 0; a < D; ++a) v[a] = x[j * D + a];
  const T wt = c ? c[j] : T(1);
  uint32_t b = b0, mul = 1;
  double offs[3] = {0, 0, 0};
  packed = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    int cell;
    map_component(v[a], m, a, cell, offs[a]);
    packed |= (uint32_t)cell << (a * m.cell_bits);
    // cell / tile[a] without an integer division: a shift, else the
    // float64 quotient corrected by one step (cell < 2^31: off by one at most)
    int t;
    if (m.tshift[a] >= 0) {
      t = cell >> m.tshift[a];
    } else {
      t = (int)((double)cell * m.itile[a]);
      if (t * m.tile[a] > cell)
        --t;
      else if ((t + 1) * m.tile[a] <= cell)
        ++t;
    }
    t = t < m.ntile[a] ? t : m.ntile[a] - 1;
    b += mul * (uint32_t)t;
    mul *= (uint32_t)m.ntile[a];
  }
  r.x = (T)offs[0];
  r.y = (T)offs[1];
  r.z = (T)offs[2];
  r.w = wt;
  return b;
}

template <typename T>
__device__ __forceinline__ uint32_t point_record(const double* __restrict__ x, int64_t j, const MapDev& m,
                                                 uint32_t b0, const T* __restrict__ c, uint32_t& packed,
                                                 typename Vec4<T>::type& r) {
  if (m.d == 3) return point_record_d<T, 3>(x, j, m, b0, c, packed, r);
  if (m.d == 2) return point_record_d<T, 2>(x, j, m, b0, c, packed, r);
  return point_record_d<T, 1>(x, j, m, b0, c, packed, r);
}

// bin key of every point (pattern * nbins + tile of its first support cell)
// and, in the original order, its record and packed first cell, which the
// sort carries to the sorted order
template <typename T>
__global__ void k_bin_records(const double* __restrict__ x, int64_t n, MapDev m, uint32_t* __restrict__ key,
                              uint32_t* __restrict__ idx, const uint32_t* __restrict__ pat, uint32_t nbins,
                              const T* __restrict__ c, uint32_t* __restrict__ cell_out,
                              typename Vec4<T>::type* __restrict__ rec_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t packed;
  typename Vec4<T>::type r;
  key[j] = point_record(x, j, m, pat ? pat[j] * nbins : 0u, c, packed, r);
  idx[j] = (uint32_t)j;
  cell_out[j] = packed;
  rec_out[j] = r;
}

// ---------------------------------------------------------------------------
// Direct binning: one counting pass when all bins of the plan (batch * nbins)
// fit PB_MAXB and the plan is not segmented.
//   k_bin_tiles   per tile of PB_TILE points: the tile's bin counts
//                 (tile-major, cnt[tile * nb + bin], so every later access is
//                 coalesced); nothing per point is stored
//   k_bin_groups  per group of gsz tiles and bin: exclusive prefix of the
//                 counts over the group's tiles, in place; the group total
//                 to gsum[group * nb + bin]
//   k_bin_bases   one workgroup: exclusive prefix of the group totals over
//                 the groups (in place) and of the bin totals over the bins
//                 (the bin starts)
//   k_bin_place   per tile: running offsets start + group prefix + tile
//                 prefix; each wavefront ranks its quarter of the tile
//                 (rounds of 64 points in index order, one ballot per key
//                 bit, a running count per bin in LDS), then every point
//                 goes to offset + earlier wavefronts' counts + rank (two
//                 barriers per tile); each point is mapped again (reading
//                 the coordinates twice costs less than writing and reading
//                 back records in the original order) and its record, cell
//                 and index go to their sorted places
// The result is the stable sort by key (index order inside every bin) that
// the radix sort gives, without its extra key and index arrays.
// ---------------------------------------------------------------------------
constexpr int PB_T = 256;
constexpr int PB_ROUNDS = 4;  // 1024-point tiles: placement 27 -> 24 us at N = 2^20 (2048-point tiles)
constexpr int PB_TILE = PB_T * PB_ROUNDS;
constexpr int PB_MAXB = 2048;
constexpr int PB_GMAX = 32;  // at most this many tile groups

constexpr int PB_CT = 512;  // threads of k_bin_tiles: PB_TILE / PB_CT points each (loads in flight)

template <typename T>
__global__ void __launch_bounds__(PB_CT)
    k_bin_tiles(const double* __restrict__ x, int64_t n, MapDev m, const uint32_t* __restrict__ pat,
                uint32_t nbins, int nb, uint32_t* __restrict__ cnt) {
  constexpr int PER = PB_TILE / PB_CT;
  __shared__ uint32_t h[PB_MAXB];
  const int t = threadIdx.x;
  for (int b = t; b < nb; b += PB_CT) h[b] = 0;
  const int64_t base = (int64_t)blockIdx.x * PB_TILE;
  uint32_t k[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int64_t j = base + r * PB_CT + t;
    uint32_t packed;
    typename Vec4<T>::type rc;
    if (j < n) k[r] = point_record<T>(x, j, m, pat ? pat[j] * nbins : 0u, nullptr, packed, rc);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int64_t j = base + r * PB_CT + t;
    if (j < n) atomicAdd(&h[k[r]], 1u);
  }
  __syncthreads();
  for (int b = t; b < nb; b += PB_CT) cnt[(int64_t)blockIdx.x * nb + b] = h[b];
}

// grid (ceil(nb / PB_T), groups); gsz a multiple of 16
__global__ void __launch_bounds__(PB_T)
    k_bin_groups(uint32_t* __restrict__ cnt, int ntiles, int nb, int gsz, uint32_t* __restrict__ gsum) {
  const int b = blockIdx.x * PB_T + threadIdx.x;
  if (b >= nb) return;
  const int t0 = blockIdx.y * gsz;
  const int t1 = t0 + gsz < ntiles ? t0 + gsz : ntiles;
  uint32_t s = 0;
  for (int c0 = t0; c0 < t1; c0 += 16) {
    uint32_t v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = c0 + i < t1 ? cnt[(int64_t)(c0 + i) * nb + b] : 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (c0 + i < t1) cnt[(int64_t)(c0 + i) * nb + b] = s;
      s += v[i];
    }
  }
  gsum[(int64_t)blockIdx.y * nb + b] = s;
}

constexpr int PB_BT = PB_MAXB / 2;  // threads of k_bin_bases: two bins each

__global__ void __launch_bounds__(PB_BT)
    k_bin_bases(uint32_t* __restrict__ gsum, int ngroups, int nb, int64_t n, int* __restrict__ start) {
  __shared__ uint32_t tot[PB_MAXB];
  __shared__ uint32_t wsum[PB_BT / 64];
  const int t = threadIdx.x;
  for (int b = t; b < PB_MAXB; b += PB_BT) {
    uint32_t acc = 0;
    if (b < nb) {
      for (int g0 = 0; g0 < ngroups; g0 += 16) {
        uint32_t v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = g0 + i < ngroups ? gsum[(g0 + i) * nb + b] : 0u;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          if (g0 + i < ngroups) gsum[(g0 + i) * nb + b] = acc;
          acc += v[i];
        }
      }
    }
    tot[b] = acc;
  }
  __syncthreads();
  // exclusive scan of the bin totals, two consecutive bins per thread:
  // wavefront scan with lane shifts, then the wavefront totals
  const uint32_t a0 = tot[2 * t], a1 = tot[2 * t + 1];
  const int lane = t & 63;
  uint32_t inc = a0 + a1;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(inc, off, 64);
    if (lane >= off) inc += u;
  }
  if (lane == 63) wsum[t >> 6] = inc;
  __syncthreads();
  uint32_t ex = inc - (a0 + a1);
  for (int w = 0; w < (t >> 6); ++w) ex += wsum[w];
  if (2 * t < nb) start[2 * t] = (int)ex;
  if (2 * t + 1 < nb) start[2 * t + 1] = (int)(ex + a0);
  if (t == 0) start[nb] = (int)n;
}

template <typename T>
__global__ void __launch_bounds__(PB_T)
    k_bin_place(const double* __restrict__ x, MapDev m, const uint32_t* __restrict__ pat, uint32_t nbins,
                const T* __restrict__ c, int64_t n, int nb, int dbits,
                const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ gsum, int gsz,
                const int* __restrict__ start, uint32_t* __restrict__ idx, uint32_t* __restrict__ cell_out,
                typename Vec4<T>::type* __restrict__ rec_out) {
  __shared__ uint32_t run[PB_MAXB];
  __shared__ uint32_t wc[PB_T / 64][PB_MAXB];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  // workgroups go round-robin to the 8 XCDs; each XCD takes a contiguous
  // range of tiles, so the runs a bin receives from neighbouring tiles (which
  // share cache lines of the outputs) are written through one L2
  const int nt = gridDim.x, xq = nt >> 3, xr = nt & 7, xc = blockIdx.x & 7;
  const int tile = xc * xq + (xc < xr ? xc : xr) + (blockIdx.x >> 3);
  const int64_t grow = (int64_t)(tile / gsz) * nb, trow = (int64_t)tile * nb;
  for (int b = t; b < nb; b += PB_T) {
    run[b] = (uint32_t)start[b] + gsum[grow + b] + cnt[trow + b];
#pragma unroll
    for (int w = 0; w < PB_T / 64; ++w) wc[w][b] = 0;
  }
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
  // wavefront wv owns points [base, base + 64 PB_ROUNDS) of the tile, in
  // rounds of 64; all of them are mapped up front (the same key, cell and
  // record k_bin_tiles counted: the map has no contractions)
  const int64_t base = (int64_t)tile * PB_TILE + (int64_t)wv * (64 * PB_ROUNDS);
  uint32_t ks[PB_ROUNDS], cls[PB_ROUNDS], rk[PB_ROUNDS];
  typename Vec4<T>::type rcs[PB_ROUNDS];
#pragma unroll
  for (int r = 0; r < PB_ROUNDS; ++r) {
    const int64_t j = base + r * 64 + lane;
    ks[r] = 0;
    cls[r] = 0;
    if (j < n) ks[r] = point_record(x, j, m, pat ? pat[j] * nbins : 0u, c, cls[r], rcs[r]);
  }
  __syncthreads();
  // rank of every point among the wavefront's earlier points of its bin:
  // equal-key lanes by one ballot per key bit, earlier rounds by the
  // wavefront's running count wc[wv][bin] (LDS operations of one wavefront
  // complete in order, so no barrier is needed)
#pragma unroll
  for (int r = 0; r < PB_ROUNDS; ++r) {
    const bool valid = base + r * 64 + lane < n;
    const uint32_t k = ks[r];
    uint64_t same = __ballot(valid);
    for (int b = 0; b < dbits; ++b) {
      const bool bit = (k >> b) & 1u;
      const uint64_t mk = __ballot(bit);
      same &= bit ? mk : ~mk;
    }
    const uint32_t rank = (uint32_t)__popcll(same & lt);
    const uint32_t before = valid ? wc[wv][k] : 0u;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (valid && rank == 0) wc[wv][k] = before + (uint32_t)__popcll(same);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    rk[r] = before + rank;
  }
  __syncthreads();
  // place: the tile's running offset, the earlier wavefronts' counts, the rank
#pragma unroll
  for (int r = 0; r < PB_ROUNDS; ++r) {
    const int64_t j = base + r * 64 + lane;
    if (j < n) {
      const uint32_t k = ks[r];
      uint32_t pos = run[k] + rk[r];
      for (int w = 0; w < wv; ++w) pos += wc[w][k];
      if (pos >= (uint32_t)n) pos = (uint32_t)n - 1u;  // unreachable when the counts hold
      idx[pos] = (uint32_t)j;
      cell_out[pos] = cls[r];
      rec_out[pos] = rcs[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Segmented direct binning: the points are stored pattern by pattern and each
// pattern's own bins fit PB_MAXB.  Tiles of up to PB_TILE points never
// straddle a pattern (tile t: points [tbeg[t], tbeg[t + 1]) of pattern
// tpat[t]), so counts, prefixes and places are per pattern over its local
// bins; the result is the same stable order as the segmented radix sort.
//   k_bin_tiles_seg  per tile: the tile's local bin counts
//                    (cnt[tile * nbins + bin]) from the mapped coordinates
//   k_bin_scan_seg   per pattern: exclusive prefix of each local bin's
//                    counts over the pattern's tiles (in place), then of the
//                    bin totals over the bins: bin_start
//   k_bin_place_seg  per tile: maps the points again, ranks them over
//                    local bins and writes each bin's run from LDS
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(PB_CT)
    k_bin_tiles_seg(const double* __restrict__ x, MapDev m, uint32_t nbins, const int* __restrict__ tbeg,
                    uint32_t* __restrict__ cnt) {
  constexpr int PER = PB_TILE / PB_CT;
  __shared__ uint32_t h[PB_MAXB];
  const int t = threadIdx.x;
  const int nb = (int)nbins;
  for (int b = t; b < nb; b += PB_CT) h[b] = 0;
  const int j0 = tbeg[blockIdx.x], j1 = tbeg[blockIdx.x + 1];
  // local bins (the tile lies inside one pattern): key offset 0
  uint32_t k[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int j = j0 + r * PB_CT + t;
    uint32_t packed;
    typename Vec4<T>::type rc;
    if (j < j1) k[r] = point_record<T>(x, j, m, 0u, nullptr, packed, rc);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int j = j0 + r * PB_CT + t;
    if (j < j1) atomicAdd(&h[k[r]], 1u);
  }
  __syncthreads();
  for (int b = t; b < nb; b += PB_CT) cnt[(int64_t)blockIdx.x * nb + b] = h[b];
}

constexpr int PB_SGT = 1024;  // threads of k_bin_scan_seg: two local bins each at most
static_assert(PB_MAXB <= 2 * PB_SGT, "a pattern's bins are scanned two per thread");

// grid: patterns; ptile[p] .. ptile[p + 1] the pattern's tiles, pstart[p] its
// first point; start[npat * nbins] = n is written by the last pattern
__global__ void __launch_bounds__(PB_SGT)
    k_bin_scan_seg(uint32_t* __restrict__ cnt, const int* __restrict__ ptile, const int* __restrict__ pstart,
                   int nbins, int npat, int64_t n, int* __restrict__ start) {
  __shared__ uint32_t wsum[PB_SGT / 64];
  const int p = blockIdx.x, t = threadIdx.x;
  const int t0 = ptile[p], t1 = ptile[p + 1];
  // the running prefix of local bins b = 2t and 2t + 1 over the tiles, 16
  // loads in flight per bin
  constexpr int U = 16;
  uint32_t tot[2] = {0u, 0u};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int b = 2 * t + q;
    if (b >= nbins) continue;
    uint32_t off = 0;
    for (int tt = t0; tt < t1; tt += U) {
      uint32_t v[U];
#pragma unroll
      for (int i = 0; i < U; ++i) v[i] = tt + i < t1 ? cnt[(int64_t)(tt + i) * nbins + b] : 0u;
#pragma unroll
      for (int i = 0; i < U; ++i) {
        if (tt + i < t1) cnt[(int64_t)(tt + i) * nbins + b] = off;
        off += v[i];
      }
    }
    tot[q] = off;
  }
  // exclusive scan of the bin totals (wavefront scan, then the wavefront
  // totals)
  const int lane = t & 63;
  uint32_t inc = tot[0] + tot[1];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[t >> 6] = inc;
  __syncthreads();
  uint32_t ex = inc - (tot[0] + tot[1]);
  for (int w = 0; w < (t >> 6); ++w) ex += wsum[w];
  const int base = pstart[p];
  if (2 * t < nbins) start[(int64_t)p * nbins + 2 * t] = base + (int)ex;
  if (2 * t + 1 < nbins) start[(int64_t)p * nbins + 2 * t + 1] = base + (int)(ex + tot[0]);
  if (p == npat - 1 && t == 0) start[(int64_t)npat * nbins] = (int)n;
}

// k_bin_place over the tile's local bins, with the tile's records staged in
// LDS in bin order before they go out: consecutive threads then write each
// bin's run (a few records per bin and tile) to consecutive places, instead
// of every lane writing to a different partial line.  Dynamic LDS layout:
// records [PB_TILE], cells, indices and destinations [PB_TILE] each, then
// run, run0, lst [nbins] each and wc [PB_T / 64][nbins].
template <typename R4>
__host__ __device__ constexpr size_t place_seg_lds_bytes(int nbins) {
  return (size_t)PB_TILE * (sizeof(R4) + 3 * 4) + (size_t)(3 + PB_T / 64) * nbins * 4;
}

template <typename T>
__global__ void __launch_bounds__(PB_T)
    k_bin_place_seg(const double* __restrict__ x, MapDev m, const T* __restrict__ c, const int* __restrict__ tbeg,
                    const int* __restrict__ tpat, int nbins, int dbits, const uint32_t* __restrict__ cnt,
                    const int* __restrict__ start, uint32_t* __restrict__ idx, uint32_t* __restrict__ cell_out,
                    typename Vec4<T>::type* __restrict__ rec_out) {
  using R4 = typename Vec4<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  R4* srec = reinterpret_cast<R4*>(smem);
  uint32_t* scl = reinterpret_cast<uint32_t*>(srec + PB_TILE);
  uint32_t* sidx = scl + PB_TILE;
  uint32_t* sdst = sidx + PB_TILE;
  uint32_t* run = sdst + PB_TILE;  // [nbins] next place of each bin
  uint32_t* run0 = run + nbins;    // [nbins] the tile's first place of each bin
  uint32_t* lst = run0 + nbins;    // [nbins] the bin's first slot in the staged tile
  uint32_t* wc = lst + nbins;      // [PB_T / 64][nbins] per-wavefront counts of a round
  __shared__ uint32_t wsum[PB_T / 64];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  // XCD-contiguous tile ranges (as k_bin_place)
  const int nt = gridDim.x, xq = nt >> 3, xr = nt & 7, xc = blockIdx.x & 7;
  const int tile = xc * xq + (xc < xr ? xc : xr) + (blockIdx.x >> 3);
  const int64_t trow = (int64_t)tile * nbins;
  const int64_t sbase = (int64_t)tpat[tile] * nbins;
  for (int b = t; b < nbins; b += PB_T) {
    const uint32_t r0 = (uint32_t)start[sbase + b] + cnt[trow + b];
    run[b] = r0;
    run0[b] = r0;
#pragma unroll
    for (int w = 0; w < PB_T / 64; ++w) wc[w * nbins + b] = 0;
  }
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
  const int j0 = tbeg[tile], j1 = tbeg[tile + 1];
  uint32_t ks[PB_ROUNDS], cls[PB_ROUNDS], dst[PB_ROUNDS];
  R4 rcs[PB_ROUNDS];
#pragma unroll
  for (int r = 0; r < PB_ROUNDS; ++r) {
    const int j = j0 + r * PB_T + t;
    ks[r] = 0;
    cls[r] = 0;
    // mapped again: the local bin k_bin_tiles_seg counted (the map has no
    // contractions), the record and the packed cell
    if (j < j1) ks[r] = point_record(x, j, m, 0u, c, cls[r], rcs[r]);
  }
  // each point's place: rounds of 256 points in index order, ranks among
  // equal bins by ballots
#pragma unroll
  for (int r = 0; r < PB_ROUNDS; ++r) {
    const int j = j0 + r * PB_T + t;
    const bool valid = j < j1;
    const uint32_t k = ks[r];
    uint64_t same = __ballot(valid);
    for (int b = 0; b < dbits; ++b) {
      const bool bit = (k >> b) & 1u;
      const uint64_t mk = __ballot(bit);
      same &= bit ? mk : ~mk;
    }
    const int rank = __popcll(same & lt);
    if (valid && rank == 0) wc[wv * nbins + k] = (uint32_t)__popcll(same);
    __syncthreads();
    dst[r] = 0;
    if (valid) {
      uint32_t pos = run[k] + (uint32_t)rank;
      for (int w = 0; w < wv; ++w) pos += wc[w * nbins + k];
      dst[r] = pos;
    }
    __syncthreads();
    if (valid && rank == 0) {
      atomicAdd(&run[k], wc[wv * nbins + k]);
      wc[wv * nbins + k] = 0;
    }
    __syncthreads();
  }
  // the tile's bin counts (run - run0), their exclusive prefix over the
  // bins: thread t takes bins [t * per, (t + 1) * per)
  const int per = (nbins + PB_T - 1) / PB_T;
  uint32_t loc = 0;
  for (int b = t * per; b < min(nbins, (t + 1) * per); ++b) loc += run[b] - run0[b];
  uint32_t inc = loc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  uint32_t ex = inc - loc;
  for (int w = 0; w < wv; ++w) ex += wsum[w];
  for (int b = t * per; b < min(nbins, (t + 1) * per); ++b) {
    lst[b] = ex;
    ex += run[b] - run0[b];
  }
  __syncthreads();
  // stage in bin order, then write the runs with consecutive threads
#pragma unroll
  for (int r = 0; r < PB_ROUNDS; ++r) {
    const int j = j0 + r * PB_T + t;
    if (j < j1) {
      const uint32_t k = ks[r];
      const uint32_t s2 = lst[k] + (dst[r] - run0[k]);
      srec[s2] = rcs[r];
      scl[s2] = cls[r];
      sidx[s2] = (uint32_t)j;
      sdst[s2] = dst[r];
    }
  }
  __syncthreads();
  const int np = j1 - j0;
  for (int q = t; q < np; q += PB_T) {
    const uint32_t pos = sdst[q];
    idx[pos] = sidx[q];
    cell_out[pos] = scl[q];
    rec_out[pos] = srec[q];
  }
}

template <typename K>
__global__ void k_bin_start(const K* __restrict__ keys, int64_t n, int nbins, int* __restrict__ start) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int)keys[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[b] = (int)lo;
}

static MapDev make_map_dev(const GridGeom& g, const EsParams& es, const CoordMap& cm) {
  MapDev m;
  m.d = g.d;
  m.wrap = cm.wrap;
  m.w = es.w;
  for (int a = 0; a < 3; ++a) {
    m.L[a] = cm.L[a];
    m.shift[a] = cm.shift[a];
    m.scale[a] = cm.scale[a];
    m.nf[a] = g.nf[a];
    m.tile[a] = g.tile[a];
    m.ntile[a] = g.ntile[a];
    m.tshift[a] = -1;
    m.itile[a] = 1.0 / g.tile[a];
    for (int k = 0; k < 30; ++k)
      if (g.tile[a] == (1 << k)) m.tshift[a] = k;
  }
  m.cell_bits = g.cell_bits;
  return m;
}

template <typename T>
void SortedPoints<T>::init(const GridGeom& g, const EsParams& e, int64_t npts) {
  if (npts <= 0) throw InvalidArgument("number of points must be positive");
  if (npts >= (1LL << 31)) throw InvalidArgument("too many points (limit 2^31 - 1)");
  geom = g;
  es = e;
  n = npts;
  const int64_t tb = (int64_t)g.nbins * g.batch;
  idx.allocate(n);
  bin_start.allocate(tb + 1);
  scell.allocate(n);
  rec.allocate(n);
  // FRESCO_DIRECT_BINNING=0: always the radix sort (the same order)
  // interpolation layout: FRESCO_INTERP_LPP=8 (eight lanes per point instead
  // of four), FRESCO_INTERP_WGS (target workgroup count, default 2048),
  // FRESCO_INTERP_UNITS=0 (row-wise region staging instead of 16-byte units)
  {
    const char* e = std::getenv("FRESCO_INTERP_LPP");
    interp_lpp = e && std::atoi(e) == 8 ? 8 : 4;
    e = std::getenv("FRESCO_INTERP_WGS");
    interp_wgs = e ? std::max(1, std::atoi(e)) : 2048;
    e = std::getenv("FRESCO_INTERP_UNITS");
    interp_units = !e || std::atoi(e) != 0;
  }
  const char* env = std::getenv("FRESCO_DIRECT_BINNING");
  direct = tb <= PB_MAXB && !(env && env[0] == '0');
  if (direct) {
    ntiles_ = (int)((n + PB_TILE - 1) / PB_TILE);
    gsz_ = 16 * (int)((ntiles_ + 16 * PB_GMAX - 1) / (16 * PB_GMAX));
    ngroups_ = (ntiles_ + gsz_ - 1) / gsz_;
    tcnt_.allocate((size_t)ntiles_ * tb);
    gsum_.allocate((size_t)ngroups_ * tb);
  } else {
    init_sorter();
  }
}

template <typename T>
void SortedPoints<T>::init_sorter() {
  direct = false;
  // keys and the records in the original order: the radix sort's input
  // (the direct passes map the points twice instead)
  scell_in.allocate(n);
  rec_in.allocate(n);
  key.allocate(n);
  idx_in.allocate(n);
  key_sorted.allocate(n);
  sorter.init(n, geom.key_bits);
}

template <typename T>
void SortedPoints<T>::set_segments(const std::vector<int64_t>& starts) {
  if (starts.size() < 2 || starts.front() != 0 || starts.back() != n)
    throw InvalidArgument("point segments must cover the points");
  if ((int64_t)starts.size() - 1 != geom.batch) throw InvalidArgument("one point segment per pattern");
  // FRESCO_DIRECT_BINNING=0 keeps the segmented radix sort (the same order)
  const char* env = std::getenv("FRESCO_DIRECT_BINNING");
  if (geom.nbins <= PB_MAXB && !(env && env[0] == '0')) {
    std::vector<int> tb, tp, pt, ps;
    for (size_t p = 0; p + 1 < starts.size(); ++p) {
      if (starts[p + 1] < starts[p]) throw InvalidArgument("point segments must be ordered");
      pt.push_back((int)tp.size());
      ps.push_back((int)starts[p]);
      for (int64_t b = starts[p]; b < starts[p + 1]; b += PB_TILE) {
        tb.push_back((int)b);
        tp.push_back((int)p);
      }
    }
    pt.push_back((int)tp.size());
    ps.push_back((int)n);
    tb.push_back((int)n);
    direct = false;
    seg_direct_ = true;
    seg_ntiles_ = (int)tp.size();
    seg_tbeg_.allocate(tb.size());
    seg_tbeg_.upload(tb.data(), tb.size());
    seg_tpat_.allocate(std::max<size_t>(1, tp.size()));
    if (!tp.empty()) seg_tpat_.upload(tp.data(), tp.size());
    seg_ptile_.allocate(pt.size());
    seg_ptile_.upload(pt.data(), pt.size());
    seg_pstart_.allocate(ps.size());
    seg_pstart_.upload(ps.data(), ps.size());
    tcnt_.allocate(std::max<size_t>(1, (size_t)seg_ntiles_ * geom.nbins));
    return;
  }
  seg_direct_ = false;
  if (direct) init_sorter();
  sorter.set_segments(starts, geom.nbins);
}

template <typename T>
void SortedPoints<T>::set(const double* x, const CoordMap& cm, const T* c, hipStream_t s, const uint32_t* pat) {
  MapDev m = make_map_dev(geom, es, cm);
  const unsigned thr = 256;
  if ((geom.batch > 1) != (pat != nullptr)) throw InvalidArgument("pattern indices needed exactly for batched plans");
  const int tb = geom.nbins * geom.batch;
  if (direct) {
    int dbits = 1;
    while ((1 << dbits) < tb) ++dbits;
    hipLaunchKernelGGL(k_bin_tiles<T>, dim3(ntiles_), dim3(PB_CT), 0, s, x, n, m, pat, (uint32_t)geom.nbins, tb,
                       tcnt_.get());
    hipLaunchKernelGGL(k_bin_groups, dim3((tb + PB_T - 1) / PB_T, ngroups_), dim3(PB_T), 0, s, tcnt_.get(), ntiles_,
                       tb, gsz_, gsum_.get());
    hipLaunchKernelGGL(k_bin_bases, dim3(1), dim3(PB_BT), 0, s, gsum_.get(), ngroups_, tb, n, bin_start.get());
    hipLaunchKernelGGL(k_bin_place<T>, dim3(ntiles_), dim3(PB_T), 0, s, x, m, pat, (uint32_t)geom.nbins, c, n, tb,
                       dbits, tcnt_.get(), gsum_.get(), gsz_, bin_start.get(), idx.get(), scell.get(), rec.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    return;
  }
  if (seg_direct_) {
    int dbits = 1;
    while ((1 << dbits) < geom.nbins) ++dbits;
    hipLaunchKernelGGL(k_bin_tiles_seg<T>, dim3(seg_ntiles_), dim3(PB_CT), 0, s, x, m, (uint32_t)geom.nbins,
                       seg_tbeg_.get(), tcnt_.get());
    hipLaunchKernelGGL(k_bin_scan_seg, dim3(geom.batch), dim3(PB_SGT), 0, s, tcnt_.get(), seg_ptile_.get(),
                       seg_pstart_.get(), geom.nbins, geom.batch, n, bin_start.get());
    const size_t lds = place_seg_lds_bytes<typename Vec4<T>::type>(geom.nbins);
    hipLaunchKernelGGL(k_bin_place_seg<T>, dim3(seg_ntiles_), dim3(PB_T), lds, s, x, m, c, seg_tbeg_.get(), seg_tpat_.get(), geom.nbins, dbits, tcnt_.get(), bin_start.get(),
                       idx.get(), scell.get(), rec.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    return;
  }
  // records computed once, in the original (coalesced) order, and carried to
  // the sorted order by the sort's last pass
  hipLaunchKernelGGL(k_bin_records<T>, dim3(blocks_for(n, thr)), dim3(thr), 0, s, x, n, m, key.get(), idx_in.get(),
                     pat, (uint32_t)geom.nbins, c, scell_in.get(), rec_in.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  SortPayload pl;
  pl.rec_in = rec_in.get();
  pl.rec_out = rec.get();
  pl.rec_words = (int)(sizeof(typename Vec4<T>::type) / 4);
  pl.cell_in = scell_in.get();
  pl.cell_out = scell.get();
  sorter.sort(key.get(), idx_in.get(), key_sorted.get(), idx.get(), s, &pl);
  hipLaunchKernelGGL(k_bin_start<uint32_t>, dim3(blocks_for(tb + 1, thr)), dim3(thr), 0, s, key_sorted.get(), n, tb,
                     bin_start.get());
  FRESCO_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// wavefront primitives
// ---------------------------------------------------------------------------

__device__ __forceinline__ float lane_read(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_read(double v, int lane) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ unsigned lane_read_u(unsigned v, int lane) {
  return (unsigned)__builtin_amdgcn_readlane((int)v, lane);
}

// v from lane src (any lane pattern): ds_bpermute
__device__ __forceinline__ float lane_get(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_get(double v, int src) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b & 0xffffffffLL));
  int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// DPP move (out-of-range lanes read 0)
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wavefront; the result is valid in lane 63.
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_shr:4, row_shr:8,
// row_bcast:15, row_bcast:31 (GFX9 DPP).
template <typename T>
__device__ __forceinline__ T wave_sum63(T v) {
  v += dpp_f<0xb1>(v);
  v += dpp_f<0x4e>(v);
  v += dpp_f<0x114>(v);
  v += dpp_f<0x118>(v);
  v += dpp_f<0x142>(v);
  v += dpp_f<0x143>(v);
  return v;
}

// out[j] = value of lane (lane & ~7) + j, j < 8 (ds_swizzle bit mode:
// and_mask 0x18, or_mask j)
template <int J>
__device__ __forceinline__ int swz8(int v) {
  return __builtin_amdgcn_ds_swizzle(v, 0x18 | (J << 5));
}
template <typename T>
__device__ __forceinline__ void group8_bcast(T v, T* out);
template <>
__device__ __forceinline__ void group8_bcast<float>(float v, float* out) {
  const int b = __float_as_int(v);
  out[0] = __int_as_float(swz8<0>(b));
  out[1] = __int_as_float(swz8<1>(b));
  out[2] = __int_as_float(swz8<2>(b));
  out[3] = __int_as_float(swz8<3>(b));
  out[4] = __int_as_float(swz8<4>(b));
  out[5] = __int_as_float(swz8<5>(b));
  out[6] = __int_as_float(swz8<6>(b));
  out[7] = __int_as_float(swz8<7>(b));
}
template <>
__device__ __forceinline__ void group8_bcast<double>(double v, double* out) {
  const long long bb = __double_as_longlong(v);
  const int lo = (int)(bb & 0xffffffffLL), hi = (int)(bb >> 32);
  int l[8], h[8];
  l[0] = swz8<0>(lo); h[0] = swz8<0>(hi);
  l[1] = swz8<1>(lo); h[1] = swz8<1>(hi);
  l[2] = swz8<2>(lo); h[2] = swz8<2>(hi);
  l[3] = swz8<3>(lo); h[3] = swz8<3>(hi);
  l[4] = swz8<4>(lo); h[4] = swz8<4>(hi);
  l[5] = swz8<5>(lo); h[5] = swz8<5>(hi);
  l[6] = swz8<6>(lo); h[6] = swz8<6>(hi);
  l[7] = swz8<7>(lo); h[7] = swz8<7>(hi);
  for (int j = 0; j < 8; ++j) out[j] = __longlong_as_double(((long long)h[j] << 32) | (unsigned int)l[j]);
}

// ---------------------------------------------------------------------------
// spreading: one wavefront owns one tile (LDS), owner-computes
// ---------------------------------------------------------------------------

// halo tile of the input-driven 3D spreader (k_spread3_halo): 16 x TY x 8
// bins (TY = 16, or 8 for the FP32 grids) plus the one-sided halo of a kernel
// up to 8 wide
constexpr int HW = 8;
constexpr int HS_X = 24;  // row stride: (ja, jb) lane pattern conflict-free for 4- and 8-byte cells
constexpr int HP_Z = 8 + HW - 1;
// window tile of the ring spreader (k_spread3_halo_ring): [15 rows][16
// planes][16 columns] (plane 15 zero), two per bin slice
constexpr int HR_WT = 15 * 16 * 16;
template <int TY>
struct HaloShape {
  static constexpr int HP_Y = TY + HW - 1;
  // plane stride in rows: HS_X * HS_Y cells is a multiple of 64 words, so
  // the 8 z-cells of a lane pair up into read2st64 / write2st64 instructions
  static constexpr int HS_Y = TY == 16 ? 24 : 16;
  static constexpr int ELEMS = HS_X * HS_Y * HP_Z;
};

template <typename T>
struct SpreadDev {
  int nf[3], tile[3], ntile[3];
  int cbits;
  int w;
  int sx;        // LDS row stride (tile[0] + pad)
  int maxc[3];   // candidate bins per dimension (table width)
  int cstride;   // table entries per dimension block
  EsDev<T> k;
};

template <int WMAX>
struct Lanes {
  static constexpr int SH = WMAX == 8 ? 3 : 4;
};

// The wavefront walks the candidate bins of its tile (the bins whose points
// can reach the tile) in a fixed order, 64 candidates at a time; the ones
// whose support touches the tile are processed one after another.  For each
// point, lanes 0..D*WMAX-1 evaluate the D*w kernel values once; lane
// (ja, jb) then takes w_x[ja] w_y[jb] c by ds_bpermute and adds it, times
// w_z[jc] (broadcast by readlane), into LDS with a non-returning LDS add.
// One wavefront per tile and in-order LDS instructions make the plain
// read-modify-write safe and the summation order of every cell fixed
// (bitwise reproducible); each fine-grid cell is written to HBM exactly once,
// no atomics.  Used for 1D / 2D and for 3D grids too small for the
// register-tile spreader.
// NC = 1: real strengths (the record's weight); NC = 2: complex strengths
// str[2 idx[j] + {0,1}] (original point order) times the record's weight,
// accumulated on an interleaved complex grid.
template <typename T, int D, int WMAX, int NC = 1>
__global__ void __launch_bounds__(64)
    k_spread(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
             const int* __restrict__ bin_start, const int* __restrict__ cand, SpreadDev<T> a,
             T* __restrict__ grid, const T* __restrict__ str = nullptr, const uint32_t* __restrict__ sidx = nullptr) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tileb = reinterpret_cast<T*>(smem_raw);
  const int lane = threadIdx.x;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  constexpr int SH = Lanes<WMAX>::SH;
  constexpr int NCH = D == 1 ? 1 : (WMAX * WMAX + 63) / 64;

  // batched plans: pattern blockIdx.y has its own bins and grid; sliced
  // launches: slice blockIdx.z takes the z-th of gridDim.z contiguous parts
  // of each candidate bin's points and writes its own partial grid
  bin_start += (int64_t)blockIdx.y * a.ntile[0] * a.ntile[1] * a.ntile[2];
  grid += ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * a.nf[0] * a.nf[1] * a.nf[2] * NC;
  const int b = blockIdx.x;
  int t[3];
  t[0] = b % a.ntile[0];
  t[1] = (b / a.ntile[0]) % a.ntile[1];
  t[2] = b / (a.ntile[0] * a.ntile[1]);
  int t0[3], tl[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    t0[q] = t[q] * a.tile[q];
    tl[q] = min(a.tile[q], a.nf[q] - t0[q]);
  }
  const int ty = a.tile[1];
  const int zstride = ty * a.sx;
  const int tile_elems = a.sx * (D >= 2 ? a.tile[1] : 1) * (D >= 3 ? a.tile[2] : 1);
  for (int e = lane; e < tile_elems * NC; e += 64) tileb[e] = T(0);
  __syncthreads();

  // kernel-value lane of this thread: dim kd, index kj
  const int kd = lane >> SH;
  const int kj = lane & (WMAX - 1);

  const int nc0 = a.maxc[0];
  const int nc1 = D >= 2 ? a.maxc[1] : 1;
  const int nc2 = D >= 3 ? a.maxc[2] : 1;
  for (int c2 = 0; c2 < nc2; ++c2) {
    const int bz = D >= 3 ? cand[2 * a.cstride + t[2] * a.maxc[2] + c2] : 0;
    if (bz < 0) continue;
    for (int c1 = 0; c1 < nc1; ++c1) {
      const int by = D >= 2 ? cand[1 * a.cstride + t[1] * a.maxc[1] + c1] : 0;
      if (by < 0) continue;
      for (int c0 = 0; c0 < nc0; ++c0) {
        const int bx = cand[t[0] * a.maxc[0] + c0];
        if (bx < 0) continue;
        const int bin = bx + a.ntile[0] * (by + a.ntile[1] * bz);
        const int beg0 = bin_start[bin];
        const int len = bin_start[bin + 1] - beg0;
        const int beg = beg0 + (int)(((int64_t)len * blockIdx.z) / gridDim.z);
        const int end = beg0 + (int)(((int64_t)len * (blockIdx.z + 1)) / gridDim.z);
        for (int base = beg; base < end; base += 64) {
          const int j = base + lane;
          uint32_t sc = 0;
          typename Vec4<T>::type r;
          r.x = r.y = r.z = r.w = T(0);
          T sre = T(0), sim = T(0);
          bool touch = false;
          if (j < end) {
            sc = scell[j];
            r = rec[j];
            if (NC == 2) {
              const uint32_t o = sidx[j];
              sre = str[2 * (size_t)o] * r.w;
              sim = str[2 * (size_t)o + 1] * r.w;
            }
            touch = true;
#pragma unroll
            for (int q = 0; q < D; ++q) {
              int s = (int)((sc >> (q * a.cbits)) & cmask);
              int u = s - t0[q];
              if (u < 0) u += a.nf[q];
              touch = touch && (u < tl[q] || u + w > a.nf[q]);
            }
          }
          unsigned long long m = __ballot(touch);
          while (m) {
            const int p = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t scp = lane_read_u(sc, p);
            const T cw = NC == 2 ? T(1) : lane_read(r.w, p);
            const T cre = NC == 2 ? lane_read(sre, p) : T(0);
            const T cim = NC == 2 ? lane_read(sim, p) : T(0);
            const T offx = lane_read(r.x, p);
            const T offy = D >= 2 ? lane_read(r.y, p) : T(0);
            const T offz = D >= 3 ? lane_read(r.z, p) : T(0);
            // kernel values: lane (kd, kj) holds phi((kj - off_kd) * 2/w)
            const T offk = kd == 0 ? offx : (kd == 1 ? offy : offz);
            const T kv = (kd < D && kj < w) ? es_phi<T>((T(kj) - offk) * a.k.two_over_w, a.k) : T(0);
            int s[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) s[q] = q < D ? (int)((scp >> (q * a.cbits)) & cmask) : 0;
            int uz0 = 0;
            if (D == 3) {
              uz0 = s[2] - t0[2];
              if (uz0 < 0) uz0 += a.nf[2];
            }
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
              const int pr = lane + 64 * ch;
              const int ja = D == 1 ? pr : (pr & (WMAX - 1));
              const int jb = D == 1 ? 0 : (pr >> SH);
              bool act = ja < w && (D == 1 || jb < w);
              int ux = s[0] + ja - t0[0];
              if (ux >= a.nf[0]) ux -= a.nf[0];
              if (ux < 0) ux += a.nf[0];
              act = act && ux < tl[0];
              int uy = 0;
              if (D >= 2) {
                uy = s[1] + jb - t0[1];
                if (uy >= a.nf[1]) uy -= a.nf[1];
                if (uy < 0) uy += a.nf[1];
                act = act && uy < tl[1];
              }
              const T wx = lane_get(kv, ja < WMAX ? ja : 0);
              T v = wx * cw;
              if (D >= 2) v *= lane_get(kv, WMAX + (jb < WMAX ? jb : 0));
              T* rowp = tileb + (uy * a.sx + ux) * NC;
              if (D == 3) {
#pragma unroll
                for (int jc = 0; jc < WMAX; ++jc) {
                  if (jc < w) {
                    int uz = uz0 + jc;
                    if (uz >= a.nf[2]) uz -= a.nf[2];
                    const T wz = lane_read(kv, 2 * WMAX + jc);
                    if (uz < tl[2] && act) {
                      if (NC == 1) {
                        rowp[uz * zstride] += v * wz;
                      } else {
                        rowp[uz * zstride * 2] += v * wz * cre;
                        rowp[uz * zstride * 2 + 1] += v * wz * cim;
                      }
                    }
                  }
                }
              } else if (act) {
                if (NC == 1) {
                  rowp[0] += v;
                } else {
                  rowp[0] += v * cre;
                  rowp[1] += v * cim;
                }
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  const int nx = tl[0];
  const int ny = D >= 2 ? tl[1] : 1;
  const int nz = D >= 3 ? tl[2] : 1;
  const int tot = nx * ny * nz;
  for (int e = lane; e < tot; e += 64) {
    int ux = e % nx;
    int rr = e / nx;
    int uy = rr % ny;
    int uz = rr / ny;
    int64_t gidx = ((int64_t)(t0[2] + uz) * a.nf[1] + (t0[1] + uy)) * a.nf[0] + (t0[0] + ux);
    if (NC == 1) {
      grid[gidx] = tileb[(uz * ty + uy) * a.sx + ux];
    } else {
      grid[2 * gidx] = tileb[2 * ((uz * ty + uy) * a.sx + ux)];
      grid[2 * gidx + 1] = tileb[2 * ((uz * ty + uy) * a.sx + ux) + 1];
    }
  }
}

// (register-tile 3D spreader is defined after the interpolation structs)

template <typename T>
void Spreader<T>::init(const GridGeom& g, const EsParams& e) {
  geom = g;
  es = e;
  if (e.w > 16) throw InvalidArgument("kernel width > 16");
  if (g.d >= 1 && e.w > 8 && g.d == 1 && e.w > 64) throw InvalidArgument("kernel width too large");
  // candidate bins per tile and dimension: bins holding a first cell s with
  // [s, s+w) (periodic) intersecting the tile.
  std::vector<std::vector<std::vector<int>>> lists(3);
  int maxtile = std::max(g.ntile[0], std::max(g.ntile[1], g.ntile[2]));
  for (int a = 0; a < 3; ++a) {
    maxc[a] = 1;
    lists[a].resize(g.ntile[a]);
    for (int t = 0; t < g.ntile[a]; ++t) {
      std::vector<int>& L = lists[a][t];
      if (a >= g.d) {
        L.push_back(0);
        continue;
      }
      int t0 = t * g.tile[a];
      int tl = std::min(g.tile[a], g.nf[a] - t0);
      for (int s = t0 - e.w + 1; s <= t0 + tl - 1; ++s) {
        int sm = ((s % g.nf[a]) + g.nf[a]) % g.nf[a];
        int bb = std::min(sm / g.tile[a], g.ntile[a] - 1);
        if (std::find(L.begin(), L.end(), bb) == L.end()) L.push_back(bb);
      }
      maxc[a] = std::max<int>(maxc[a], (int)L.size());
    }
  }
  cand_stride = maxtile * std::max(maxc[0], std::max(maxc[1], maxc[2]));
  std::vector<int> host(3 * cand_stride, -1);
  for (int a = 0; a < 3; ++a)
    for (int t = 0; t < g.ntile[a]; ++t)
      for (size_t i = 0; i < lists[a][t].size(); ++i) host[a * cand_stride + t * maxc[a] + i] = lists[a][t][i];
  cand.allocate(host.size());
  cand.upload(host.data(), host.size());
  lds_bytes = (int)spread_lds_bytes(g, (int)sizeof(T), e.w);
  if (lds_bytes > 160 * 1024) throw InvalidArgument("spreader tile exceeds LDS");
  // register-tile spreader on a grid with few tiles: slice each tile's
  // points over several wavefronts (about 8 per CU in total)
  parts = 1;
  if (g.d == 3 && g.tile[0] == 16 && g.tile[1] == 16 && g.tile[2] == 8 && g.nf[1] >= 16 + e.w - 1)
    parts = std::max(1, std::min(16, (8192 + g.nbins - 1) / g.nbins));
  // LDS-tile spreader (1D / 2D) on few tiles: slice every tile's points
  // over several wavefronts the same way
  lparts = 1;
  if (g.d <= 2 && (int64_t)g.nbins * g.batch < 1024)
    lparts = (int)std::max<int64_t>(1, std::min<int64_t>(16, (2048 + (int64_t)g.nbins * g.batch - 1) /
                                                                ((int64_t)g.nbins * g.batch)));
  const size_t pb = std::max((size_t)parts * (size_t)g.ncells(), lparts > 1 ? (size_t)lparts * g.ncells() * g.batch : 0);
  if (pb > (size_t)g.ncells() || parts > 1) part_buf.allocate(pb);
  // input-driven halo spreader: 3D, w <= 8, 16 x 16 x 8 bins, a halo tile
  // no longer than the grid in any dimension
  halo_ok = g.d == 3 && e.w <= 8 && g.tile[0] == 16 && (g.tile[1] == 16 || g.tile[1] == 8) && g.tile[2] == 8 &&
            g.nf[0] >= 23 && g.nf[1] >= g.tile[1] + 7 && g.nf[2] >= 15;
  // 2D input-driven halo spreader (real strengths, w <= 8): rows of
  // tile + 7 cells, stride a multiple of 8 with an odd factor
  halo2_ok = g.d == 2 && e.w <= 8 && g.tile[0] + 7 <= g.nf[0] && g.tile[1] + 7 <= g.nf[1] && g.tile[0] + 7 < 256 &&
             g.tile[1] + 7 < 256;
  if (halo2_ok) {
    int hsx = ((g.tile[0] + 7 + 7) / 8) * 8;
    if ((hsx / 8) % 2 == 0) hsx += 8;
    halo2_hs = hsx;
    halo2_rows = g.tile[1] + 7;
    std::vector<int> tabs;
    for (int a = 0; a < 2; ++a) {
      std::vector<int> t(4 * (size_t)g.nf[a], -1);
      const int reach = g.tile[a] + 7;
      for (int b = 0; b < g.ntile[a]; ++b) {
        const int t0 = b * g.tile[a];
        const int tl = std::min(g.tile[a], g.nf[a] - t0);
        const int ulim = std::min(reach, tl + e.w - 1);
        for (int u = 0; u < ulim; ++u) {
          const int c = (t0 + u) % g.nf[a];
          int k = 0;
          while (k < 4 && t[4 * c + k] >= 0) ++k;
          if (k == 4) throw RuntimeFailure("halo merge table overflow");
          t[4 * c + k] = (b << 8) | u;
        }
      }
      tabs.insert(tabs.end(), t.begin(), t.end());
    }
    halo_tab.allocate(tabs.size());
    halo_tab.upload(tabs.data(), tabs.size());
    const int64_t wg = (int64_t)g.nbins * g.batch;
    halo2_parts = (int)std::max<int64_t>(1, std::min<int64_t>(16, (2048 + wg - 1) / wg));
    // FP32 strengths on 16 x 16 bins: the matrix-core spreader
    // (FRESCO_SPREAD2=halo: the LDS one)
    const char* env2 = std::getenv("FRESCO_SPREAD2");
    halo2_mfma = sizeof(T) == 4 && g.tile[0] == 16 && g.tile[1] == 16 && halo2_hs == 24 &&
                 !(env2 && std::string(env2) == "halo");
    halo_buf.allocate((size_t)wg * halo2_parts * (size_t)halo2_hs * halo2_rows);
  }
  if (halo_ok) {
    std::vector<int> tabs;
    for (int a = 0; a < 3; ++a) {
      std::vector<int> t(4 * (size_t)g.nf[a], -1);
      const int reach[3] = {g.tile[0] + 7, g.tile[1] + 7, g.tile[2] + 7};
      for (int b = 0; b < g.ntile[a]; ++b) {
        const int t0 = b * g.tile[a];
        const int tl = std::min(g.tile[a], g.nf[a] - t0);
        const int ulim = std::min(reach[a], tl + e.w - 1);
        for (int u = 0; u < ulim; ++u) {
          const int c = (t0 + u) % g.nf[a];
          int k = 0;
          while (k < 4 && t[4 * c + k] >= 0) ++k;
          if (k == 4) throw RuntimeFailure("halo merge table overflow");
          t[4 * c + k] = (b << 6) | u;
        }
      }
      tabs.insert(tabs.end(), t.begin(), t.end());
    }
    {
      // x windows of the ring spreader: window 2 b + h covers columns
      // [16 b + 8 h, 16 b + 8 h + 16) and holds the points whose first x cell
      // is in its first 8 columns
      std::vector<int> t(4 * (size_t)g.nf[0], -1);
      for (int b = 0; b < g.ntile[0]; ++b) {
        const int tl = std::min(g.tile[0], g.nf[0] - b * g.tile[0]);
        for (int hw = 0; hw < 2; ++hw) {
          const int first = std::min(8, tl - 8 * hw);
          if (first <= 0) continue;
          const int ulim = std::min(15, first + e.w - 1);
          for (int u = 0; u < ulim; ++u) {
            const int c = (b * g.tile[0] + 8 * hw + u) % g.nf[0];
            int k = 0;
            while (k < 4 && t[4 * c + k] >= 0) ++k;
            if (k == 4) throw RuntimeFailure("halo merge table overflow");
            t[4 * c + k] = ((2 * b + hw) << 6) | u;
          }
        }
      }
      tabs.insert(tabs.end(), t.begin(), t.end());
    }
    halo_tab.allocate(tabs.size());
    halo_tab.upload(tabs.data(), tabs.size());
    // slices per bin: more slices, more wavefronts in flight but more halo
    // tiles to write and merge; the autotuner picks among these
    max_halo_parts = std::max(1, std::min(16, (16384 + g.nbins - 1) / g.nbins));
    const int helems = g.tile[1] == 8 ? HaloShape<8>::ELEMS : HaloShape<16>::ELEMS;
    // the ring spreader writes 2 window tiles per slice, at least one slice
    const size_t per_bin =
        std::max((size_t)max_halo_parts * helems, (size_t)std::max(2, max_halo_parts & ~1) * HR_WT);
    halo_buf.allocate((size_t)g.nbins * per_bin);
    mode = HALO;
    halo_parts = std::min(4, max_halo_parts);
    tuned = false;
    // FRESCO_SPREAD3=reg | halo[:<slices>] | halo4[:<slices>] | halom[:<parts>] |
    // halor[:<parts>] pins the choice (reproducible rounding across processes)
    // instead of timing the candidates; halom and halor (FP32 strengths,
    // 16 x 8 x 8 bins; parts = 2 x windows per slice) fall back to halo4
    // otherwise
    if (const char* env = std::getenv("FRESCO_SPREAD3")) {
      const std::string v(env);
      if (v == "reg") {
        mode = REG;
      } else if (v.rfind("halom", 0) == 0 || v.rfind("halor", 0) == 0) {
        // two x windows per slice, so parts is even
        const bool mm = sizeof(T) == 4 && g.tile[0] == 16 && g.tile[1] == 8 && g.tile[2] == 8;
        mode = !mm ? HALO4 : v[4] == 'm' ? HALOM : HALOR;
        if (v.size() > 6) halo_parts = std::max(1, std::min(max_halo_parts, std::atoi(v.c_str() + 6)));
        if (mode != HALO4) halo_parts = std::max(2, std::min(max_halo_parts & ~1, (halo_parts + 1) & ~1));
      } else if (v.rfind("halo4", 0) == 0) {
        mode = HALO4;
        if (v.size() > 6) halo_parts = std::max(1, std::min(max_halo_parts, std::atoi(v.c_str() + 6)));
      } else if (v.rfind("halo", 0) == 0) {
        mode = HALO;
        if (v.size() > 5) halo_parts = std::max(1, std::min(max_halo_parts, std::atoi(v.c_str() + 5)));
      }
      tuned = true;
      tune_source = 3;
    }
  }
}

// ---------------------------------------------------------------------------
// autotuning cache: the 3D spreader picked for a plan shape, so that separate
// processes spread with the same kernel (and round identically).  Entries
// are lines "<key> <mode> <slices>"; the first line of a key wins.  Lookup:
// the user cache ($FRESCO_AUTOTUNE_CACHE, else
// $XDG_CACHE_HOME/fresco_amd/spread3.txt or ~/.cache/fresco_amd/spread3.txt;
// FRESCO_AUTOTUNE_CACHE=off disables it), then the table shipped beside the
// library (tuning/spread3_mi355x.txt).  New picks go to the user cache.
// ---------------------------------------------------------------------------

namespace {
std::mutex g_tune_mu;

std::string user_tune_path() {
  if (const char* e = std::getenv("FRESCO_AUTOTUNE_CACHE")) {
    const std::string v(e);
    if (v.empty() || v == "off" || v == "0") return "";
    return v;
  }
  std::string dir;
  if (const char* x = std::getenv("XDG_CACHE_HOME")) dir = std::string(x);
  else if (const char* h = std::getenv("HOME")) dir = std::string(h) + "/.cache";
  if (dir.empty()) return "";
  return dir + "/fresco_amd/spread3.txt";
}

std::string shipped_tune_path() {
  Dl_info info;
  if (!dladdr(reinterpret_cast<void*>(&shipped_tune_path), &info) || !info.dli_fname) return "";
  std::string so(info.dli_fname);
  const size_t slash = so.rfind('/');
  return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tuning/spread3_mi355x.txt";
}

bool tune_lookup(const std::string& path, const std::string& key, int* mode, int* parts) {
  if (path.empty()) return false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string k;
    int m = -1, pp = 0;
    if (!(ls >> k >> m >> pp) || k != key) continue;
    *mode = m;
    *parts = pp;
    return true;
  }
  return false;
}

void tune_store(const std::string& path, const std::string& key, int mode, int parts) {
  if (path.empty()) return;
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    // create the directory chain (best effort)
    std::string dir = path.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') (void)::mkdir(dir.substr(0, i).c_str(), 0755);
  }
  if (FILE* f = std::fopen(path.c_str(), "a")) {
    std::fprintf(f, "%s %d %d\n", key.c_str(), mode, parts);
    std::fclose(f);
  }
}
}  // namespace

template <typename T>
std::string Spreader<T>::tune_key(int64_t npts) const {
  std::ostringstream k;
  k << "spread3:" << (sizeof(T) == 4 ? "f32" : "f64") << ":nf" << geom.nf[0] << "x" << geom.nf[1] << "x" << geom.nf[2]
    << ":tile" << geom.tile[0] << "x" << geom.tile[1] << "x" << geom.tile[2] << ":w" << es.w << ":batch" << geom.batch
    << ":n" << npts;
  return k.str();
}

// Time every spreader configuration once on the caller's sorted points (a
// warm-up launch, then one launch between HIP events) and keep the fastest;
// a cached pick for the same plan shape is taken without timing.
// Runs on the first spread of a plan whose grid admits both 3D spreaders.
template <typename T>
void Spreader<T>::autotune(const SortedPoints<T>& p, T* grid, hipStream_t s) const {
  const std::string key = tune_key(p.n);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    int m = -1, pp = 0;
    if (tune_lookup(user_tune_path(), key, &m, &pp) || tune_lookup(shipped_tune_path(), key, &m, &pp)) {
      const bool halo_mode = m == HALO || m == HALO4 ||
                             ((m == HALOM || m == HALOR) && sizeof(T) == 4 && geom.tile[0] == 16 &&
                              geom.tile[1] == 8 && geom.tile[2] == 8 && pp % 2 == 0);
      if ((m == REG && geom.tile[1] == 16) || (halo_mode && pp >= 1 && pp <= max_halo_parts)) {
        mode = m;
        halo_parts = pp;
        tuned = true;
        tune_source = 1;
        return;
      }
    }
  }
  tune_source = 2;
  autotune_timed(p, grid, s);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_store(user_tune_path(), key, mode, halo_parts);
}

template <typename T>
void Spreader<T>::autotune_timed(const SortedPoints<T>& p, T* grid, hipStream_t s) const {
  struct Cand {
    int mode, parts;
  };
  std::vector<Cand> cands;
  if (geom.tile[0] == 16 && geom.tile[1] == 16 && geom.tile[2] == 8) cands.push_back({REG, 1});
  for (int sl : {1, 2, 3, 4, 6, 8, 12, 16})
    if (sl <= max_halo_parts) cands.push_back({HALO, sl});
  for (int sl : {1, 2, 3, 4})
    if (sl <= max_halo_parts) cands.push_back({HALO4, sl});
  if (sizeof(T) == 4 && geom.tile[0] == 16 && geom.tile[1] == 8 && geom.tile[2] == 8)
    for (int sl : {1, 2, 3})
      if (2 * sl <= max_halo_parts) {
        cands.push_back({HALOM, 2 * sl});
        cands.push_back({HALOR, 2 * sl});
      }
  hipEvent_t e0, e1;
  FRESCO_HIP_CHECK(hipEventCreate(&e0));
  FRESCO_HIP_CHECK(hipEventCreate(&e1));
  float best = 1e30f;
  Cand pick = cands[0];
  tuned = true;  // the launches below dispatch on mode / halo_parts directly
  const unsigned prof = prof_mask();
  prof_set(false);
  for (const Cand& c : cands) {
    mode = c.mode;
    halo_parts = c.parts;
    spread(p, grid, s);
    // fastest of three timed launches (one launch is noisy at this scale)
    float ms = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      FRESCO_HIP_CHECK(hipEventRecord(e0, s));
      spread(p, grid, s);
      FRESCO_HIP_CHECK(hipEventRecord(e1, s));
      FRESCO_HIP_CHECK(hipEventSynchronize(e1));
      float t = 0.f;
      FRESCO_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
      ms = std::min(ms, t);
    }
    if (ms < best) {
      best = ms;
      pick = c;
    }
  }
  prof_set_mask(prof);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  mode = pick.mode;
  halo_parts = pick.parts;
}

template <typename T>
__global__ void k_sum_parts(const T* __restrict__ part, int parts, int64_t n, T* __restrict__ grid);
template <typename T>
static bool launch_spread2_halo(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s);

template <typename T, int D, int WMAX>
static void launch_spread(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s) {
  SpreadDev<T> a;
  const GridGeom& g = sp.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.maxc[q] = sp.maxc[q];
  }
  a.cbits = g.cell_bits;
  a.w = sp.es.w;
  a.sx = D >= 2 ? lds_row_stride(g.tile[0], (int)sizeof(T), sp.es.w) : g.tile[0];
  a.cstride = sp.cand_stride;
  a.k.beta = (T)sp.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(sp.es.beta * M_LOG2E) : (T)sp.es.beta;
  a.k.two_over_w = (T)(2.0 / sp.es.w);
  const int lp = sp.lparts;
  {
    ProfScope ps(ST_SPREAD, s);
    hipLaunchKernelGGL((k_spread<T, D, WMAX>), dim3(g.nbins, g.batch, lp), dim3(64), sp.lds_bytes, s, p.scell.get(),
                       p.rec.get(), p.bin_start.get(), sp.cand.get(), a, lp > 1 ? sp.part_buf.get() : grid);
    FRESCO_HIP_CHECK(hipGetLastError());
  }
  if (lp > 1) {
    ProfScope pm(ST_MERGE, s);
    const size_t n = (size_t)g.ncells() * g.batch;
    hipLaunchKernelGGL(k_sum_parts<T>, dim3((unsigned)std::min<size_t>(blocks_for(n, 256), 16384)), dim3(256), 0, s,
                       sp.part_buf.get(), lp, (int64_t)n, grid);
    FRESCO_HIP_CHECK(hipGetLastError());
  }
}

template <typename T>
static bool launch_spread3_reg(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s);
template <typename T>
static bool launch_spread3_halo(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s);

template <typename T, int D, int WMAX>
static void launch_spread_c(const Spreader<T>& sp, const SortedPoints<T>& p, const T* str, T* grid, hipStream_t s) {
  SpreadDev<T> a;
  const GridGeom& g = sp.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.maxc[q] = sp.maxc[q];
  }
  a.cbits = g.cell_bits;
  a.w = sp.es.w;
  a.sx = D >= 2 ? lds_row_stride(g.tile[0], (int)sizeof(T), sp.es.w) : g.tile[0];
  a.cstride = sp.cand_stride;
  a.k.beta = (T)sp.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(sp.es.beta * M_LOG2E) : (T)sp.es.beta;
  a.k.two_over_w = (T)(2.0 / sp.es.w);
  const size_t lds = 2 * (size_t)sp.lds_bytes;
  if (lds > 160 * 1024) throw InvalidArgument("complex spreading tile exceeds LDS");
  const int cp = sp.cparts;
  hipLaunchKernelGGL((k_spread<T, D, WMAX, 2>), dim3(g.nbins, 1, cp), dim3(64), lds, s, p.scell.get(), p.rec.get(),
                     p.bin_start.get(), sp.cand.get(), a, cp > 1 ? sp.part_buf_c.get() : grid, str, p.idx.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  if (cp > 1) {
    const size_t n = 2 * (size_t)g.ncells();
    hipLaunchKernelGGL(k_sum_parts<T>, dim3((unsigned)std::min<size_t>(blocks_for(n, 256), 16384)), dim3(256), 0, s,
                       sp.part_buf_c.get(), cp, (int64_t)n, grid);
    FRESCO_HIP_CHECK(hipGetLastError());
  }
}

template <typename T>
void Spreader<T>::plan_complex_slices(const SortedPoints<T>& p) {
  if (geom.batch != 1) return;
  std::vector<int> st((size_t)geom.nbins + 1);
  FRESCO_HIP_CHECK(hipMemcpy(st.data(), p.bin_start.get(), st.size() * sizeof(int), hipMemcpyDeviceToHost));
  int busy = 0, mx = 0;
  for (int b = 0; b < geom.nbins; ++b) {
    const int c = st[b + 1] - st[b];
    busy += c > 0;
    mx = std::max(mx, c);
  }
  // about 2048 wavefronts over the tiles that hold points, slices of >= 32
  // points
  int cp = busy > 0 ? (2048 + busy - 1) / busy : 1;
  cp = std::max(1, std::min({cp, 16, mx / 32}));
  if (cp > 1) part_buf_c.allocate((size_t)cp * 2 * (size_t)geom.ncells());
  cparts = cp;
}

template <typename T>
void Spreader<T>::spread_complex(const SortedPoints<T>& p, const Cplx<T>* str, Cplx<T>* grid, hipStream_t s) const {
  const T* st = reinterpret_cast<const T*>(str);
  T* gr = reinterpret_cast<T*>(grid);
  const bool small = es.w <= 8;
  if (geom.d == 3) {
    small ? launch_spread_c<T, 3, 8>(*this, p, st, gr, s) : launch_spread_c<T, 3, 16>(*this, p, st, gr, s);
  } else if (geom.d == 2) {
    small ? launch_spread_c<T, 2, 8>(*this, p, st, gr, s) : launch_spread_c<T, 2, 16>(*this, p, st, gr, s);
  } else {
    launch_spread_c<T, 1, 16>(*this, p, st, gr, s);
  }
}

template <typename T>
void Spreader<T>::spread(const SortedPoints<T>& p, T* grid, hipStream_t s) const {
  const int d = geom.d;
  const bool small = es.w <= 8;
  if (d == 3 && halo_ok && !tuned) autotune(p, grid, s);
  if (d == 3) {
    if (launch_spread3_halo<T>(*this, p, grid, s)) return;
    if (launch_spread3_reg<T>(*this, p, grid, s)) return;
    small ? launch_spread<T, 3, 8>(*this, p, grid, s) : launch_spread<T, 3, 16>(*this, p, grid, s);
  } else if (d == 2) {
    if (launch_spread2_halo<T>(*this, p, grid, s)) return;
    small ? launch_spread<T, 2, 8>(*this, p, grid, s) : launch_spread<T, 2, 16>(*this, p, grid, s);
  } else {
    launch_spread<T, 1, 16>(*this, p, grid, s);
  }
}

// ---------------------------------------------------------------------------
// interpolation: the grid region of one bin staged in LDS, one wavefront per
// point, DPP wavefront reduction
// ---------------------------------------------------------------------------

template <typename T>
struct InterpDev {
  int nf[3], tile[3], ntile[3];
  int cbits;
  int w;
  int rx, ry, rz;  // staged region extent (tile + w - 1)
  int srx;         // LDS row stride of the region
  int ostride, ocomp;
  double dscale[3];
  EsDev<T> k;
};

template <typename T, int D, int WMAX, bool DERIV>
__global__ void __launch_bounds__(256)
    k_interp(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
             const uint32_t* __restrict__ idx, const int* __restrict__ bin_start, const T* __restrict__ grid,
             InterpDev<T> a, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* reg = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  constexpr int SH = Lanes<WMAX>::SH;
  constexpr int NCH = D == 1 ? 1 : (WMAX * WMAX + 63) / 64;

  bin_start += (int64_t)blockIdx.y * a.ntile[0] * a.ntile[1] * a.ntile[2];  // batched plans
  grid += (int64_t)blockIdx.y * a.nf[0] * a.nf[1] * a.nf[2];
  const int b = blockIdx.x;
  const int beg = bin_start[b];
  const int end = bin_start[b + 1];
  if (beg == end) return;  // uniform across the workgroup
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  const int ry = D >= 2 ? a.ry : 1;
  const int rz = D >= 3 ? a.rz : 1;
  const int tot = a.rx * ry * rz;
  for (int e = tid; e < tot; e += 256) {
    int lx = e % a.rx;
    int rr = e / a.rx;
    int ly = rr % ry;
    int lz = rr / ry;
    const int gx = (t0[0] + lx) % a.nf[0];  // the region may wrap (and repeat
    const int gy = (t0[1] + ly) % a.nf[1];  // cells on grids smaller than
    const int gz = (t0[2] + lz) % a.nf[2];  // tile + w - 1)
    reg[(lz * ry + ly) * a.srx + lx] = grid[((int64_t)gz * a.nf[1] + gy) * a.nf[0] + gx];
  }
  __syncthreads();

  const int kd = lane >> SH;
  const int kj = lane & (WMAX - 1);
  const int zs = ry * a.srx;
  for (int k = beg + wave; k < end; k += 4) {
    const uint32_t sc = scell[k];
    const typename Vec4<T>::type r = rec[k];
    int l0[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int s = q < D ? (int)((sc >> (q * a.cbits)) & cmask) : 0;
      int u = s - t0[q];
      if (u < 0) u += a.nf[q];  // first cell lies in this bin's tile
      l0[q] = q < D ? u : 0;
    }
    const T offk = kd == 0 ? r.x : (kd == 1 ? r.y : r.z);
    T kv = T(0), kdv = T(0);
    if (kd < D && kj < w) es_phi_dphi<T>((T(kj) - offk) * a.k.two_over_w, a.k, kv, kdv);
    T c0 = T(0), c1 = T(0), c2 = T(0);
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      const int pr = lane + 64 * ch;
      const int ja = D == 1 ? pr : (pr & (WMAX - 1));
      const int jb = D == 1 ? 0 : (pr >> SH);
      const bool act = ja < w && (D == 1 || jb < w);
      const int jas = act ? ja : 0;  // inactive lanes read a valid address
      const int jbs = act ? jb : 0;
      const T px = lane_get(kv, jas);
      const T dx = DERIV ? lane_get(kdv, jas) : T(0);
      const T py = D >= 2 ? lane_get(kv, WMAX + jbs) : T(1);
      const T dy = (DERIV && D >= 2) ? lane_get(kdv, WMAX + jbs) : T(0);
      const T* rowp = reg + (l0[1] + jbs) * a.srx + (l0[0] + jas);
      T s0 = T(0), s1 = T(0);
      if (D == 3) {
        const T* colp = rowp + l0[2] * zs;
#pragma unroll
        for (int jc = 0; jc < WMAX; ++jc) {
          if (jc < w) {
            const T wz = lane_read(kv, 2 * WMAX + jc);
            const T v = act ? colp[jc * zs] : T(0);
            s0 += v * wz;
            if (DERIV) s1 += v * lane_read(kdv, 2 * WMAX + jc);
          }
        }
      } else {
        s0 = act ? rowp[0] : T(0);
      }
      if (DERIV) {
        c0 += dx * py * s0;
        if (D >= 2) c1 += px * dy * s0;
        if (D >= 3) c2 += px * py * s1;
      } else {
        c0 += px * py * s0;
      }
    }
    c0 = wave_sum63(c0);
    if (DERIV && D >= 2) c1 = wave_sum63(c1);
    if (DERIV && D >= 3) c2 = wave_sum63(c2);
    if (lane == 63) {
      const int64_t j = idx[k];
      const double cw = (double)r.w;
      if (DERIV) {
        out[j * a.ostride + 0] = cw * a.dscale[0] * (double)c0;
        if (D >= 2) out[j * a.ostride + 1] = cw * a.dscale[1] * (double)c1;
        if (D >= 3) out[j * a.ostride + 2] = cw * a.dscale[2] * (double)c2;
      } else {
        out[j * a.ostride + a.ocomp] = cw * (double)c0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3D spreading, output-stationary: one wavefront owns a 16 x 16 x 8 tile held
// in registers.  Lane (x_l = lane % 16, zq = lane / 16) owns the cells
// x = x_l, z = 2 zq + {0, 1}, y = 0..15 (32 accumulators).  For every point
// whose support touches the tile, lanes 0..3*WMAX-1 evaluate the 3w kernel
// values once; each lane takes its x and z weights by ds_bpermute and the 16
// y weights (shifted to the tile) from a small per-wavefront LDS table read by
// broadcast, and accumulates up to 32 FMAs in registers (the half of the
// y-column the support misses is skipped).  No LDS read-modify-write, no atomics;
// every cell sums its points in one fixed order (bitwise reproducible), and
// is written to HBM once.  Requires nf[1] >= 16 + w - 1 (so the tile sees
// one contiguous piece of any support along y).
// Small grids have few tiles: gridDim.y > 1 splits every candidate bin's
// point range into gridDim.y contiguous slices, wavefront (tile, part) writes
// its partial tile to grid + part * ncells, and k_sum_parts adds the parts
// in a fixed order (still bitwise reproducible).
// ---------------------------------------------------------------------------

constexpr int RT_X = 16, RT_Y = 16, RT_Z = 8;

// grid[i] = sum_{p < parts} part[p * n + i], in order p = 0, 1, ...
template <typename T>
__global__ void k_sum_parts(const T* __restrict__ part, int parts, int64_t n, T* __restrict__ grid) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T acc = part[i];
    for (int p = 1; p < parts; ++p) acc += part[(int64_t)p * n + i];
    grid[i] = acc;
  }
}

template <typename T, int WMAX>
__global__ void __launch_bounds__(64)
    k_spread3_reg(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                  const int* __restrict__ bin_start, const int* __restrict__ cand, SpreadDev<T> a,
                  T* __restrict__ grid) {
  constexpr int TAB = 2 * RT_Y + WMAX;
  __shared__ T tab[64];
  const int lane = threadIdx.x;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  constexpr int SH = Lanes<WMAX>::SH;
  const int x_l = lane & 15;
  const int zq = lane >> 4;

  const int b = blockIdx.x;
  int t[3];
  t[0] = b % a.ntile[0];
  t[1] = (b / a.ntile[0]) % a.ntile[1];
  t[2] = b / (a.ntile[0] * a.ntile[1]);
  int t0[3], tl[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    t0[q] = t[q] * a.tile[q];
    tl[q] = min(a.tile[q], a.nf[q] - t0[q]);
  }
  T acc0[RT_Y], acc1[RT_Y];
#pragma unroll
  for (int r = 0; r < RT_Y; ++r) {
    acc0[r] = T(0);
    acc1[r] = T(0);
  }
  const int kd = lane >> SH;
  const int kj = lane & (WMAX - 1);
  const int z0 = 2 * zq, z1 = 2 * zq + 1;
  // table lane i holds the weight of y-kernel index i - RT_Y
  const int tj = lane - RT_Y;
  const bool tvalid = tj >= 0 && tj < w && lane < TAB;

  for (int c2 = 0; c2 < a.maxc[2]; ++c2) {
    const int bz = cand[2 * a.cstride + t[2] * a.maxc[2] + c2];
    if (bz < 0) continue;
    for (int c1 = 0; c1 < a.maxc[1]; ++c1) {
      const int by = cand[1 * a.cstride + t[1] * a.maxc[1] + c1];
      if (by < 0) continue;
      for (int c0 = 0; c0 < a.maxc[0]; ++c0) {
        const int bx = cand[t[0] * a.maxc[0] + c0];
        if (bx < 0) continue;
        const int bin = bx + a.ntile[0] * (by + a.ntile[1] * bz);
        const int beg0 = bin_start[bin];
        const int len = bin_start[bin + 1] - beg0;
        const int beg = beg0 + (int)(((int64_t)len * blockIdx.y) / gridDim.y);
        const int end = beg0 + (int)(((int64_t)len * (blockIdx.y + 1)) / gridDim.y);
        for (int base = beg; base < end; base += 64) {
          const int j = base + lane;
          uint32_t sc = 0;
          typename Vec4<T>::type rr;
          rr.x = rr.y = rr.z = rr.w = T(0);
          bool touch = false;
          if (j < end) {
            sc = scell[j];
            rr = rec[j];
            touch = true;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              int s = (int)((sc >> (q * a.cbits)) & cmask);
              int u = s - t0[q];
              if (u < 0) u += a.nf[q];
              touch = touch && (u < tl[q] || u + w > a.nf[q]);
            }
          }
          unsigned long long m = __ballot(touch);
          while (m) {
            const int p = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t scp = lane_read_u(sc, p);
            const T cw = lane_read(rr.w, p);
            const T ox = lane_read(rr.x, p), oy = lane_read(rr.y, p), oz = lane_read(rr.z, p);
            const T offk = kd == 0 ? ox : (kd == 1 ? oy : oz);
            const T kv = (kd < 3 && kj < w) ? es_phi<T>((T(kj) - offk) * a.k.two_over_w, a.k) : T(0);
            // first cell relative to the tile start, wrapped into
            // [-(w-1), tile) (wavefront-uniform, scalar arithmetic)
            int e3[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              int dq = (int)((scp >> (q * a.cbits)) & cmask) - t0[q];
              if (dq < 0) dq += a.nf[q];
              e3[q] = dq > a.nf[q] - w ? dq - a.nf[q] : dq;
            }
            // x weight of this lane's column
            const int jx = x_l - e3[0];
            const bool vx = (unsigned)jx < (unsigned)w && x_l < tl[0];
            const T wxv = lane_get(kv, vx ? jx : 0);
            const T wx = vx ? wxv * cw : T(0);
            // z weights of this lane's two cells
            const int jz0 = z0 - e3[2];
            const int jz1 = jz0 + 1;
            const bool vz0 = (unsigned)jz0 < (unsigned)w && z0 < tl[2];
            const bool vz1 = (unsigned)jz1 < (unsigned)w && z1 < tl[2];
            const T wz0v = lane_get(kv, 2 * WMAX + (vz0 ? jz0 : 0));
            const T wz1v = lane_get(kv, 2 * WMAX + (vz1 ? jz1 : 0));
            const T a0 = vz0 ? wx * wz0v : T(0);
            const T a1 = vz1 ? wx * wz1v : T(0);
            // y weights shifted to the tile, W(r) = kv_y[r - ey], through a
            // per-wavefront LDS table read by broadcast
            const int ey = e3[1];
            const T tv = lane_get(kv, WMAX + (tvalid ? tj : 0));
            tab[lane] = tvalid ? tv : T(0);
            __builtin_amdgcn_wave_barrier();  // table written (in-order LDS) before it is read
            const T* wrow = tab + (RT_Y - ey);
            // the support covers y = ey .. ey + w - 1: skip a half of the
            // column that it misses (wavefront-uniform branches)
            if (ey < RT_Y / 2) {
#pragma unroll
              for (int r = 0; r < RT_Y / 2; ++r) {
                const T wy = wrow[r];
                acc0[r] += a0 * wy;
                acc1[r] += a1 * wy;
              }
            }
            if (ey + w > RT_Y / 2) {
#pragma unroll
              for (int r = RT_Y / 2; r < RT_Y; ++r) {
                const T wy = wrow[r];
                acc0[r] += a0 * wy;
                acc1[r] += a1 * wy;
              }
            }
            __builtin_amdgcn_wave_barrier();  // reads done before the next point rewrites the table
          }
        }
      }
    }
  }
  if (gridDim.y > 1) grid += (int64_t)blockIdx.y * a.nf[0] * a.nf[1] * a.nf[2];
  if (x_l < tl[0]) {
    const int64_t gx = t0[0] + x_l;
#pragma unroll
    for (int r = 0; r < RT_Y; ++r) {
      if (r < tl[1]) {
        const int64_t row = (int64_t)(t0[1] + r) * a.nf[0] + gx;
        if (z0 < tl[2]) grid[(int64_t)(t0[2] + z0) * a.nf[1] * a.nf[0] + row] = acc0[r];
        if (z1 < tl[2]) grid[(int64_t)(t0[2] + z1) * a.nf[1] * a.nf[0] + row] = acc1[r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3D interpolation, lane per point: the bin's grid region is staged in LDS by
// one wavefront; each lane evaluates its point's 3w kernel values (and
// derivatives) and sums its w^3 cells from LDS.  No cross-lane traffic.
// ---------------------------------------------------------------------------

constexpr int LANE_INTERP_THREADS = 128;

template <typename T, int WMAX, bool DERIV>
__global__ void __launch_bounds__(LANE_INTERP_THREADS)
    k_interp3_lane(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                   const uint32_t* __restrict__ idx, const int* __restrict__ bin_start, const T* __restrict__ grid,
                   InterpDev<T> a, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* reg = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const int b = blockIdx.x;
  const int beg = bin_start[b];
  const int end = bin_start[b + 1];
  if (beg == end) return;
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  const int tot = a.rx * a.ry * a.rz;
  for (int e = tid; e < tot; e += LANE_INTERP_THREADS) {
    int lx = e % a.rx;
    int rr = e / a.rx;
    int ly = rr % a.ry;
    int lz = rr / a.ry;
    const int gx = (t0[0] + lx) % a.nf[0];
    const int gy = (t0[1] + ly) % a.nf[1];
    const int gz = (t0[2] + lz) % a.nf[2];
    reg[(lz * a.ry + ly) * a.srx + lx] = grid[((int64_t)gz * a.nf[1] + gy) * a.nf[0] + gx];
  }
  __syncthreads();
  const int zs = a.ry * a.srx;
  for (int k = beg + tid; k < end; k += LANE_INTERP_THREADS) {
    const uint32_t sc = scell[k];
    const typename Vec4<T>::type r = rec[k];
    int l0[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int u = (int)((sc >> (q * a.cbits)) & cmask) - t0[q];
      if (u < 0) u += a.nf[q];
      l0[q] = u;
    }
    T wx[WMAX], dwx[WMAX], wy[WMAX], dwy[WMAX];
#pragma unroll
    for (int jj = 0; jj < WMAX; ++jj) {
      wx[jj] = dwx[jj] = wy[jj] = dwy[jj] = T(0);
      if (jj < w) {
        if (DERIV) {
          es_phi_dphi<T>((T(jj) - r.x) * a.k.two_over_w, a.k, wx[jj], dwx[jj]);
          es_phi_dphi<T>((T(jj) - r.y) * a.k.two_over_w, a.k, wy[jj], dwy[jj]);
        } else {
          wx[jj] = es_phi<T>((T(jj) - r.x) * a.k.two_over_w, a.k);
          wy[jj] = es_phi<T>((T(jj) - r.y) * a.k.two_over_w, a.k);
        }
      }
    }
    T gx = T(0), gy = T(0), gz = T(0);
    const T* col = reg + (l0[2] * a.ry + l0[1]) * a.srx + l0[0];
#pragma unroll 1
    for (int jc = 0; jc < w; ++jc) {
      T wz, dwz = T(0);
      if (DERIV)
        es_phi_dphi<T>((T(jc) - r.z) * a.k.two_over_w, a.k, wz, dwz);
      else
        wz = es_phi<T>((T(jc) - r.z) * a.k.two_over_w, a.k);
      T sx0 = T(0), sx1 = T(0), sy = T(0);
#pragma unroll
      for (int jb = 0; jb < WMAX; ++jb) {
        if (jb < w) {
          const T* row = col + jb * a.srx;
          T t0v = T(0), t1v = T(0);
#pragma unroll
          for (int ja = 0; ja < WMAX; ++ja) {
            if (ja < w) {
              const T g = row[ja];
              t0v += g * wx[ja];
              if (DERIV) t1v += g * dwx[ja];
            }
          }
          sx0 += t0v * wy[jb];
          if (DERIV) {
            sx1 += t1v * wy[jb];
            sy += t0v * dwy[jb];
          }
        }
      }
      if (DERIV) {
        gx += sx1 * wz;
        gy += sy * wz;
        gz += sx0 * dwz;
      } else {
        gx += sx0 * wz;
      }
      col += zs;
    }
    const int64_t j = idx[k];
    const double cw = (double)r.w;
    if (DERIV) {
      out[j * a.ostride + 0] = cw * a.dscale[0] * (double)gx;
      out[j * a.ostride + 1] = cw * a.dscale[1] * (double)gy;
      out[j * a.ostride + 2] = cw * a.dscale[2] * (double)gz;
    } else {
      out[j * a.ostride + a.ocomp] = cw * (double)gx;
    }
  }
}

// ---------------------------------------------------------------------------
// 3D spreading, input-driven with a one-sided halo: bins are keyed by the
// first cell of each point's support, so the support of every point of a
// 16 x 16 x 8 bin lies in the bin's (16+7) x (16+7) x (8+7) halo tile.  One
// wavefront per (bin, slice) accumulates its points into that tile in LDS
// with plain read-modify-write (a wavefront's LDS operations complete in
// order, and the 64 lanes of one point touch 64 distinct cells), two points
// per iteration sharing the kernel evaluation (lanes 0-23 point A, lanes
// 32-55 point B).  Every point is handled exactly once.  The tiles go to HBM
// and k_merge_halo sums, for every fine-grid cell, the tiles that cover it
// in a fixed order (bitwise reproducible, no atomics).
// ---------------------------------------------------------------------------


template <typename T, int TY>
__global__ void __launch_bounds__(64)
    k_spread3_halo(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                   const int* __restrict__ bin_start, SpreadDev<T> a, T* __restrict__ halo) {
  constexpr int HS_Y = HaloShape<TY>::HS_Y;
  constexpr int HT_ELEMS = HaloShape<TY>::ELEMS;
  __shared__ __align__(16) T tile[HT_ELEMS];
  const int lane = threadIdx.x;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const int b = blockIdx.x;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.y) / gridDim.y);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.y + 1)) / gridDim.y);
  T* out = halo + ((int64_t)b * gridDim.y + blockIdx.y) * HT_ELEMS;
  for (int e = lane; e < HT_ELEMS; e += 64) tile[e] = T(0);
  if (beg == end) {
    for (int e = lane; e < HT_ELEMS; e += 64) out[e] = T(0);
    return;
  }
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  __syncthreads();

  const int hb = lane >> 5;               // kernel-value half: 0 point A, 1 point B
  const int kd = (lane & 31) >> 3;        // dimension of this lane's kernel value
  const int kj = lane & 7;                // kernel index
  const bool kact = kd < 3 && kj < w;
  const int ja = lane & 7, jb = lane >> 3;  // this lane's (x, y) cell of a support
  for (int base = beg; base < end; base += 64) {
    const int cnt = min(64, end - base);
    uint32_t sc = 0;
    typename Vec4<T>::type rr;
    rr.x = rr.y = rr.z = rr.w = T(0);
    if (lane < cnt) {
      sc = scell[base + lane];
      rr = rec[base + lane];
    }
    for (int p = 0; p < cnt; p += 2) {
      const int q = p + 1 < cnt ? p + 1 : p;  // odd tail: B repeats A with weight 0
      const uint32_t sa = lane_read_u(sc, p), sb = lane_read_u(sc, q);
      const T cA = lane_read(rr.w, p);
      const T cB = p + 1 < cnt ? lane_read(rr.w, q) : T(0);
      const T oAx = lane_read(rr.x, p), oAy = lane_read(rr.y, p), oAz = lane_read(rr.z, p);
      const T oBx = lane_read(rr.x, q), oBy = lane_read(rr.y, q), oBz = lane_read(rr.z, q);
      const T offA = kd == 0 ? oAx : (kd == 1 ? oAy : oAz);
      const T offB = kd == 0 ? oBx : (kd == 1 ? oBy : oBz);
      const T off = hb ? offB : offA;
      const T kv = kact ? es_phi<T>((T(kj) - off) * a.k.two_over_w, a.k) : T(0);
      // (x, y) weights of this lane's cell, for A and B
      const T vA = lane_get(kv, ja) * lane_get(kv, 8 + jb) * cA;
      const T vB = lane_get(kv, 32 + ja) * lane_get(kv, 40 + jb) * cB;
      int uA[3], uB[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uA[d] = (int)((sa >> (d * a.cbits)) & cmask) - t0[d];
        uB[d] = (int)((sb >> (d * a.cbits)) & cmask) - t0[d];
      }
      T* ca = tile + (uA[2] * HS_Y + uA[1] + jb) * HS_X + uA[0] + ja;
      T* cb = tile + (uB[2] * HS_Y + uB[1] + jb) * HS_X + uB[0] + ja;
      T wzA[HW], wzB[HW];
#pragma unroll
      for (int jc = 0; jc < HW; ++jc) {
        wzA[jc] = lane_read(kv, 16 + jc);
        wzB[jc] = lane_read(kv, 48 + jc);
      }
#pragma unroll
      for (int jc = 0; jc < HW; ++jc) ca[jc * HS_Y * HS_X] += vA * wzA[jc];
#pragma unroll
      for (int jc = 0; jc < HW; ++jc) cb[jc * HS_Y * HS_X] += vB * wzB[jc];
    }
  }
  __syncthreads();
  for (int e = lane; e < HT_ELEMS; e += 64) out[e] = tile[e];
}

// The same spreading with NW wavefronts per halo tile (launched with 4 and
// 32-point chunks: measured best of NW 4 / 8 and CH 32 / 64 on the 3D
// N = 2^20 configuration): wavefront q owns the tile's
// z-planes congruent to q mod NW, so of every point's 8 support planes it
// adds exactly 8 / NW (offsets j1 = (q - z0) mod NW, j1 + NW, ...) and no two
// wavefronts ever touch one cell (plain read-modify-write stays race-free).
// Per chunk of CH points the workgroup evaluates the kernel values once into
// LDS (thread = point x kernel index, records loaded a chunk ahead) and
// forms the 64 (x, y) products c * wx * wy of every point;
// a wavefront then reads its lane's product and multiplies by its z values.
// More wavefronts per LDS tile than the one-wavefront kernel; the products,
// their order and the per-cell point order are the same, so the halo tiles
// are bitwise equal.
template <typename T, int TY, int NW, int CH>
__global__ void __launch_bounds__(64 * NW)
    k_spread3_halo_mw(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                      const int* __restrict__ bin_start, SpreadDev<T> a, T* __restrict__ halo) {
  constexpr int NT = 64 * NW;
  constexpr int KR = CH * 8 / NT;  // (point, kernel index) pairs per thread
  static_assert(KR * NT == CH * 8, "chunk must be a multiple of NT / 8 points");
  constexpr int NP = 8 / NW;   // planes per point and wavefront
  constexpr int HS_Y = HaloShape<TY>::HS_Y;
  constexpr int HT_ELEMS = HaloShape<TY>::ELEMS;
  constexpr int PLANE = HS_X * HS_Y;
  __shared__ __align__(16) T tile[HT_ELEMS];
  __shared__ T wzs[CH][HW];                  // z kernel values
  __shared__ __align__(16) T wxy[CH][64];     // c * wx[ja] * wy[jb] at lane ja + 8 jb
  __shared__ uint32_t sch[CH];               // packed first cells
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int q = tid >> 6;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const int b = blockIdx.x;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.y) / gridDim.y);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.y + 1)) / gridDim.y);
  T* out = halo + ((int64_t)b * gridDim.y + blockIdx.y) * HT_ELEMS;
  if (beg == end) {
    for (int e = tid; e < HT_ELEMS; e += NT) out[e] = T(0);
    return;
  }
  for (int e = tid; e < HT_ELEMS; e += NT) tile[e] = T(0);
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  const int ja = lane & 7, jb = lane >> 3;
  const int loff = jb * HS_X + ja;  // this lane's (x, y) cell of a support
  const int kj = tid & 7;
  typename Vec4<T>::type rn[KR];
  uint32_t scn = 0;
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int kp = k * (NT / 8) + (tid >> 3);
    rn[k].x = rn[k].y = rn[k].z = rn[k].w = T(0);
    if (beg + kp < end) rn[k] = rec[beg + kp];
  }
  if (tid < CH && beg + tid < end) scn = scell[beg + tid];
  for (int c0 = beg; c0 < end; c0 += CH) {
    const int cnt = min(CH, end - c0);
    const uint32_t scc = scn;
    if (tid < CH && c0 + CH + tid < end) scn = scell[c0 + CH + tid];
    // thread (point kp, index kj): its x, y, z kernel values; the 8 threads
    // of a point share the x values by swizzle and thread kj writes row
    // jb = kj of the point's products, (wx * wy) * c
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int kp = k * (NT / 8) + (tid >> 3);
      const typename Vec4<T>::type r = rn[k];
      if (c0 + CH + kp < end) rn[k] = rec[c0 + CH + kp];
      const bool on = kj < w && kp < cnt;
      const T wxv = on ? es_phi<T>((T(kj) - r.x) * a.k.two_over_w, a.k) : T(0);
      const T wyv = on ? es_phi<T>((T(kj) - r.y) * a.k.two_over_w, a.k) : T(0);
      const T wzv = on ? es_phi<T>((T(kj) - r.z) * a.k.two_over_w, a.k) : T(0);
      T wx8[8];
      group8_bcast<T>(wxv, wx8);
      if (kp < cnt) {
        T* row = &wxy[kp][8 * kj];
#pragma unroll
        for (int j = 0; j < 8; ++j) row[j] = wx8[j] * wyv * r.w;
        wzs[kp][kj] = wzv;
      }
    }
    if (tid < cnt) sch[tid] = scc;
    __syncthreads();
    // lanes of every wavefront, one point each: the tile offset of the
    // point's plane j1 and the z values this wavefront applies
    int base[CH / 64 + 1];
    T zv[CH / 64 + 1][NP];
#pragma unroll
    for (int h = 0; h * 64 < CH; ++h) {
      const int pp = h * 64 + lane;
      base[h] = 0;
#pragma unroll
      for (int m = 0; m < NP; ++m) zv[h][m] = T(0);
      if (pp < cnt) {
        const uint32_t sc = sch[pp];
        const int ux = (int)(sc & cmask) - t0[0];
        const int uy = (int)((sc >> a.cbits) & cmask) - t0[1];
        const int uz = (int)((sc >> (2 * a.cbits)) & cmask) - t0[2];
        const int j1 = (q - uz) & (NW - 1);
        base[h] = ((uz + j1) * HS_Y + uy) * HS_X + ux;
#pragma unroll
        for (int m = 0; m < NP; ++m) zv[h][m] = wzs[pp][j1 + m * NW];
      }
    }
    // Every LDS load above completes here, so the point loop below waits
    // once per point (for its own reads) instead of also waiting for the
    // previous point's write at a merged loop-entry state.
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
#pragma unroll
    for (int h = 0; h * 64 < CH; ++h) {
      const int lim = min(cnt - h * 64, 64);
      for (int i = 0; i < lim; ++i) {
        const int bp = __builtin_amdgcn_readlane(base[h], i);
        T za[NP];
#pragma unroll
        for (int m = 0; m < NP; ++m) za[m] = lane_read(zv[h][m], i);
        const T v = wxy[h * 64 + i][lane];
        T* c = tile + bp + loff;
#pragma unroll
        for (int m = 0; m < NP; ++m) c[m * NW * PLANE] += v * za[m];
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < HT_ELEMS; e += NT) out[e] = tile[e];
}

// The halo spreading of FP32 strengths on the matrix cores (16 x 8 x 8
// bins).  A bin slice's points split by the x cell their support starts in
// into two windows of the halo tile, columns [8 h, 8 h + 16) (h = 0, 1: the
// bin's left or right 8 columns), each the work of one wavefront that holds
// the window's 15 planes in accumulator registers.  Plane z of a window is
//   D_z[y][j] = sum_p  wy_p[y - uy_p]  *  wz_p[z - uz_p] c_p wx_p[8 h + j - ux_p]
// with the kernel values zero outside [0, 8): a 16 x 16 product over the
// points (K, 4 per v_mfma_f32_16x16x4f32), A = the points' y values placed at
// their first y cell, B = their z-weighted x values placed at their first x
// cell.  The window's points are ranked by the z of their first cell, each
// run padded to whole chunks of 8 slots (empty slots weigh 0), so the 4
// points of a product share it and add to the 8 planes uz .. uz + 7: 8
// products per 4 points and no read-modify-write of a tile while points are
// added.  Ranking is stable (ballot ranks in index order), so the summation
// order is fixed and the result bitwise reproducible.
constexpr int HM_CH = 8;      // slots per chunk (one key, two products; runs padded to it)
constexpr int HM_RS = 48;     // floats per slot record: ay[16], bx[16], wz[8]
constexpr int HM_NK = 8;      // keys: uz

typedef float hm_f4 __attribute__((ext_vector_type(4)));

// f(integral_constant<int, I>) for I = B .. E - 1, in order
template <int B, int E, typename F>
__device__ __forceinline__ void hm_static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    hm_static_for<B + 1, E>(f);
  }
}

// Order of a wavefront's own LDS accesses (they execute in program order):
// keeps the compiler from moving LDS reads and writes across this point.  The
// ranking of a window's points touches only that window's arrays, so its
// steps need no workgroup barrier.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

template <int UZ>
__device__ __forceinline__ void hm_products(hm_f4 (&acc)[HP_Z], float av, float bx, const float4& zlo,
                                            const float4& zhi) {
  const float wz[8] = {zlo.x, zlo.y, zlo.z, zlo.w, zhi.x, zhi.y, zhi.z, zhi.w};
#pragma unroll
  for (int m = 0; m < 8; ++m)
    acc[UZ + m] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bx * wz[m], acc[UZ + m], 0, 0, 0);
}

// Both windows of a (bin, slice) run in one workgroup of two wavefronts
// (wavefront h = window h) that share one staged copy of the slice's
// records and first cells (every point is read from memory once), rank the
// slice in rounds of up to WIN points (each window's z runs padded to whole
// chunks once per round), read the operands of chunk c + 1 from LDS while
// the products of chunk c run, and at the end add the two windows into one
// halo tile in LDS (the right window's columns, then the left window's
// added) that the workgroup writes whole: one tile per (bin, slice), which
// the merge sums like the LDS spreaders' tiles.
template <int WIN>
__global__ void __launch_bounds__(128)
    k_spread3_halo_mfma2(const uint32_t* __restrict__ scell, const float4* __restrict__ rec,
                         const int* __restrict__ bin_start, SpreadDev<float> a, float* __restrict__ halo) {
  constexpr int HS_Y = HaloShape<8>::HS_Y;
  constexpr int ELEMS = HaloShape<8>::ELEMS;
  static_assert(HS_Y == 16, "the 16 rows of a product are the tile's y rows");
  static_assert(WIN % 128 == 0, "a round is staged by 128 threads");
  constexpr int MAXSLOT = WIN + (HM_CH - 1) * HM_NK;
  constexpr int WPL = WIN / 64;   // round points per lane (ranking, per wavefront)
  constexpr int SPL = WIN / 128;  // round points per thread (staging)
  // the round's records and first cells, and each window's slot order;
  // after the last round the same bytes hold the summed halo tile
  constexpr int TILE = HP_Z * HS_Y * HS_X;
  constexpr int STAGE = WIN * (16 + 4) + 2 * MAXSLOT * 2;
  __shared__ __align__(16) unsigned char smem[STAGE > TILE * 4 ? STAGE : TILE * 4];
  float4* wrec = reinterpret_cast<float4*>(smem);
  uint32_t* wsc = reinterpret_cast<uint32_t*>(smem + WIN * 16);
  uint16_t* perm = reinterpret_cast<uint16_t*>(smem + WIN * 20);
  __shared__ __align__(16) float recs[2][2][HM_CH * HM_RS];  // [window][buffer]: kernel values of two chunks
  __shared__ int kend[2][HM_NK];
  __shared__ int run[2][HM_NK];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int h = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.x;
  const int nsl = gridDim.y, sl = blockIdx.y;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * sl) / nsl);
  const int end = beg0 + (int)(((int64_t)len * (sl + 1)) / nsl);
  float* out = halo + ((int64_t)b * nsl + sl) * ELEMS;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  auto rel = [&](uint32_t sc, int q) {
    int u = (int)((sc >> (q * a.cbits)) & cmask) - t0[q];
    if (u < 0) u += a.nf[q];
    return u;
  };
  const int i16 = lane & 15, k4 = lane >> 4;
  const int kj = lane & 7;
  const uint64_t lt = (1ull << lane) - 1ull;
  uint16_t* wperm = perm + h * MAXSLOT;
  int* wrun = run[h];
  int* wkend = kend[h];
  hm_f4 acc[HP_Z];
#pragma unroll
  for (int z = 0; z < HP_Z; ++z) acc[z] = hm_f4{0.f, 0.f, 0.f, 0.f};
  for (int ws = beg; ws < end; ws += WIN) {
    const int wn = min(WIN, end - ws);
    // the thread index, opaque here: the staging and ranking addresses are
    // formed in every round instead of being kept in registers across the
    // products (a slice rarely takes more than one round)
    int tq = tid;
    asm volatile("" : "+v"(tq));
    const int lq = tq & 63;
    {
      float4 wr[SPL];
      uint32_t wk[SPL];
#pragma unroll
      for (int u = 0; u < SPL; ++u) {
        const int e = u * 128 + tq;
        wk[u] = e < wn ? scell[ws + e] : 0u;
        wr[u] = e < wn ? rec[ws + e] : float4{0.f, 0.f, 0.f, 0.f};
      }
      __syncthreads();  // both windows are done with the previous round's records
#pragma unroll
      for (int u = 0; u < SPL; ++u) {
        wrec[u * 128 + tq] = wr[u];
        wsc[u * 128 + tq] = wk[u];
      }
    }
    if (lane < HM_NK) wrun[lane] = 0;
    __syncthreads();
    // key: z of the first cell, for the points of this wavefront's window; -1 others
    int wkey[WPL];
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
      const int e = u * 64 + lq;
      const uint32_t sc = wsc[e];
      const bool mine = e < wn && ((rel(sc, 0) >> 3) & 1) == h;
      wkey[u] = mine ? (rel(sc, 2) & 7) : -1;
    }
#pragma unroll
    for (int u = 0; u < WPL; ++u)
      if (wkey[u] >= 0) atomicAdd(&wrun[wkey[u]], 1);
    wave_lds_order();
    const int cnt = lane < HM_NK ? wrun[lane] : 0;
    const int pad = (cnt + HM_CH - 1) / HM_CH * HM_CH;
    int inc = pad;
#pragma unroll
    for (int off = 1; off < HM_NK; off <<= 1) {
      const int u = __builtin_amdgcn_ds_bpermute(max(lq - off, 0) << 2, inc);
      if (lq >= off) inc += u;
    }
    const int nslot = __builtin_amdgcn_readlane(inc, HM_NK - 1);
    const int st = inc - pad;
    wave_lds_order();
    if (lane < HM_NK) {
      wrun[lane] = st;
      wkend[lane] = st + pad;
    }
    for (int q = lq; q < nslot; q += 64) wperm[q] = 0xFFFF;
    wave_lds_order();
    // stable placement, 64 points a round per wavefront
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
      if (u * 64 >= wn) break;
      const int key = wkey[u];
      const bool valid = key >= 0;
      uint64_t same = __ballot(valid);
#pragma unroll
      for (int bt = 0; bt < 3; ++bt) {
        const bool bit = (key >> bt) & 1;
        const uint64_t mk = __ballot(bit);
        same &= bit ? mk : ~mk;
      }
      const int rank = __popcll(same & lt);
      if (valid) wperm[min(wrun[key] + rank, MAXSLOT - 1)] = (uint16_t)(u * 64 + lq);
      wave_lds_order();
      if (valid && rank == 0) wrun[key] += __popcll(same);
      wave_lds_order();
    }
    // slot -> point: the permutation entry of chunk c + 3 is read one chunk
    // before the record of that chunk, so neither waits on the other
    uint32_t scn;
    float4 rcn;
    int pslot, praw;
    auto fetch = [&](int c0) {
      pslot = c0 + (lane >> 3);
      praw = (int)wperm[min(pslot, MAXSLOT - 1)];
    };
    auto load = [&]() {
      const bool empty = pslot >= nslot || praw >= wn;
      const int pq = empty ? 0 : praw;
      const uint32_t sc = wsc[pq];
      rcn = wrec[pq];
      scn = empty ? 0xFFFFFFFFu : sc;
    };
    auto values = [&](float* rb) {
      float* rr = rb + (lane >> 3) * HM_RS;
      const float4 r = rcn;
      const bool empty = scn == 0xFFFFFFFFu;
      const int dx = rel(scn, 0) & 7;
      const int dy = rel(scn, 1) & 7;
      const bool on = kj < a.w;
      float wx = es_phi<float>((float(kj) - r.x) * a.k.two_over_w, a.k);
      float wy = es_phi<float>((float(kj) - r.y) * a.k.two_over_w, a.k);
      float wz = es_phi<float>((float(kj) - r.z) * a.k.two_over_w, a.k);
      wx = on && !empty ? r.w * wx : 0.f;
      wy = on ? wy : 0.f;
      wz = on ? wz : 0.f;
      rr[kj + dy] = wy;
      rr[(kj + dy + 8) & 15] = 0.f;
      rr[16 + kj + dx] = wx;
      rr[16 + ((kj + dx + 8) & 15)] = 0.f;
      rr[32 + kj] = wz;
    };
    // the operands of one chunk: two products of 4 slots
    struct Ops {
      float av0, bx0, av1, bx1;
      float4 zl0, zh0, zl1, zh1;
    };
    auto operands = [&](const float* rb) {
      const float* r0 = rb + k4 * HM_RS;
      const float* r1 = r0 + 4 * HM_RS;
      Ops o;
      o.av0 = r0[i16];
      o.bx0 = r0[16 + i16];
      o.zl0 = *reinterpret_cast<const float4*>(r0 + 32);
      o.zh0 = *reinterpret_cast<const float4*>(r0 + 36);
      o.av1 = r1[i16];
      o.bx1 = r1[16 + i16];
      o.zl1 = *reinterpret_cast<const float4*>(r1 + 32);
      o.zh1 = *reinterpret_cast<const float4*>(r1 + 36);
      return o;
    };
    int kb[HM_NK + 1];
    kb[0] = 0;
#pragma unroll
    for (int K = 0; K < HM_NK; ++K) kb[K + 1] = __builtin_amdgcn_readfirstlane(wkend[K]);
    auto rbuf = [&](int i) { return &recs[0][0][0] + (2 * h + i) * (HM_CH * HM_RS); };
    fetch(0);
    load();
    values(rbuf(0));
    fetch(HM_CH);
    load();
    fetch(2 * HM_CH);
    Ops cur = operands(rbuf(0));
    int c0 = 0, cb = 0;
    // one wavefront per window: its LDS accesses complete in program order,
    // so the chunk buffers need no barrier
    hm_static_for<0, HM_NK>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      for (; c0 < kb[K + 1]; c0 += HM_CH, cb ^= 1) {
        values(rbuf(cb ^ 1));
        load();
        fetch(c0 + 3 * HM_CH);
        const Ops nxt = operands(rbuf(cb ^ 1));
        hm_products<K>(acc, cur.av0, cur.bx0, cur.zl0, cur.zh0);
        hm_products<K>(acc, cur.av1, cur.bx1, cur.zl1, cur.zh1);
        cur = nxt;
      }
    });
    __syncthreads();
  }
  // register r of lane (k4, i16) holds row y = 4 k4 + r, column x = 8 h +
  // i16 of every plane; rows 0..14 carry points (row 15 stays zero).  The
  // tile [plane][row][24 columns] is the right window's columns 8..23 and the
  // left window's 0..7, then the left window's 8..15 added (column 23 is 0).
  float* tile = reinterpret_cast<float*>(smem);
  constexpr int RS = 15 * HS_X;  // floats per plane in LDS
  const int xo = 8 * h + i16;
  if (h == 1 || i16 < 8) {
#pragma unroll
    for (int z = 0; z < HP_Z; ++z)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < 3 || k4 < 3) tile[z * RS + (4 * k4 + r) * HS_X + xo] = acc[z][r];
  }
  __syncthreads();
  if (h == 0 && i16 >= 8) {
#pragma unroll
    for (int z = 0; z < HP_Z; ++z)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < 3 || k4 < 3) tile[z * RS + (4 * k4 + r) * HS_X + xo] += acc[z][r];
  }
  __syncthreads();
  // rows 0..14 of each plane to the halo tile (row stride HS_X, plane stride
  // HS_Y rows), 16-byte units
  constexpr int UPR = HS_X / 4;
  constexpr int NU = HP_Z * 15 * UPR;
  for (int u = tid; u < NU; u += 128) {
    const int row = u / UPR, xu = u - row * UPR;
    const int z = row / 15, y = row - z * 15;
    reinterpret_cast<float4*>(out + (z * HS_Y + y) * HS_X)[xu] = reinterpret_cast<const float4*>(tile + row * HS_X)[xu];
  }
}

// The matrix-core halo spreader with the y rows in a ring (FP32 strengths,
// 16 x 8 x 8 bins): like k_spread3_halo_mfma2 a two-wavefront workgroup per
// (bin, slice), wavefront h = x window h (columns [8 h, 8 h + 16)), but a
// window's points are ranked by (uy, uz >> 1), the y and z of their first
// cell, each run padded to whole chunks of 8 slots (empty slots weigh 0),
// two groups of 4.  The 4 points of a group share uy, so their y supports
// are the same 8 rows, and a product's 16 rows are 2 planes x those 8 rows:
//   D_P[(dz, r)][j] += sum_p  wy_p[(r - uy) mod 8] wz_p[2P + dz - uz_p]  *  c_p wx_p[8 h + j - ux_p]
// for the plane pairs P = uz >> 1 .. (uz >> 1) + 4 (5 products per group: a
// point's 8 planes span 4 pairs, or 5 for odd uz).  Ring row r holds the tile
// row y = r (mod 8) of [uy, uy + 8): once the runs of uy are done, row uy is
// complete and is written out from the registers, and its ring row starts
// over as row uy + 8.  Half of every product's multiply-adds carry points (a
// quarter in the two-window spreader).  Each window writes its own tile
// [15 rows][16 planes][16 columns]; the merge sums the windows of
// neighbouring bins (stride 8 in x) like halo tiles.  Slices of more than WIN
// points take rounds, the rounds after the first adding to the tile.
// Ranking is stable (ballot ranks in index order): the summation order is
// fixed and the result bitwise reproducible.
constexpr int HR_NK = 32;  // keys: (uy, uz >> 1)
constexpr int HR_CH = 8;   // slots per chunk: two groups of 4 (the K of a product), kernel values evaluated together
constexpr int HR_RS = 48;  // floats per slot record: y[8] (ring order), x[16], z even [24..28], z odd [32..36]

template <int P0>
__device__ __forceinline__ void hr_products(hm_f4 (&acc)[8], float ay, float bx, const float4& z4, float z5) {
  const float z[5] = {z4.x, z4.y, z4.z, z4.w, z5};
#pragma unroll
  for (int i = 0; i < 5; ++i)
    acc[P0 + i] = __builtin_amdgcn_mfma_f32_16x16x4f32(ay * z[i], bx, acc[P0 + i], 0, 0, 0);
}

// ring row R of every plane pair to tile row `row` (< 15), then cleared.
// Row m = 8 dz + R of a product sits in register R % 4 of the lanes with
// k4 = 2 dz + R / 4.  The row's 16 planes x 16 columns go through LDS (buf,
// 256 floats) and out as one contiguous 1 KB store (whole cache lines).
template <int R>
__device__ __forceinline__ void hr_flush(hm_f4 (&acc)[8], float* __restrict__ out, float* __restrict__ buf, int row,
                                         int lane, bool add) {
  constexpr int REG = R & 3;
  const int k4 = lane >> 4, i16 = lane & 15;
  const bool mine = (k4 & 1) == (R >> 2);
  // every lane writes the ring row, its own value or the one of lane l ^ 16
  // (lanes l and l ^ 16 write the same value): no divergent branch around
  // the accumulator reads
  float* b = buf + (k4 >> 1) * 16 + i16;
#pragma unroll
  for (int P = 0; P < 8; ++P) {
    const float own = acc[P][REG];
    const float other = __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(own), 0x401F));
    b[P * 32] = mine ? own : other;
  }
#pragma unroll
  for (int P = 0; P < 8; ++P) acc[P][REG] = mine ? 0.f : acc[P][REG];
  float4 v = reinterpret_cast<const float4*>(buf)[lane];
  float4* o = reinterpret_cast<float4*>(out + row * 256) + lane;
  if (add) {
    const float4 q = *o;
    v.x += q.x;
    v.y += q.y;
    v.z += q.z;
    v.w += q.w;
  }
  *o = v;
}

template <int WIN>
__global__ void __launch_bounds__(128)
    k_spread3_halo_ring(const uint32_t* __restrict__ scell, const float4* __restrict__ rec,
                        const int* __restrict__ bin_start, SpreadDev<float> a, float* __restrict__ halo) {
  static_assert(WIN % 128 == 0, "a round is staged by 128 threads");
  constexpr int MAXSLOT = WIN + (HR_CH - 1) * HR_NK;
  constexpr int WPL = WIN / 64;   // round points per lane (ranking, per wavefront)
  constexpr int SPL = WIN / 128;  // round points per thread (staging)
  __shared__ __align__(16) float4 wrec[WIN];
  __shared__ uint32_t wsc[WIN];
  __shared__ uint16_t perm[2][MAXSLOT];
  __shared__ __align__(16) float recs[2][2][HR_CH * HR_RS];  // [window][buffer]: kernel values of two chunks
  __shared__ int kend[2][HR_NK];
  __shared__ int run[2][HR_NK];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int h = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: the window tile base is scalar
  const int b = blockIdx.x;
  const int nsl = gridDim.y, sl = blockIdx.y;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * sl) / nsl);
  const int end = beg0 + (int)(((int64_t)len * (sl + 1)) / nsl);
  float* out = halo + ((int64_t)(b * nsl + sl) * 2 + h) * HR_WT;
  if (beg == end) {
    for (int e = lane; e < HR_WT; e += 64) out[e] = 0.f;
    return;
  }
  const uint32_t cmask = (1u << a.cbits) - 1u;
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  auto rel = [&](uint32_t sc, int q) {
    int u = (int)((sc >> (q * a.cbits)) & cmask) - t0[q];
    if (u < 0) u += a.nf[q];
    return u;
  };
  const int i16 = lane & 15, k4 = lane >> 4;
  const int kj = lane & 7;
  const uint64_t lt = (1ull << lane) - 1ull;
  uint16_t* wperm = perm[h];
  int* wrun = run[h];
  int* wkend = kend[h];
  hm_f4 acc[8];
#pragma unroll
  for (int P = 0; P < 8; ++P) acc[P] = hm_f4{0.f, 0.f, 0.f, 0.f};
  for (int ws = beg; ws < end; ws += WIN) {
    const int wn = min(WIN, end - ws);
    // a later round adds to the tile the earlier ones wrote (the same lanes
    // read back the addresses they stored: their stores complete first)
    const bool add = ws != beg;
    if (add) __threadfence();
    // the thread index, opaque here: the staging and ranking addresses are
    // formed in every round instead of being kept in registers across the
    // products (a slice rarely takes more than one round)
    int tq = tid;
    asm volatile("" : "+v"(tq));
    const int lq = tq & 63;
    {
      float4 wr[SPL];
      uint32_t wk[SPL];
#pragma unroll
      for (int u = 0; u < SPL; ++u) {
        const int e = u * 128 + tq;
        wk[u] = e < wn ? scell[ws + e] : 0u;
        wr[u] = e < wn ? rec[ws + e] : float4{0.f, 0.f, 0.f, 0.f};
      }
      __syncthreads();  // both windows are done with the previous round's records
#pragma unroll
      for (int u = 0; u < SPL; ++u) {
        wrec[u * 128 + tq] = wr[u];
        wsc[u * 128 + tq] = wk[u];
      }
    }
    if (lane < HR_NK) wrun[lane] = 0;
    __syncthreads();
    // key (uy, uz >> 1) for the points of this wavefront's window; -1 others
    int wkey[WPL];
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
      const int e = u * 64 + lq;
      const uint32_t sc = wsc[e];
      const bool mine = e < wn && ((rel(sc, 0) >> 3) & 1) == h;
      wkey[u] = mine ? ((rel(sc, 1) & 7) << 2) | ((rel(sc, 2) & 7) >> 1) : -1;
    }
#pragma unroll
    for (int u = 0; u < WPL; ++u)
      if (wkey[u] >= 0) atomicAdd(&wrun[wkey[u]], 1);
    wave_lds_order();
    const int cnt = lane < HR_NK ? wrun[lane] : 0;
    const int pad = (cnt + HR_CH - 1) / HR_CH * HR_CH;
    int inc = pad;
#pragma unroll
    for (int off = 1; off < HR_NK; off <<= 1) {
      const int u = __builtin_amdgcn_ds_bpermute(max(lq - off, 0) << 2, inc);
      if (lq >= off) inc += u;
    }
    const int nslot = __builtin_amdgcn_readlane(inc, HR_NK - 1);
    const int st = inc - pad;
    wave_lds_order();
    if (lane < HR_NK) {
      wrun[lane] = st;
      wkend[lane] = st + pad;
    }
    for (int q = lq; q < nslot; q += 64) wperm[q] = 0xFFFF;
    wave_lds_order();
    // stable placement, 64 points a round per wavefront
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
      if (u * 64 >= wn) break;
      const int key = wkey[u];
      const bool valid = key >= 0;
      uint64_t same = __ballot(valid);
#pragma unroll
      for (int bt = 0; bt < 5; ++bt) {
        const bool bit = (key >> bt) & 1;
        const uint64_t mk = __ballot(bit);
        same &= bit ? mk : ~mk;
      }
      const int rank = __popcll(same & lt);
      if (valid) wperm[min(wrun[key] + rank, MAXSLOT - 1)] = (uint16_t)(u * 64 + lq);
      wave_lds_order();
      if (valid && rank == 0) wrun[key] += __popcll(same);
      wave_lds_order();
    }
    // slot -> point: the permutation entry of chunk c + 3 is read one chunk
    // before the record of that chunk, so neither waits on the other
    uint32_t scn;
    float4 rcn;
    bool emp;
    int pslot, praw;
    auto fetch = [&](int c0) {
      pslot = c0 + (lane >> 3);
      praw = (int)wperm[min(pslot, MAXSLOT - 1)];
    };
    // (nothing here waits for the loads: values() uses them a chunk later)
    auto load = [&]() {
      emp = pslot >= nslot || praw >= wn;
      const int pq = emp ? 0 : praw;
      scn = wsc[pq];
      rcn = wrec[pq];
    };
    // lane (slot, kj): the slot's y value kj at its ring row, x value kj at
    // its window column, z value kj at position kj + (uz & 1) of the parity
    // split z list (positions 0..9, the two not written are zero)
    auto values = [&](float* rb) {
      float* rr = rb + (lane >> 3) * HR_RS;
      const float4 r = rcn;
      const bool empty = emp;
      const int dx = rel(scn, 0) & 7;
      const int uy = rel(scn, 1) & 7;
      const int par = rel(scn, 2) & 1;
      const bool on = kj < a.w;
      float wx = es_phi<float>((float(kj) - r.x) * a.k.two_over_w, a.k);
      float wy = es_phi<float>((float(kj) - r.y) * a.k.two_over_w, a.k);
      float wz = es_phi<float>((float(kj) - r.z) * a.k.two_over_w, a.k);
      wx = on && !empty ? r.w * wx : 0.f;
      wy = on ? wy : 0.f;
      wz = on ? wz : 0.f;
      rr[(kj + uy) & 7] = wy;
      rr[8 + kj + dx] = wx;
      rr[8 + ((kj + dx + 8) & 15)] = 0.f;
      const int pos = kj + par;
      rr[24 + 8 * (pos & 1) + (pos >> 1)] = wz;
      // the two zeros: position 9 (lane 0) and position 0 (odd uz) or 8 (even
      // uz) (lane 1); the other lanes write to the record's unused tail
      rr[kj == 0 ? 36 : kj == 1 ? (par ? 24 : 28) : 40 + kj] = 0.f;
    };
    struct Ops {
      float ay, bx;
      float4 z4;
      float z5;
    };
    auto rbuf = [&](int i) { return &recs[0][0][0] + (2 * h + i) * (HR_CH * HR_RS); };
    // the operands of group gg: A row i16 = (dz, r), B column i16
    auto operands = [&](int gg) {
      const float* rr = rbuf((gg >> 1) & 1) + (4 * (gg & 1) + k4) * HR_RS;
      const float* zp = rr + 24 + 8 * (i16 >> 3);
      Ops o;
      o.ay = rr[i16 & 7];
      o.bx = rr[8 + i16];
      o.z4 = *reinterpret_cast<const float4*>(zp);
      o.z5 = zp[4];
      return o;
    };
    int ce[HR_NK];
#pragma unroll
    for (int K = 0; K < HR_NK; ++K) ce[K] = __builtin_amdgcn_readfirstlane(wkend[K]) / HR_CH;
    fetch(0);
    load();
    values(rbuf(0));
    fetch(HR_CH);
    load();
    fetch(2 * HR_CH);
    // chunk c = groups 2c (operands in oa) and 2c + 1 (in ob): each set is
    // read a group or more before its products
    Ops oa = operands(0), ob;
    int c = 0;
    // one wavefront per window: its LDS accesses complete in program order,
    // so the chunk buffers need no barrier
    hm_static_for<0, 8>([&](auto rc) {
      constexpr int R = decltype(rc)::value;
      hm_static_for<0, 4>([&](auto jc) {
        constexpr int J = decltype(jc)::value;
        for (; c < ce[R * 4 + J]; ++c) {
          ob = operands(2 * c + 1);
          // values of chunk c + 1, record of chunk c + 2, permutation entry
          // of chunk c + 3
          values(rbuf((c + 1) & 1));
          load();
          fetch((c + 3) * HR_CH);
          hr_products<J>(acc, oa.ay, oa.bx, oa.z4, oa.z5);
          oa = operands(2 * c + 2);
          hr_products<J>(acc, ob.ay, ob.bx, ob.z4, ob.z5);
        }
      });
      // the chunk buffer not holding chunk c is free here
      hr_flush<R>(acc, out, rbuf((c + 1) & 1), R, lane, add);
    });
    // rows 8..14 (ring rows 0..6); ring row 7 (tile row 15) stays empty
    hm_static_for<0, 7>([&](auto rc) {
      constexpr int R = decltype(rc)::value;
      hr_flush<R>(acc, out, rbuf((c + 1) & 1), 8 + R, lane, add);
    });
  }
}

// tabs[a][c * 4 + i] = (bin coordinate << 6) | u of the i-th halo tile
// covering fine-grid coordinate c along dimension a, -1 terminated
template <typename T, int TY>
__global__ void k_merge_halo(const T* __restrict__ halo, const int* __restrict__ tabs, int parts, SpreadDev<T> a,
                             T* __restrict__ grid) {
  constexpr int HS_Y = HaloShape<TY>::HS_Y;
  constexpr int HT_ELEMS = HaloShape<TY>::ELEMS;
  const int64_t n = (int64_t)a.nf[0] * a.nf[1] * a.nf[2];
  const int* tx = tabs;
  const int* ty = tabs + 4 * a.nf[0];
  const int* tz = ty + 4 * a.nf[1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int gx = (int)(i % a.nf[0]);
    const int gy = (int)((i / a.nf[0]) % a.nf[1]);
    const int gz = (int)(i / ((int64_t)a.nf[0] * a.nf[1]));
    T acc = T(0);
    for (int iz = 0; iz < 4; ++iz) {
      const int ez = tz[gz * 4 + iz];
      if (ez < 0) break;
      for (int iy = 0; iy < 4; ++iy) {
        const int ey = ty[gy * 4 + iy];
        if (ey < 0) break;
        for (int ix = 0; ix < 4; ++ix) {
          const int ex = tx[gx * 4 + ix];
          if (ex < 0) break;
          const int64_t bin = (ex >> 6) + (int64_t)a.ntile[0] * ((ey >> 6) + (int64_t)a.ntile[1] * (ez >> 6));
          const int64_t off = ((ez & 63) * HS_Y + (ey & 63)) * HS_X + (ex & 63);
          for (int s = 0; s < parts; ++s) acc += halo[(bin * parts + s) * HT_ELEMS + off];
        }
      }
    }
    grid[i] = acc;
  }
}

// k_merge_halo for P slices per bin: a cell covered by at most two tiles per
// dimension (every cell unless a partial last tile sits next to the wrap)
// issues all its up to 8 P halo loads at once and sums them in the same order
// as k_merge_halo (z, y, x entries, then slices), so the grid is bitwise the
// same; 32-bit offsets (the host checks the halo buffer size).
template <typename T, int TY, int P>
__global__ void __launch_bounds__(256)
    k_merge_halo_p(const T* __restrict__ halo, const int* __restrict__ tabs, SpreadDev<T> a, T* __restrict__ grid) {
  constexpr int HS_Y = HaloShape<TY>::HS_Y;
  constexpr int HT = HaloShape<TY>::ELEMS;
  const int nfx = a.nf[0], nfy = a.nf[1];
  const int64_t n = (int64_t)nfx * nfy * a.nf[2];
  const int4* tx = reinterpret_cast<const int4*>(tabs);
  const int4* ty = tx + nfx;
  const int4* tz = ty + nfy;
  const int nt0 = a.ntile[0], nt01 = a.ntile[0] * a.ntile[1];
  // workgroups go round-robin to the 8 XCDs; each XCD takes a contiguous
  // range of cells (a slab of z planes), so the tiles it reads are shared in
  // its own L2
  const int nbk = gridDim.x, xq = nbk >> 3, xr = nbk & 7, xc = blockIdx.x & 7;
  const int blk = xc * xq + (xc < xr ? xc : xr) + (blockIdx.x >> 3);
  for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int gx = (int)(i % nfx);
    const int64_t r = i / nfx;
    const int gy = (int)(r % nfy);
    const int gz = (int)(r / nfy);
    const int4 ex = tx[gx], ey = ty[gy], ez = tz[gz];
    T acc = T(0);
    if (ex.z < 0 && ey.z < 0 && ez.z < 0) {
      // per dimension: (bin part, cell part) of entries 0 and 1
      const int bx[2] = {ex.x >> 6, ex.y >> 6}, ux[2] = {ex.x & 63, ex.y & 63};
      const int by[2] = {(ey.x >> 6) * nt0, (ey.y >> 6) * nt0};
      const int uy[2] = {(ey.x & 63) * HS_X, (ey.y & 63) * HS_X};
      const int bz[2] = {(ez.x >> 6) * nt01, (ez.y >> 6) * nt01};
      const int uz[2] = {(ez.x & 63) * (HS_Y * HS_X), (ez.y & 63) * (HS_Y * HS_X)};
      const bool vx1 = ex.y >= 0, vy1 = ey.y >= 0, vz1 = ez.y >= 0;
      T v[8][P];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int iz = c >> 2, iy = (c >> 1) & 1, ix = c & 1;
        const bool ok = (ix == 0 || vx1) && (iy == 0 || vy1) && (iz == 0 || vz1);
        const int base = (bx[ix] + by[iy] + bz[iz]) * (P * HT) + ux[ix] + uy[iy] + uz[iz];
#pragma unroll
        for (int sl = 0; sl < P; ++sl) {
          v[c][sl] = ok ? halo[base + sl * HT] : T(0);
        }
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int iz = c >> 2, iy = (c >> 1) & 1, ix = c & 1;
        if ((ix == 0 || vx1) && (iy == 0 || vy1) && (iz == 0 || vz1)) {
#pragma unroll
          for (int sl = 0; sl < P; ++sl) acc += v[c][sl];
        }
      }
    } else {
      const int exa[4] = {ex.x, ex.y, ex.z, ex.w}, eya[4] = {ey.x, ey.y, ey.z, ey.w};
      const int eza[4] = {ez.x, ez.y, ez.z, ez.w};
      for (int iz = 0; iz < 4 && eza[iz] >= 0; ++iz)
        for (int iy = 0; iy < 4 && eya[iy] >= 0; ++iy)
          for (int ix = 0; ix < 4 && exa[ix] >= 0; ++ix) {
            const int bin = (exa[ix] >> 6) + nt0 * (eya[iy] >> 6) + nt01 * (eza[iz] >> 6);
            const int off = ((eza[iz] & 63) * HS_Y + (eya[iy] & 63)) * HS_X + (exa[ix] & 63);
            for (int sl = 0; sl < P; ++sl) acc += halo[(bin * P + sl) * HT + off];
          }
    }
    grid[i] = acc;
  }
}

// Merge of the ring spreader's window tiles: per (bin, slice) two windows
// [15 rows][16 planes][16 columns] at x offsets 0 and 8 of the bin.  Windows
// follow each other with stride 8 in x as the bins do in y and z, so a cell
// is covered by at most two windows, two bin rows and two bin planes (every
// cell unless a partial last tile sits next to the wrap); its up to 8 P
// loads are issued together and summed in a fixed order (z, y, x entries,
// then slices).  The x entries come from the window table (window index
// 2 bx + h); 32-bit offsets (the host checks the halo buffer size).
template <int P>
__global__ void __launch_bounds__(256)
    k_merge_win(const float* __restrict__ halo, const int* __restrict__ tabs, SpreadDev<float> a,
                float* __restrict__ grid) {
  constexpr int WT = HR_WT;
  constexpr int BT = 2 * P * WT;  // floats per bin
  const int nfx = a.nf[0], nfy = a.nf[1], nfz = a.nf[2];
  const int64_t n = (int64_t)nfx * nfy * nfz;
  const int4* ty = reinterpret_cast<const int4*>(tabs) + nfx;
  const int4* tz = ty + nfy;
  const int4* tw = tz + nfz;
  const int nt0 = a.ntile[0], nt01 = a.ntile[0] * a.ntile[1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int gx = (int)(i % nfx);
    const int64_t r = i / nfx;
    const int gy = (int)(r % nfy);
    const int gz = (int)(r / nfy);
    const int4 ex = tw[gx], ey = ty[gy], ez = tz[gz];
    float acc = 0.f;
    if (ex.z < 0 && ey.z < 0 && ez.z < 0) {
      const int ox[2] = {(ex.x >> 7) * BT + ((ex.x >> 6) & 1) * WT + (ex.x & 63),
                         (ex.y >> 7) * BT + ((ex.y >> 6) & 1) * WT + (ex.y & 63)};
      const int oy[2] = {(ey.x >> 6) * nt0 * BT + (ey.x & 63) * 256, (ey.y >> 6) * nt0 * BT + (ey.y & 63) * 256};
      const int oz[2] = {(ez.x >> 6) * nt01 * BT + (ez.x & 63) * 16, (ez.y >> 6) * nt01 * BT + (ez.y & 63) * 16};
      const bool vx1 = ex.y >= 0, vy1 = ey.y >= 0, vz1 = ez.y >= 0;
      float v[8][P];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int iz = c >> 2, iy = (c >> 1) & 1, ix = c & 1;
        const bool ok = (ix == 0 || vx1) && (iy == 0 || vy1) && (iz == 0 || vz1);
        const int base = ox[ix] + oy[iy] + oz[iz];
#pragma unroll
        for (int s = 0; s < P; ++s) v[c][s] = ok ? halo[base + s * 2 * WT] : 0.f;
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int iz = c >> 2, iy = (c >> 1) & 1, ix = c & 1;
        if ((ix == 0 || vx1) && (iy == 0 || vy1) && (iz == 0 || vz1)) {
#pragma unroll
          for (int s = 0; s < P; ++s) acc += v[c][s];
        }
      }
    } else {
      const int exa[4] = {ex.x, ex.y, ex.z, ex.w}, eya[4] = {ey.x, ey.y, ey.z, ey.w};
      const int eza[4] = {ez.x, ez.y, ez.z, ez.w};
      for (int iz = 0; iz < 4 && eza[iz] >= 0; ++iz)
        for (int iy = 0; iy < 4 && eya[iy] >= 0; ++iy)
          for (int ix = 0; ix < 4 && exa[ix] >= 0; ++ix) {
            const int bin = (exa[ix] >> 7) + nt0 * (eya[iy] >> 6) + nt01 * (eza[iz] >> 6);
            const int off = ((exa[ix] >> 6) & 1) * WT + ((eya[iy] & 63) * 16 + (eza[iz] & 63)) * 16 + (exa[ix] & 63);
            for (int s = 0; s < P; ++s) acc += halo[bin * BT + s * 2 * WT + off];
          }
    }
    grid[i] = acc;
  }
}

// ---------------------------------------------------------------------------
// 2D input-driven spreading (w <= 8, real strengths): a bin's supports lie in
// its (tile + 7)^2 halo tile; one wavefront per (bin, slice, pattern) adds its
// points there in LDS (lane (ja, jb) = cell (x + ja, y + jb) of the support;
// rows of stride hs, a multiple of 8 with an odd factor so the lane pattern is
// bank-conflict-free), two points per iteration sharing the kernel
// evaluation; every point is handled once.  k_merge_halo2 then sums, per
// fine cell, the halo tiles covering it in a fixed order (reproducible).
// ---------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(64)
    k_spread2_halo(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                   const int* __restrict__ bin_start, SpreadDev<T> a, int hs, int hrows, T* __restrict__ halo) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const int helems = hs * hrows;
  const int lane = threadIdx.x;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const int nbins = a.ntile[0] * a.ntile[1];
  bin_start += (int64_t)blockIdx.y * nbins;
  const int b = blockIdx.x;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.z) / gridDim.z);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.z + 1)) / gridDim.z);
  T* out = halo + (((int64_t)blockIdx.y * nbins + b) * gridDim.z + blockIdx.z) * helems;
  for (int e = lane; e < helems; e += 64) tile[e] = T(0);
  if (beg == end) {
    for (int e = lane; e < helems; e += 64) out[e] = T(0);
    return;
  }
  const int t0x = (b % a.ntile[0]) * a.tile[0];
  const int t0y = (b / a.ntile[0]) * a.tile[1];
  __syncthreads();
  const int hb = lane >> 5;         // kernel-value half: 0 point A, 1 point B
  const int kd = (lane & 31) >> 3;  // dimension of this lane's kernel value
  const int kj = lane & 7;
  const bool kact = kd < 2 && kj < w;
  const int ja = lane & 7, jb = lane >> 3;
  for (int base = beg; base < end; base += 64) {
    const int cnt = min(64, end - base);
    uint32_t sc = 0;
    typename Vec4<T>::type rr;
    rr.x = rr.y = rr.z = rr.w = T(0);
    if (lane < cnt) {
      sc = scell[base + lane];
      rr = rec[base + lane];
    }
    for (int p = 0; p < cnt; p += 2) {
      const int q = p + 1 < cnt ? p + 1 : p;  // odd tail: B repeats A with weight 0
      const uint32_t sa = lane_read_u(sc, p), sb = lane_read_u(sc, q);
      const T cA = lane_read(rr.w, p);
      const T cB = p + 1 < cnt ? lane_read(rr.w, q) : T(0);
      const T oAx = lane_read(rr.x, p), oAy = lane_read(rr.y, p);
      const T oBx = lane_read(rr.x, q), oBy = lane_read(rr.y, q);
      const T off = hb ? (kd == 0 ? oBx : oBy) : (kd == 0 ? oAx : oAy);
      const T kv = kact ? es_phi<T>((T(kj) - off) * a.k.two_over_w, a.k) : T(0);
      const T vA = lane_get(kv, ja) * lane_get(kv, 8 + jb) * cA;
      const T vB = lane_get(kv, 32 + ja) * lane_get(kv, 40 + jb) * cB;
      int uAx = (int)(sa & cmask) - t0x, uAy = (int)((sa >> a.cbits) & cmask) - t0y;
      int uBx = (int)(sb & cmask) - t0x, uBy = (int)((sb >> a.cbits) & cmask) - t0y;
      tile[(uAy + jb) * hs + uAx + ja] += vA;
      tile[(uBy + jb) * hs + uBx + ja] += vB;
    }
  }
  __syncthreads();
  for (int e = lane; e < helems; e += 64) out[e] = tile[e];
}

// The 2D halo spreading of FP32 strengths on the matrix cores (16 x 16 bins,
// w <= 8), one wavefront per (bin, pattern, slice).  The 23 x 23 halo tile
// is covered by four 16 x 16 windows [8 hx, 8 hx + 16) x [8 hy, 8 hy + 16)
// (hx, hy = 0, 1: the bin's halves the first cell lies in); a window is
//   D[y][x] = sum_p  wy_p[y - dy_p]  *  c_p wx_p[x - dx_p]
// (dx, dy = the first cell's offset in its window, kernel values zero
// outside [0, 8)): one v_mfma_f32_16x16x4f32 per 4 points, the four windows
// held in accumulator registers.  The slice's points are ranked by window
// (stable ballot ranks, runs padded to whole chunks of 8 slots with empty
// slots of weight 0), the kernel values of a chunk evaluated in LDS records
// while the previous chunk's products run.  At the end the windows are added
// into an LDS tile in a fixed order and the tile is written out: the sum
// order is fixed, the result bitwise reproducible.
constexpr int H2_WIN = 384;  // points ranked (and staged in LDS) per window of a slice
constexpr int H2_CH = 8;     // slots per chunk (one key, two products)
constexpr int H2_RS = 48;    // floats per slot record: ay[16], bx[16]
constexpr int H2_NK = 4;     // keys: hx + 2 hy
constexpr int H2_HS = 24;    // halo row stride of 16 x 16 bins

__global__ void __launch_bounds__(64)
    k_spread2_halo_mfma(const uint32_t* __restrict__ scell, const float4* __restrict__ rec,
                        const int* __restrict__ bin_start, SpreadDev<float> a, int hrows, float* __restrict__ halo) {
  constexpr int MAXSLOT = H2_WIN + (H2_CH - 1) * H2_NK;
  constexpr int WPL = H2_WIN / 64;
  __shared__ __align__(16) float recs[2][H2_CH * H2_RS];
  __shared__ __align__(16) float4 wrec[H2_WIN];
  __shared__ uint32_t wsc[H2_WIN];
  __shared__ uint16_t perm[MAXSLOT];
  __shared__ int kend[H2_NK];
  __shared__ int run[H2_NK];
  __shared__ float tile[24 * H2_HS];
  const int lane = threadIdx.x;
  const int helems = H2_HS * hrows;
  const int nbins = a.ntile[0] * a.ntile[1];
  bin_start += (int64_t)blockIdx.y * nbins;
  const int b = blockIdx.x;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.z) / gridDim.z);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.z + 1)) / gridDim.z);
  float* out = halo + (((int64_t)blockIdx.y * nbins + b) * gridDim.z + blockIdx.z) * helems;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const int t0[2] = {(b % a.ntile[0]) * a.tile[0], (b / a.ntile[0]) * a.tile[1]};
  auto rel = [&](uint32_t sc, int q) {
    int u = (int)((sc >> (q * a.cbits)) & cmask) - t0[q];
    if (u < 0) u += a.nf[q];
    return u;
  };
  const int i16 = lane & 15, k4 = lane >> 4;
  const int kj = lane & 7;
  const uint64_t lt = (1ull << lane) - 1ull;
  hm_f4 acc[H2_NK];
#pragma unroll
  for (int k = 0; k < H2_NK; ++k) acc[k] = hm_f4{0.f, 0.f, 0.f, 0.f};
  for (int ws = beg; ws < end; ws += H2_WIN) {
    const int wn = min(H2_WIN, end - ws);
    int wkey[WPL];
    {
      float4 wr[WPL];
#pragma unroll
      for (int u = 0; u < WPL; ++u) {
        const int e = u * 64 + lane;
        wkey[u] = e < wn ? (int)scell[ws + e] : 0;
        wr[u] = e < wn ? rec[ws + e] : float4{0.f, 0.f, 0.f, 0.f};
      }
      __syncthreads();  // the previous window's chunks are done with wrec / wsc
#pragma unroll
      for (int u = 0; u < WPL; ++u) {
        wrec[u * 64 + lane] = wr[u];
        wsc[u * 64 + lane] = (uint32_t)wkey[u];
      }
    }
    if (lane < H2_NK) run[lane] = 0;
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
      const uint32_t sc = (uint32_t)wkey[u];
      wkey[u] = u * 64 + lane < wn ? ((rel(sc, 0) >> 3) & 1) + 2 * ((rel(sc, 1) >> 3) & 1) : -1;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < WPL; ++u)
      if (wkey[u] >= 0) atomicAdd(&run[wkey[u]], 1);
    __syncthreads();
    const int cnt = lane < H2_NK ? run[lane] : 0;
    const int pad = (cnt + H2_CH - 1) / H2_CH * H2_CH;
    int inc = pad;
#pragma unroll
    for (int off = 1; off < H2_NK; off <<= 1) {
      const int u = __shfl_up(inc, off, 64);
      if (lane >= off) inc += u;
    }
    const int nslot = __builtin_amdgcn_readfirstlane(__shfl(inc, H2_NK - 1, 64));
    const int st = inc - pad;
    __syncthreads();
    if (lane < H2_NK) {
      run[lane] = st;
      kend[lane] = st + pad;
    }
    for (int q = lane; q < nslot; q += 64) perm[q] = 0xFFFF;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < WPL; ++u) {
      if (u * 64 >= wn) break;
      const int key = wkey[u];
      const bool valid = key >= 0;
      uint64_t same = __ballot(valid);
#pragma unroll
      for (int bt = 0; bt < 2; ++bt) {
        const bool bit = (key >> bt) & 1;
        const uint64_t mk = __ballot(bit);
        same &= bit ? mk : ~mk;
      }
      const int rank = __popcll(same & lt);
      if (valid) perm[run[key] + rank] = (uint16_t)(u * 64 + lane);
      __syncthreads();
      if (valid && rank == 0) run[key] += __popcll(same);
      __syncthreads();
    }
    // branch-free chunk preparation, as in k_spread3_halo_mfma
    uint32_t scn;
    float4 rcn;
    auto load = [&](int c0) {
      const int q = c0 + (lane >> 3);
      const int praw = (int)perm[min(q, MAXSLOT - 1)];
      const bool empty = q >= nslot || praw >= wn;
      const int pq = empty ? 0 : praw;
      const uint32_t sc = wsc[pq];
      rcn = wrec[pq];
      scn = empty ? 0xFFFFFFFFu : sc;
    };
    auto values = [&](float* rb) {
      float* rr = rb + (lane >> 3) * H2_RS;
      const float4 r = rcn;
      const bool empty = scn == 0xFFFFFFFFu;
      const int dx = rel(scn, 0) & 7;
      const int dy = rel(scn, 1) & 7;
      const bool on = kj < a.w;
      float wx = es_phi<float>((float(kj) - r.x) * a.k.two_over_w, a.k);
      float wy = es_phi<float>((float(kj) - r.y) * a.k.two_over_w, a.k);
      wx = on && !empty ? r.w * wx : 0.f;
      wy = on ? wy : 0.f;
      rr[kj + dy] = wy;
      rr[(kj + dy + 8) & 15] = 0.f;
      rr[16 + kj + dx] = wx;
      rr[16 + ((kj + dx + 8) & 15)] = 0.f;
    };
    int kb[H2_NK + 1];
    kb[0] = 0;
#pragma unroll
    for (int K = 0; K < H2_NK; ++K) kb[K + 1] = __builtin_amdgcn_readfirstlane(kend[K]);
    load(0);
    values(recs[0]);
    load(H2_CH);
    int c0 = 0, cb = 0;
    hm_static_for<0, H2_NK>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      for (; c0 < kb[K + 1]; c0 += H2_CH, cb ^= 1) {
        const float* r0 = recs[cb] + k4 * H2_RS;
        const float* r1 = r0 + 4 * H2_RS;
        const float av0 = r0[i16], bx0 = r0[16 + i16];
        const float av1 = r1[i16], bx1 = r1[16 + i16];
        values(recs[cb ^ 1]);
        load(c0 + 2 * H2_CH);
        acc[K] = __builtin_amdgcn_mfma_f32_16x16x4f32(av0, bx0, acc[K], 0, 0, 0);
        acc[K] = __builtin_amdgcn_mfma_f32_16x16x4f32(av1, bx1, acc[K], 0, 0, 0);
      }
    });
    __syncthreads();
  }
  // the four windows into the tile, in key order: register r of lane
  // (k4, i16) holds row 8 hy + 4 k4 + r, column 8 hx + i16
  for (int e = lane; e < 24 * H2_HS; e += 64) tile[e] = 0.f;
  __syncthreads();
#pragma unroll
  for (int K = 0; K < H2_NK; ++K) {
    const int hx = K & 1, hy = K >> 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[(8 * hy + 4 * k4 + r) * H2_HS + 8 * hx + i16] += acc[K][r];
    __syncthreads();
  }
  for (int e = lane; e < helems; e += 64) out[e] = tile[e];
}

// per fine cell of every pattern: the sum of the halo tiles covering it
// (tabs as k_merge_halo's, x then y), slices in order
template <typename T>
__global__ void k_merge_halo2(const T* __restrict__ halo, const int* __restrict__ tabs, int parts, int hs,
                              int helems, SpreadDev<T> a, int batch, T* __restrict__ grid) {
  const int64_t nc = (int64_t)a.nf[0] * a.nf[1];
  const int64_t n = nc * batch;
  const int nbins = a.ntile[0] * a.ntile[1];
  const int* tx = tabs;
  const int* ty = tabs + 4 * a.nf[0];
  // workgroups go round-robin to the 8 XCDs; each XCD takes a contiguous
  // range of cells, so the tiles it reads are shared in its own L2
  const int nbk = gridDim.x, xq = nbk >> 3, xr = nbk & 7, xc = blockIdx.x & 7;
  const int blk = xc * xq + (xc < xr ? xc : xr) + (blockIdx.x >> 3);
  for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pat = i / nc;
    const int64_t c = i - pat * nc;
    const int gx = (int)(c % a.nf[0]);
    const int gy = (int)(c / a.nf[0]);
    T acc = T(0);
    const int4 ex4 = reinterpret_cast<const int4*>(tx)[gx], ey4 = reinterpret_cast<const int4*>(ty)[gy];
    if (parts == 1 && ex4.z < 0 && ey4.z < 0) {
      // at most two tiles per dimension and one slice: the up to 4 loads
      // issued together, summed in the loop's order (y entries, then x)
      const int exa[2] = {ex4.x, ex4.y}, eya[2] = {ey4.x, ey4.y};
      T v[2][2];
#pragma unroll
      for (int iy = 0; iy < 2; ++iy)
#pragma unroll
        for (int ix = 0; ix < 2; ++ix) {
          const bool ok = eya[iy] >= 0 && exa[ix] >= 0;
          const int64_t bin = pat * nbins + (exa[ix] >> 8) + (int64_t)a.ntile[0] * (eya[iy] >> 8);
          const int64_t off = (int64_t)(eya[iy] & 255) * hs + (exa[ix] & 255);
          v[iy][ix] = ok ? halo[bin * helems + off] : T(0);
        }
#pragma unroll
      for (int iy = 0; iy < 2; ++iy)
#pragma unroll
        for (int ix = 0; ix < 2; ++ix)
          if (eya[iy] >= 0 && exa[ix] >= 0) acc += v[iy][ix];
      grid[i] = acc;
      continue;
    }
    for (int iy = 0; iy < 4; ++iy) {
      const int ey = ty[gy * 4 + iy];
      if (ey < 0) break;
      for (int ix = 0; ix < 4; ++ix) {
        const int ex = tx[gx * 4 + ix];
        if (ex < 0) break;
        const int64_t bin = pat * nbins + (ex >> 8) + (int64_t)a.ntile[0] * (ey >> 8);
        const int64_t off = (int64_t)(ey & 255) * hs + (ex & 255);
        for (int sl = 0; sl < parts; ++sl) acc += halo[(bin * parts + sl) * helems + off];
      }
    }
    grid[i] = acc;
  }
}

template <typename T>
static bool launch_spread2_halo(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s) {
  if (!sp.halo2_ok) return false;
  SpreadDev<T> a;
  const GridGeom& g = sp.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.maxc[q] = sp.maxc[q];
  }
  a.cbits = g.cell_bits;
  a.w = sp.es.w;
  a.sx = sp.halo2_hs;
  a.cstride = sp.cand_stride;
  a.k.beta = (T)sp.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(sp.es.beta * M_LOG2E) : (T)sp.es.beta;
  a.k.two_over_w = (T)(2.0 / sp.es.w);
  const int hs = sp.halo2_hs, hrows = sp.halo2_rows;
  const int helems = hs * hrows;
  {
    ProfScope ps(ST_SPREAD, s);
    bool done = false;
    if constexpr (sizeof(T) == 4) {
      if (sp.halo2_mfma) {
        hipLaunchKernelGGL(k_spread2_halo_mfma, dim3(g.nbins, g.batch, sp.halo2_parts), dim3(64), 0, s,
                           p.scell.get(), p.rec.get(), p.bin_start.get(), a, hrows, sp.halo_buf.get());
        done = true;
      }
    }
    if (!done)
      hipLaunchKernelGGL(k_spread2_halo<T>, dim3(g.nbins, g.batch, sp.halo2_parts), dim3(64),
                         sizeof(T) * (size_t)helems, s, p.scell.get(), p.rec.get(), p.bin_start.get(), a, hs, hrows,
                         sp.halo_buf.get());
    FRESCO_HIP_CHECK(hipGetLastError());
  }
  ProfScope pm(ST_MERGE, s);
  const size_t n = (size_t)g.ncells() * g.batch;
  hipLaunchKernelGGL(k_merge_halo2<T>, dim3((unsigned)std::min<size_t>(blocks_for(n, 256), 16384)), dim3(256), 0, s,
                     sp.halo_buf.get(), sp.halo_tab.get(), sp.halo2_parts, hs, helems, a, g.batch, grid);
  FRESCO_HIP_CHECK(hipGetLastError());
  return true;
}

template <typename T>
static bool launch_spread3_halo(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s) {
  if (!sp.halo_ok || sp.mode == Spreader<T>::REG) return false;
  SpreadDev<T> a;
  const GridGeom& g = sp.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.maxc[q] = sp.maxc[q];
  }
  a.cbits = g.cell_bits;
  a.w = sp.es.w;
  a.sx = HS_X;
  a.cstride = sp.cand_stride;
  a.k.beta = (T)sp.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(sp.es.beta * M_LOG2E) : (T)sp.es.beta;
  a.k.two_over_w = (T)(2.0 / sp.es.w);
  const bool t8 = g.tile[1] == 8;
  {
    ProfScope ps(ST_SPREAD, s);
    const dim3 grd(g.nbins, sp.halo_parts);
#define FRESCO_HALO_MW(TY_, CH_)                                                                          \
  hipLaunchKernelGGL((k_spread3_halo_mw<T, TY_, 4, CH_>), grd, dim3(256), 0, s, p.scell.get(), p.rec.get(), \
                     p.bin_start.get(), a, sp.halo_buf.get())
    bool done = false;
    if constexpr (sizeof(T) == 4) {
      if ((sp.mode == Spreader<T>::HALOM || sp.mode == Spreader<T>::HALOR) && t8) {
        const size_t need = (size_t)g.nbins * (sp.halo_parts / 2) *
                            (sp.mode == Spreader<T>::HALOM ? (size_t)HaloShape<8>::ELEMS : 2 * (size_t)HR_WT);
        if (sp.halo_parts < 2 || need > sp.halo_buf.size())
          throw RuntimeFailure("matrix-core halo spreader: halo buffer too small for the slices");
      }
      if (sp.mode == Spreader<T>::HALOM && t8) {
        // one tile per (bin, slice): halo_parts counts 2 windows per slice
        const dim3 grd2(g.nbins, sp.halo_parts / 2);
        hipLaunchKernelGGL(k_spread3_halo_mfma2<1024>, grd2, dim3(128), 0, s, p.scell.get(), p.rec.get(),
                           p.bin_start.get(), a, sp.halo_buf.get());
        done = true;
      } else if (sp.mode == Spreader<T>::HALOR && t8) {
        // one tile per (bin, slice, window)
        const dim3 grd2(g.nbins, sp.halo_parts / 2);
        hipLaunchKernelGGL(k_spread3_halo_ring<1024>, grd2, dim3(128), 0, s, p.scell.get(), p.rec.get(),
                           p.bin_start.get(), a, sp.halo_buf.get());
        done = true;
      }
    }
    if (done) {
    } else if (sp.mode == Spreader<T>::HALO4) {
      if (t8)
        FRESCO_HALO_MW(8, 32);
      else
        FRESCO_HALO_MW(16, 32);
    } else if (t8) {
      hipLaunchKernelGGL((k_spread3_halo<T, 8>), grd, dim3(64), 0, s, p.scell.get(), p.rec.get(), p.bin_start.get(),
                         a, sp.halo_buf.get());
    } else {
      hipLaunchKernelGGL((k_spread3_halo<T, 16>), grd, dim3(64), 0, s, p.scell.get(), p.rec.get(), p.bin_start.get(),
                         a, sp.halo_buf.get());
    }
#undef FRESCO_HALO_MW
    FRESCO_HIP_CHECK(hipGetLastError());
  }
  ProfScope pm(ST_MERGE, s);
  const size_t n = (size_t)g.ncells();
  const dim3 grd((unsigned)std::min<size_t>(blocks_for(n, 256), 16384));
  if constexpr (sizeof(T) == 4) {
    if (sp.mode == Spreader<T>::HALOR && t8) {
      const int sls = sp.halo_parts / 2;
      if ((size_t)g.nbins * sp.halo_parts * HR_WT >= (1u << 31))
        throw RuntimeFailure("halo merge: window tiles exceed 32-bit offsets");
      switch (sls) {
        case 1: hipLaunchKernelGGL(k_merge_win<1>, grd, dim3(256), 0, s, sp.halo_buf.get(), sp.halo_tab.get(), a, grid); break;
        case 2: hipLaunchKernelGGL(k_merge_win<2>, grd, dim3(256), 0, s, sp.halo_buf.get(), sp.halo_tab.get(), a, grid); break;
        case 3: hipLaunchKernelGGL(k_merge_win<3>, grd, dim3(256), 0, s, sp.halo_buf.get(), sp.halo_tab.get(), a, grid); break;
        case 4: hipLaunchKernelGGL(k_merge_win<4>, grd, dim3(256), 0, s, sp.halo_buf.get(), sp.halo_tab.get(), a, grid); break;
        default: throw RuntimeFailure("halo merge: ring spreader slices must be 1..4");
      }
      FRESCO_HIP_CHECK(hipGetLastError());
      return true;
    }
  }
  const size_t hel = t8 ? HaloShape<8>::ELEMS : HaloShape<16>::ELEMS;
  // halo tiles per bin (the matrix-core spreader counts 2 windows per tile)
  const int tiles = sp.mode == Spreader<T>::HALOM && t8 ? sp.halo_parts / 2 : sp.halo_parts;
  if (tiles <= 4 && (size_t)g.nbins * tiles * hel < (1u << 31)) {
#define FRESCO_MERGE_P(TY_, P_)                                                                              \
  hipLaunchKernelGGL((k_merge_halo_p<T, TY_, P_>), grd, dim3(256), 0, s, sp.halo_buf.get(), sp.halo_tab.get(), a, \
                     grid)
    switch (tiles * (t8 ? 1 : -1)) {
      case 1: FRESCO_MERGE_P(8, 1); break;
      case 2: FRESCO_MERGE_P(8, 2); break;
      case 3: FRESCO_MERGE_P(8, 3); break;
      case 4: FRESCO_MERGE_P(8, 4); break;
      case -1: FRESCO_MERGE_P(16, 1); break;
      case -2: FRESCO_MERGE_P(16, 2); break;
      case -3: FRESCO_MERGE_P(16, 3); break;
      default: FRESCO_MERGE_P(16, 4); break;
    }
#undef FRESCO_MERGE_P
    FRESCO_HIP_CHECK(hipGetLastError());
    return true;
  }
  if (t8)
    hipLaunchKernelGGL((k_merge_halo<T, 8>), grd, dim3(256), 0, s, sp.halo_buf.get(), sp.halo_tab.get(), tiles, a,
                       grid);
  else
    hipLaunchKernelGGL((k_merge_halo<T, 16>), grd, dim3(256), 0, s, sp.halo_buf.get(), sp.halo_tab.get(), tiles, a,
                       grid);
  FRESCO_HIP_CHECK(hipGetLastError());
  return true;
}

template <typename T, int WMAX>
static void launch_spread3_reg_w(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s) {
  SpreadDev<T> a;
  const GridGeom& g = sp.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.maxc[q] = sp.maxc[q];
  }
  a.cbits = g.cell_bits;
  a.w = sp.es.w;
  a.sx = 0;
  a.cstride = sp.cand_stride;
  a.k.beta = (T)sp.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(sp.es.beta * M_LOG2E) : (T)sp.es.beta;
  a.k.two_over_w = (T)(2.0 / sp.es.w);
  const int parts = sp.parts;
  T* out = parts > 1 ? sp.part_buf.get() : grid;
  {
    ProfScope ps(ST_SPREAD, s);
    hipLaunchKernelGGL((k_spread3_reg<T, WMAX>), dim3(g.nbins, parts), dim3(64), 0, s, p.scell.get(), p.rec.get(),
                       p.bin_start.get(), sp.cand.get(), a, out);
    FRESCO_HIP_CHECK(hipGetLastError());
  }
  if (parts > 1) {
    ProfScope pm(ST_MERGE, s);
    const size_t nc4 = (size_t)g.ncells();
    hipLaunchKernelGGL(k_sum_parts<T>, dim3((unsigned)std::min<size_t>(blocks_for(nc4, 256), 16384)), dim3(256), 0,
                       s, out, parts, (int64_t)nc4, grid);
    FRESCO_HIP_CHECK(hipGetLastError());
  }
}

template <typename T>
static bool launch_spread3_reg(const Spreader<T>& sp, const SortedPoints<T>& p, T* grid, hipStream_t s) {
  const GridGeom& g = sp.geom;
  if (g.tile[0] != RT_X || g.tile[1] != RT_Y || g.tile[2] != RT_Z) return false;
  if (g.nf[1] < RT_Y + sp.es.w - 1) return false;
  if (sp.es.w <= 8)
    launch_spread3_reg_w<T, 8>(sp, p, grid, s);
  else
    launch_spread3_reg_w<T, 16>(sp, p, grid, s);
  return true;
}

template <typename T, int WMAX, bool DERIV>
static void launch_interp3_lane(const SortedPoints<T>& p, const T* grid, const double* dscale, double* out,
                                int ostride, int ocomp, hipStream_t s) {
  InterpDev<T> a;
  const GridGeom& g = p.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.dscale[q] = dscale ? dscale[q] : 0.0;
  }
  a.cbits = g.cell_bits;
  a.w = p.es.w;
  a.rx = g.tile[0] + a.w - 1;
  a.ry = g.tile[1] + a.w - 1;
  a.rz = g.tile[2] + a.w - 1;
  a.srx = a.rx;
  a.ostride = ostride;
  a.ocomp = ocomp;
  a.k.beta = (T)p.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(p.es.beta * M_LOG2E) : (T)p.es.beta;
  a.k.two_over_w = (T)(2.0 / p.es.w);
  size_t lds = sizeof(T) * (size_t)a.rx * a.ry * a.rz;
  if (lds > 160 * 1024) throw InvalidArgument("interpolation region exceeds LDS");
  hipLaunchKernelGGL((k_interp3_lane<T, WMAX, DERIV>), dim3(g.nbins), dim3(LANE_INTERP_THREADS), lds, s,
                     p.scell.get(), p.rec.get(), p.idx.get(), p.bin_start.get(), grid, a, out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// 3D interpolation, eight lanes per point (w <= 8): the bin's grid region is
// staged in LDS by a 256-thread workgroup (tile + 7 cells per dimension);
// lane ja of an 8-lane group sums the 8 x 8 cells of its x offset (weights w_y w_z and their derivatives, computed
// per lane), multiplies by its own x weight, and the group reduces with three
// DPP steps (result in the group's last lane).  Consecutive lanes read
// consecutive LDS words.
// ---------------------------------------------------------------------------

constexpr int G8_THREADS = 256;
constexpr int G8_SORT_MAX = 1024;  // slice points reordered in LDS

// out[j] = value j of the quad (lane j % 4 holds values j % 4 and j % 4 + 4
// in v2[0], v2[1]): DPP quad_perm broadcasts
// (no lane of a quad_perm is out of range: the move needs no old value)
template <int J>
__device__ __forceinline__ float quad_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), J | (J << 2) | (J << 4) | (J << 6), 0xf, 0xf, true));
}
template <int J>
__device__ __forceinline__ double quad_lane(double v) {
  const long long b = __double_as_longlong(v);
  constexpr int C = J | (J << 2) | (J << 4) | (J << 6);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), C, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), C, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ void quad_bcast8(const T* v2, T* out) {
  out[0] = quad_lane<0>(v2[0]);
  out[1] = quad_lane<1>(v2[0]);
  out[2] = quad_lane<2>(v2[0]);
  out[3] = quad_lane<3>(v2[0]);
  out[4] = quad_lane<0>(v2[1]);
  out[5] = quad_lane<1>(v2[1]);
  out[6] = quad_lane<2>(v2[1]);
  out[7] = quad_lane<3>(v2[1]);
}

// the sum of a quad's values, complete in its last lane (and every lane)
template <typename T>
__device__ __forceinline__ T sum4_last(T v) {
  v += dpp_f<0xb1>(v);  // quad_perm [1,0,3,2]
  v += dpp_f<0x4e>(v);  // quad_perm [2,3,0,1]
  return v;
}

template <typename T>
__device__ __forceinline__ T sum8_last(T v) {
  v += dpp_f<0xb1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4e>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x114>(v);  // row_shr:4 (lanes 8g+4..8g+7 add the lower quad)
  return v;
}

// SRX, RY != 0: the region's row stride and rows per plane fixed at compile
// time (the 16 x 8 x 8 and 16 x 16 x 8 bins), so the 64 cell reads of a
// point become immediate offsets of one address (paired LDS reads, no
// per-read address arithmetic); 0: taken from a
template <typename T, bool DERIV, int SRX = 0, int RY = 0, int LPP = 8>
__global__ void __launch_bounds__(G8_THREADS)
    k_interp3_g8(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                 const uint32_t* __restrict__ idx, const int* __restrict__ bin_start, const T* __restrict__ grid,
                 InterpDev<T> a, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* reg = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const int b = blockIdx.x;
  // gridDim.y > 1: the bin's points are split into contiguous slices, one per
  // workgroup (each stages the same region)
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.y) / gridDim.y);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.y + 1)) / gridDim.y);
  if (beg == end) return;
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  // the slice's packed first cells for the counting sort below, loaded
  // together with the region (one memory latency for both)
  const int npts = end - beg;
  const bool sorted = npts <= G8_SORT_MAX;
  __shared__ uint16_t perm[G8_SORT_MAX];
  __shared__ int kcnt[128];
  uint32_t skey[G8_SORT_MAX / G8_THREADS];
#pragma unroll
  for (int u = 0; u < G8_SORT_MAX / G8_THREADS; ++u) {
    const int e = tid + u * G8_THREADS;
    skey[u] = sorted && e < npts ? scell[beg + e] : 0u;
  }
  if (tid < 128) kcnt[tid] = 0;
  // region rows of up to 32 cells: thread (row slot, x)
  {
    const int lx = tid & 31;
    int gx = t0[0] + lx;
    if (gx >= a.nf[0]) gx -= a.nf[0];
    if (gx >= a.nf[0]) gx -= a.nf[0];
    auto src = [&](int row) {
      const int ly = row % (RY ? RY : a.ry);
      const int lz = row / (RY ? RY : a.ry);
      int gy = t0[1] + ly;
      if (gy >= a.nf[1]) gy -= a.nf[1];
      if (gy >= a.nf[1]) gy -= a.nf[1];
      int gz = t0[2] + lz;
      if (gz >= a.nf[2]) gz -= a.nf[2];
      if (gz >= a.nf[2]) gz -= a.nf[2];
      return ((int64_t)gz * a.nf[1] + gy) * a.nf[0] + gx;
    };
    if constexpr (SRX == 24) {
      // rows of 24 cells (23 + one that lands in the row's padding) moved as
      // 16-byte units: the grid row starts at a multiple of the unit (the
      // launcher checks nf[0] and the tile width), so a unit never straddles
      // the periodic wrap in x; 32-bit offsets (the launcher checks the grid
      // size); all of a thread's loads are issued before its LDS stores
      using V = typename Vec4<float>::type;
      constexpr int EPU = 16 / sizeof(T);  // cells per unit
      constexpr int UPR = 24 / EPU;        // units per row
      constexpr int NU = RY * 15 * UPR;
      constexpr int PER = (NU + G8_THREADS - 1) / G8_THREADS;
      const int nu0 = a.nf[0] / EPU, tu0 = t0[0] / EPU;
      const int plane = a.nf[0] * a.nf[1];
      V v[PER];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int u = tid + q * G8_THREADS;
        const int row = u / UPR, xu = u - row * UPR;
        const int ly = row % RY, lz = row / RY;
        int gy = t0[1] + ly;
        if (gy >= a.nf[1]) gy -= a.nf[1];
        int gz = t0[2] + lz;
        if (gz >= a.nf[2]) gz -= a.nf[2];
        int gu = tu0 + xu;
        if (gu >= nu0) gu -= nu0;
        const int off = gz * plane + gy * a.nf[0] + gu * EPU;
        v[q] = (NU % G8_THREADS == 0 || u < NU) ? *reinterpret_cast<const V*>(grid + off) : V{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int u = tid + q * G8_THREADS;
        if (NU % G8_THREADS == 0 || u < NU) *reinterpret_cast<V*>(reg + u * EPU) = v[q];
      }
    } else if constexpr (RY != 0) {
      // fixed region (RY x 15 rows, the launcher checks rz): every load of
      // a thread is issued before its LDS stores (one memory latency per
      // workgroup instead of one per row)
      constexpr int NROWS = RY * 15;
      constexpr int PER = (NROWS + G8_THREADS / 32 - 1) / (G8_THREADS / 32);
      T v[PER];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int row = (tid >> 5) + q * (G8_THREADS / 32);
        v[q] = (row < NROWS && lx < a.rx) ? grid[src(row)] : T(0);
      }
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int row = (tid >> 5) + q * (G8_THREADS / 32);
        if (row < NROWS && lx < a.rx) reg[row * SRX + lx] = v[q];
      }
    } else {
      const int nrows = a.ry * a.rz;
      for (int row = tid >> 5; row < nrows; row += G8_THREADS / 32)
        if (lx < a.rx) reg[row * a.srx + lx] = grid[src(row)];
    }
  }
  __syncthreads();
  const int ja = tid & (LPP - 1);
  const int grp = tid / LPP;
  const int srx = SRX ? SRX : a.srx;
  const int zs = (RY ? RY : a.ry) * srx;
  // The 8 points one wavefront interpolates together read 8-cell windows of
  // the region; at unrelated (y, z) rows those windows collide in the LDS
  // banks.  Points taken in the order of their first (z, y) cell share their
  // rows, and then every read of the wavefront hits distinct banks or the
  // same words.  Each point's result does not depend on the order, so the
  // slice is reordered by a counting sort in LDS (slices up to G8_SORT_MAX
  // points; longer ones keep the bin order).
  if (sorted) {
    int key[G8_SORT_MAX / G8_THREADS];
#pragma unroll
    for (int u = 0; u < G8_SORT_MAX / G8_THREADS; ++u) {
      const int e = tid + u * G8_THREADS;
      key[u] = -1;
      if (e < npts) {
        const uint32_t sc = skey[u];
        const int uy = (int)((sc >> a.cbits) & cmask) - t0[1];
        const int uz = (int)((sc >> (2 * a.cbits)) & cmask) - t0[2];
        key[u] = ((uz & 7) << 4) | (uy & 15);
        atomicAdd(&kcnt[key[u]], 1);
      }
    }
    __syncthreads();
    if (tid < 64) {
      // exclusive prefix over the 128 counts, two per lane
      const int c0 = kcnt[2 * tid], c1 = kcnt[2 * tid + 1];
      int inc = c0 + c1;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off, 64);
        if (tid >= off) inc += u;
      }
      const int ex = inc - (c0 + c1);
      kcnt[2 * tid] = ex;
      kcnt[2 * tid + 1] = ex + c0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < G8_SORT_MAX / G8_THREADS; ++u)
      if (key[u] >= 0) perm[atomicAdd(&kcnt[key[u]], 1)] = (uint16_t)(tid + u * G8_THREADS);
    __syncthreads();
  }
  auto point_at = [&](int pos) { return pos < npts ? beg + (sorted ? (int)perm[pos] : pos) : beg; };
  // every group runs the same number of iterations (DPP needs all lanes)
  const int iters = (npts + (G8_THREADS / LPP) - 1) / (G8_THREADS / LPP);
  // the next iteration's point is loaded while this one is interpolated
  int pn = grp;
  int kq0 = point_at(pn);
  uint32_t sc_n = scell[kq0];
  typename Vec4<T>::type r_n = rec[kq0];
  uint32_t j_n = idx[kq0];
  for (int it = 0; it < iters; ++it) {
    const bool valid = pn < npts;
    const uint32_t sc = sc_n;
    const typename Vec4<T>::type r = r_n;
    const uint32_t jout = j_n;
    pn += G8_THREADS / LPP;
    if (it + 1 < iters) {
      const int kq = point_at(pn);
      sc_n = scell[kq];
      r_n = rec[kq];
      j_n = idx[kq];
    }
    int l0[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int u = (int)((sc >> (q * a.cbits)) & cmask) - t0[q];
      if (u < 0) u += a.nf[q];
      l0[q] = u;
    }
    if constexpr (LPP == 4) {
      // four lanes per point: lane ja takes the x offsets ja and ja + 4 and
      // evaluates (phi, phi') of indices ja and ja + 4 along x, y and z; the
      // quad shares the y and z values by DPP quad broadcasts and reduces
      // with two DPP steps
      T wx2[2] = {T(0), T(0)}, dwx2[2] = {T(0), T(0)}, wy2[2] = {T(0), T(0)}, dwy2[2] = {T(0), T(0)};
      T wz2[2] = {T(0), T(0)}, dwz2[2] = {T(0), T(0)};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int jj = ja + 4 * h;
        if (jj < w) {
          if (DERIV) {
            es_phi_dphi<T>((T(jj) - r.x) * a.k.two_over_w, a.k, wx2[h], dwx2[h]);
            es_phi_dphi<T>((T(jj) - r.y) * a.k.two_over_w, a.k, wy2[h], dwy2[h]);
            es_phi_dphi<T>((T(jj) - r.z) * a.k.two_over_w, a.k, wz2[h], dwz2[h]);
          } else {
            wx2[h] = es_phi<T>((T(jj) - r.x) * a.k.two_over_w, a.k);
            wy2[h] = es_phi<T>((T(jj) - r.y) * a.k.two_over_w, a.k);
            wz2[h] = es_phi<T>((T(jj) - r.z) * a.k.two_over_w, a.k);
          }
        }
      }
      T wy[8], dwy[8], wz[8], dwz[8];
      quad_bcast8<T>(wy2, wy);
      quad_bcast8<T>(wz2, wz);
      if (DERIV) {
        quad_bcast8<T>(dwy2, dwy);
        quad_bcast8<T>(dwz2, dwz);
      }
      const T* col = reg + (l0[2] * (RY ? RY : a.ry) + l0[1]) * srx + l0[0] + ja;
      T sA[2] = {T(0), T(0)}, sB[2] = {T(0), T(0)}, sC[2] = {T(0), T(0)};
#pragma unroll
      for (int jc = 0; jc < 8; ++jc) {
        T pA[2] = {T(0), T(0)}, pB[2] = {T(0), T(0)};
        // the plane's 16 cells are read in one batch (the compiler barrier
        // keeps the reads together: 8 paired reads in flight instead of 2)
        T gv[8][2];
#pragma unroll
        for (int jb = 0; jb < 8; ++jb)
#pragma unroll
          for (int h = 0; h < 2; ++h) gv[jb][h] = col[jc * zs + jb * srx + 4 * h];
        asm volatile("" ::: "memory");
#pragma unroll
        for (int jb = 0; jb < 8; ++jb) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const T g = gv[jb][h];
            pA[h] += g * wy[jb];
            if (DERIV) pB[h] += g * dwy[jb];
          }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          sA[h] += pA[h] * wz[jc];
          if (DERIV) {
            sB[h] += pB[h] * wz[jc];
            sC[h] += pA[h] * dwz[jc];
          }
        }
      }
      T vx, vy = T(0), vz = T(0);
      if (DERIV) {
        vx = sum4_last(dwx2[0] * sA[0] + dwx2[1] * sA[1]);
        vy = sum4_last(wx2[0] * sB[0] + wx2[1] * sB[1]);
        vz = sum4_last(wx2[0] * sC[0] + wx2[1] * sC[1]);
      } else {
        vx = sum4_last(wx2[0] * sA[0] + wx2[1] * sA[1]);
      }
      if (ja == 3 && valid) {
        const int64_t j = jout;
        const double cw = (double)r.w;
        if (DERIV) {
          out[j * a.ostride + 0] = cw * a.dscale[0] * (double)vx;
          out[j * a.ostride + 1] = cw * a.dscale[1] * (double)vy;
          out[j * a.ostride + 2] = cw * a.dscale[2] * (double)vz;
        } else {
          out[j * a.ostride + a.ocomp] = cw * (double)vx;
        }
      }
    } else {
    // lane ja evaluates (phi, phi') of index ja along x, y and z; the group
    // shares the y and z values by ds_swizzle (broadcast of lane j of the
    // 8-lane group)
    T wxa = T(0), dwxa = T(0), wya = T(0), dwya = T(0), wza = T(0), dwza = T(0);
    if (ja < w) {
      if (DERIV) {
        es_phi_dphi<T>((T(ja) - r.x) * a.k.two_over_w, a.k, wxa, dwxa);
        es_phi_dphi<T>((T(ja) - r.y) * a.k.two_over_w, a.k, wya, dwya);
        es_phi_dphi<T>((T(ja) - r.z) * a.k.two_over_w, a.k, wza, dwza);
      } else {
        wxa = es_phi<T>((T(ja) - r.x) * a.k.two_over_w, a.k);
        wya = es_phi<T>((T(ja) - r.y) * a.k.two_over_w, a.k);
        wza = es_phi<T>((T(ja) - r.z) * a.k.two_over_w, a.k);
      }
    }
    T wy[8], dwy[8], wz[8], dwz[8];
    group8_bcast<T>(wya, wy);
    group8_bcast<T>(wza, wz);
    if (DERIV) {
      group8_bcast<T>(dwya, dwy);
      group8_bcast<T>(dwza, dwz);
    }
    const T* col = reg + (l0[2] * (RY ? RY : a.ry) + l0[1]) * srx + l0[0] + ja;
    // all 8 x 8 cells are read (weights beyond w are zero and the region
    // is staged 8 - 1 cells wider than the tile), so the loads can be batched
    T sA = T(0), sB = T(0), sC = T(0);
#pragma unroll
    for (int jc = 0; jc < 8; ++jc) {
      T pA = T(0), pB = T(0);
#pragma unroll
      for (int jb = 0; jb < 8; ++jb) {
        const T g = col[jc * zs + jb * srx];
        pA += g * wy[jb];
        if (DERIV) pB += g * dwy[jb];
      }
      sA += pA * wz[jc];
      if (DERIV) {
        sB += pB * wz[jc];
        sC += pA * dwz[jc];
      }
    }
    T vx, vy = T(0), vz = T(0);
    if (DERIV) {
      vx = sum8_last(dwxa * sA);
      vy = sum8_last(wxa * sB);
      vz = sum8_last(wxa * sC);
    } else {
      vx = sum8_last(wxa * sA);
    }
    if (ja == 7 && valid) {
      const int64_t j = jout;
      const double cw = (double)r.w;
      if (DERIV) {
        out[j * a.ostride + 0] = cw * a.dscale[0] * (double)vx;
        out[j * a.ostride + 1] = cw * a.dscale[1] * (double)vy;
        out[j * a.ostride + 2] = cw * a.dscale[2] * (double)vz;
      } else {
        out[j * a.ostride + a.ocomp] = cw * (double)vx;
      }
    }
    }
  }
}

template <typename T, bool DERIV>
static void launch_interp3_g8(const SortedPoints<T>& p, const T* grid, const double* dscale, double* out,
                              int ostride, int ocomp, hipStream_t s) {
  InterpDev<T> a;
  const GridGeom& g = p.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.dscale[q] = dscale ? dscale[q] : 0.0;
  }
  a.cbits = g.cell_bits;
  a.w = p.es.w;
  a.rx = g.tile[0] + 8 - 1;
  a.ry = g.tile[1] + 8 - 1;
  a.rz = g.tile[2] + 8 - 1;
  a.srx = a.rx;
  a.ostride = ostride;
  a.ocomp = ocomp;
  a.k.beta = (T)p.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(p.es.beta * M_LOG2E) : (T)p.es.beta;
  a.k.two_over_w = (T)(2.0 / p.es.w);
  size_t lds = sizeof(T) * (size_t)a.rx * a.ry * a.rz;
  if (lds > 160 * 1024) throw InvalidArgument("interpolation region exceeds LDS");
  // enough workgroups to fill the GPU: bins split into slices when there
  // are few, ~2048 workgroups (2 per bin for the 1183 bins of a 100^3 grid;
  // four lanes per point: 0.128 / 0.080 / 0.081 / 0.088 / 0.096 ms for 1 /
  // 2 / 3 / 4 / 6 slices per bin; FRESCO_INTERP_WGS overrides the target)
  const int want = p.interp_wgs;
  const int chunks = std::max(1, std::min(8, (want + g.nbins - 1) / g.nbins));
  const dim3 grd(g.nbins, chunks);
  // 16-byte row units: whole units per grid row, the tile's first cell on a
  // unit, one wrap per dimension at most, 32-bit offsets, an aligned grid
  const bool units_on = p.interp_units;
  const int epu = 16 / (int)sizeof(T);
  const bool units = units_on && a.rx == 23 && a.rz == 15 && (a.ry == 15 || a.ry == 23) && g.nf[0] % epu == 0 &&
                     g.tile[0] % epu == 0 && g.nf[0] >= 24 && g.nf[1] >= a.ry && g.nf[2] >= 15 &&
                     g.ncells() < (1ll << 31) && ((uintptr_t)grid % 16) == 0;
  const int lpp = p.interp_lpp;
  if (units) {
    lds = sizeof(T) * (size_t)24 * a.ry * a.rz;
    a.srx = 24;
    if (a.ry == 15 && lpp == 4)
      hipLaunchKernelGGL((k_interp3_g8<T, DERIV, 24, 15, 4>), grd, dim3(G8_THREADS), lds, s, p.scell.get(),
                         p.rec.get(), p.idx.get(), p.bin_start.get(), grid, a, out);
    else if (a.ry == 15)
      hipLaunchKernelGGL((k_interp3_g8<T, DERIV, 24, 15>), grd, dim3(G8_THREADS), lds, s, p.scell.get(),
                         p.rec.get(), p.idx.get(), p.bin_start.get(), grid, a, out);
    else
      hipLaunchKernelGGL((k_interp3_g8<T, DERIV, 24, 23>), grd, dim3(G8_THREADS), lds, s, p.scell.get(),
                         p.rec.get(), p.idx.get(), p.bin_start.get(), grid, a, out);
  } else if (a.srx == 23 && a.ry == 15 && a.rx == 23 && a.rz == 15)
    hipLaunchKernelGGL((k_interp3_g8<T, DERIV, 23, 15>), grd, dim3(G8_THREADS), lds, s, p.scell.get(), p.rec.get(),
                       p.idx.get(), p.bin_start.get(), grid, a, out);
  else if (a.srx == 23 && a.ry == 23 && a.rx == 23 && a.rz == 15)
    hipLaunchKernelGGL((k_interp3_g8<T, DERIV, 23, 23>), grd, dim3(G8_THREADS), lds, s, p.scell.get(), p.rec.get(),
                       p.idx.get(), p.bin_start.get(), grid, a, out);
  else
    hipLaunchKernelGGL((k_interp3_g8<T, DERIV>), grd, dim3(G8_THREADS), lds, s, p.scell.get(), p.rec.get(),
                       p.idx.get(), p.bin_start.get(), grid, a, out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// 3D interpolation for kernels 9 to 16 wide (the FP64 grids of eps <= 1e-8),
// eight lanes per point: lane ja of an 8-lane group takes the x offsets ja
// and ja + 8 (the second weighted 0 where ja + 8 >= W), sums their W x W
// cells with the y and z weights shared by ds_swizzle, and the group reduces
// with three DPP steps, as k_interp3_g8.  The bin's region (tile + W - 1
// cells per dimension) is staged in LDS by the 256-thread workgroup in
// batches of row loads; W is a template parameter so that the weight arrays
// stay in registers.
// ---------------------------------------------------------------------------

template <typename T, bool DERIV, int W>
__global__ void __launch_bounds__(G8_THREADS)
    k_interp3_g16(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                  const uint32_t* __restrict__ idx, const int* __restrict__ bin_start, const T* __restrict__ grid,
                  InterpDev<T> a, double* __restrict__ out) {
  static_assert(W > 8 && W <= 16, "eight lanes cover two x offsets each");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* reg = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const int b = blockIdx.x;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.y) / gridDim.y);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.y + 1)) / gridDim.y);
  if (beg == end) return;
  int t0[3];
  t0[0] = (b % a.ntile[0]) * a.tile[0];
  t0[1] = ((b / a.ntile[0]) % a.ntile[1]) * a.tile[1];
  t0[2] = (b / (a.ntile[0] * a.ntile[1])) * a.tile[2];
  const int rx = a.rx, ry = a.ry, srx = a.srx;
  const int zs = ry * srx;
  {
    // thread (row slot, x): rows of up to 32 cells, 8 rows per pass, every
    // load of a batch issued before its LDS stores
    constexpr int BATCH = 8;
    const int lx = tid & 31;
    int gx = t0[0] + lx;
    if (gx >= a.nf[0]) gx -= a.nf[0];
    if (gx >= a.nf[0]) gx -= a.nf[0];
    const int nrows = ry * a.rz;
    for (int r0 = tid >> 5; r0 < nrows; r0 += BATCH * (G8_THREADS / 32)) {
      T v[BATCH];
#pragma unroll
      for (int q = 0; q < BATCH; ++q) {
        const int row = r0 + q * (G8_THREADS / 32);
        v[q] = T(0);
        if (row < nrows && lx < rx) {
          int gy = t0[1] + row % ry;
          if (gy >= a.nf[1]) gy -= a.nf[1];
          if (gy >= a.nf[1]) gy -= a.nf[1];
          int gz = t0[2] + row / ry;
          if (gz >= a.nf[2]) gz -= a.nf[2];
          if (gz >= a.nf[2]) gz -= a.nf[2];
          v[q] = grid[((int64_t)gz * a.nf[1] + gy) * a.nf[0] + gx];
        }
      }
#pragma unroll
      for (int q = 0; q < BATCH; ++q) {
        const int row = r0 + q * (G8_THREADS / 32);
        if (row < nrows && lx < rx) reg[row * srx + lx] = v[q];
      }
    }
  }
  __syncthreads();
  const int ja = tid & 7;
  const int grp = tid >> 3;
  const bool two = ja + 8 < W;
  const int npts = end - beg;
  const int iters = (npts + (G8_THREADS / 8) - 1) / (G8_THREADS / 8);
  int kn = beg + grp;
  uint32_t sc_n = scell[kn < end ? kn : beg];
  typename Vec4<T>::type r_n = rec[kn < end ? kn : beg];
  uint32_t j_n = idx[kn < end ? kn : beg];
  for (int it = 0; it < iters; ++it) {
    const int k = kn;
    const bool valid = k < end;
    const uint32_t sc = sc_n;
    const typename Vec4<T>::type r = r_n;
    const uint32_t jout = j_n;
    kn = k + G8_THREADS / 8;
    if (it + 1 < iters) {
      const int kq = kn < end ? kn : beg;
      sc_n = scell[kq];
      r_n = rec[kq];
      j_n = idx[kq];
    }
    int l0[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int u = (int)((sc >> (q * a.cbits)) & cmask) - t0[q];
      if (u < 0) u += a.nf[q];
      l0[q] = u;
    }
    // kernel values of indices ja (set 0) and ja + 8 (set 1)
    T wx[2], dwx[2], wyl[2], dwyl[2], wzl[2], dwzl[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = ja + 8 * h;
      wx[h] = dwx[h] = wyl[h] = dwyl[h] = wzl[h] = dwzl[h] = T(0);
      if (j < W) {
        if (DERIV) {
          es_phi_dphi<T>((T(j) - r.x) * a.k.two_over_w, a.k, wx[h], dwx[h]);
          es_phi_dphi<T>((T(j) - r.y) * a.k.two_over_w, a.k, wyl[h], dwyl[h]);
          es_phi_dphi<T>((T(j) - r.z) * a.k.two_over_w, a.k, wzl[h], dwzl[h]);
        } else {
          wx[h] = es_phi<T>((T(j) - r.x) * a.k.two_over_w, a.k);
          wyl[h] = es_phi<T>((T(j) - r.y) * a.k.two_over_w, a.k);
          wzl[h] = es_phi<T>((T(j) - r.z) * a.k.two_over_w, a.k);
        }
      }
    }
    T wy[16], wz[16], dwy[16], dwz[16];
    group8_bcast<T>(wyl[0], wy);
    group8_bcast<T>(wyl[1], wy + 8);
    group8_bcast<T>(wzl[0], wz);
    group8_bcast<T>(wzl[1], wz + 8);
    if (DERIV) {
      group8_bcast<T>(dwyl[0], dwy);
      group8_bcast<T>(dwyl[1], dwy + 8);
      group8_bcast<T>(dwzl[0], dwz);
      group8_bcast<T>(dwzl[1], dwz + 8);
    }
    const T* col0 = reg + (l0[2] * ry + l0[1]) * srx + l0[0] + ja;
    // the second column of lanes without one reads a staged cell (weight 0)
    const T* col1 = two ? col0 + 8 : col0;
    T sA[2] = {T(0), T(0)}, sB[2] = {T(0), T(0)}, sC[2] = {T(0), T(0)};
#pragma unroll
    for (int jc = 0; jc < W; ++jc) {
      T pA[2] = {T(0), T(0)}, pB[2] = {T(0), T(0)};
#pragma unroll
      for (int jb = 0; jb < W; ++jb) {
        const T g0 = col0[jc * zs + jb * srx];
        const T g1 = col1[jc * zs + jb * srx];
        pA[0] += g0 * wy[jb];
        pA[1] += g1 * wy[jb];
        if (DERIV) {
          pB[0] += g0 * dwy[jb];
          pB[1] += g1 * dwy[jb];
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        sA[h] += pA[h] * wz[jc];
        if (DERIV) {
          sB[h] += pB[h] * wz[jc];
          sC[h] += pA[h] * dwz[jc];
        }
      }
    }
    T vx, vy = T(0), vz = T(0);
    if (DERIV) {
      vx = sum8_last(dwx[0] * sA[0] + dwx[1] * sA[1]);
      vy = sum8_last(wx[0] * sB[0] + wx[1] * sB[1]);
      vz = sum8_last(wx[0] * sC[0] + wx[1] * sC[1]);
    } else {
      vx = sum8_last(wx[0] * sA[0] + wx[1] * sA[1]);
    }
    if (ja == 7 && valid) {
      const int64_t j = jout;
      const double cw = (double)r.w;
      if (DERIV) {
        out[j * a.ostride + 0] = cw * a.dscale[0] * (double)vx;
        out[j * a.ostride + 1] = cw * a.dscale[1] * (double)vy;
        out[j * a.ostride + 2] = cw * a.dscale[2] * (double)vz;
      } else {
        out[j * a.ostride + a.ocomp] = cw * (double)vx;
      }
    }
  }
}

// true if the region of a kernel w wide fits LDS (then launched)
template <typename T, bool DERIV>
static bool launch_interp3_g16(const SortedPoints<T>& p, const T* grid, const double* dscale, double* out,
                               int ostride, int ocomp, hipStream_t s) {
  InterpDev<T> a;
  const GridGeom& g = p.geom;
  const int w = p.es.w;
  if (w <= 8 || w > 16) return false;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.dscale[q] = dscale ? dscale[q] : 0.0;
  }
  a.cbits = g.cell_bits;
  a.w = w;
  a.rx = g.tile[0] + w - 1;
  a.ry = g.tile[1] + w - 1;
  a.rz = g.tile[2] + w - 1;
  // lanes with a second column read x offsets up to W - 1 < rx
  a.srx = a.rx;
  a.ostride = ostride;
  a.ocomp = ocomp;
  a.k.beta = (T)p.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(p.es.beta * M_LOG2E) : (T)p.es.beta;
  a.k.two_over_w = (T)(2.0 / w);
  const size_t lds = sizeof(T) * (size_t)a.srx * a.ry * a.rz;
  if (lds > 160 * 1024 || a.rx > 32 || a.nf[0] < a.rx || a.nf[1] < a.ry || a.nf[2] < a.rz) return false;
  const int chunks = std::max(1, std::min(8, (4096 + g.nbins - 1) / g.nbins));
  const dim3 grd(g.nbins, chunks);
#define FRESCO_G16(W_)                                                                                    \
  case W_:                                                                                                \
    hipLaunchKernelGGL((k_interp3_g16<T, DERIV, W_>), grd, dim3(G8_THREADS), lds, s, p.scell.get(),      \
                       p.rec.get(), p.idx.get(), p.bin_start.get(), grid, a, out);                        \
    break;
  switch (w) {
    FRESCO_G16(9)
    FRESCO_G16(10)
    FRESCO_G16(11)
    FRESCO_G16(12)
    FRESCO_G16(13)
    FRESCO_G16(14)
    FRESCO_G16(15)
    FRESCO_G16(16)
  }
#undef FRESCO_G16
  FRESCO_HIP_CHECK(hipGetLastError());
  return true;
}

// ---------------------------------------------------------------------------
// 2D interpolation, eight lanes per point (w <= 8): the bin's (tile + 7)^2
// grid region is staged in LDS; lane ja of an 8-lane group sums the 8 cells
// of its x offset with the y weights (and their derivatives) shared by
// ds_swizzle, multiplies by its x weight (or its derivative), and the group
// reduces with three DPP steps.  blockIdx.y: pattern of a batched plan;
// blockIdx.z: contiguous slice of the bin's points.
// ---------------------------------------------------------------------------

template <typename T, bool DERIV>
__global__ void __launch_bounds__(G8_THREADS)
    k_interp2_g8(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
                 const uint32_t* __restrict__ idx, const int* __restrict__ bin_start, const T* __restrict__ grid,
                 InterpDev<T> a, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* reg = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  bin_start += (int64_t)blockIdx.y * a.ntile[0] * a.ntile[1];
  grid += (int64_t)blockIdx.y * a.nf[0] * a.nf[1];
  const int b = blockIdx.x;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.z) / gridDim.z);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.z + 1)) / gridDim.z);
  if (beg == end) return;  // uniform across the workgroup
  const int t0x = (b % a.ntile[0]) * a.tile[0];
  const int t0y = (b / a.ntile[0]) * a.tile[1];
  // region rows of up to 32 cells, thread (row slot, x); every load of a
  // thread is issued before its LDS stores (one memory latency per pass of
  // four rows per thread instead of one per cell); the region may wrap
  {
    const int lx = tid & 31;
    int gx = t0x + lx;
    while (gx >= a.nf[0]) gx -= a.nf[0];
    constexpr int RPP = G8_THREADS / 32;  // rows per pass
    for (int r0 = 0; r0 < a.ry; r0 += 4 * RPP) {
      T v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ly = r0 + q * RPP + (tid >> 5);
        v[q] = T(0);
        if (ly < a.ry && lx < a.rx) {
          int gy = t0y + ly;
          while (gy >= a.nf[1]) gy -= a.nf[1];
          v[q] = grid[(int64_t)gy * a.nf[0] + gx];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ly = r0 + q * RPP + (tid >> 5);
        if (ly < a.ry && lx < a.rx) reg[ly * a.srx + lx] = v[q];
      }
    }
  }
  __syncthreads();
  const int ja = tid & 7;
  const int grp = tid >> 3;
  const int npts = end - beg;
  const int iters = (npts + (G8_THREADS / 8) - 1) / (G8_THREADS / 8);
  int kn = beg + grp;
  uint32_t sc_n = scell[kn < end ? kn : beg];
  typename Vec4<T>::type r_n = rec[kn < end ? kn : beg];
  uint32_t j_n = idx[kn < end ? kn : beg];
  for (int it = 0; it < iters; ++it) {
    const int k = kn;
    const bool valid = k < end;
    const uint32_t sc = sc_n;
    const typename Vec4<T>::type r = r_n;
    const uint32_t jout = j_n;
    kn = k + G8_THREADS / 8;
    if (it + 1 < iters) {
      const int kq = kn < end ? kn : beg;
      sc_n = scell[kq];
      r_n = rec[kq];
      j_n = idx[kq];
    }
    int ux = (int)(sc & cmask) - t0x;
    if (ux < 0) ux += a.nf[0];
    int uy = (int)((sc >> a.cbits) & cmask) - t0y;
    if (uy < 0) uy += a.nf[1];
    T wxa = T(0), dwxa = T(0), wya = T(0), dwya = T(0);
    if (ja < w) {
      if (DERIV) {
        es_phi_dphi<T>((T(ja) - r.x) * a.k.two_over_w, a.k, wxa, dwxa);
        es_phi_dphi<T>((T(ja) - r.y) * a.k.two_over_w, a.k, wya, dwya);
      } else {
        wxa = es_phi<T>((T(ja) - r.x) * a.k.two_over_w, a.k);
        wya = es_phi<T>((T(ja) - r.y) * a.k.two_over_w, a.k);
      }
    }
    T wy[8], dwy[8];
    group8_bcast<T>(wya, wy);
    if (DERIV) group8_bcast<T>(dwya, dwy);
    // all 8 x 8 cells are read (weights beyond w are zero and the region is
    // staged 8 - 1 cells wider than the tile), so the loads can be batched
    const T* col = reg + uy * a.srx + ux + ja;
    T sA = T(0), sB = T(0);
#pragma unroll
    for (int jb = 0; jb < 8; ++jb) {
      const T g = col[jb * a.srx];
      sA += g * wy[jb];
      if (DERIV) sB += g * dwy[jb];
    }
    T vx, vy = T(0);
    if (DERIV) {
      vx = sum8_last(dwxa * sA);
      vy = sum8_last(wxa * sB);
    } else {
      vx = sum8_last(wxa * sA);
    }
    if (ja == 7 && valid) {
      const int64_t j = jout;
      const double cw = (double)r.w;
      if (DERIV) {
        out[j * a.ostride + 0] = cw * a.dscale[0] * (double)vx;
        out[j * a.ostride + 1] = cw * a.dscale[1] * (double)vy;
      } else {
        out[j * a.ostride + a.ocomp] = cw * (double)vx;
      }
    }
  }
}

// 2D interpolation of FP32 grids (w <= 8), one lane per point: the lane
// evaluates the point's 8 x and 8 y kernel values (and derivatives) itself
// and contracts its 8 x 8 cells from the staged region, x first (rows
// T[y] = sum_x g wx, and with dwx), then y; no cross-lane traffic.  In 2D the
// contraction is short (64 cells), so the eight-lane layout's shared
// broadcasts and reductions cost more than the redundancy they save.
template <bool DERIV>
__global__ void __launch_bounds__(256)
    k_interp2_lane1(const uint32_t* __restrict__ scell, const float4* __restrict__ rec,
                    const uint32_t* __restrict__ idx, const int* __restrict__ bin_start, const float* __restrict__ grid,
                    InterpDev<float> a, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* reg = reinterpret_cast<float*>(smem_raw);
  const int tid = threadIdx.x;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  bin_start += (int64_t)blockIdx.y * a.ntile[0] * a.ntile[1];
  grid += (int64_t)blockIdx.y * a.nf[0] * a.nf[1];
  const int b = blockIdx.x;
  const int beg0 = bin_start[b];
  const int len = bin_start[b + 1] - beg0;
  const int beg = beg0 + (int)(((int64_t)len * blockIdx.z) / gridDim.z);
  const int end = beg0 + (int)(((int64_t)len * (blockIdx.z + 1)) / gridDim.z);
  if (beg == end) return;  // uniform across the workgroup
  const int t0x = (b % a.ntile[0]) * a.tile[0];
  const int t0y = (b / a.ntile[0]) * a.tile[1];
  {
    const int lx = tid & 31;
    int gx = t0x + lx;
    while (gx >= a.nf[0]) gx -= a.nf[0];
    for (int r0 = 0; r0 < a.ry; r0 += 32) {
      float v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ly = r0 + q * 8 + (tid >> 5);
        v[q] = 0.f;
        if (ly < a.ry && lx < a.rx) {
          int gy = t0y + ly;
          while (gy >= a.nf[1]) gy -= a.nf[1];
          v[q] = grid[(int64_t)gy * a.nf[0] + gx];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ly = r0 + q * 8 + (tid >> 5);
        if (ly < a.ry && lx < a.rx) reg[ly * a.srx + lx] = v[q];
      }
    }
  }
  __syncthreads();
  const bool on[8] = {0 < a.w, 1 < a.w, 2 < a.w, 3 < a.w, 4 < a.w, 5 < a.w, 6 < a.w, 7 < a.w};
  for (int k = beg + tid; k < end; k += 256) {
    const uint32_t sc = scell[k];
    const float4 r = rec[k];
    const uint32_t jout = idx[k];
    int ux = (int)(sc & cmask) - t0x;
    if (ux < 0) ux += a.nf[0];
    int uy = (int)((sc >> a.cbits) & cmask) - t0y;
    if (uy < 0) uy += a.nf[1];
    float wx[8], dwx[8], wy[8], dwy[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      wx[j] = dwx[j] = wy[j] = dwy[j] = 0.f;
      if (DERIV) {
        float v0, d0, v1, d1;
        es_phi_dphi<float>((float(j) - r.x) * a.k.two_over_w, a.k, v0, d0);
        es_phi_dphi<float>((float(j) - r.y) * a.k.two_over_w, a.k, v1, d1);
        wx[j] = on[j] ? v0 : 0.f;
        dwx[j] = on[j] ? d0 : 0.f;
        wy[j] = on[j] ? v1 : 0.f;
        dwy[j] = on[j] ? d1 : 0.f;
      } else {
        const float v0 = es_phi<float>((float(j) - r.x) * a.k.two_over_w, a.k);
        const float v1 = es_phi<float>((float(j) - r.y) * a.k.two_over_w, a.k);
        wx[j] = on[j] ? v0 : 0.f;
        wy[j] = on[j] ? v1 : 0.f;
      }
    }
    const float* base = reg + uy * a.srx + ux;
    float vx = 0.f, vy = 0.f;
#pragma unroll
    for (int y = 0; y < 8; ++y) {
      const float* row = base + y * a.srx;
      float t1 = 0.f, t2 = 0.f;
#pragma unroll
      for (int x = 0; x < 8; ++x) {
        const float g = row[x];
        t1 += g * wx[x];
        if (DERIV) t2 += g * dwx[x];
      }
      if (DERIV) {
        vx += wy[y] * t2;
        vy += dwy[y] * t1;
      } else {
        vx += wy[y] * t1;
      }
    }
    const int64_t j = jout;
    const double cw = (double)r.w;
    if (DERIV) {
      out[j * a.ostride + 0] = cw * a.dscale[0] * (double)vx;
      out[j * a.ostride + 1] = cw * a.dscale[1] * (double)vy;
    } else {
      out[j * a.ostride + a.ocomp] = cw * (double)vx;
    }
  }
}

template <typename T, bool DERIV>
static bool launch_interp2_g8(const SortedPoints<T>& p, const T* grid, const double* dscale, double* out,
                              int ostride, int ocomp, hipStream_t s) {
  InterpDev<T> a;
  const GridGeom& g = p.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.dscale[q] = dscale ? dscale[q] : 0.0;
  }
  a.cbits = g.cell_bits;
  a.w = p.es.w;
  a.rx = g.tile[0] + 8 - 1;
  a.ry = g.tile[1] + 8 - 1;
  a.rz = 1;
  a.srx = a.rx;
  a.ostride = ostride;
  a.ocomp = ocomp;
  a.k.beta = (T)p.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(p.es.beta * M_LOG2E) : (T)p.es.beta;
  a.k.two_over_w = (T)(2.0 / p.es.w);
  const size_t lds = sizeof(T) * (size_t)a.rx * a.ry;
  if (lds > 64 * 1024 || a.rx > 32) return false;
  const int64_t wg = (int64_t)g.nbins * g.batch;
  const int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(8, (2048 + wg - 1) / wg));
  if constexpr (sizeof(T) == 4) {
    // FRESCO_INTERP2=g8: the eight-lane kernel instead
    const char* env = std::getenv("FRESCO_INTERP2");
    if (!(env && std::string(env) == "g8")) {
      hipLaunchKernelGGL((k_interp2_lane1<DERIV>), dim3(g.nbins, g.batch, chunks), dim3(256), lds, s,
                         p.scell.get(), p.rec.get(), p.idx.get(), p.bin_start.get(), grid, a, out);
      FRESCO_HIP_CHECK(hipGetLastError());
      return true;
    }
  }
  hipLaunchKernelGGL((k_interp2_g8<T, DERIV>), dim3(g.nbins, g.batch, chunks), dim3(G8_THREADS), lds, s,
                     p.scell.get(), p.rec.get(), p.idx.get(), p.bin_start.get(), grid, a, out);
  FRESCO_HIP_CHECK(hipGetLastError());
  return true;
}

template <typename T, int D, int WMAX>
static void launch_interp(const SortedPoints<T>& p, const T* grid, bool deriv, const double* dscale, double* out,
                          int ostride, int ocomp, hipStream_t s) {
  InterpDev<T> a;
  const GridGeom& g = p.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.dscale[q] = dscale ? dscale[q] : 0.0;
  }
  a.cbits = g.cell_bits;
  a.w = p.es.w;
  a.rx = g.tile[0] + a.w - 1;
  a.ry = D >= 2 ? g.tile[1] + a.w - 1 : 1;
  a.rz = D >= 3 ? g.tile[2] + a.w - 1 : 1;
  a.srx = D >= 2 ? lds_row_stride(a.rx, (int)sizeof(T), a.w) : a.rx;
  a.ostride = ostride;
  a.ocomp = ocomp;
  a.k.beta = (T)p.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(p.es.beta * M_LOG2E) : (T)p.es.beta;
  a.k.two_over_w = (T)(2.0 / p.es.w);
  size_t lds = interp_lds_bytes(g, (int)sizeof(T), a.w);
  if (lds > 160 * 1024) throw InvalidArgument("interpolation region exceeds LDS; use a smaller tile");
  dim3 grd(g.nbins, g.batch), blk(256);
  if (deriv)
    hipLaunchKernelGGL((k_interp<T, D, WMAX, true>), grd, blk, lds, s, p.scell.get(), p.rec.get(), p.idx.get(),
                       p.bin_start.get(), grid, a, out);
  else
    hipLaunchKernelGGL((k_interp<T, D, WMAX, false>), grd, blk, lds, s, p.scell.get(), p.rec.get(), p.idx.get(),
                       p.bin_start.get(), grid, a, out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

template <typename T>
void interpolate(const SortedPoints<T>& p, const T* grid, bool deriv, const double* dscale, double* out,
                 int ostride, int ocomp, hipStream_t s) {
  const int d = p.geom.d;
  const bool small = p.es.w <= 8;
  if (d == 3) {
    // lane per point, except the FP64 derivative with w > 8 (its 4 x 16 FP64
    // weight arrays would spill): wavefront per point there
    const bool fits = p.geom.tile[0] + 8 - 1 <= 32;
    if (small && fits) {
      deriv ? launch_interp3_g8<T, true>(p, grid, dscale, out, ostride, ocomp, s)
            : launch_interp3_g8<T, false>(p, grid, dscale, out, ostride, ocomp, s);
    } else if (small) {
      deriv ? launch_interp3_lane<T, 8, true>(p, grid, dscale, out, ostride, ocomp, s)
            : launch_interp3_lane<T, 8, false>(p, grid, dscale, out, ostride, ocomp, s);
    } else if (!deriv) {
      if (!launch_interp3_g16<T, false>(p, grid, dscale, out, ostride, ocomp, s))
        launch_interp3_lane<T, 16, false>(p, grid, dscale, out, ostride, ocomp, s);
    } else if (launch_interp3_g16<T, true>(p, grid, dscale, out, ostride, ocomp, s)) {
    } else if (sizeof(T) == 4) {
      launch_interp3_lane<T, 16, true>(p, grid, dscale, out, ostride, ocomp, s);
    } else {
      launch_interp<T, 3, 16>(p, grid, deriv, dscale, out, ostride, ocomp, s);
    }
  } else if (d == 2) {
    if (small && (deriv ? launch_interp2_g8<T, true>(p, grid, dscale, out, ostride, ocomp, s)
                        : launch_interp2_g8<T, false>(p, grid, dscale, out, ostride, ocomp, s)))
      return;
    small ? launch_interp<T, 2, 8>(p, grid, deriv, dscale, out, ostride, ocomp, s)
          : launch_interp<T, 2, 16>(p, grid, deriv, dscale, out, ostride, ocomp, s);
  } else {
    launch_interp<T, 1, 16>(p, grid, deriv, dscale, out, ostride, ocomp, s);
  }
}

// ---------------------------------------------------------------------------
// complex-valued interpolation (type-2 of complex modes): one wavefront per
// point, lanes over the (x, y) kernel pairs, loop over z, cells read from the
// complex grid in HBM/L2; DPP reduction of the real and imaginary parts.
// ---------------------------------------------------------------------------

template <typename T, int D, int WMAX>
__global__ void __launch_bounds__(256)
    k_interp_c(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
               const uint32_t* __restrict__ idx, const T* __restrict__ grid, InterpDev<T> a, int64_t n,
               double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n) return;  // wavefront-uniform
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  constexpr int SH = Lanes<WMAX>::SH;
  constexpr int NCH = D == 1 ? 1 : (WMAX * WMAX + 63) / 64;
  const uint32_t sc = scell[k];
  const typename Vec4<T>::type r = rec[k];
  int s[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) s[q] = q < D ? (int)((sc >> (q * a.cbits)) & cmask) : 0;
  const int kd = lane >> SH;
  const int kj = lane & (WMAX - 1);
  const T offk = kd == 0 ? r.x : (kd == 1 ? r.y : r.z);
  const T kv = (kd < D && kj < w) ? es_phi<T>((T(kj) - offk) * a.k.two_over_w, a.k) : T(0);
  T cr = T(0), ci = T(0);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int pr = lane + 64 * ch;
    const int ja = D == 1 ? pr : (pr & (WMAX - 1));
    const int jb = D == 1 ? 0 : (pr >> SH);
    const bool act = ja < w && (D == 1 || jb < w);
    const T wx = lane_get(kv, act ? ja : 0);
    const T wy = D >= 2 ? lane_get(kv, WMAX + (act ? jb : 0)) : T(1);
    int gx = s[0] + ja;
    if (gx >= a.nf[0]) gx -= a.nf[0];
    int gy = D >= 2 ? s[1] + jb : 0;
    if (D >= 2 && gy >= a.nf[1]) gy -= a.nf[1];
    T sr = T(0), si = T(0);
    if (D == 3) {
#pragma unroll
      for (int jc = 0; jc < WMAX; ++jc) {
        if (jc < w) {
          const T wz = lane_read(kv, 2 * WMAX + jc);
          int gz = s[2] + jc;
          if (gz >= a.nf[2]) gz -= a.nf[2];
          if (act) {
            const int64_t c = ((int64_t)gz * a.nf[1] + gy) * a.nf[0] + gx;
            sr += grid[2 * c] * wz;
            si += grid[2 * c + 1] * wz;
          }
        }
      }
    } else if (act) {
      const int64_t c = (int64_t)gy * a.nf[0] + gx;
      sr = grid[2 * c];
      si = grid[2 * c + 1];
    }
    if (act) {
      cr += wx * wy * sr;
      ci += wx * wy * si;
    }
  }
  cr = wave_sum63(cr);
  ci = wave_sum63(ci);
  if (lane == 63) {
    const int64_t j = idx[k];
    out[2 * j] = (double)r.w * (double)cr;
    out[2 * j + 1] = (double)r.w * (double)ci;
  }
}

template <typename T, int D, int WMAX>
static void launch_interp_c(const SortedPoints<T>& p, const Cplx<T>* grid, double* out, hipStream_t s) {
  InterpDev<T> a;
  const GridGeom& g = p.geom;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = g.nf[q];
    a.tile[q] = g.tile[q];
    a.ntile[q] = g.ntile[q];
    a.dscale[q] = 0.0;
  }
  a.cbits = g.cell_bits;
  a.w = p.es.w;
  a.rx = a.ry = a.rz = a.srx = 0;
  a.ostride = 2;
  a.ocomp = 0;
  a.k.beta = (T)p.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(p.es.beta * M_LOG2E) : (T)p.es.beta;
  a.k.two_over_w = (T)(2.0 / p.es.w);
  hipLaunchKernelGGL((k_interp_c<T, D, WMAX>), dim3(blocks_for(p.n, 4)), dim3(256), 0, s, p.scell.get(), p.rec.get(),
                     p.idx.get(), reinterpret_cast<const T*>(grid), a, p.n, out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

template <typename T>
void interpolate_complex(const SortedPoints<T>& p, const Cplx<T>* grid, double* out, hipStream_t s) {
  const bool small = p.es.w <= 8;
  if (p.geom.d == 3) {
    small ? launch_interp_c<T, 3, 8>(p, grid, out, s) : launch_interp_c<T, 3, 16>(p, grid, out, s);
  } else if (p.geom.d == 2) {
    small ? launch_interp_c<T, 2, 8>(p, grid, out, s) : launch_interp_c<T, 2, 16>(p, grid, out, s);
  } else {
    launch_interp_c<T, 1, 16>(p, grid, out, s);
  }
}

// ---------------------------------------------------------------------------
// rocFFT
// ---------------------------------------------------------------------------

static void rocfft_check(rocfft_status st, const char* what) {
  if (st != rocfft_status_success)
    throw HipFailure(std::string("rocFFT ") + what + " failed with status " + std::to_string((int)st));
}

static void rocfft_init_once() {
  static std::once_flag flag;
  std::call_once(flag, []() { rocfft_check(rocfft_setup(), "setup"); });
}

FftPlan::~FftPlan() {
  if (plan_) rocfft_plan_destroy(plan_);
  if (info_) rocfft_execution_info_destroy(info_);
}

void FftPlan::create(Kind k, bool dbl, int d, const int* len, int batch) {
  rocfft_init_once();
  size_t lengths[3];
  for (int a = 0; a < d; ++a) lengths[a] = (size_t)len[a];
  rocfft_transform_type tt = k == R2C       ? rocfft_transform_type_real_forward
                             : k == C2R     ? rocfft_transform_type_real_inverse
                             : k == C2C_FWD ? rocfft_transform_type_complex_forward
                                            : rocfft_transform_type_complex_inverse;
  rocfft_check(rocfft_plan_create(&plan_, rocfft_placement_notinplace, tt,
                                  dbl ? rocfft_precision_double : rocfft_precision_single, (size_t)d, lengths,
                                  (size_t)std::max(1, batch), nullptr),
               "plan_create");
  size_t wb = 0;
  rocfft_check(rocfft_plan_get_work_buffer_size(plan_, &wb), "get_work_buffer_size");
  rocfft_check(rocfft_execution_info_create(&info_), "execution_info_create");
  if (wb > 0) {
    work_.allocate(wb);
    rocfft_check(rocfft_execution_info_set_work_buffer(info_, work_.get(), wb), "set_work_buffer");
  }
}

void FftPlan::exec(void* in, void* out, hipStream_t s) {
  rocfft_check(rocfft_execution_info_set_stream(info_, s), "set_stream");
  void* ib[1] = {in};
  void* ob[1] = {out};
  rocfft_check(rocfft_execute(plan_, ib, ob, info_), "execute");
}

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_fill_zero(T* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = T{};
}

template <typename T>
void fill_zero(T* p, size_t n, hipStream_t s) {
  if (n == 0) return;
  unsigned blocks = (unsigned)std::min<size_t>(blocks_for(n, 256), 8192);
  hipLaunchKernelGGL(k_fill_zero<T>, dim3(blocks), dim3(256), 0, s, p, n);
  FRESCO_HIP_CHECK(hipGetLastError());
}

__global__ void k_reduce_partials(const double* __restrict__ part, int n, double* __restrict__ out,
                                  double* __restrict__ out_host) {
  block_sum_partials(part, n, out, out_host);
}

void reduce_partials(const double* partials, int n, double* out, hipStream_t s, double* out_host) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, partials, n, out, out_host);
  FRESCO_HIP_CHECK(hipGetLastError());
}

// explicit instantiations
template class SortedPoints<float>;
template class SortedPoints<double>;
template class Spreader<float>;
template class Spreader<double>;
template void interpolate<float>(const SortedPoints<float>&, const float*, bool, const double*, double*, int, int,
                                 hipStream_t);
template void interpolate<double>(const SortedPoints<double>&, const double*, bool, const double*, double*, int,
                                  int, hipStream_t);
template void interpolate_complex<float>(const SortedPoints<float>&, const Cplx<float>*, double*, hipStream_t);
template void interpolate_complex<double>(const SortedPoints<double>&, const Cplx<double>*, double*, hipStream_t);
template void fill_zero<float>(float*, size_t, hipStream_t);
template void fill_zero<double>(double*, size_t, hipStream_t);
template void fill_zero<unsigned int>(unsigned int*, size_t, hipStream_t);
template void fill_zero<Cplx<float>>(Cplx<float>*, size_t, hipStream_t);
template void fill_zero<Cplx<double>>(Cplx<double>*, size_t, hipStream_t);

}  // namespace fresco_hip
