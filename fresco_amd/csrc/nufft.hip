// Non-uniform FFT kernels for gfx950: point binning/sorting, owner-computes
// LDS-tiled spreading, wavefront-per-point interpolation, rocFFT plans.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "nufft.hpp"

namespace fresco_hip {

// ---------------------------------------------------------------------------
// ES kernel parameters and deconvolution factors (host, float64)
// ---------------------------------------------------------------------------

EsParams es_params(double eps, double sigma) {
  if (!(eps > 0)) throw InvalidArgument("eps must be > 0");
  if (!(sigma > 1.0)) throw InvalidArgument("upsampling factor sigma must be > 1");
  EsParams p;
  int w;
  double bw;
  if (sigma == 2.0) {
    w = (int)std::ceil(-std::log10(eps / 10.0));
    w = std::max(2, w);
    bw = 2.30;
    if (w == 2) bw = 2.20;
    if (w == 3) bw = 2.26;
    if (w == 4) bw = 2.38;
  } else {
    w = (int)std::ceil(-std::log(eps) / (M_PI * std::sqrt(1.0 - 1.0 / sigma)));
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

GridGeom make_geom(int d, const int* nf, int w, int elem_bytes) {
  GridGeom g;
  g.d = d;
  for (int a = 0; a < 3; ++a) g.nf[a] = a < d ? nf[a] : 1;
  // power-of-two tiles; shrink until there are enough tiles to fill 256 CUs.
  int t3 = 16, t2 = 64, t1 = 1024;
  for (int a = 0; a < 3; ++a) g.tile[a] = 1;
  if (d == 3) {
    for (int a = 0; a < 3; ++a) g.tile[a] = t3;
    auto ntiles = [&]() {
      long long t = 1;
      for (int a = 0; a < 3; ++a) t *= (g.nf[a] + g.tile[a] - 1) / g.tile[a];
      return t;
    };
    while (ntiles() < 1024 && g.tile[2] > 4) {
      for (int a = 2; a >= 0; --a)
        if (g.tile[a] > 4) {
          g.tile[a] /= 2;
          break;
        }
    }
  } else if (d == 2) {
    g.tile[0] = g.tile[1] = t2;
    auto ntiles = [&]() {
      return (long long)((g.nf[0] + g.tile[0] - 1) / g.tile[0]) * ((g.nf[1] + g.tile[1] - 1) / g.tile[1]);
    };
    while (ntiles() < 1024 && g.tile[1] > 8) {
      if (g.tile[1] >= g.tile[0])
        g.tile[1] /= 2;
      else
        g.tile[0] /= 2;
    }
  } else {
    g.tile[0] = t1;
    while ((g.nf[0] + g.tile[0] - 1) / g.tile[0] < 512 && g.tile[0] > 32) g.tile[0] /= 2;
  }
  for (int a = 0; a < 3; ++a) {
    g.tile[a] = std::min(g.tile[a], g.nf[a]);
    g.ntile[a] = (g.nf[a] + g.tile[a] - 1) / g.tile[a];
  }
  g.nbins = g.ntile[0] * g.ntile[1] * g.ntile[2];
  g.key_bits = ceil_log2(g.nbins + 1);
  g.cell_bits = d == 3 ? 10 : (d == 2 ? 16 : 31);
  for (int a = 0; a < d; ++a)
    if (g.nf[a] > (1 << g.cell_bits) - 1)
      throw InvalidArgument("fine grid too large for the packed cell index (" + std::to_string(g.nf[a]) + ")");
  (void)w;
  (void)elem_bytes;
  return g;
}

// ---------------------------------------------------------------------------
// point binning and sorting
// ---------------------------------------------------------------------------

struct MapDev {
  int d, wrap, w;
  double L[3], shift[3], scale[3];
  int nf[3], tile[3], ntile[3];
  int cell_bits;
};

__device__ __forceinline__ void map_component(double v, const MapDev& m, int a, int& cell, double& off) {
  double xs = m.wrap ? (v - round(v / m.L[a]) + m.shift[a]) * m.scale[a] : v * m.scale[a];
  double t = xs * (1.0 / (2.0 * M_PI));
  t -= floor(t);
  double gg = t * (double)m.nf[a];
  double i0 = floor(gg);
  double f = gg - i0;
  double sf = ceil(f - 0.5 * m.w);
  off = f - sf;  // in [w/2 - 1, w/2]
  long long ic = (long long)(isfinite(i0) ? i0 : 0.0) + (long long)(isfinite(sf) ? sf : 0.0);
  ic %= m.nf[a];
  if (ic < 0) ic += m.nf[a];
  cell = (int)ic;
}

__global__ void k_bin_keys(const double* __restrict__ x, int64_t n, MapDev m, uint32_t* __restrict__ key,
                           uint32_t* __restrict__ idx) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t b = 0, mul = 1;
  for (int a = 0; a < m.d; ++a) {
    int cell;
    double off;
    map_component(x[j * m.d + a], m, a, cell, off);
    int t = cell / m.tile[a];
    t = t < m.ntile[a] ? t : m.ntile[a] - 1;
    b += mul * (uint32_t)t;
    mul *= (uint32_t)m.ntile[a];
  }
  key[j] = b;
  idx[j] = (uint32_t)j;
}

__global__ void k_bin_start(const uint32_t* __restrict__ keys, int64_t n, int nbins, int* __restrict__ start) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < (uint32_t)b)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[b] = (int)lo;
}

template <typename T>
__global__ void k_permute(const double* __restrict__ x, int64_t n, MapDev m, const uint32_t* __restrict__ idx,
                          const T* __restrict__ c, uint32_t* __restrict__ scell,
                          typename Vec4<T>::type* __restrict__ rec) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t j = idx[k];
  double offs[3] = {0, 0, 0};
  uint32_t packed = 0;
  for (int a = 0; a < m.d; ++a) {
    int cell;
    map_component(x[(int64_t)j * m.d + a], m, a, cell, offs[a]);
    packed |= (uint32_t)cell << (a * m.cell_bits);
  }
  typename Vec4<T>::type r;
  r.x = (T)offs[0];
  r.y = (T)offs[1];
  r.z = (T)offs[2];
  r.w = c ? c[j] : T(1);
  scell[k] = packed;
  rec[k] = r;
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
  key.allocate(n);
  key_sorted.allocate(n);
  idx_in.allocate(n);
  idx.allocate(n);
  bin_start.allocate(g.nbins + 1);
  scell.allocate(n);
  rec.allocate(n);
  size_t bytes = 0;
  FRESCO_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key.get(), key_sorted.get(), idx_in.get(),
                                                      idx.get(), (int)n, 0, g.key_bits));
  sort_tmp_bytes = std::max<size_t>(bytes, 16);
  sort_tmp.allocate(sort_tmp_bytes);
}

template <typename T>
void SortedPoints<T>::set(const double* x, const CoordMap& cm, const T* c, hipStream_t s) {
  MapDev m = make_map_dev(geom, es, cm);
  const unsigned thr = 256;
  hipLaunchKernelGGL(k_bin_keys, dim3(blocks_for(n, thr)), dim3(thr), 0, s, x, n, m, key.get(), idx_in.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  size_t bytes = sort_tmp_bytes;
  FRESCO_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp.get(), bytes, key.get(), key_sorted.get(),
                                                      idx_in.get(), idx.get(), (int)n, 0, geom.key_bits, s));
  hipLaunchKernelGGL(k_bin_start, dim3(blocks_for(geom.nbins + 1, thr)), dim3(thr), 0, s, key_sorted.get(), n,
                     geom.nbins, bin_start.get());
  FRESCO_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_permute<T>, dim3(blocks_for(n, thr)), dim3(thr), 0, s, x, n, m, idx.get(), c, scell.get(),
                     rec.get());
  FRESCO_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// spreading: owner-computes tiles in LDS
// ---------------------------------------------------------------------------

template <typename T>
struct SpreadDev {
  int nf[3], tile[3], ntile[3];
  int cbits;
  int w;
  int sx;        // LDS row stride (tile[0] + pad)
  int slab;      // cells of the slowest tile dimension owned by each wave
  int maxc[3];   // candidate bins per dimension (table width)
  int cstride;   // table entries per dimension block
  EsDev<T> k;
};

template <int WMAX>
struct Lanes {
  static constexpr int SH = WMAX == 8 ? 3 : 4;
};

template <typename T, int D>
__device__ __forceinline__ T rec_comp(const typename Vec4<T>::type& r, int a) {
  return a == 0 ? r.x : (a == 1 ? r.y : r.z);
}

// One workgroup owns one tile.  Candidate points (those binned in tiles whose
// points can reach this tile) are staged 64 at a time: wave 0 compacts the
// ones whose support touches the tile, all threads tabulate their kernel
// values in LDS, then every wave adds the contributions falling in its slab of
// the tile (slabs along the slowest dimension are disjoint, so LDS updates
// need no atomics and each cell sees the points in one fixed order).
template <typename T, int D, int WMAX, int NW>
__global__ void __launch_bounds__(64 * NW)
    k_spread(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
             const int* __restrict__ bin_start, const int* __restrict__ cand, SpreadDev<T> a,
             T* __restrict__ grid) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;

  int b = blockIdx.x;
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
  const int tile_elems = a.sx * (D >= 2 ? a.tile[1] : 1) * (D >= 3 ? a.tile[2] : 1);
  T* tileb = reinterpret_cast<T*>(smem_raw);
  T* kv = tileb + tile_elems;                  // [64][D][WMAX]
  T* pcw = kv + 64 * D * WMAX;                 // [64]
  uint32_t* ps = reinterpret_cast<uint32_t*>(pcw + 64);  // [64]
  __shared__ int s_np;

  for (int e = tid; e < tile_elems; e += 64 * NW) tileb[e] = T(0);

  const int slo = wave * a.slab;
  const int shi = min(slo + a.slab, tl[D - 1]);

  const int nc0 = a.maxc[0];
  const int nc1 = D >= 2 ? a.maxc[1] : 1;
  const int nc2 = D >= 3 ? a.maxc[2] : 1;
  for (int c2 = 0; c2 < nc2; ++c2) {
    int bz = D >= 3 ? cand[2 * a.cstride + t[2] * a.maxc[2] + c2] : 0;
    if (bz < 0) continue;
    for (int c1 = 0; c1 < nc1; ++c1) {
      int by = D >= 2 ? cand[1 * a.cstride + t[1] * a.maxc[1] + c1] : 0;
      if (by < 0) continue;
      for (int c0 = 0; c0 < nc0; ++c0) {
        int bx = cand[t[0] * a.maxc[0] + c0];
        if (bx < 0) continue;
        int bin = bx + a.ntile[0] * (by + a.ntile[1] * bz);
        int beg = bin_start[bin];
        int end = bin_start[bin + 1];
        for (int base = beg; base < end; base += 64) {
          __syncthreads();  // previous batch fully consumed
          if (wave == 0) {
            int j = base + lane;
            bool touch = false;
            uint32_t sc = 0;
            T cw = T(0);
            typename Vec4<T>::type r;
            if (j < end) {
              sc = scell[j];
              r = rec[j];
              cw = r.w;
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
            int pos = __popcll(m & ((1ull << lane) - 1ull));
            if (touch) {
              ps[pos] = sc;
              pcw[pos] = cw;
#pragma unroll
              for (int q = 0; q < D; ++q) {
                T off = rec_comp<T, D>(r, q);
                T* dst = kv + (pos * D + q) * WMAX;
#pragma unroll
                for (int jj = 0; jj < WMAX; ++jj)
                  dst[jj] = jj < w ? es_phi<T>((T(jj) - off) * a.k.two_over_w, a.k) : T(0);
              }
            }
            if (lane == 0) s_np = __popcll(m);
          }
          __syncthreads();
          const int np = s_np;
          for (int p = 0; p < np; ++p) {
            const uint32_t sc = ps[p];
            const T cw = pcw[p];
            const T* kvp = kv + p * D * WMAX;
            int s[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) s[q] = q < D ? (int)((sc >> (q * a.cbits)) & cmask) : 0;
            if (D == 3) {
              // wave-uniform z range inside this wave's slab
              int uz0 = s[2] - t0[2];
              if (uz0 < 0) uz0 += a.nf[2];
              bool any = false;
              for (int jc = 0; jc < w; ++jc) {
                int uz = uz0 + jc;
                if (uz >= a.nf[2]) uz -= a.nf[2];
                any = any || (uz >= slo && uz < shi);
              }
              if (!any) continue;
              constexpr int SH = Lanes<WMAX>::SH;
              constexpr int NCH = (WMAX * WMAX + 63) / 64;
#pragma unroll
              for (int ch = 0; ch < NCH; ++ch) {
                int pr = lane + 64 * ch;
                int ja = pr & (WMAX - 1);
                int jb = pr >> SH;
                bool act = ja < w && jb < w;
                int ux = s[0] + ja - t0[0];
                if (ux >= a.nf[0]) ux -= a.nf[0];
                if (ux < 0) ux += a.nf[0];
                int uy = s[1] + jb - t0[1];
                if (uy >= a.nf[1]) uy -= a.nf[1];
                if (uy < 0) uy += a.nf[1];
                act = act && ux < tl[0] && uy < tl[1];
                T v = kvp[ja] * kvp[WMAX + (jb < WMAX ? jb : 0)] * cw;
                T* rowp = tileb + uy * a.sx + ux;
#pragma unroll
                for (int jc = 0; jc < WMAX; ++jc) {
                  if (jc < w) {
                    int uz = uz0 + jc;
                    if (uz >= a.nf[2]) uz -= a.nf[2];
                    if (uz >= slo && uz < shi && act) rowp[uz * ty * a.sx] += v * kvp[2 * WMAX + jc];
                  }
                }
              }
            } else if (D == 2) {
              constexpr int SH = Lanes<WMAX>::SH;
              constexpr int NCH = (WMAX * WMAX + 63) / 64;
#pragma unroll
              for (int ch = 0; ch < NCH; ++ch) {
                int pr = lane + 64 * ch;
                int ja = pr & (WMAX - 1);
                int jb = pr >> SH;
                bool act = ja < w && jb < w;
                int ux = s[0] + ja - t0[0];
                if (ux >= a.nf[0]) ux -= a.nf[0];
                if (ux < 0) ux += a.nf[0];
                int uy = s[1] + jb - t0[1];
                if (uy >= a.nf[1]) uy -= a.nf[1];
                if (uy < 0) uy += a.nf[1];
                act = act && ux < tl[0] && uy >= slo && uy < shi;
                if (act) tileb[uy * a.sx + ux] += kvp[ja] * kvp[WMAX + jb] * cw;
              }
            } else {
              int ja = lane;
              bool act = ja < w;
              int ux = s[0] + ja - t0[0];
              if (ux >= a.nf[0]) ux -= a.nf[0];
              if (ux < 0) ux += a.nf[0];
              act = act && ux >= slo && ux < shi;
              if (act) tileb[ux] += kvp[ja < WMAX ? ja : 0] * cw;
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
  for (int e = tid; e < tot; e += 64 * NW) {
    int ux = e % nx;
    int r = e / nx;
    int uy = r % ny;
    int uz = r / ny;
    int64_t gidx = ((int64_t)(t0[2] + uz) * a.nf[1] + (t0[1] + uy)) * a.nf[0] + (t0[0] + ux);
    grid[gidx] = tileb[(uz * ty + uy) * a.sx + ux];
  }
}

template <typename T>
void Spreader<T>::init(const GridGeom& g, const EsParams& e) {
  geom = g;
  es = e;
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
  int sx = g.tile[0] + (g.d >= 2 ? 8 : 0);
  int ny = g.d >= 2 ? g.tile[1] : 1;
  int nz = g.d >= 3 ? g.tile[2] : 1;
  int wmax = e.w <= 8 ? 8 : 16;
  lds_bytes = (int)(sizeof(T) * ((size_t)sx * ny * nz + 64 * g.d * wmax + 64) + 64 * sizeof(uint32_t));
  nwaves = g.d == 1 ? 1 : 4;
  if (lds_bytes > 160 * 1024) throw InvalidArgument("spreader tile exceeds LDS");
}

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
  a.sx = g.tile[0] + (D >= 2 ? 8 : 0);
  a.cstride = sp.cand_stride;
  a.k.beta = (T)sp.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(sp.es.beta * M_LOG2E) : (T)sp.es.beta;
  a.k.two_over_w = (T)(2.0 / sp.es.w);
  if (D == 1) {
    a.slab = g.tile[0];
    hipLaunchKernelGGL((k_spread<T, D, WMAX, 1>), dim3(g.nbins), dim3(64), sp.lds_bytes, s, p.scell.get(),
                       p.rec.get(), p.bin_start.get(), sp.cand.get(), a, grid);
  } else {
    a.slab = (g.tile[D - 1] + 3) / 4;
    hipLaunchKernelGGL((k_spread<T, D, WMAX, 4>), dim3(g.nbins), dim3(256), sp.lds_bytes, s, p.scell.get(),
                       p.rec.get(), p.bin_start.get(), sp.cand.get(), a, grid);
  }
  FRESCO_HIP_CHECK(hipGetLastError());
}

template <typename T>
void Spreader<T>::spread(const SortedPoints<T>& p, T* grid, hipStream_t s) const {
  const int d = geom.d;
  const bool small = es.w <= 8;
  if (d == 3) {
    small ? launch_spread<T, 3, 8>(*this, p, grid, s) : launch_spread<T, 3, 16>(*this, p, grid, s);
  } else if (d == 2) {
    small ? launch_spread<T, 2, 8>(*this, p, grid, s) : launch_spread<T, 2, 16>(*this, p, grid, s);
  } else {
    small ? launch_spread<T, 1, 8>(*this, p, grid, s) : launch_spread<T, 1, 16>(*this, p, grid, s);
  }
}

// ---------------------------------------------------------------------------
// interpolation: one wavefront per point
// ---------------------------------------------------------------------------

template <typename T>
struct InterpDev {
  int64_t n;
  int nf[3];
  int cbits;
  int w;
  int ostride, ocomp;
  double dscale[3];
  EsDev<T> k;
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

template <typename T, int D, int WMAX, bool DERIV>
__global__ void __launch_bounds__(256)
    k_interp(const uint32_t* __restrict__ scell, const typename Vec4<T>::type* __restrict__ rec,
             const uint32_t* __restrict__ idx, const T* __restrict__ grid, InterpDev<T> a,
             double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= a.n) return;  // wave-uniform
  const int w = a.w;
  const uint32_t cmask = (1u << a.cbits) - 1u;
  const uint32_t sc = scell[k];
  const typename Vec4<T>::type r = rec[k];
  int s[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) s[q] = q < D ? (int)((sc >> (q * a.cbits)) & cmask) : 0;
  T off[3] = {r.x, r.y, r.z};

  // kernel values along the slowest dimension: lane j < w holds entry j
  T zv = T(0), zd = T(0);
  if (D == 3 && lane < w) es_phi_dphi<T>((T(lane) - off[2]) * a.k.two_over_w, a.k, zv, zd);

  T c0 = T(0), c1 = T(0), c2 = T(0);
  constexpr int SH = Lanes<WMAX>::SH;
  constexpr int NCH = D == 1 ? 1 : (WMAX * WMAX + 63) / 64;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    int pr = lane + 64 * ch;
    int ja = D == 1 ? pr : (pr & (WMAX - 1));
    int jb = D == 1 ? 0 : (pr >> SH);
    bool act = ja < w && (D == 1 || jb < w);
    T px, dx, py = T(1), dy = T(0);
    es_phi_dphi<T>((T(ja) - off[0]) * a.k.two_over_w, a.k, px, dx);
    if (D >= 2) es_phi_dphi<T>((T(jb) - off[1]) * a.k.two_over_w, a.k, py, dy);
    int gx = s[0] + ja;
    if (gx >= a.nf[0]) gx -= a.nf[0];
    int gy = D >= 2 ? s[1] + jb : 0;
    if (D >= 2 && gy >= a.nf[1]) gy -= a.nf[1];
    T s0 = T(0), s1 = T(0);
    if (D == 3) {
#pragma unroll
      for (int jc = 0; jc < WMAX; ++jc) {
        if (jc < w) {
          T wv = __shfl(zv, jc, 64);
          T wd = __shfl(zd, jc, 64);
          int gz = s[2] + jc;
          if (gz >= a.nf[2]) gz -= a.nf[2];
          if (act) {
            T v = grid[((int64_t)gz * a.nf[1] + gy) * a.nf[0] + gx];
            s0 += v * wv;
            s1 += v * wd;
          }
        }
      }
    } else {
      if (act) s0 = grid[(int64_t)gy * a.nf[0] + gx];
    }
    if (!act) {
      s0 = T(0);
      s1 = T(0);
    }
    if (DERIV) {
      c0 += dx * py * s0;
      if (D >= 2) c1 += px * dy * s0;
      if (D >= 3) c2 += px * py * s1;
    } else {
      c0 += px * py * s0;
    }
  }
  c0 = wave_sum(c0);
  if (DERIV && D >= 2) c1 = wave_sum(c1);
  if (DERIV && D >= 3) c2 = wave_sum(c2);
  if (lane == 0) {
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

template <typename T, int D, int WMAX>
static void launch_interp(const SortedPoints<T>& p, const T* grid, bool deriv, const double* dscale, double* out,
                          int ostride, int ocomp, hipStream_t s) {
  InterpDev<T> a;
  a.n = p.n;
  for (int q = 0; q < 3; ++q) {
    a.nf[q] = p.geom.nf[q];
    a.dscale[q] = dscale ? dscale[q] : 0.0;
  }
  a.cbits = p.geom.cell_bits;
  a.w = p.es.w;
  a.ostride = ostride;
  a.ocomp = ocomp;
  a.k.beta = (T)p.es.beta;
  a.k.bexp = sizeof(T) == 4 ? (T)(p.es.beta * M_LOG2E) : (T)p.es.beta;
  a.k.two_over_w = (T)(2.0 / p.es.w);
  dim3 grd(blocks_for(p.n, 4)), blk(256);
  if (deriv)
    hipLaunchKernelGGL((k_interp<T, D, WMAX, true>), grd, blk, 0, s, p.scell.get(), p.rec.get(), p.idx.get(), grid,
                       a, out);
  else
    hipLaunchKernelGGL((k_interp<T, D, WMAX, false>), grd, blk, 0, s, p.scell.get(), p.rec.get(), p.idx.get(),
                       grid, a, out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

template <typename T>
void interpolate(const SortedPoints<T>& p, const T* grid, bool deriv, const double* dscale, double* out,
                 int ostride, int ocomp, hipStream_t s) {
  const int d = p.geom.d;
  const bool small = p.es.w <= 8;
  if (d == 3) {
    small ? launch_interp<T, 3, 8>(p, grid, deriv, dscale, out, ostride, ocomp, s)
          : launch_interp<T, 3, 16>(p, grid, deriv, dscale, out, ostride, ocomp, s);
  } else if (d == 2) {
    small ? launch_interp<T, 2, 8>(p, grid, deriv, dscale, out, ostride, ocomp, s)
          : launch_interp<T, 2, 16>(p, grid, deriv, dscale, out, ostride, ocomp, s);
  } else {
    launch_interp<T, 1, 16>(p, grid, deriv, dscale, out, ostride, ocomp, s);
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

void FftPlan::create(Kind k, bool dbl, int d, const int* len) {
  rocfft_init_once();
  size_t lengths[3];
  for (int a = 0; a < d; ++a) lengths[a] = (size_t)len[a];
  rocfft_transform_type tt = k == R2C       ? rocfft_transform_type_real_forward
                             : k == C2R     ? rocfft_transform_type_real_inverse
                             : k == C2C_FWD ? rocfft_transform_type_complex_forward
                                            : rocfft_transform_type_complex_inverse;
  rocfft_check(rocfft_plan_create(&plan_, rocfft_placement_notinplace, tt,
                                  dbl ? rocfft_precision_double : rocfft_precision_single, (size_t)d, lengths, 1,
                                  nullptr),
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

__global__ void k_reduce_partials(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double ss[256];
  __shared__ double cc[256];
  double sum = 0.0, comp = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    double y = part[i] - comp;
    double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  ss[threadIdx.x] = sum;
  cc[threadIdx.x] = comp;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      // two compensated partial sums: add with a TwoSum and keep the error
      double a = ss[threadIdx.x], b = ss[threadIdx.x + h];
      double s2 = a + b;
      double bb = s2 - a;
      double err = (a - (s2 - bb)) + (b - bb);
      ss[threadIdx.x] = s2;
      cc[threadIdx.x] = cc[threadIdx.x] + cc[threadIdx.x + h] - err;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = ss[0] - cc[0];
}

void reduce_partials(const double* partials, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, partials, n, out);
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
template void fill_zero<float>(float*, size_t, hipStream_t);
template void fill_zero<double>(double*, size_t, hipStream_t);
template void fill_zero<Cplx<float>>(Cplx<float>*, size_t, hipStream_t);
template void fill_zero<Cplx<double>>(Cplx<double>*, size_t, hipStream_t);

}  // namespace fresco_hip
