// Non-uniform FFT building blocks on gfx950:
//   * SortedPoints : coordinate map -> fine-grid cell + offset, bin (tile) key,
//                    stable radix sort by bin, bin offsets, sorted point records.
//   * Spreader     : type-1 spreading, one wavefront per owner-computes tile
//                    accumulated in LDS with in-order non-returning LDS adds
//                    (each fine-grid cell written to HBM exactly once, no
//                    global atomics, fixed summation order -> reproducible).
//   * interpolate  : type-2 interpolation from the bin's grid region staged
//                    in LDS, one wavefront per point, DPP reduction; either
//                    the value or the gradient through the kernel derivative.
//   * FftPlan      : rocFFT plan + work buffer + execution info.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"
#include "es_kernel.hpp"
#include "sort.hpp"

struct rocfft_plan_t;
struct rocfft_execution_info_t;

namespace fresco_hip {

template <typename T>
struct Cplx {
  T re, im;
};

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

// Coordinate map applied to each component of a point before it is placed on
// the fine grid:  wrap=1: x = (p - round(p/L) + shift) * scale   (the
// reference's map, nuwu.hpp:160,219);  wrap=0: x = p * scale.
// The fine-grid coordinate is then g = frac(x / 2pi) * nf.
struct CoordMap {
  int wrap = 1;
  double L[3] = {1, 1, 1};
  double shift[3] = {0, 0, 0};
  double scale[3] = {1, 1, 1};
};

struct GridGeom {
  int d = 3;
  int nf[3] = {1, 1, 1};     // fine grid cells per dimension
  int tile[3] = {1, 1, 1};   // spreader tile (= bin) size per dimension
  int ntile[3] = {1, 1, 1};  // tiles per dimension
  int nbins = 1;   // bins per pattern
  int batch = 1;   // independent patterns, each with its own grid and bins
  int key_bits = 1;
  int cell_bits = 10;  // bits per dimension of the packed first-cell word
  int64_t ncells() const { return int64_t(nf[0]) * nf[1] * nf[2]; }
};

// Choose tiles for a fine grid (power-of-two tiles, enough of them to fill
// the GPU, LDS budget respected).
GridGeom make_geom(int d, const int* nf, int w, int elem_bytes, int batch = 1);
int lds_row_stride(int n, int elem_bytes, int w);
size_t spread_lds_bytes(const GridGeom& g, int elem_bytes, int w);
size_t interp_lds_bytes(const GridGeom& g, int elem_bytes, int w);

// Sorted non-uniform point set shared by the spreader and the interpolator.
template <typename T>
class SortedPoints {
 public:
  using R4 = typename Vec4<T>::type;
  void init(const GridGeom& g, const EsParams& es, int64_t n);
  // batched plans whose points are stored pattern by pattern: pattern p owns
  // points [starts[p], starts[p+1]); the bin sort then runs one pass per
  // pattern instead of sorting the combined keys
  void set_segments(const std::vector<int64_t>& starts);
  // x: device float64 [n*d]; c: device T[n] weights (may be null -> 1);
  // pat: device pattern index of every point (geom.batch > 1), else null.
  void set(const double* x, const CoordMap& m, const T* c, hipStream_t s, const uint32_t* pat = nullptr);

  int64_t n = 0;
  GridGeom geom;
  EsParams es;
  DeviceBuffer<uint32_t> key, key_sorted, idx_in, idx;  // idx: sorted -> original
  // direct: one counting pass over tiles (all bins of the plan fit one
  // 11-bit digit, no segments); else the radix sort over the keys
  bool direct = false;
  int interp_lpp = 4, interp_wgs = 2048;  // 3D interpolation layout (set at init)
  bool interp_units = true;
  BinSort sorter;                                       // stable LSD radix sort (sort.hip)
  DeviceBuffer<int> bin_start;                          // [batch * nbins + 1]
  DeviceBuffer<uint32_t> scell;                         // packed first cell per sorted point
  DeviceBuffer<R4> rec;                                 // (off_x, off_y, off_z, c) per sorted point
  DeviceBuffer<uint32_t> scell_in;                      // the same in the original order (radix sort
  DeviceBuffer<R4> rec_in;                              // input; key too: allocated for that path only)

 private:
  void init_sorter();
  int ntiles_ = 0, gsz_ = 16, ngroups_ = 0;
  DeviceBuffer<uint32_t> tcnt_, gsum_;  // direct binning: per-tile counts, per-group totals
  // segmented direct binning (points stored pattern by pattern, each
  // pattern's bins fit one counting pass): tiles within patterns
  bool seg_direct_ = false;
  int seg_ntiles_ = 0;
  DeviceBuffer<int> seg_tbeg_, seg_tpat_, seg_ptile_, seg_pstart_;
};

template <typename T>
class Spreader {
 public:
  void init(const GridGeom& g, const EsParams& es);
  // grid: device T[ncells], fully overwritten.
  void spread(const SortedPoints<T>& p, T* grid, hipStream_t s) const;
  // complex strengths str[n] in original point order (times the record
  // weight) onto a complex grid[ncells], fully overwritten.
  void spread_complex(const SortedPoints<T>& p, const Cplx<T>* str, Cplx<T>* grid, hipStream_t s) const;
  // 3D spreader configuration for these points: a cached pick for the plan
  // shape (user cache, then the shipped table), else the fastest of the
  // candidates timed here (then stored in the user cache)
  void autotune(const SortedPoints<T>& p, T* grid, hipStream_t s) const;
  void autotune_timed(const SortedPoints<T>& p, T* grid, hipStream_t s) const;
  std::string tune_key(int64_t npts) const;
  // where the 3D spreader choice came from: 0 none (single candidate),
  // 1 tuning cache or table, 2 timed here, 3 pinned by FRESCO_SPREAD3
  mutable int tune_source = 0;
  // complex spreading of a point set that stays fixed (the k points of
  // UwNU / NUwNU): point-range slices per tile chosen from its bin counts
  // (one host read, at set-up)
  void plan_complex_slices(const SortedPoints<T>& p);

  GridGeom geom;
  EsParams es;
  int maxc[3] = {1, 1, 1};
  DeviceBuffer<int> cand;  // [3][maxtile][MAXC] candidate bins per tile, -1 padded
  int cand_stride = 0;     // entries per dimension block
  int lds_bytes = 0;
  int parts = 1;             // register-tile spreader: point-range slices per tile
  int lparts = 1;            // LDS-tile spreader (1D / 2D, real): point-range slices per tile
  int cparts = 1;            // the same for complex spreading (plan_complex_slices)
  DeviceBuffer<T> part_buf_c;  // [cparts][ncells][2]
  DeviceBuffer<T> part_buf;  // [parts][ncells] / [lparts][batch][ncells] partial grids
  // input-driven halo spreader (3D, w <= 8)
  // HALO4: the halo spreader with four wavefronts per tile (z-planes split
  // by residue mod 4); HALOM: FP32 halo spreading as matrix products of the
  // points that share a z support, two x windows per tile; HALOR: matrix
  // products of the points that share a y support, the y rows kept in a
  // ring, one tile per x window
  enum Mode { REG = 0, HALO = 1, HALO4 = 2, HALOM = 3, HALOR = 4 };
  mutable int mode = REG;
  bool halo_ok = false;
  mutable bool tuned = true;   // false: time the candidates on the first spread
  mutable int halo_parts = 1;  // point-range slices per bin (HALOM, HALOR: 2 windows per slice)
  int max_halo_parts = 1;
  DeviceBuffer<T> halo_buf;    // [nbins][halo_parts][halo tile]
  DeviceBuffer<int> halo_tab;  // per-dimension coverage tables for the merge
  // input-driven 2D spreader (real strengths, w <= 8)
  bool halo2_ok = false;
  int halo2_hs = 0, halo2_rows = 0;  // halo tile row stride and rows
  int halo2_parts = 1;               // point-range slices per bin
  bool halo2_mfma = false;           // FP32 16 x 16 bins: matrix-core spreader
};

// Interpolation of a real fine grid at the sorted points.
//  deriv=false: out[idx*ostride + ocomp] = c * sum_l b_l psi(l - g)
//  deriv=true : out[idx*ostride + a] = c * dscale[a] * sum_l b_l dpsi_a(l - g), a < d
template <typename T>
void interpolate(const SortedPoints<T>& p, const T* grid, bool deriv, const double* dscale,
                 double* out, int ostride, int ocomp, hipStream_t s);

// Complex-valued interpolation: out[idx] = c * sum_l b_l psi(l - g) (complex,
// float64 interleaved, original point order).
template <typename T>
void interpolate_complex(const SortedPoints<T>& p, const Cplx<T>* grid, double* out, hipStream_t s);

// rocFFT plan for one transform on a d-dimensional grid (lengths fastest first).
class FftPlan {
 public:
  enum Kind { R2C, C2R, C2C_FWD, C2C_INV };
  FftPlan() = default;
  ~FftPlan();
  FftPlan(const FftPlan&) = delete;
  FftPlan& operator=(const FftPlan&) = delete;
  // batch > 1: that many transforms, packed one after another
  void create(Kind k, bool dbl, int d, const int* lengths_fastest_first, int batch = 1);
  void exec(void* in, void* out, hipStream_t s);

 private:
  rocfft_plan_t* plan_ = nullptr;
  rocfft_execution_info_t* info_ = nullptr;
  DeviceBuffer<unsigned char> work_;
};

// out[i] = 0 over n elements of T (graph-capture safe fill).
template <typename T>
void fill_zero(T* p, size_t n, hipStream_t s);

// Deterministic FP64 sum of partials[n] -> *out (device), Kahan-compensated,
// and -> *out_host (device-visible pinned host memory) when not null.
void reduce_partials(const double* partials, int n, double* out, hipStream_t s, double* out_host = nullptr);

// The same compensated, fixed-order sum by the first NT threads of one
// workgroup (every thread of it calls it, blockDim >= NT; NT a power of two):
// part[n] -> *out, and *out_host (device-visible pinned host memory) when
// not null.  The order depends on NT only, so every caller uses 256.
template <int NT = 256>
__device__ __forceinline__ void block_sum_partials(const double* part, int n, double* out, double* out_host) {
  __shared__ double ss[NT];
  __shared__ double cc[NT];
  if ((int)threadIdx.x < NT) {
    double sum = 0.0, comp = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
      double y = part[i] - comp;
      double t = sum + y;
      comp = (t - sum) - y;
      sum = t;
    }
    ss[threadIdx.x] = sum;
    cc[threadIdx.x] = comp;
  }
  __syncthreads();
  for (int h = NT / 2; h >= 1; h >>= 1) {
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
  if (threadIdx.x == 0) {
    const double e = ss[0] - cc[0];
    out[0] = e;
    if (out_host) {
      out_host[0] = e;
      __threadfence_system();
    }
  }
}

// true in the last of nblocks workgroups to arrive (each after its thread 0
// wrote its partial); that workgroup then reads every partial.  The ticket
// (device memory, zero before the first launch) is reset by the last one, so
// the launch can be repeated or replayed in a graph.
__device__ __forceinline__ bool last_block_arrived(unsigned* ticket, unsigned nblocks) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = t == nblocks - 1;
    if (last) {
      __threadfence();
      *ticket = 0u;
    }
  }
  __syncthreads();
  return last;
}

}  // namespace fresco_hip
