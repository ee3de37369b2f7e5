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
#include <cstdlib>
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

// The same contraction on the matrix cores (FP32): v_mfma_f32_16x16x4f32
// takes A[m][q] (m = i, q = j) in lane m + 16 q and B[q][n] (n = k) in lane
// n + 16 q, and returns D[m][n] in register m % 4 of lane n + 16 (m / 4).  A
// complex product is four real ones (two for a real input or a real output):
//   Re += Ar Wr - Ai Wi,  Im += Ar Wi + Ai Wr.
// A workgroup (4 wavefronts) stages a [32 x J] slab of the input and the
// [32 x J] twiddle rows in LDS (rows of MF_JP(J) entries, J padded with zeros
// to a multiple of 4) and computes a 32 x 32 block of outputs, a 16 x 16 tile
// per wavefront.
typedef float mf_f4 __attribute__((ext_vector_type(4)));

constexpr int MF_T = 256;
constexpr int MF_TILE = 32;

// LDS row length in entries: J padded to a multiple of 4 and to 4 mod 32, so
// that the 16 rows x 4 entries of one operand read spread over the banks
__host__ __device__ constexpr int mf_jp(int J) { return ((J + 31) / 32) * 32 + 4; }

// TR: the matrix roles swapped (M = k, N = i: the output tile's lanes run
// along i, for outputs whose unit stride is i)
template <bool IN_REAL, bool OUT_REAL, bool TR>
__global__ void __launch_bounds__(MF_T)
    k_band_mfma(const void* __restrict__ in_, int64_t so, int64_t si, int64_t sj, int I, int J,
                const Cplx<float>* __restrict__ W, int K, void* __restrict__ out_, int64_t to, int64_t ti,
                int64_t tk, Cplx<float>* __restrict__ zero_out) {
  using IT = typename std::conditional<IN_REAL, float, Cplx<float>>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int JP = mf_jp(J);
  IT* slab = reinterpret_cast<IT*>(smem);  // [32][JP]
  Cplx<float>* ws = reinterpret_cast<Cplx<float>*>(smem + (((size_t)MF_TILE * JP * sizeof(IT) + 15) & ~(size_t)15));
  const IT* in = static_cast<const IT*>(in_);
  const int o = blockIdx.x;
  const int i0 = blockIdx.y * MF_TILE, k0 = blockIdx.z * MF_TILE;
  const int tid = threadIdx.x;
  const int64_t ob = (int64_t)o * so;
  const int J4 = (J + 3) & ~3;
  // staging: the unit-stride index runs over consecutive threads; entries
  // j in [J, J4) and rows beyond I / K are zero; every load of a thread is
  // issued before its LDS stores (one memory latency per workgroup)
  constexpr int PER = MF_TILE * BD_JMAX / MF_T;
  {
    IT v[PER];
    Cplx<float> wv[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + u * MF_T;
      v[u] = IT{};
      wv[u] = Cplx<float>{0.f, 0.f};
      if (e < MF_TILE * J4) {
        int r, j;
        if (sj == 1) {
          r = e / J4;
          j = e - r * J4;
        } else {
          j = e / MF_TILE;
          r = e - j * MF_TILE;
        }
        if (i0 + r < I && j < J) v[u] = in[ob + (int64_t)(i0 + r) * si + (int64_t)j * sj];
        const int rw = e / J4, jw = e - rw * J4;
        if (k0 + rw < K && jw < J) wv[u] = W[(int64_t)(k0 + rw) * J + jw];
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + u * MF_T;
      if (e < MF_TILE * J4) {
        int r, j;
        if (sj == 1) {
          r = e / J4;
          j = e - r * J4;
        } else {
          j = e / MF_TILE;
          r = e - j * MF_TILE;
        }
        slab[r * JP + j] = v[u];
        const int rw = e / J4, jw = e - rw * J4;
        ws[rw * JP + jw] = wv[u];
      }
    }
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  const int wi = wv & 1, wk = wv >> 1;
  const int m = lane & 15, q = lane >> 4;
  // both operands are read as [16 rows][4 j] blocks: row m of the input slab
  // and row m of the twiddles (A and B of the product, in either order)
  const IT* arow = slab + (wi * 16 + m) * JP + q;
  const Cplx<float>* brow = ws + (wk * 16 + m) * JP + q;
  mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
  auto mf = [](float a, float b, mf_f4 c) {
    return TR ? __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, c, 0, 0, 0)
              : __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  };
  for (int j = 0; j < J4; j += 4) {
    const IT av = arow[j];
    const Cplx<float> bv = brow[j];
    if constexpr (IN_REAL) {
      cr = mf(av, bv.re, cr);
      ci = mf(av, bv.im, ci);
    } else {
      cr = mf(av.re, bv.re, cr);
      cr = mf(av.im, -bv.im, cr);
      if constexpr (!OUT_REAL) {
        ci = mf(av.re, bv.im, ci);
        ci = mf(av.im, bv.re, ci);
      }
    }
  }
  // D[m'][n'] in register m' % 4 of lane n' + 16 (m' / 4): (m', n') = (i, k),
  // or (k, i) for TR
  const int kr = TR ? 4 * q : m;
  const int ir = TR ? m : 4 * q;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + wi * 16 + ir + (TR ? 0 : r);
    const int k = k0 + wk * 16 + kr + (TR ? r : 0);
    if (i < I && k < K) {
      const int64_t at = (int64_t)o * to + (int64_t)i * ti + (int64_t)k * tk;
      if constexpr (OUT_REAL)
        static_cast<float*>(out_)[at] = cr[r];
      else
        static_cast<Cplx<float>*>(out_)[at] = Cplx<float>{cr[r], ci[r]};
      if (zero_out) zero_out[at] = Cplx<float>{0.f, 0.f};
    }
  }
}

template <typename T, bool IN_REAL, bool OUT_REAL, bool LANE_K, int NPT, bool TR = false>
void launch_pass(const void* in, int64_t so, int64_t si, int64_t sj, int O, int I, int J, const Cplx<T>* W, int K,
                 void* out, int64_t to, int64_t ti, int64_t tk, hipStream_t s, Cplx<T>* zero_out = nullptr) {
  using IT = typename std::conditional<IN_REAL, T, Cplx<T>>::type;
  constexpr int ITILE = LANE_K ? 8 * NPT : 32;
  constexpr int KTILE = LANE_K ? 32 : 8 * NPT;
  const size_t lds = (((size_t)J * ITILE * sizeof(IT) + 15) & ~(size_t)15) + (size_t)KTILE * (J | 1) * sizeof(Cplx<T>);
  if (lds > 160 * 1024 || J > BD_JMAX)
    throw InvalidArgument("band transform pass: contracted axis over " + std::to_string(BD_JMAX) + " or LDS exceeded");
  if constexpr (std::is_same<T, float>::value) {
    if (J <= BD_JMAX) {
      const size_t mlds = (((size_t)MF_TILE * mf_jp(J) * sizeof(IT) + 15) & ~(size_t)15) +
                          (size_t)MF_TILE * mf_jp(J) * sizeof(Cplx<float>);
      const dim3 mgrd(O, (I + MF_TILE - 1) / MF_TILE, (K + MF_TILE - 1) / MF_TILE);
      hipLaunchKernelGGL((k_band_mfma<IN_REAL, OUT_REAL, TR>), mgrd, dim3(MF_T), mlds, s, in, so, si, sj, I, J, W, K, out,
                         to, ti, tk, zero_out);
      FRESCO_HIP_CHECK(hipGetLastError());
      return;
    }
  }
  const dim3 grd(O, (I + ITILE - 1) / ITILE, (K + KTILE - 1) / KTILE);
  hipLaunchKernelGGL((k_band_pass<T, IN_REAL, OUT_REAL, LANE_K, NPT>), grd, dim3(BD_T), lds, s, in, so, si, sj, I, J,
                     W, K, out, to, ti, tk, zero_out);
  FRESCO_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Fused y / z passes of one kx plane on the matrix cores (FP32), one
// workgroup of 4 wavefronts per (kx, 16-wide tile of the other output index):
//   forward  T[z][ky] = sum_y U1[kx][z][y] Wy[ky][y]      (ky tile in LDS)
//            S[kz][ky][kx] = sum_z Wz[kz][z] T[z][ky]
//   inverse  Z1[kz][y] = sum_ky X[kz][ky][kx] Wyi[y][ky]  (y tile in LDS)
//            Y2[kx][z][y] = sum_kz Wzi[z][kz] Z1[kz][y]
// The contraction over the short index runs first (the inverse contracts
// ky before kz), the intermediate stays in LDS, and the A operand of every
// 16 x 16 tile (one 8-byte value per lane and 4-deep step) is loaded for all
// its steps before the first product.
// ---------------------------------------------------------------------------

constexpr int FZ_MAXS = 32;  // 4-deep steps of a contraction (indices up to 128)

// one 16 x 16 complex tile: cr/ci += A B with A[m][q] = av[s] (lane m + 16 q,
// step s) and B[q][n] = bp[(4 s + q) * bstride + n-th column offset]
__device__ __forceinline__ void mf_tile(const Cplx<float>* av, int nsteps, const Cplx<float>* bcol, int bstep,
                                        mf_f4& cr, mf_f4& ci) {
#pragma unroll
  for (int st = 0; st < FZ_MAXS; ++st) {
    if (st < nsteps) {
      const Cplx<float> a = av[st];
      const Cplx<float> b = bcol[st * bstep];
      cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.re, cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.im, -b.im, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.im, ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.im, b.re, ci, 0, 0, 0);
    }
  }
}

// U1 [nb][nz][ny], Wy [nk][ny], Wz [nk][nz] -> S [nk][nk][nb] (kz, ky, kx);
// zero_out (optional, same layout as S) cleared where S is written
__global__ void __launch_bounds__(256)
    k_band_fwd_yz(const Cplx<float>* __restrict__ U1, int ny, int nz, int nb, int nk,
                  const Cplx<float>* __restrict__ Wy, const Cplx<float>* __restrict__ Wz, Cplx<float>* __restrict__ S,
                  Cplx<float>* __restrict__ zero_out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int kx = blockIdx.x;
  const int ky0 = blockIdx.y * 16;
  const int ny4 = (ny + 3) & ~3, nz4 = (nz + 3) & ~3;
  const int JPY = mf_jp(ny);
  Cplx<float>* wyl = reinterpret_cast<Cplx<float>*>(smem);  // [16][JPY]: Wy rows ky0..ky0+15
  Cplx<float>* T = wyl + 16 * JPY;                           // [nz4 rounded to 16][16]: z x ky
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = lane & 15, q = lane >> 4;
  {
    constexpr int PER = 16 * 128 / 256;
    Cplx<float> v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + 256 * u;
      const int r = e / ny4, y = e - r * ny4;
      v[u] = Cplx<float>{0.f, 0.f};
      if (e < 16 * ny4 && ky0 + r < nk && y < ny) v[u] = Wy[(int64_t)(ky0 + r) * ny + y];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + 256 * u;
      const int r = e / ny4, y = e - r * ny4;
      if (e < 16 * ny4) wyl[r * JPY + y] = v[u];
    }
  }
  __syncthreads();
  // step A: wavefront wv takes z tiles wv, wv + 4, ...
  const int nzt = (nz + 15) / 16;
  const int sy = ny4 / 4;
  const Cplx<float>* u1 = U1 + (int64_t)kx * nz * ny;
  for (int zt = wv; zt < nzt; zt += 4) {
    const int z = zt * 16 + m;
    Cplx<float> av[FZ_MAXS];
#pragma unroll
    for (int st = 0; st < FZ_MAXS; ++st) {
      const int y = 4 * st + q;
      av[st] = Cplx<float>{0.f, 0.f};
      if (st < sy && z < nz && y < ny) av[st] = u1[(int64_t)z * ny + y];
    }
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    // B[q][n] = Wy[ky0 + n][4 st + q]: lane n + 16 q reads row n
    mf_tile(av, sy, wyl + m * JPY + q, 4, cr, ci);
#pragma unroll
    for (int r = 0; r < 4; ++r) T[(zt * 16 + 4 * q + r) * 16 + m] = Cplx<float>{cr[r], ci[r]};
  }
  __syncthreads();
  // step B: wavefront wv takes kz tiles wv, wv + 4, ...
  const int nkt = (nk + 15) / 16;
  const int sz = nz4 / 4;
  for (int kt = wv; kt < nkt; kt += 4) {
    const int kz = kt * 16 + m;
    Cplx<float> av[FZ_MAXS];
#pragma unroll
    for (int st = 0; st < FZ_MAXS; ++st) {
      const int z = 4 * st + q;
      av[st] = Cplx<float>{0.f, 0.f};
      if (st < sz && kz < nk && z < nz) av[st] = Wz[(int64_t)kz * nz + z];
    }
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    // B[q][n] = T[4 st + q][n]
    mf_tile(av, sz, T + q * 16 + m, 64, cr, ci);
    const int ky = ky0 + m;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int kzo = kt * 16 + 4 * q + r;
      if (kzo < nk && ky < nk) {
        const int64_t at = ((int64_t)kzo * nk + ky) * nb + kx;
        S[at] = Cplx<float>{cr[r], ci[r]};
        if (zero_out) zero_out[at] = Cplx<float>{0.f, 0.f};
      }
    }
  }
}

// X [nk][nk][nb] (kz, ky, kx), Wyi [ny][nk], Wzi [nz][nk] -> Y2 [nb][nz][ny]
__global__ void __launch_bounds__(256)
    k_band_inv_zy(const Cplx<float>* __restrict__ X, int ny, int nz, int nb, int nk,
                  const Cplx<float>* __restrict__ Wyi, const Cplx<float>* __restrict__ Wzi,
                  Cplx<float>* __restrict__ Y2) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int kx = blockIdx.x;
  const int y0 = blockIdx.y * 16;
  const int nk4 = (nk + 3) & ~3;
  const int JPK = mf_jp(nk);
  Cplx<float>* wyl = reinterpret_cast<Cplx<float>*>(smem);  // [16][JPK]: Wyi rows y0..y0+15
  Cplx<float>* Z1 = wyl + 16 * JPK;                          // [nk4 rounded to 16][16]: kz x y
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = lane & 15, q = lane >> 4;
  {
    constexpr int PER = 16 * 128 / 256;
    Cplx<float> v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + 256 * u;
      const int r = e / nk4, j = e - r * nk4;
      v[u] = Cplx<float>{0.f, 0.f};
      if (e < 16 * nk4 && y0 + r < ny && j < nk) v[u] = Wyi[(int64_t)(y0 + r) * nk + j];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + 256 * u;
      const int r = e / nk4, j = e - r * nk4;
      if (e < 16 * nk4) wyl[r * JPK + j] = v[u];
    }
  }
  // Z1 rows beyond nk stay zero for the kz steps that run past them
  const int nkt = (nk + 15) / 16;
  for (int e = tid; e < nkt * 16 * 16; e += 256) Z1[e] = Cplx<float>{0.f, 0.f};
  __syncthreads();
  const int sk = nk4 / 4;
  // step A: Z1[kz][y] over ky; wavefront wv takes kz tiles wv, wv + 4, ...
  for (int kt = wv; kt < nkt; kt += 4) {
    const int kz = kt * 16 + m;
    Cplx<float> av[FZ_MAXS];
#pragma unroll
    for (int st = 0; st < FZ_MAXS; ++st) {
      const int ky = 4 * st + q;
      av[st] = Cplx<float>{0.f, 0.f};
      if (st < sk && kz < nk && ky < nk) av[st] = X[((int64_t)kz * nk + ky) * nb + kx];
    }
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    // B[q][n] = Wyi[y0 + n][4 st + q]
    mf_tile(av, sk, wyl + m * JPK + q, 4, cr, ci);
#pragma unroll
    for (int r = 0; r < 4; ++r) Z1[(kt * 16 + 4 * q + r) * 16 + m] = Cplx<float>{cr[r], ci[r]};
  }
  __syncthreads();
  // step B: Y2[z][y] over kz; wavefront wv takes z tiles wv, wv + 4, ...
  const int nzt = (nz + 15) / 16;
  for (int zt = wv; zt < nzt; zt += 4) {
    const int z = zt * 16 + m;
    Cplx<float> av[FZ_MAXS];
#pragma unroll
    for (int st = 0; st < FZ_MAXS; ++st) {
      const int kz = 4 * st + q;
      av[st] = Cplx<float>{0.f, 0.f};
      if (st < sk && z < nz && kz < nk) av[st] = Wzi[(int64_t)z * nk + kz];
    }
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    // B[q][n] = Z1[4 st + q][n]
    mf_tile(av, sk, Z1 + q * 16 + m, 64, cr, ci);
    const int y = y0 + m;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int zo = zt * 16 + 4 * q + r;
      if (zo < nz && y < ny) Y2[((int64_t)kx * nz + zo) * ny + y] = Cplx<float>{cr[r], ci[r]};
    }
  }
}

// W[k][j] = wt(j') exp(sign 2 pi i k' j' / n) with k' = k + koff, j' = j + joff,
// stored at [k * stride + j] of a zero-padded [prow][stride] array (prow >= K,
// stride >= J; stride = J, prow = K: the plain [K][J] matrix)
template <typename T>
void twiddles(DeviceBuffer<Cplx<T>>& buf, int K, int koff, int J, int joff, int n, int sign, bool weight_kx,
              int prow = 0, int stride = 0) {
  if (prow == 0) prow = K;
  if (stride == 0) stride = J;
  std::vector<Cplx<T>> h((size_t)prow * stride, Cplx<T>{T(0), T(0)});
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < J; ++j) {
      const long long kk = k + koff, jj = j + joff;
      long long r = (kk * jj) % n;
      if (r < 0) r += n;
      const double ang = sign * 2.0 * M_PI * (double)r / (double)n;
      const double wt = weight_kx ? (jj == 0 ? 1.0 : 2.0) : 1.0;
      h[(size_t)k * stride + j] = Cplx<T>{(T)(wt * std::cos(ang)), (T)(wt * std::sin(ang))};
    }
  buf.allocate(h.size());
  buf.upload(h.data(), h.size());
}



// ---------------------------------------------------------------------------
// Fused FP32 form (three launches per energy + gradient evaluation, the loss
// inside the middle one):
//   k_band_plane_fwd  per z plane: U1[y][kx] = sum_x g[z][y][x] Wx[kx][x],
//                     then U2[z][ky][kx] = sum_y Wy[ky][y] U1[y][kx]
//   k_band_zloss      per tile of 16 band columns (ky, kx): S[kz][col] =
//                     sum_z Wz[kz][z] U2[z][col], the loss of the pairs whose
//                     representative lies in the tile and their pullback
//                     values X (in LDS), then Y1[z][col] = sum_kz Wzi[z][kz]
//                     X[kz][col]; the kx = 0 columns ky and -ky share a tile,
//                     so every X a tile needs is written in it
//   k_band_plane_inv  per z plane: Y2[y][kx] = sum_ky Wyi[y][ky] Y1[z][ky][kx],
//                     then grid[z][y][x] = Re sum_kx Wxi[x][kx] Y2[y][kx]
// Every operand a workgroup reads more than once is staged in LDS with all
// of its loads issued together (one memory latency per stage); products are
// v_mfma_f32_16x16x4f32 tiles as in k_band_mfma (a complex product is four
// real ones, two for a real input or output).
// ---------------------------------------------------------------------------

constexpr int BP_T = 512;   // threads of the plane kernels (8 wavefronts)
constexpr int BP_YG = 4;    // y tiles per inverse plane workgroup
constexpr int BZ_T = 512;   // threads of the z-pass + loss kernel (8 wavefronts)
constexpr int BZ_COLS = 16;  // band columns per z-pass workgroup
constexpr int BZ_KROWS = 80;  // kz rows of the z-pass workgroup's S / X tiles (band <= 32)

// LDS row lengths, in complex entries, for the two read patterns of a
// 16 x 16 x 4 product (each half wavefront reads 256 bytes):
//   B operand, lane (n, q) reads [4 s + q][n]: rows 16 mod 32 entries apart
__host__ __device__ constexpr int bd_bstride(int cols) {
  return ((cols + 15) / 16 * 16) % 32 == 0 ? (cols + 15) / 16 * 16 + 16 : (cols + 15) / 16 * 16;
}
//   A operand, lane (m, q) reads [m][4 s + q]: rows 2 mod 32 entries apart
__host__ __device__ constexpr int bd_astride(int cols) { return (cols + 31) / 32 * 32 + 2; }

__device__ __forceinline__ void cmfma(const Cplx<float>& a, const Cplx<float>& b, mf_f4& cr, mf_f4& ci) {
  cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.re, cr, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.im, -b.im, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.im, ci, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.im, b.re, ci, 0, 0, 0);
}

// Row lengths of the operand arrays (complex entries unless noted); every
// contracted index is zero-padded to a multiple of 16 (4 steps of 4), so the
// product loops read 4 steps' operands at once without bounds checks.
struct BandShape {
  int nx, ny, nz, nb, nk;
  int nyR, nxR, nzR, nbR, nkR;  // rounded up to 16
  int PX;                       // plane row (floats)
  int SWX;                      // forward plane: Wx [nbR][SWX] (a workgroup stages one 16-row tile)
  int SWZI;                     // z pass: Wzi [nzR][SWZI]
  int SYT, SY2, SWYI, SWXI;     // inverse plane: Y1 [nkR][SYT], Y2 [nyR][SY2], Wyi [nyR][SWYI], Wxi [nxR][SWXI]
  __host__ __device__ BandShape(int nx_, int ny_, int nz_, int nb_, int nk_)
      : nx(nx_), ny(ny_), nz(nz_), nb(nb_), nk(nk_) {
    nyR = (ny + 15) / 16 * 16;
    nxR = (nx + 15) / 16 * 16;
    nzR = (nz + 15) / 16 * 16;
    nbR = (nb + 15) / 16 * 16;
    nkR = (nk + 15) / 16 * 16;
    PX = mf_jp(nx);
    SWX = bd_astride(nx);
    SWZI = bd_astride(nk);
    SYT = bd_bstride(nb);
    SY2 = bd_astride(nb);
    SWYI = bd_astride(nk);
    SWXI = bd_astride(nb);
  }
  __host__ __device__ size_t fwd_lds() const { return (size_t)nyR * PX * 4 + (size_t)nyR * 16 * 8 + (size_t)16 * SWX * 8; }
  __host__ __device__ size_t inv_lds() const {
    return (size_t)nkR * SYT * 8 + (size_t)BP_YG * 16 * (SY2 + SWYI) * 8 + (size_t)nxR * SWXI * 8;
  }
};

// n 16-byte units, batches of NT x U (each batch's loads issued together)
template <int NT, int U>
__device__ __forceinline__ void copy_units(void* dst_, const void* __restrict__ src_, int n) {
  float4* dst = static_cast<float4*>(dst_);
  const float4* src = static_cast<const float4*>(src_);
  for (int b0 = 0; b0 < n; b0 += NT * U) {
    // clamped, unconditional loads keep the batch in registers
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = src[min(b0 + u * NT + (int)threadIdx.x, n - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = b0 + u * NT + (int)threadIdx.x;
      if (e < n) dst[e] = v[u];
    }
  }
}

// grid [nz][ny][nx] (real), Wx padded [nbR][SWX], Wy [nk][ny] -> U2 [nz][nk][nb]
__global__ void __launch_bounds__(BP_T)
    k_band_plane_fwd(const float* __restrict__ grid, BandShape sh, const Cplx<float>* __restrict__ Wxp,
                     const Cplx<float>* __restrict__ Wy, Cplx<float>* __restrict__ U2) {
  extern __shared__ __align__(16) unsigned char smem[];
  // workgroup (z, nt): plane z, kx tile nt (the x pass of one kx tile
  // feeds the y pass of the same tile only)
  const int z = blockIdx.x, nt = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = lane & 15, q = lane >> 4;
  const int nx = sh.nx, ny = sh.ny, nb = sh.nb, nk = sh.nk;
  const int nyT = sh.nyR / 16, nkT = sh.nkR / 16;
  const int PX = sh.PX, SB = 16, SW = sh.SWX;
  float* plane = reinterpret_cast<float*>(smem);                                 // [nyR][PX]
  Cplx<float>* u1 = reinterpret_cast<Cplx<float>*>(smem + (size_t)sh.nyR * PX * 4);  // [nyR][16]
  Cplx<float>* wx = u1 + (size_t)sh.nyR * SB;                                      // [16][SW]: tile nt
  {
    // the plane: one wavefront per row, lanes over columns (zero padding),
    // every load issued before the stores; the padded twiddles as units
    constexpr int MAXR = 128 / (BP_T / 64), NCH = 3;
    const float* gp = grid + (int64_t)z * ny * nx;
    float v[MAXR][NCH];
#pragma unroll
    for (int i = 0; i < MAXR; ++i)
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const int r = wv + i * (BP_T / 64), cc = lane + 64 * c;
        v[i][c] = (r < ny && cc < nx) ? gp[r * nx + cc] : 0.f;
      }
    copy_units<BP_T, 8>(wx, Wxp + (size_t)nt * 16 * SW, 16 * SW / 2);
#pragma unroll
    for (int i = 0; i < MAXR; ++i)
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const int r = wv + i * (BP_T / 64), cc = lane + 64 * c;
        if (r < sh.nyR && cc < PX) plane[r * PX + cc] = v[i][c];
      }
  }
  __syncthreads();
  // x: U1[y][kx], A = plane rows (y), B[x][kx] = Wx[kx][x]; 4 steps a round
  const int KX4 = (nx + 15) / 16;
  for (int mt = wv; mt < nyT; mt += BP_T / 64) {
    const float* arow = plane + (mt * 16 + m) * PX + q;
    const Cplx<float>* brow = wx + m * SW + q;
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    for (int s4 = 0; s4 < KX4; ++s4) {
      float a[4];
      Cplx<float> b[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = arow[16 * s4 + 4 * j];
        b[j] = brow[16 * s4 + 4 * j];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b[j].re, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b[j].im, ci, 0, 0, 0);
      }
    }
    // D[y][kx]: register r of lane (kx = m, q) is row 4 q + r
#pragma unroll
    for (int r = 0; r < 4; ++r) u1[(mt * 16 + 4 * q + r) * SB + m] = Cplx<float>{cr[r], ci[r]};
  }
  __syncthreads();
  // y: U2[ky][kx], A[ky][y] = Wy (all of a tile's steps loaded first), B = U1
  const int KY = (ny + 3) / 4;
  for (int mt = wv; mt < nkT; mt += BP_T / 64) {
    const int ky = mt * 16 + m;
    Cplx<float> aw[FZ_MAXS];
#pragma unroll
    for (int st = 0; st < FZ_MAXS; ++st) {
      const int y = 4 * st + q;
      aw[st] = (st < KY && ky < nk && y < ny) ? Wy[(int64_t)ky * ny + y] : Cplx<float>{0.f, 0.f};
    }
    const Cplx<float>* bcol = u1 + q * SB + m;
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s4 = 0; s4 < FZ_MAXS / 4; ++s4) {
      if (4 * s4 < KY) {
        Cplx<float> b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = bcol[(16 * s4 + 4 * j) * SB];
#pragma unroll
        for (int j = 0; j < 4; ++j) cmfma(aw[4 * s4 + j], b[j], cr, ci);
      }
    }
    const int kx = nt * 16 + m;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int kyo = mt * 16 + 4 * q + r;
      if (kyo < nk && kx < nb) U2[((int64_t)z * nk + kyo) * nb + kx] = Cplx<float>{cr[r], ci[r]};
    }
  }
}

// U2 [nz][nk][nb] -> S (compact band spectrum), the tile's pair energies and,
// MODE 1, Y1 [nz][nk][nb]; tcol [tiles][16]: column ky * nb + kx, or -1;
// Wzip: Wzi padded [nzR][SWZI]
template <int MODE>
__global__ void __launch_bounds__(BZ_T)
    k_band_zloss(const Cplx<float>* __restrict__ U2, BandShape sh, const int* __restrict__ tcol,
                 const Cplx<float>* __restrict__ Wz, const Cplx<float>* __restrict__ Wzip, PairDev pd,
                 BandLossDev bl, double norm, double gscale, Cplx<float>* __restrict__ S,
                 Cplx<float>* __restrict__ Y1, double* __restrict__ partial) {
  __shared__ __align__(16) Cplx<float> ut[128][BZ_COLS];          // U2 columns [z][slot]
  __shared__ __align__(16) Cplx<float> sx[2][BZ_KROWS][BZ_COLS];  // S and X [kz][slot]
  extern __shared__ __align__(16) unsigned char smem[];           // Wzi [nzR][SWZI] (MODE 1)
  __shared__ double red[BZ_T];
  __shared__ int cols[BZ_COLS];
  const int tile = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = lane & 15, q = lane >> 4;
  const int nz = sh.nz, nk = sh.nk;
  const int64_t plane = (int64_t)nk * sh.nb;
  const int SW = sh.SWZI;
  Cplx<float>* wzi = reinterpret_cast<Cplx<float>*>(smem);
  // the tables of each thread's first pair and the forward z pass's first
  // twiddle tile are loaded before anything waits
  const int p0 = bl.pstart[tile], p1 = bl.pstart[tile + 1];
  const int KZ = (nz + 3) / 4, nkT = sh.nkR / 16;
  PairVals pv0{};
  int lc0 = 0, mr0 = -1;
  if (p0 + tid < p1) {
    pv0 = pair_vals(pd, p0 + tid);
    lc0 = bl.loc[p0 + tid];
    mr0 = bl.mir[p0 + tid];
  }
  Cplx<float> aw[FZ_MAXS];
  auto load_wz = [&](int mt) {
    const int kz = mt * 16 + m;
#pragma unroll
    for (int st = 0; st < FZ_MAXS; ++st) {
      const int zz = 4 * st + q;
      aw[st] = (st < KZ && kz < nk && zz < nz) ? Wz[(int64_t)kz * nz + zz] : Cplx<float>{0.f, 0.f};
    }
  };
  load_wz(wv);
  {
    constexpr int PU = 128 * BZ_COLS / BZ_T;
    const int* tc = tcol + tile * BZ_COLS;
    Cplx<float> v[PU];
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int e = tid + u * BZ_T;
      const int zz = e / BZ_COLS, col = tc[e % BZ_COLS];
      v[u] = (zz < nz && col >= 0) ? U2[zz * plane + col] : Cplx<float>{0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int e = tid + u * BZ_T;
      ut[e / BZ_COLS][e % BZ_COLS] = v[u];
    }
    for (int e = tid; e < 2 * BZ_KROWS * BZ_COLS; e += BZ_T) (&sx[0][0][0])[e] = Cplx<float>{0.f, 0.f};
    if (MODE > 0) copy_units<BZ_T, 8>(wzi, Wzip, sh.nzR * SW / 2);
    if (tid < BZ_COLS) cols[tid] = tc[tid];
  }
  __syncthreads();
  // forward z: S[kz][slot], A[kz][z] = Wz (a tile's steps loaded first), B = ut
  for (int mt = wv; mt < nkT; mt += BZ_T / 64) {
    if (mt != wv) load_wz(mt);
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s4 = 0; s4 < FZ_MAXS / 4; ++s4) {
      if (4 * s4 < KZ) {
        Cplx<float> b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = ut[16 * s4 + 4 * j + q][m];
#pragma unroll
        for (int j = 0; j < 4; ++j) cmfma(aw[4 * s4 + j], b[j], cr, ci);
      }
    }
    const int col = cols[m];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int kzo = mt * 16 + 4 * q + r;
      const Cplx<float> sv{cr[r], ci[r]};
      sx[0][kzo][m] = sv;
      if (kzo < nk && col >= 0) S[kzo * plane + col] = sv;
    }
  }
  __syncthreads();
  // the loss of the tile's pairs, each thread's in index order
  double e = 0.0;
  for (int i = p0 + tid; i < p1; i += BZ_T) {
    const bool first = i == p0 + tid;
    const PairVals pv = first ? pv0 : pair_vals(pd, i);
    const int lc = first ? lc0 : bl.loc[i];
    double xr, xi;
    e += pair_loss(pv, sx[0][lc >> 4][lc & 15], norm, gscale, MODE > 0, xr, xi);
    // k = 0 moves no point: it stays off the pullback (as in k_nuwu_loss)
    if (MODE > 0 && !pv.self) {
      sx[1][lc >> 4][lc & 15] = Cplx<float>{(float)xr, (float)xi};
      const int mr = first ? mr0 : bl.mir[i];
      if (mr >= 0) sx[1][mr >> 4][mr & 15] = Cplx<float>{(float)xr, (float)(-xi)};
    }
  }
  red[tid] = e;
  __syncthreads();
  for (int hh = BZ_T / 2; hh >= 1; hh >>= 1) {
    if (tid < hh) red[tid] += red[tid + hh];
    __syncthreads();
  }
  if (tid == 0) partial[tile] = red[0];
  if (MODE > 0) {
    // inverse z: Y1[z][slot], A[z][kz] = Wzi (LDS), B = X; 4 steps a round
    const int KK4 = (nk + 15) / 16, nzT = sh.nzR / 16;
    for (int mt = wv; mt < nzT; mt += BZ_T / 64) {
      const Cplx<float>* arow = wzi + (mt * 16 + m) * SW + q;
      mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
      for (int s4 = 0; s4 < KK4; ++s4) {
        Cplx<float> a[4], b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          a[j] = arow[16 * s4 + 4 * j];
          b[j] = sx[1][16 * s4 + 4 * j + q][m];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) cmfma(a[j], b[j], cr, ci);
      }
      const int col = cols[m];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int zo = mt * 16 + 4 * q + r;
        if (zo < nz && col >= 0) Y1[zo * plane + col] = Cplx<float>{cr[r], ci[r]};
      }
    }
  }
}

// Y1 [nz][nk][nb], Wyi padded [nyR][SWYI], Wxi padded [nxR][SWXI] (kx > 0
// weighted twice) -> grid
__global__ void __launch_bounds__(BP_T)
    k_band_plane_inv(const Cplx<float>* __restrict__ Y1, BandShape sh, const Cplx<float>* __restrict__ Wyip,
                     const Cplx<float>* __restrict__ Wxip, float* __restrict__ grid, const double* __restrict__ partial,
                     int nparts, double* __restrict__ e_out, double* __restrict__ e_host) {
  extern __shared__ __align__(16) unsigned char smem[];
  // the z pass's per-tile energies, summed by the first workgroup (the
  // previous launch has completed, so they are visible without a fence)
  if (partial && blockIdx.x == 0 && blockIdx.y == 0) block_sum_partials(partial, nparts, e_out, e_host);
  // workgroup (z, g): plane z, y tiles BP_YG g .. BP_YG g + BP_YG - 1 (both
  // passes keep y as the row index, so the groups are independent)
  const int z = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = lane & 15, q = lane >> 4;
  const int nx = sh.nx, ny = sh.ny, nb = sh.nb, nk = sh.nk;
  const int yt0 = BP_YG * blockIdx.y;
  const int nyT = min(BP_YG, sh.nyR / 16 - yt0), nbT = sh.nbR / 16, nxT = sh.nxR / 16;
  const int SB = sh.SYT, SA = sh.SY2, SWY = sh.SWYI, SWX = sh.SWXI;
  Cplx<float>* yt = reinterpret_cast<Cplx<float>*>(smem);  // [nkR][SB]: Y1[z]
  Cplx<float>* y2 = yt + (size_t)sh.nkR * SB;              // [BP_YG * 16][SA]: the group's rows
  Cplx<float>* wyi = y2 + (size_t)BP_YG * 16 * SA;         // [BP_YG * 16][SWY]
  Cplx<float>* wxi = wyi + (size_t)BP_YG * 16 * SWY;       // [nxR][SWX]
  {
    // Y1[z]: one wavefront per row, lanes over columns (zero padding)
    constexpr int MAXR = 80 / (BP_T / 64);
    const Cplx<float>* yp = Y1 + (int64_t)z * nk * nb;
    Cplx<float> v[MAXR];
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
      const int r = wv + i * (BP_T / 64);
      v[i] = (r < nk && lane < nb) ? yp[r * nb + lane] : Cplx<float>{0.f, 0.f};
    }
    copy_units<BP_T, 8>(wyi, Wyip + (size_t)yt0 * 16 * SWY, nyT * 16 * SWY / 2);
    copy_units<BP_T, 8>(wxi, Wxip, sh.nxR * SWX / 2);
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
      const int r = wv + i * (BP_T / 64);
      if (r < sh.nkR && lane < SB) yt[r * SB + lane] = v[i];
    }
  }
  __syncthreads();
  // y: Y2[y][kx], A[y][ky] = Wyi, B[ky][kx] = Y1[z]; 4 steps a round
  const int KK4 = (nk + 15) / 16;
  for (int t = wv; t < nyT * nbT; t += BP_T / 64) {
    const int mt = t / nbT, nt = t - mt * nbT;
    const Cplx<float>* arow = wyi + (mt * 16 + m) * SWY + q;
    const Cplx<float>* bcol = yt + q * SB + nt * 16 + m;
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    for (int s4 = 0; s4 < KK4; ++s4) {
      Cplx<float> a[4], b[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = arow[16 * s4 + 4 * j];
        b[j] = bcol[(16 * s4 + 4 * j) * SB];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) cmfma(a[j], b[j], cr, ci);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) y2[(mt * 16 + 4 * q + r) * SA + nt * 16 + m] = Cplx<float>{cr[r], ci[r]};
  }
  __syncthreads();
  // x: grid[y][x] = Re sum_kx Y2[y][kx] Wxi[x][kx], A = Y2, B[kx][x] = Wxi[x][kx]
  float* gp = grid + ((int64_t)z * ny + yt0 * 16) * nx;
  const int KB4 = (nb + 15) / 16;
  for (int t = wv; t < nyT * nxT; t += BP_T / 64) {
    const int mt = t / nxT, nt = t - mt * nxT;
    const Cplx<float>* arow = y2 + (mt * 16 + m) * SA + q;
    const Cplx<float>* brow = wxi + (nt * 16 + m) * SWX + q;
    mf_f4 cr = {0.f, 0.f, 0.f, 0.f};
    for (int s4 = 0; s4 < KB4; ++s4) {
      Cplx<float> a[4], b[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = arow[16 * s4 + 4 * j];
        b[j] = brow[16 * s4 + 4 * j];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j].re, b[j].re, cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j].im, -b[j].im, cr, 0, 0, 0);
      }
    }
    const int x = nt * 16 + m;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int yo = mt * 16 + 4 * q + r;
      if (yt0 * 16 + yo < ny && x < nx) gp[yo * nx + x] = cr[r];
    }
  }
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
  const char* env = std::getenv("FRESCO_BANDDFT_FUSED");
  const BandShape sh(nf_[0], nf_[1], nf_[2], nb_, nk);
  fused_ = std::is_same<T, float>::value && !(env && env[0] == '0') && sh.nkR <= BZ_KROWS && sh.nzR <= 128 &&
           sh.fwd_lds() <= 160 * 1024 && sh.inv_lds() <= 160 * 1024;
  if (fused_) {
    // the twiddles the fused kernels stage, in their zero-padded LDS layouts
    twiddles(wxp_, nb_, 0, nf_[0], 0, nf_[0], -1, false, sh.nbR, sh.SWX);
    twiddles(wzip_, nf_[2], 0, nk, -band, nf_[2], 1, false, sh.nzR, sh.SWZI);
    twiddles(wyip_, nf_[1], 0, nk, -band, nf_[1], 1, false, sh.nyR, sh.SWYI);
    twiddles(wxip_, nf_[0], 0, nb_, 0, nf_[0], 1, true, sh.nxR, sh.SWXI);
    // column list: kx = 0 first, ky = 0 alone and then ky, -ky side by side
    // (never split over two tiles), then kx > 0 row by row; tiles of 16
    std::vector<int> list;
    auto push = [&](int c0, int c1) {
      if (c1 >= 0 && list.size() % 16 == 15) list.push_back(-1);
      list.push_back(c0);
      if (c1 >= 0) list.push_back(c1);
    };
    auto col = [&](int ky, int kx) { return (ky + band) * nb_ + kx; };
    push(col(0, 0), -1);
    for (int j = 1; j <= band; ++j) push(col(j, 0), col(-j, 0));
    for (int ky = -band; ky <= band; ++ky)
      for (int kx = 1; kx < nb_; ++kx) push(col(ky, kx), -1);
    while (list.size() % 16) list.push_back(-1);
    ntiles_ = (int)(list.size() / 16);
    col_pos_.assign((size_t)nk * nb_, -1);
    for (size_t i = 0; i < list.size(); ++i)
      if (list[i] >= 0) col_pos_[list[i]] = (int)i;
    tcol_.allocate(list.size());
    tcol_.upload(list.data(), list.size());
  }
}

template <typename T>
void BandDft<T>::forward_planes(const T* grid, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    const BandShape sh(nf_[0], nf_[1], nf_[2], nb_, 2 * B_ + 1);
    hipLaunchKernelGGL(k_band_plane_fwd, dim3(nf_[2], sh.nbR / 16), dim3(BP_T), sh.fwd_lds(), s, grid, sh,
                       wxp_.get(), wb_.get(), u2_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
  } else {
    throw RuntimeFailure("band transform: the fused form is FP32 only");
  }
}

template <typename T>
void BandDft<T>::zloss(int mode, const PairDev& pd, const BandLossDev& bl, double norm, double gscale, Cplx<T>* S,
                       double* partial, double* e_out, double* e_host, bool defer, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    const BandShape sh(nf_[0], nf_[1], nf_[2], nb_, 2 * B_ + 1);
    if (mode == 0) {
      hipLaunchKernelGGL(k_band_zloss<0>, dim3(ntiles_), dim3(BZ_T), 0, s, u2_.get(), sh, tcol_.get(), wc_.get(),
                         wzip_.get(), pd, bl, norm, gscale, S, u1_.get(), partial);
    } else {
      const size_t lds = sizeof(Cplx<float>) * (size_t)sh.nzR * sh.SWZI;
      hipLaunchKernelGGL(k_band_zloss<1>, dim3(ntiles_), dim3(BZ_T), lds, s, u2_.get(), sh, tcol_.get(), wc_.get(),
                         wzip_.get(), pd, bl, norm, gscale, S, u1_.get(), partial);
    }
    FRESCO_HIP_CHECK(hipGetLastError());
    if (!defer || mode == 0) reduce_partials(partial, ntiles_, e_out, s, e_host);
  } else {
    throw RuntimeFailure("band transform: the fused form is FP32 only");
  }
}

template <typename T>
void BandDft<T>::inverse_planes(T* grid, hipStream_t s, const double* partial, double* e_out, double* e_host) {
  if constexpr (std::is_same<T, float>::value) {
    const BandShape sh(nf_[0], nf_[1], nf_[2], nb_, 2 * B_ + 1);
    hipLaunchKernelGGL(k_band_plane_inv, dim3(nf_[2], (sh.nyR / 16 + BP_YG - 1) / BP_YG), dim3(BP_T), sh.inv_lds(), s,
                       u1_.get(), sh, wyip_.get(), wxip_.get(), grid, partial, ntiles_, e_out, e_host);
    FRESCO_HIP_CHECK(hipGetLastError());
  } else {
    throw RuntimeFailure("band transform: the fused form is FP32 only");
  }
}

template <typename T>
void BandDft<T>::forward(const T* grid, Cplx<T>* S, hipStream_t s, Cplx<T>* zero) {
  const int nx = nf_[0], ny = nf_[1], nz = nf_[2], nk = 2 * B_ + 1, nb = nb_;
  if constexpr (std::is_same<T, float>::value) {
    // x: grid[z][y][x] -> U1[kx][z][y] (output lanes along y); then y and z
    // of every kx plane in one launch
    launch_pass<T, true, false, true, BD_NPT, true>(grid, (int64_t)ny * nx, nx, 1, nz, ny, nx, wa_.get(), nb,
                                                    u1_.get(), ny, 1, (int64_t)nz * ny, s);
    const size_t lds = sizeof(Cplx<float>) * (16 * (size_t)mf_jp(ny) + (size_t)((nz + 15) / 16) * 16 * 16);
    hipLaunchKernelGGL(k_band_fwd_yz, dim3(nb, (nk + 15) / 16), dim3(256), lds, s, u1_.get(), ny, nz, nb, nk,
                       wb_.get(), wc_.get(), S, zero);
    FRESCO_HIP_CHECK(hipGetLastError());
    return;
  }
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
  if constexpr (std::is_same<T, float>::value) {
    // ky then kz of every kx plane in one launch: X -> Y2[kx][z][y] (in u1_)
    const size_t lds = sizeof(Cplx<float>) * (16 * (size_t)mf_jp(nk) + (size_t)((nk + 15) / 16) * 16 * 16);
    hipLaunchKernelGGL(k_band_inv_zy, dim3(nb, (ny + 15) / 16), dim3(256), lds, s, X, ny, nz, nb, nk, wbi_.get(),
                       wci_.get(), u1_.get());
    FRESCO_HIP_CHECK(hipGetLastError());
    // x: Y2[kx][z][y] -> grid[z][y][x] (real part, kx > 0 twice)
    launch_pass<T, false, true, true, BD_NPT>(u1_.get(), ny, 1, (int64_t)nz * ny, nz, ny, nb, wai_.get(), nx, grid,
                                              (int64_t)ny * nx, nx, 1, s);
    return;
  }
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
