/*
 * CPU (FP64, OpenMP) exponential-of-semicircle spreading / interpolation.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference path
 * (source/fresco/nuwu.hpp calling FINUFFT type 1 / type 2) used as the timed
 * CPU baseline and as a mid-size checker.  Built by oracle/Makefile into
 * oracle/libnufft_cpu.so and driven from oracle/nufft_cpu.py.
 *
 * Method: Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput. 41 (2019):
 * a point at fine-grid coordinate g touches cells s..s+w-1, s = ceil(g - w/2),
 * with weights phi(2 (s + j - g) / w), phi(z) = exp(beta (sqrt(1 - z^2) - 1)).
 *
 * Spreading is deterministic and race-free: points are bucketed by the first
 * cell along the slowest dimension and every thread owns a contiguous range
 * of slabs of that dimension, adding only the cells inside its range.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXW 16

static inline double es(double z, double beta) {
  double a = 1.0 - z * z;
  return a > 0.0 ? exp(beta * (sqrt(a) - 1.0)) : 0.0;
}

/* g = frac(x / 2pi) * nf ; first cell s and kernel values ker[0..w-1] */
static inline int place(double x, int nf, int w, double beta, double* ker) {
  double t = x / (2.0 * M_PI);
  t -= floor(t);
  double g = t * nf;
  double sf = ceil(g - 0.5 * w);
  for (int j = 0; j < w; ++j) ker[j] = es((sf + j - g) * 2.0 / w, beta);
  long s = (long)sf % nf;
  if (s < 0) s += nf;
  return (int)s;
}

int nufft_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Type-1 spreading of complex strengths c onto a complex grid [nz][ny][nx]
 * (interleaved re/im, x fastest).  x[a] are the coordinates of dimension a
 * (radians, any range).  The grid is overwritten. */
int nufft_spread(int d, int64_t n, const double* x0, const double* x1, const double* x2, const double* c,
                 const int* nf, int w, double beta, double* grid) {
  if (w > MAXW || d < 1 || d > 3) return 1;
  const int nx = nf[0], ny = d >= 2 ? nf[1] : 1, nz = d >= 3 ? nf[2] : 1;
  const int sd = d - 1; /* slab dimension */
  const int ns = nf[sd];
  const double* xs = sd == 0 ? x0 : (sd == 1 ? x1 : x2);
  const size_t plane = (size_t)nx * (d == 3 ? ny : 1); /* cells per slab index */
  memset(grid, 0, sizeof(double) * 2 * (size_t)nx * ny * nz);
  /* bucket points by first cell along the slab dimension */
  int* first = (int*)malloc(sizeof(int) * (size_t)n);
  int64_t* cnt = (int64_t*)calloc((size_t)ns + 1, sizeof(int64_t));
  int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!first || !cnt || !order) return 2;
  for (int64_t j = 0; j < n; ++j) {
    double ker[MAXW];
    first[j] = place(xs[j], ns, w, beta, ker);
    cnt[first[j] + 1]++;
  }
  for (int i = 0; i < ns; ++i) cnt[i + 1] += cnt[i];
  {
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    memcpy(pos, cnt, sizeof(int64_t) * (size_t)ns);
    for (int64_t j = 0; j < n; ++j) order[pos[first[j]]++] = j;
    free(pos);
  }
#pragma omp parallel
  {
    int nt = 1, tid = 0;
#ifdef _OPENMP
    nt = omp_get_num_threads();
    tid = omp_get_thread_num();
#endif
    int lo = (int)((long)ns * tid / nt), hi = (int)((long)ns * (tid + 1) / nt);
    if (lo < hi) {
      /* first cells that can reach [lo, hi): lo - w + 1 .. hi - 1 (periodic) */
      int span = hi - lo + w - 1;
      if (span > ns) span = ns; /* visit each bucket at most once */
      for (int q0 = 0; q0 < span; ++q0) {
        int sv = lo - w + 1 + q0;
        int sb = ((sv % ns) + ns) % ns;
        for (int64_t q = cnt[sb]; q < cnt[sb + 1]; ++q) {
          int64_t j = order[q];
          double kx[MAXW], ky[MAXW], kz[MAXW];
          int s0 = place(x0[j], nx, w, beta, kx);
          int s1 = d >= 2 ? place(x1[j], ny, w, beta, ky) : 0;
          int s2 = d >= 3 ? place(x2[j], nz, w, beta, kz) : 0;
          if (d < 2) ky[0] = 1.0;
          if (d < 3) kz[0] = 1.0;
          const double cr = c[2 * j], ci = c[2 * j + 1];
          const int wy = d >= 2 ? w : 1, wz = d >= 3 ? w : 1;
          for (int jc = 0; jc < wz; ++jc) {
            int iz = s2 + jc;
            if (iz >= nz) iz -= nz;
            for (int jb = 0; jb < wy; ++jb) {
              int iy = s1 + jb;
              if (iy >= ny) iy -= ny;
              int slab = d == 3 ? iz : (d == 2 ? iy : 0);
              if (d >= 2 && (slab < lo || slab >= hi)) continue;
              double wyz = ky[jb] * kz[jc];
              double* row = grid + 2 * (((size_t)iz * ny + iy) * nx);
              for (int ja = 0; ja < w; ++ja) {
                int ix = s0 + ja;
                if (ix >= nx) ix -= nx;
                if (d == 1 && (ix < lo || ix >= hi)) continue;
                double wt = kx[ja] * wyz;
                row[2 * ix] += wt * cr;
                row[2 * ix + 1] += wt * ci;
              }
            }
          }
        }
      }
    }
  }
  (void)plane;
  free(first);
  free(cnt);
  free(order);
  return 0;
}

/* Type-2 interpolation: out[j] = sum_cells grid * weights (complex). */
int nufft_interp(int d, int64_t n, const double* x0, const double* x1, const double* x2, const double* grid,
                 const int* nf, int w, double beta, double* out) {
  if (w > MAXW || d < 1 || d > 3) return 1;
  const int nx = nf[0], ny = d >= 2 ? nf[1] : 1, nz = d >= 3 ? nf[2] : 1;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j) {
    double kx[MAXW], ky[MAXW], kz[MAXW];
    int s0 = place(x0[j], nx, w, beta, kx);
    int s1 = d >= 2 ? place(x1[j], ny, w, beta, ky) : 0;
    int s2 = d >= 3 ? place(x2[j], nz, w, beta, kz) : 0;
    if (d < 2) ky[0] = 1.0;
    if (d < 3) kz[0] = 1.0;
    const int wy = d >= 2 ? w : 1, wz = d >= 3 ? w : 1;
    double sr = 0.0, si = 0.0;
    for (int jc = 0; jc < wz; ++jc) {
      int iz = s2 + jc;
      if (iz >= nz) iz -= nz;
      for (int jb = 0; jb < wy; ++jb) {
        int iy = s1 + jb;
        if (iy >= ny) iy -= ny;
        const double* row = grid + 2 * (((size_t)iz * ny + iy) * nx);
        double wyz = ky[jb] * kz[jc];
        for (int ja = 0; ja < w; ++ja) {
          int ix = s0 + ja;
          if (ix >= nx) ix -= nx;
          double wt = kx[ja] * wyz;
          sr += wt * row[2 * ix];
          si += wt * row[2 * ix + 1];
        }
      }
    }
    out[2 * j] = sr;
    out[2 * j + 1] = si;
  }
  return 0;
}
