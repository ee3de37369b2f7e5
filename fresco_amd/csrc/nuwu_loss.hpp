// Per-pair arithmetic of the NUwU loss (nuwu.hpp:260-287), shared by the
// stand-alone loss kernel (nuwu.hip) and the band transform's fused
// z-pass + loss kernel (banddft.hip).
//
// A pair is a mode k with kx >= 0 (the representative; for kx = 0 the one
// with ky > 0, or ky = 0 and kz >= 0) together with -k: the real input gives
// rho_-k = conj(rho_k), so one spectrum value serves both terms.
#pragma once

#include <hip/hip_runtime.h>

#include "nufft.hpp"

namespace fresco_hip {

struct PairDev {
  const int* h;      // half-spectrum index of k (kx >= 0 representative)
  const int* hneg;   // half-spectrum index of -k when kx == 0 and k != 0, else -1
  const int* ik;     // flat mode index of k (reference order)
  const int* imk;    // flat mode index of -k
  const int* kv;     // k components [npairs*3]
  const double* Vk;
  const double* Vm;
  const double* Sk;
  const double* Sm;
  const double* ipsi;  // 1 / Psi(k)
};

// The tables of pair i that its loss reads (loaded before the spectrum value
// is known, so the loads can be issued early).
struct PairVals {
  double ip, Sk, Vk, Sm, Vm;
  bool self;  // k = 0 alone
};
__device__ __forceinline__ PairVals pair_vals(const PairDev& pd, int i) {
  PairVals v;
  v.ip = pd.ipsi[i];
  v.Sk = pd.Sk[i];
  v.Vk = pd.Vk[i];
  v.Sm = pd.Sm[i];
  v.Vm = pd.Vm[i];
  v.self = pd.ik[i] == pd.imk[i];
  return v;
}

// Loss terms of a pair from the forward transform's value sv of its
// representative (S_k = sum_j c_j e^{+i k.x_j} Psi(k)), returned as the
// energy; with want_x, the value the inverse (complex-to-real) transform
// takes at k: X_k = conj(H_k) / Psi(k), H_k = (f_k + conj(f_-k)) / 2 with
// f = gscale g rho (xr, xi).
template <typename C>
__device__ __forceinline__ double pair_loss(const PairVals& v, C sv, double norm, double gscale, bool want_x,
                                            double& xr, double& xi) {
  const double ip = v.ip;
  const double re = (double)sv.re * ip;  // rho_k = conj(S_k) / Psi(k)
  const double im = -(double)sv.im * ip;
  const double s = (re * re + im * im) / norm;  // nuwu.hpp:262-264
  const double Sk = v.Sk, Vk = v.Vk;
  double D = s - Sk;
  if (Sk != 0) D /= Sk;  // nuwu.hpp:265-268
  double e = Vk * D * D;  // nuwu.hpp:270-271
  double gk = Vk * D;
  if (Sk != 0) gk /= Sk;  // nuwu.hpp:272-276
  double gm = gk;
  if (!v.self) {
    const double Sm = v.Sm, Vm = v.Vm;
    double Dm = s - Sm;
    if (Sm != 0) Dm /= Sm;
    e += Vm * Dm * Dm;
    gm = Vm * Dm;
    if (Sm != 0) gm /= Sm;
  }
  xr = 0.0;
  xi = 0.0;
  if (want_x) {
    double hr, hi;
    if (v.self) {
      hr = gscale * gk * re;
      hi = 0.0;
    } else {
      const double gg = 0.5 * gscale * (gk + gm);
      hr = gg * re;
      hi = gg * im;
    }
    // C2R input X_k = conj(H_k) / Psi(k): the backward transform then yields
    // sum_k H_k / Psi(k) e^{-i k.l h} (type-2 sign -1, nuwu.hpp:51)
    xr = hr * ip;
    xi = -hi * ip;
  }
  return e;
}

}  // namespace fresco_hip
