// Exponential-of-semicircle ("ES") spreading kernel of the non-uniform FFT.
//
//   phi(z) = exp(beta * (sqrt(1 - z^2) - 1)),  |z| <= 1,   0 otherwise
//
// A point at fine-grid coordinate g touches the w cells s..s+w-1 with
// s = ceil(g - w/2); cell s+j carries phi(2 (j - off) / w), off = g - s.
// Width/beta follow the published rule of Barnett, Magland & af Klinteberg,
// "A parallel non-uniform fast Fourier transform library based on an
// exponential of semicircle kernel", SIAM J. Sci. Comput. 41 (2019):
//   sigma = 2 : w = ceil(log10(10/eps)),  beta = 2.30 w (2.20/2.26/2.38 for w=2/3/4)
//   sigma != 2: w = ceil(-ln(eps) / (pi sqrt(1 - 1/sigma))),
//               beta = 0.97 pi (1 - 1/(2 sigma)) w
// The deconvolution factor of mode k on a grid of n cells is
//   Psi(k) = integral_{-w/2}^{w/2} phi(2t/w) cos(2 pi k t / n) dt,
// computed here by Gauss-Legendre quadrature in float64.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace fresco_hip {

struct EsParams {
  int w = 7;
  double beta = 16.1;
};

// extra_w widens the kernel beyond the eps rule (the kernel-derivative
// interpolation of the adjoint path amplifies the aliasing error by about the
// ratio of the fine-grid Nyquist to the mode band, one more digit in 2D/3D).
EsParams es_params(double eps, double sigma, int extra_w = 0);

// Psi(k) for k = 0..kmax on a periodic grid of n cells (float64).
std::vector<double> es_deconv(const EsParams& p, int n, int kmax);

// Continuous transform of the kernel on [-1, 1]: int phi(z) cos(u z) dz for
// each u (float64, Gauss-Legendre), the deconvolution of type-3 transforms.
std::vector<double> es_ft(const EsParams& p, const std::vector<double>& u);

// Smallest even 2,3,5-smooth integer >= max(target, 2w).
int next_smooth_even(int target, int w);

// Device-side kernel constants.  `bexp` scales the exponent: beta for
// float64 (natural exp) and beta*log2(e) for float32 (hardware exp2).
template <typename T>
struct EsDev {
  T bexp;
  T beta;
  T two_over_w;  // 2/w: kernel index offset -> z
};

__device__ __forceinline__ double es_exp(double x) { return exp(x); }
__device__ __forceinline__ float es_exp(float x) { return __builtin_amdgcn_exp2f(x); }

// float32: the hardware v_sqrt_f32 / v_rcp_f32 (about 1 ulp), far below the
// kernel's own truncation error; float64 keeps the IEEE operations.
template <typename T>
__device__ __forceinline__ T es_sqrt(T a) { return sqrt(a); }
template <>
__device__ __forceinline__ float es_sqrt<float>(float a) { return __builtin_amdgcn_sqrtf(a); }
template <typename T>
__device__ __forceinline__ T es_rcp(T a) { return T(1) / a; }
template <>
__device__ __forceinline__ float es_rcp<float>(float a) { return __builtin_amdgcn_rcpf(a); }

// phi(z)
template <typename T>
__device__ __forceinline__ T es_phi(T z, const EsDev<T>& k) {
  T a = T(1) - z * z;
  return a > T(0) ? es_exp(k.bexp * (es_sqrt(a) - T(1))) : T(0);
}

// phi(z) and d phi / d z together.  Near |z| = 1 the derivative has an
// integrable 1/sqrt singularity scaled by exp(-beta); it is returned as 0 on
// the support edge itself.
template <typename T>
__device__ __forceinline__ void es_phi_dphi(T z, const EsDev<T>& k, T& v, T& dv) {
  T a = T(1) - z * z;
  if (a > T(0)) {
    T r = es_sqrt(a);
    v = es_exp(k.bexp * (r - T(1)));
    dv = -k.beta * z * es_rcp(r) * v;
  } else {
    v = T(0);
    dv = T(0);
  }
}

}  // namespace fresco_hip
