// Derivatives of the structure-factor loss with respect to its targets S0(k)
// and weights V(k), for differentiating through the potential (the reference's
// targets are constructor constants: Sk0 / V in nuwu.hpp:25-26).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "potential.hpp"

namespace fresco_hip {

namespace {

__global__ void k_target_grads(const double* __restrict__ rho, const double* __restrict__ Sk,
                               const double* __restrict__ V, int64_t m, double norm, const double* __restrict__ nd,
                               double* __restrict__ dS0, double* __restrict__ dV) {
  const double nrm = nd ? nd[0] : norm;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const double re = rho[2 * i], im = rho[2 * i + 1];
    const double s = (re * re + im * im) / nrm;
    const double S0 = Sk[i], v = V[i];
    double D = s - S0;
    if (S0 != 0) D /= S0;
    dV[i] = D * D;
    dS0[i] = S0 != 0 ? -2.0 * v * D * s / (S0 * S0) : -2.0 * v * D;
  }
}

}  // namespace

void launch_target_grads(const double* rho, const double* Sk, const double* V, int64_t m, double norm,
                         const double* norm_dev, double* dS0, double* dV, hipStream_t s) {
  const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (m + 255) / 256));
  hipLaunchKernelGGL(k_target_grads, dim3(nb), dim3(256), 0, s, rho, Sk, V, m, norm, norm_dev, dS0, dV);
  FRESCO_HIP_CHECK(hipGetLastError());
}

}  // namespace fresco_hip
