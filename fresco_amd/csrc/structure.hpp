// Structure factor / g(r) of a point pattern and radial profiles of grids
// (fresco/utils/utils.py:16-52, 138-152); see structure.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace fresco_hip {

// x: device float64 [n*d] positions; c: device float64 [n] weights; L: host
// [d].  Sk (device float64 [nk^d], first index slowest = x) receives
// |sum_j c_j e^{i k.x_j}|^2 / n with x = 2 pi p / L and k_a = i_a - nk/2;
// gr (device, same layout, may be null) the centred real-space transform of
// Sk - 1 times prod(L) / n.  precision 0: float64 NUFFT, 1: float32.
// Synchronises s before returning.
void structure_factor(const double* x, const double* c, int64_t n, int d, const double* L, int64_t nk, double eps,
                      int precision, double* Sk, double* gr, hipStream_t s);

// data: device float64, C order, shape[d] (d <= 3); center: host [d].
// sums[b] / counts[b] (device, b = 0 .. shape[0]/2 + 1): sum and number of
// cells whose shell floor(|i - center|) is b (shells beyond shape[0]/2 fold
// into the last).  Synchronises s before returning.
void radial_profile(const double* data, int d, const int64_t* shape, const double* center, double* sums,
                    int64_t* counts, hipStream_t s);

// Limits of radial_profile_noalloc: cells of the grid, and shells
// (shape[0] / 2 + 2, so that the shell keys sort in one 11-bit pass).
constexpr int64_t RP_NOALLOC_MAX_CELLS = (int64_t)1 << 20;
constexpr int RP_NOALLOC_MAX_SHELLS = 2048;

// radial_profile without any device allocation: its buffers live in a fixed
// per-GPU scratch array of the library.  Same kernels and summation order,
// so the results equal radial_profile's.  Grids beyond the limits above are
// rejected (InvalidArgument).  Calls that overlap on one GPU race on the
// scratch array.  Synchronises s before returning.
void radial_profile_noalloc(const double* data, int d, const int64_t* shape, const double* center, double* sums,
                            int64_t* counts, hipStream_t s);

}  // namespace fresco_hip
