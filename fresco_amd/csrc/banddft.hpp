// Pruned separable DFTs between a real 3D fine grid and the band of modes
// |k_a| <= B that a structure-factor loss reads (the fine grid is sized from
// that band, so B is small against the grid: 24 of 100 cells per dimension
// for the 3D N = 2^20 configuration).  Each direction is three passes that
// contract one axis against a twiddle matrix (batched small matrix products),
// replacing a full real-to-complex / complex-to-real FFT of the grid plus the
// zero fill of its spectrum:
//   forward  grid[z][y][x] -> U1[z][y][kx] -> U2[z][ky][kx] -> S[kz][ky][kx]
//   inverse  X[kz][ky][kx] -> Y1[z][ky][kx] -> Y2[z][y][kx] -> grid[z][y][x]
// with kx in [0, B] (real input: S(-k) = conj S(k)) and ky, kz in [-B, B].
// FP32 runs on the matrix cores (v_mfma_f32_16x16x4f32) in four launches:
// the x pass, the y and z passes of every kx plane fused (the intermediate
// in LDS), and the same in reverse; FP64 keeps three VALU passes per
// direction.
// Signs and scaling are those of the FFTs they replace: forward e^{-2 pi i
// k x / n}, inverse e^{+2 pi i k x / n} unnormalised, the inverse taking the
// kx = 0 plane as it is (Hermitian by construction) and kx > 0 twice (their
// conjugate partners -k).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "nufft.hpp"
#include "nuwu_loss.hpp"

namespace fresco_hip {

// The pairs of the loss grouped by the z-pass tile of their representative
// (fused FP32 form): pairs pstart[t] .. pstart[t + 1] - 1 belong to tile t;
// loc / mir: kz index * 16 + slot of k and of -k (kx = 0, k != 0; else -1)
struct BandLossDev {
  const int* pstart;
  const int* loc;
  const int* mir;
};

template <typename T>
class BandDft {
 public:
  // nf: fine grid (x fastest); band: B
  void init(const int* nf, int band);
  // bands and grids this path takes (the passes cost ~ n^3 B against the
  // FFT's n^3 log n; contracted axes up to 128 cells)
  static bool suits(const int* nf, int band) {
    return band >= 0 && band <= 32 && nf[0] <= 128 && nf[1] <= 128 && nf[2] <= 128;
  }
  // grid (real, nf[2] x nf[1] x nf[0]) -> S (compact band spectrum); zero:
  // optional compact array cleared by the same launches (every entry is
  // written once)
  void forward(const T* grid, Cplx<T>* S, hipStream_t s, Cplx<T>* zero = nullptr);
  // X (compact band spectrum, kx = 0 plane Hermitian) -> grid
  void inverse(const Cplx<T>* X, T* grid, hipStream_t s);
  // compact layout: entries and the index of (kx >= 0, ky, kz)
  int64_t size() const { return (int64_t)nb_ * (2 * B_ + 1) * (2 * B_ + 1); }
  int64_t index(int kx, int ky, int kz) const {
    return ((int64_t)(kz + B_) * (2 * B_ + 1) + (ky + B_)) * nb_ + kx;
  }
  int band() const { return B_; }

  // Fused FP32 form, three launches per evaluation: forward_planes (x and y
  // passes of each z plane), zloss (z pass of each tile of 16 band columns
  // (ky, kx), the loss of the pairs whose representative lies in the tile,
  // mode 1: the pullback values and the inverse z pass) and inverse_planes
  // (y and x passes of each z plane).  The spectrum S is written in the
  // compact layout; the pullback values exist only inside zloss.
  // (FRESCO_BANDDFT_FUSED=0: the four-launch form.)
  bool fused() const { return fused_; }
  int ntiles() const { return ntiles_; }
  // z-pass tile and slot of band column (ky, kx); the kx = 0 columns ky and
  // -ky share a tile
  void column_slot(int ky, int kx, int* tile, int* slot) const {
    const int c = (ky + B_) * nb_ + kx;
    *tile = col_pos_[c] / 16;
    *slot = col_pos_[c] % 16;
  }
  void forward_planes(const T* grid, hipStream_t s);
  // The energy is summed from the tiles' partials (fixed order) by a
  // one-workgroup launch after the loss, or, with defer, by the first
  // workgroup of the next inverse_planes given the partials.
  void zloss(int mode, const PairDev& pd, const BandLossDev& bl, double norm, double gscale, Cplx<T>* S,
             double* partial, double* e_out, double* e_host, bool defer, hipStream_t s);
  void inverse_planes(T* grid, hipStream_t s, const double* partial = nullptr, double* e_out = nullptr,
                      double* e_host = nullptr);

 private:
  int nf_[3] = {0, 0, 0};
  int B_ = -1, nb_ = 0;
  // twiddles of the six passes, [K][J] with J the contracted length
  DeviceBuffer<Cplx<T>> wa_, wb_, wc_, wci_, wbi_, wai_;
  DeviceBuffer<Cplx<T>> u1_, u2_;  // intermediates (the inverse reuses them)
  bool fused_ = false;
  DeviceBuffer<Cplx<T>> wxp_, wzip_, wyip_, wxip_;  // fused form: twiddles in padded LDS layouts
  int ntiles_ = 0;
  std::vector<int> col_pos_;  // band column (ky + B) nb + kx -> tile * 16 + slot
  DeviceBuffer<int> tcol_;    // [tiles][16]: column, or -1
};

}  // namespace fresco_hip
