// C shim over the reference's pair potentials (source/fresco/
// inversepower_potential.hpp, wca.hpp with simple_pairwise_potential.hpp,
// cell_list_potential.hpp, cell_lists.hpp, distance.hpp), compiled where they
// lie by `make -C oracle ref` into oracle/_ref/libref_pair.so.
//
// TEST INFRASTRUCTURE ONLY: the reference's own energies / gradients for the
// GPU pair potentials' parity tests.
#include <omp.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <vector>

#include "fresco/base_potential.hpp"
#include "fresco/inversepower_potential.hpp"
#include "fresco/wca.hpp"

namespace {

// the reference's cell-list tests run single-threaded (test_cell_lists.cpp
// SetUp: omp_set_num_threads(1)); so do these calls
struct OneThread {
  int saved = omp_get_max_threads();
  OneThread() { omp_set_num_threads(1); }
  ~OneThread() { omp_set_num_threads(saved); }
};

// method: 0 cartesian, 1 periodic (all pairs), 2 periodic cell lists
template <size_t D>
fresco::BasePotential* make(int kind, int method, double p1, double p2, const std::vector<double>& r,
                            const std::vector<double>& box, double scale) {
  if (kind == 0) {
    if (method == 0) return new fresco::InversePowerCartesian<D>(p1, p2, r);
    if (method == 1) return new fresco::InversePowerPeriodic<D>(p1, p2, r, box);
    return new fresco::InversePowerPeriodicCellLists<D>(p1, p2, r, box, scale, true);
  }
  if (method == 0) return new fresco::WCA<D>(p1, p2, r);
  if (method == 1) return new fresco::WCAPeriodic<D>(p1, p2, r, box);
  return new fresco::WCAPeriodicCellLists<D>(p1, p2, r, box, scale, true);
}

}  // namespace

extern "C" {

void* ref_pair_create(int kind, int method, double p1, double p2, const double* radii, int64_t n, int d,
                      const double* box, double ncellx_scale) {
  OneThread guard;
  std::vector<double> r(radii, radii + n);
  std::vector<double> b(box, box + d);
  try {
    if (d == 1) return make<1>(kind, method, p1, p2, r, b, ncellx_scale);
    if (d == 2) return make<2>(kind, method, p1, p2, r, b, ncellx_scale);
    if (d == 3) return make<3>(kind, method, p1, p2, r, b, ncellx_scale);
  } catch (const std::exception&) {
  }
  return nullptr;
}

double ref_pair_energy(void* h, const double* x, int64_t n) {
  OneThread guard;
  std::vector<double> v(x, x + n);
  return static_cast<fresco::BasePotential*>(h)->get_energy(v);
}

double ref_pair_energy_gradient(void* h, const double* x, double* g, int64_t n) {
  OneThread guard;
  std::vector<double> v(x, x + n), gv(n);
  double e = static_cast<fresco::BasePotential*>(h)->get_energy_gradient(v, gv);
  for (int64_t i = 0; i < n; ++i) g[i] = gv[i];
  return e;
}

void ref_pair_destroy(void* h) { delete static_cast<fresco::BasePotential*>(h); }
}
