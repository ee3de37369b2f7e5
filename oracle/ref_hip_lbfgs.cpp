// The reference's own LBFGS (source/lbfgs.cpp, compiled where it lies) and
// MODIFIED_FIRE (source/fresco/modified_fire.hpp) driving a potential of
// libfresco_hip.so through include/fresco_hip_potential.hpp, the adapter a
// maintainer adds on the reference's side.  Built by `make -C oracle ref`
// into oracle/_ref/libref_hip_opt.so.
//
// TEST INFRASTRUCTURE ONLY: shows that the reference's optimizers run the GPU
// potential unchanged, and gives the device optimizers a reference trajectory
// over the very same potential evaluations.
#include <cstdint>
#include <memory>
#include <vector>

#include "fresco/lbfgs.hpp"
#include "fresco/modified_fire.hpp"
#include "fresco_hip_potential.hpp"

namespace {
struct Handle {
  std::shared_ptr<fresco::BasePotential> pot;
  std::unique_ptr<fresco::GradientOptimizer> opt;
  fresco::LBFGS* lbfgs = nullptr;
};
}  // namespace

extern "C" {

// pot: a fresco_pot* owned by the caller (the adapter does not destroy it).
// Returns NULL on failure.
void* ref_hip_lbfgs_create(void* pot, const double* x0, int64_t n, double tol, int M, double maxstep,
                           double max_f_rise, double H0, int max_iter) {
  try {
    auto* h = new Handle;
    h->pot = std::make_shared<fresco::HipPotential>(static_cast<fresco_pot*>(pot), false);
    std::vector<double> x(x0, x0 + n);
    auto* o = new fresco::LBFGS(h->pot, x, tol, M);
    o->set_H0(H0);
    o->set_maxstep(maxstep);
    o->set_max_f_rise(max_f_rise);
    o->set_max_iter(max_iter);
    h->lbfgs = o;
    h->opt.reset(o);
    return h;
  } catch (...) {
    return nullptr;
  }
}

void* ref_hip_fire_create(void* pot, const double* x0, int64_t n, double dtstart, double dtmax, double maxstep,
                          int64_t Nmin, double finc, double fdec, double fa, double astart, double tol, int stepback,
                          int max_iter) {
  try {
    auto* h = new Handle;
    h->pot = std::make_shared<fresco::HipPotential>(static_cast<fresco_pot*>(pot), false);
    std::vector<double> x(x0, x0 + n);
    auto* o = new fresco::MODIFIED_FIRE(h->pot, x, dtstart, dtmax, maxstep, (size_t)Nmin, finc, fdec, fa, astart,
                                        tol, stepback != 0);
    o->set_max_iter(max_iter);
    h->opt.reset(o);
    return h;
  } catch (...) {
    return nullptr;
  }
}

// 0 on success, 1 if the optimizer threw
int ref_hip_one_iteration(void* p) {
  try {
    static_cast<Handle*>(p)->opt->one_iteration();
    return 0;
  } catch (...) {
    return 1;
  }
}

int ref_hip_run(void* p, int niter) {
  try {
    static_cast<Handle*>(p)->opt->run(niter);
    return 0;
  } catch (...) {
    return 1;
  }
}

// stop_criterion_satisfied (optimizer.hpp:195-201): evaluates x0 first if
// nothing has been evaluated yet; -1 if the optimizer threw
int ref_hip_stop(void* p) {
  try {
    return static_cast<Handle*>(p)->opt->stop_criterion_satisfied() ? 1 : 0;
  } catch (...) {
    return -1;
  }
}

void ref_hip_state(void* p, double* f, double* rms, int* niter, int* nfev, double* x, double* g) {
  auto* h = static_cast<Handle*>(p);
  *f = h->opt->get_f();
  *rms = h->opt->get_rms();
  *niter = h->opt->get_niter();
  *nfev = h->opt->get_nfev();
  const std::vector<double> xv = h->opt->get_x(), gv = h->opt->get_g();
  for (size_t i = 0; i < xv.size(); ++i) {
    x[i] = xv[i];
    g[i] = gv[i];
  }
}

void ref_hip_destroy(void* p) { delete static_cast<Handle*>(p); }

}  // extern "C"
