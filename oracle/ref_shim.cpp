// C shim over the reference's own optimizer sources, compiled where they lie
// (source/lbfgs.cpp, source/modified_fire.cpp, source/fresco/*.hpp) by
// `make -C oracle ref` into oracle/_ref/libref_opt.so.
//
// TEST INFRASTRUCTURE ONLY: lets tests drive the reference LBFGS / modified
// FIRE (source/fresco/lbfgs.hpp, modified_fire.hpp) with a potential given as
// a C callback, to pin the optimizer oracle and the device optimizers.
#include <cstdint>
#include <memory>
#include <vector>

#include "fresco/base_potential.hpp"
#include "fresco/lbfgs.hpp"
#include "fresco/modified_fire.hpp"

extern "C" typedef double (*ref_eg_cb)(const double* x, double* g, int64_t n, void* user);

namespace {
class CallbackPotential : public fresco::BasePotential {
 public:
  CallbackPotential(ref_eg_cb f, void* u, int64_t n) : f_(f), u_(u), n_(n) {}
  double get_energy(const std::vector<double>& x) override {
    std::vector<double> g(x.size());
    return f_(x.data(), g.data(), (int64_t)x.size(), u_);
  }
  double get_energy_gradient(const std::vector<double>& x, std::vector<double>& g) override {
    return f_(x.data(), g.data(), (int64_t)x.size(), u_);
  }

 private:
  ref_eg_cb f_;
  void* u_;
  int64_t n_;
};

struct Handle {
  std::shared_ptr<fresco::BasePotential> pot;
  std::unique_ptr<fresco::GradientOptimizer> opt;
  fresco::LBFGS* lbfgs = nullptr;
};
}  // namespace

extern "C" {

void* ref_lbfgs_create(ref_eg_cb f, void* u, const double* x0, int64_t n, double tol, int M, double maxstep,
                       double max_f_rise, double H0, int rel, int max_iter) {
  auto* h = new Handle;
  h->pot = std::make_shared<CallbackPotential>(f, u, n);
  std::vector<double> x(x0, x0 + n);
  auto* o = new fresco::LBFGS(h->pot, x, tol, M);
  o->set_H0(H0);
  o->set_maxstep(maxstep);
  o->set_max_f_rise(max_f_rise);
  o->set_max_iter(max_iter);
  if (rel) o->set_use_relative_f(1);
  h->lbfgs = o;
  h->opt.reset(o);
  return h;
}

void* ref_fire_create(ref_eg_cb f, void* u, const double* x0, int64_t n, double dtstart, double dtmax, double maxstep,
                      int64_t Nmin, double finc, double fdec, double fa, double astart, double tol, int stepback,
                      int max_iter) {
  auto* h = new Handle;
  h->pot = std::make_shared<CallbackPotential>(f, u, n);
  std::vector<double> x(x0, x0 + n);
  auto* o = new fresco::MODIFIED_FIRE(h->pot, x, dtstart, dtmax, maxstep, (size_t)Nmin, finc, fdec, fa, astart, tol,
                                      stepback != 0);
  o->set_max_iter(max_iter);
  h->opt.reset(o);
  return h;
}

void ref_opt_run(void* p, int niter) {
  auto* h = static_cast<Handle*>(p);
  if (niter < 0)
    h->opt->run();
  else
    h->opt->run(niter);
}

void ref_opt_one_iteration(void* p) { static_cast<Handle*>(p)->opt->one_iteration(); }

int ref_opt_stop(void* p) { return static_cast<Handle*>(p)->opt->stop_criterion_satisfied() ? 1 : 0; }

void ref_opt_state(void* p, double* f, double* rms, int* niter, int* nfev, double* H0, double* x, double* g) {
  auto* h = static_cast<Handle*>(p);
  *f = h->opt->get_f();
  *rms = h->opt->get_rms();
  *niter = h->opt->get_niter();
  *nfev = h->opt->get_nfev();
  *H0 = h->lbfgs ? h->lbfgs->get_H0() : 0.0;
  std::vector<double> xv = h->opt->get_x(), gv = h->opt->get_g();
  for (size_t i = 0; i < xv.size(); ++i) {
    x[i] = xv[i];
    g[i] = gv[i];
  }
}

void ref_opt_reset(void* p, const double* x0, int64_t n) {
  std::vector<double> x(x0, x0 + n);
  static_cast<Handle*>(p)->opt->reset(x);
}

void ref_opt_destroy(void* p) { delete static_cast<Handle*>(p); }

}  // extern "C"
