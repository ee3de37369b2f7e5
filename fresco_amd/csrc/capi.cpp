// extern "C" boundary of libfresco_hip.so (declared in include/fresco_hip.h).
// Exceptions never cross the ABI: they become status codes plus a
// thread-local message, mirroring Cython's `except +` mapping of the
// reference (std::invalid_argument -> ValueError, others -> RuntimeError).
#include <hip/hip_runtime.h>

#include <atomic>
#include <exception>
#include <memory>
#include <string>

#include "../../include/fresco_hip.h"
#include "common.hpp"
#include "optim.hpp"
#include "potential.hpp"
#include "prof.hpp"
#include "structure.hpp"

struct fresco_pot {
  std::shared_ptr<fresco_hip::Potential> impl;
};

struct fresco_opt {
  std::unique_ptr<fresco_hip::GradientOptimizer> impl;
  fresco_hip::LBFGS* lbfgs = nullptr;
};

namespace fresco_hip {

namespace {
thread_local std::string g_last_error;
std::atomic<int> g_fault_count{0};
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

bool consume_fault() {
  int c = g_fault_count.load();
  while (c > 0) {
    if (g_fault_count.compare_exchange_weak(c, c - 1)) return true;
  }
  return false;
}

}  // namespace fresco_hip

namespace {

template <typename F>
int guarded(F&& f) {
  try {
    f();
    fresco_hip::set_last_error("");
    return FRESCO_OK;
  } catch (const fresco_hip::InvalidArgument& e) {
    fresco_hip::set_last_error(e.what());
    return FRESCO_E_INVALID;
  } catch (const fresco_hip::HipFailure& e) {
    fresco_hip::set_last_error(e.what());
    return FRESCO_E_HIP;
  } catch (const std::exception& e) {
    fresco_hip::set_last_error(e.what());
    return FRESCO_E_RUNTIME;
  } catch (...) {
    fresco_hip::set_last_error("unknown C++ exception");
    return FRESCO_E_RUNTIME;
  }
}

fresco_hip::Potential* checked(fresco_pot* p) {
  if (!p || !p->impl) throw fresco_hip::InvalidArgument("null potential handle");
  return p->impl.get();
}

fresco_hip::GradientOptimizer* checked(fresco_opt* o) {
  if (!o || !o->impl) throw fresco_hip::InvalidArgument("null optimizer handle");
  return o->impl.get();
}

}  // namespace

extern "C" {

const char* fresco_last_error(void) { return fresco_hip::g_last_error.c_str(); }

const char* fresco_version(void) { return "fresco_hip 0.1.0 (gfx950)"; }

int fresco_set_fault(int count) {
  fresco_hip::g_fault_count.store(count < 0 ? 0 : count);
  return FRESCO_OK;
}

int fresco_nuwu_create(const double* radii, int64_t n, int d, double K, const double* Sk, const double* V, int64_t m,
                       const double* L, double eps, int precision, double sigma, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    if (!radii || !Sk || !V || !L) throw fresco_hip::InvalidArgument("null array argument");
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_nuwu(radii, n, d, K, Sk, V, m, L, eps, precision, sigma);
    *out = h.release();
  });
}

int fresco_uwu_create(const int64_t* shape, int d, const double* Sk, const double* V, int64_t m, int precision,
                      int exact_grad, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    if (!shape || !Sk || !V) throw fresco_hip::InvalidArgument("null array argument");
    if (precision < 0 || precision > 2) throw fresco_hip::InvalidArgument("unknown precision");
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_uwu(shape, d, Sk, V, m, precision, exact_grad);
    *out = h.release();
  });
}

int fresco_uwnu_create(const int64_t* shape, int d, const double* K, int64_t nkd, const double* Sk, const double* V,
                       int64_t nk, const double* L, double eps, int precision, double sigma, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    if (!shape || !K || !Sk || !V || !L) throw fresco_hip::InvalidArgument("null array argument");
    if (precision < 0 || precision > 2) throw fresco_hip::InvalidArgument("unknown precision");
    if (!(eps > 0)) throw fresco_hip::InvalidArgument("eps must be positive");
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_uwnu(shape, d, K, nkd, Sk, V, nk, L, eps, precision, sigma);
    *out = h.release();
  });
}

int fresco_nuwnu_create(const double* radii, int64_t n, int d, const double* K, int64_t nkd, const double* Sk,
                        const double* V, int64_t m, const double* L, double eps, int precision, int method,
                        double sigma, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    if (!radii || !K || !Sk || !V || !L) throw fresco_hip::InvalidArgument("null array argument");
    if (precision < 0 || precision > 2) throw fresco_hip::InvalidArgument("unknown precision");
    if (method == 0 && !(eps > 0)) throw fresco_hip::InvalidArgument("eps must be positive");
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_nuwnu(radii, n, d, K, nkd, Sk, V, m, L, eps, precision, method, sigma);
    *out = h.release();
  });
}

int fresco_pair_create(int kind, double p1, double eps, const double* radii, int64_t n, int d, const double* box,
                       int method, int use_cell_lists, double ncellx_scale, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    if (!radii) throw fresco_hip::InvalidArgument("null array argument");
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_pair(kind, p1, eps, radii, n, d, box, method, use_cell_lists, ncellx_scale);
    *out = h.release();
  });
}

int fresco_nuwu_batch_create(int npat, const int64_t* npts, const double* radii, int d, const int64_t* nmodes,
                             const double* Sk, const double* V, const double* L, double eps, int precision,
                             double sigma, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    if (!npts || !radii || !nmodes || !Sk || !V || !L) throw fresco_hip::InvalidArgument("null array argument");
    if (precision < 0 || precision > 2) throw fresco_hip::InvalidArgument("unknown precision");
    if (!(eps > 0)) throw fresco_hip::InvalidArgument("eps must be positive");
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_nuwu_batch(npat, npts, radii, d, nmodes, Sk, V, L, eps, precision, sigma);
    *out = h.release();
  });
}

int fresco_pot_pattern_energies(fresco_pot* pot, double* energies, int npat) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!energies) throw fresco_hip::InvalidArgument("null pointer");
    p->pattern_energies(energies, npat, nullptr);
  });
}

int fresco_combined_create(fresco_pot* const* pots, int count, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    if (!pots || count <= 0) throw fresco_hip::InvalidArgument("no potentials");
    std::vector<std::shared_ptr<fresco_hip::Potential>> v;
    for (int i = 0; i < count; ++i) {
      checked(pots[i]);
      v.push_back(pots[i]->impl);
    }
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_combined(std::move(v));
    *out = h.release();
  });
}

int fresco_callback_create(int64_t ndof, fresco_eg_callback cb, void* user, fresco_pot** out) {
  return guarded([&]() {
    if (!out) throw fresco_hip::InvalidArgument("out is null");
    *out = nullptr;
    auto h = std::make_unique<fresco_pot>();
    h->impl = fresco_hip::make_callback(ndof, cb, user);
    *out = h.release();
  });
}

int fresco_lbfgs_create(fresco_pot* pot, const double* x0, int64_t n, double tol, int M, fresco_opt** out) {
  return guarded([&]() {
    if (!out || !x0) throw fresco_hip::InvalidArgument("null pointer");
    *out = nullptr;
    checked(pot);
    auto h = std::make_unique<fresco_opt>();
    auto* o = new fresco_hip::LBFGS(pot->impl, x0, n, tol, M);
    h->impl.reset(o);
    h->lbfgs = o;
    *out = h.release();
  });
}

int fresco_fire_create(fresco_pot* pot, const double* x0, int64_t n, double dtstart, double dtmax, double maxstep,
                       int64_t Nmin, double finc, double fdec, double fa, double astart, double tol, int stepback,
                       fresco_opt** out) {
  return guarded([&]() {
    if (!out || !x0) throw fresco_hip::InvalidArgument("null pointer");
    *out = nullptr;
    checked(pot);
    auto h = std::make_unique<fresco_opt>();
    h->impl.reset(new fresco_hip::ModifiedFire(pot->impl, x0, n, dtstart, dtmax, maxstep, Nmin, finc, fdec, fa,
                                               astart, tol, stepback != 0));
    *out = h.release();
  });
}

int fresco_opt_set_param(fresco_opt* opt, int param, double value) {
  return guarded([&]() {
    auto* o = checked(opt);
    switch (param) {
      case FRESCO_OPT_TOL: o->set_tol(value); break;
      case FRESCO_OPT_MAXSTEP: o->set_maxstep(value); break;
      case FRESCO_OPT_MAX_ITER: o->set_max_iter((int)value); break;
      case FRESCO_OPT_IPRINT: o->set_iprint((int)value); break;
      case FRESCO_OPT_VERBOSITY: o->set_verbosity((int)value); break;
      case FRESCO_OPT_USE_GRAPH: o->set_use_graph(value != 0.0); break;
      case FRESCO_OPT_LBFGS_FORM:
        if (!opt->lbfgs) throw fresco_hip::InvalidArgument("parameter only applies to LBFGS");
        opt->lbfgs->set_form((int)value);
        break;
      case FRESCO_OPT_H0:
      case FRESCO_OPT_MAX_F_RISE:
      case FRESCO_OPT_USE_RELATIVE_F:
        if (!opt->lbfgs) throw fresco_hip::InvalidArgument("parameter only applies to LBFGS");
        if (param == FRESCO_OPT_H0) opt->lbfgs->set_H0(value);
        if (param == FRESCO_OPT_MAX_F_RISE) opt->lbfgs->set_max_f_rise(value);
        if (param == FRESCO_OPT_USE_RELATIVE_F) opt->lbfgs->set_use_relative_f((int)value);
        break;
      default: throw fresco_hip::InvalidArgument("unknown optimizer parameter " + std::to_string(param));
    }
  });
}

int fresco_opt_run(fresco_opt* opt, int niter) {
  return guarded([&]() {
    auto* o = checked(opt);
    if (niter < 0)
      o->run();
    else
      o->run(niter);
  });
}

int fresco_opt_one_iteration(fresco_opt* opt) {
  return guarded([&]() { checked(opt)->one_iteration(); });
}

int fresco_opt_reset(fresco_opt* opt, const double* x0, int64_t n) {
  return guarded([&]() {
    auto* o = checked(opt);
    if (!x0) throw fresco_hip::InvalidArgument("null pointer");
    if (n != o->size())
      throw fresco_hip::InvalidArgument(
          "The number of degrees of freedom (x0.size()) cannot change when calling reset()");
    o->reset(x0);
  });
}

int fresco_opt_set_func_gradient(fresco_opt* opt, double f, const double* grad, int64_t n) {
  return guarded([&]() {
    auto* o = checked(opt);
    if (!grad) throw fresco_hip::InvalidArgument("null pointer");
    if (n != o->size()) throw fresco_hip::InvalidArgument("the gradient has the wrong size");
    o->set_func_gradient(f, grad);
  });
}

int fresco_opt_stop_criterion(fresco_opt* opt, int* satisfied) {
  return guarded([&]() { *satisfied = checked(opt)->stop_criterion_satisfied() ? 1 : 0; });
}

int fresco_opt_graph_stats(fresco_opt* opt, int64_t* captures, int64_t* replays) {
  return guarded([&]() {
    if (!captures || !replays) throw fresco_hip::InvalidArgument("null pointer");
    checked(opt)->graph_stats(captures, replays);
  });
}

int fresco_opt_get_state(fresco_opt* opt, double* f, double* rms, int* niter, int* nfev, int* maxiter, double* H0) {
  return guarded([&]() {
    auto* o = checked(opt);
    if (f) *f = o->get_f();
    if (rms) *rms = o->get_rms();
    if (niter) *niter = o->get_niter();
    if (nfev) *nfev = o->get_nfev();
    if (maxiter) *maxiter = o->get_maxiter();
    if (H0) *H0 = o->get_H0();
  });
}

int fresco_opt_get_x(fresco_opt* opt, double* x) {
  return guarded([&]() {
    if (!x) throw fresco_hip::InvalidArgument("null pointer");
    checked(opt)->get_x(x);
  });
}

int fresco_opt_get_g(fresco_opt* opt, double* g) {
  return guarded([&]() {
    if (!g) throw fresco_hip::InvalidArgument("null pointer");
    checked(opt)->get_g(g);
  });
}

int fresco_opt_destroy(fresco_opt* opt) {
  return guarded([&]() { delete opt; });
}

int fresco_pot_energy(fresco_pot* pot, const double* x, double* energy, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!x || !energy) throw fresco_hip::InvalidArgument("null pointer");
    *energy = p->energy(x, static_cast<hipStream_t>(stream));
  });
}

int fresco_pot_energy_gradient(fresco_pot* pot, const double* x, double* grad, double* energy, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!x || !grad || !energy) throw fresco_hip::InvalidArgument("null pointer");
    *energy = p->energy_gradient(x, grad, static_cast<hipStream_t>(stream));
  });
}

int fresco_pot_energy_gradient_stream(fresco_pot* pot, const double* x, double* grad, double* energy,
                                      void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!x || !grad || !energy) throw fresco_hip::InvalidArgument("null pointer");
    *energy = p->energy_gradient_stream(x, grad, static_cast<hipStream_t>(stream));
  });
}

int fresco_pot_energy_keep(fresco_pot* pot, const double* x, double* energy, uint64_t* token, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!x || !energy || !token) throw fresco_hip::InvalidArgument("null pointer");
    uint64_t t = 0;
    *token = p->energy_keep(x, energy, &t, static_cast<hipStream_t>(stream)) ? t : 0;
  });
}

int fresco_pot_gradient_kept(fresco_pot* pot, uint64_t token, double* grad, int* ok, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!ok) throw fresco_hip::InvalidArgument("null pointer");
    *ok = token != 0 && p->gradient_kept(token, grad, static_cast<hipStream_t>(stream)) ? 1 : 0;
  });
}

int fresco_pot_get_rho(fresco_pot* pot, double* rho, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!rho) throw fresco_hip::InvalidArgument("null pointer");
    p->get_rho(rho, static_cast<hipStream_t>(stream));
  });
}

int fresco_pot_backward(fresco_pot* pot, double* dS0, double* dV, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!dS0 || !dV) throw fresco_hip::InvalidArgument("null pointer");
    p->target_grads(dS0, dV, static_cast<hipStream_t>(stream));
  });
}

int fresco_pot_set_targets(fresco_pot* pot, const double* Sk, const double* V, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!Sk) throw fresco_hip::InvalidArgument("null pointer");
    p->set_targets(Sk, V, static_cast<hipStream_t>(stream));
  });
}

int fresco_pot_spread_config(fresco_pot* pot, int* mode, int* parts) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!mode || !parts) throw fresco_hip::InvalidArgument("null pointer");
    p->spread_config(mode, parts);
  });
}

int fresco_pot_spread_source(fresco_pot* pot, int* source) {
  return guarded([&]() {
    if (!source) throw fresco_hip::InvalidArgument("null pointer");
    *source = checked(pot)->spread_source();
  });
}

int fresco_pot_ndof(fresco_pot* pot, int64_t* ndof) {
  return guarded([&]() { *ndof = checked(pot)->ndof(); });
}

int fresco_pot_nmodes(fresco_pot* pot, int64_t* nmodes) {
  return guarded([&]() { *nmodes = checked(pot)->nmodes(); });
}

int fresco_pot_plan_info(fresco_pot* pot, int64_t* nf, int* w, double* beta, int* tile) {
  return guarded([&]() { checked(pot)->plan_info(nf, w, beta, tile); });
}

int fresco_prof_enable(int on) {
  return guarded([&]() { fresco_hip::prof_set(on != 0); });
}

int fresco_prof_enable_stages(unsigned mask) {
  return guarded([&]() { fresco_hip::prof_set_mask(mask); });
}

int fresco_prof_sample_period(int period) {
  return guarded([&]() {
    if (period < 1) throw fresco_hip::InvalidArgument("profiler sample period must be at least 1");
    fresco_hip::prof_set_period(period);
  });
}

int fresco_prof_reset(void) {
  return guarded([&]() { fresco_hip::prof_reset(); });
}

int fresco_prof_get(int stage, const char** name, double* total_ms, int64_t* count) {
  return guarded([&]() {
    if (stage < 0 || stage >= fresco_hip::ST_COUNT) throw fresco_hip::InvalidArgument("stage out of range");
    // every recorded pair complete (evaluations may return before their
    // streams drain)
    FRESCO_HIP_CHECK(hipDeviceSynchronize());
    fresco_hip::prof_collect();
    long long c = 0;
    fresco_hip::prof_get(stage, total_ms, &c);
    *count = c;
    if (name) *name = fresco_hip::stage_name(stage);
  });
}

int fresco_structure_factor(const double* x, const double* c, int64_t n, int d, const double* L, int64_t nk,
                            double eps, int precision, double* Sk, double* gr, void* stream) {
  return guarded([&]() {
    if (!L) throw fresco_hip::InvalidArgument("null pointer");
    fresco_hip::structure_factor(x, c, n, d, L, nk, eps, precision, Sk, gr, static_cast<hipStream_t>(stream));
  });
}

int fresco_radial_profile(const double* data, int d, const int64_t* shape, const double* center, double* sums,
                          int64_t* counts, void* stream) {
  return guarded([&]() {
    if (!shape || !center) throw fresco_hip::InvalidArgument("null pointer");
    fresco_hip::radial_profile(data, d, shape, center, sums, counts, static_cast<hipStream_t>(stream));
  });
}

int fresco_radial_profile_noalloc(const double* data, int d, const int64_t* shape, const double* center,
                                  double* sums, int64_t* counts, void* stream) {
  return guarded([&]() {
    if (!shape || !center) throw fresco_hip::InvalidArgument("null pointer");
    fresco_hip::radial_profile_noalloc(data, d, shape, center, sums, counts, static_cast<hipStream_t>(stream));
  });
}

int fresco_pot_destroy(fresco_pot* pot) {
  return guarded([&]() { delete pot; });
}

}  // extern "C"
