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
#include "potential.hpp"
#include "prof.hpp"

struct fresco_pot {
  std::unique_ptr<fresco_hip::Potential> impl;
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

int fresco_pot_get_rho(fresco_pot* pot, double* rho, void* stream) {
  return guarded([&]() {
    auto* p = checked(pot);
    if (!rho) throw fresco_hip::InvalidArgument("null pointer");
    p->get_rho(rho, static_cast<hipStream_t>(stream));
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

int fresco_prof_reset(void) {
  return guarded([&]() { fresco_hip::prof_reset(); });
}

int fresco_prof_get(int stage, const char** name, double* total_ms, int64_t* count) {
  return guarded([&]() {
    if (stage < 0 || stage >= fresco_hip::ST_COUNT) throw fresco_hip::InvalidArgument("stage out of range");
    long long c = 0;
    fresco_hip::prof_get(stage, total_ms, &c);
    *count = c;
    if (name) *name = fresco_hip::stage_name(stage);
  });
}

int fresco_pot_destroy(fresco_pot* pot) {
  return guarded([&]() { delete pot; });
}

}  // extern "C"
