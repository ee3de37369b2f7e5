// HIP-event stage profiler (see prof.hpp).
#include "prof.hpp"

#include <mutex>
#include <vector>

#include "common.hpp"

namespace fresco_hip {

namespace {
struct Pending {
  int st;
  hipEvent_t a, b;
};
std::mutex g_mu;
bool g_on = false;
unsigned g_mask = 0;
std::vector<hipEvent_t> g_pool;
std::vector<Pending> g_pending;
hipEvent_t g_open[ST_COUNT] = {};
int g_period = 1;
long long g_seen[ST_COUNT] = {};
double g_ms[ST_COUNT] = {};
long long g_cnt[ST_COUNT] = {};

hipEvent_t get_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  FRESCO_HIP_CHECK(hipEventCreate(&e));
  return e;
}
}  // namespace

const char* stage_name(int st) {
  static const char* names[ST_COUNT] = {"bin_sort", "spread", "fft_fwd", "zero_fill", "loss", "fft_inv", "interp", "spread_merge"};
  return (st >= 0 && st < ST_COUNT) ? names[st] : "";
}

bool prof_enabled() { return g_on; }

unsigned prof_mask() { return g_mask; }

bool prof_stage_on(int st) { return g_on && ((g_mask >> st) & 1u); }

void prof_set(bool on) { prof_set_mask(on ? (1u << ST_COUNT) - 1u : 0u); }

void prof_set_mask(unsigned m) {
  std::lock_guard<std::mutex> l(g_mu);
  g_mask = m & ((1u << ST_COUNT) - 1u);
  g_on = g_mask != 0;
}

void prof_set_period(int k) {
  std::lock_guard<std::mutex> l(g_mu);
  g_period = k < 1 ? 1 : k;
}

bool prof_take(int st) {
  std::lock_guard<std::mutex> l(g_mu);
  return g_seen[st]++ % g_period == 0;
}

void prof_reset() {
  std::lock_guard<std::mutex> l(g_mu);
  for (int i = 0; i < ST_COUNT; ++i) {
    g_ms[i] = 0;
    g_cnt[i] = 0;
    g_seen[i] = 0;
  }
}

void prof_begin(int st, hipStream_t s) {
  std::lock_guard<std::mutex> l(g_mu);
  hipEvent_t e = get_event();
  FRESCO_HIP_CHECK(hipEventRecord(e, s));
  g_open[st] = e;
}

void prof_end(int st, hipStream_t s) {
  std::lock_guard<std::mutex> l(g_mu);
  hipEvent_t e = get_event();
  FRESCO_HIP_CHECK(hipEventRecord(e, s));
  g_pending.push_back({st, g_open[st], e});
  g_open[st] = nullptr;
}

void prof_collect() {
  std::lock_guard<std::mutex> l(g_mu);
  // pairs whose end event has not completed stay pending (their events are
  // not reused)
  std::vector<Pending> keep;
  for (auto& p : g_pending) {
    if (hipEventQuery(p.b) != hipSuccess) {
      keep.push_back(p);
      continue;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      g_ms[p.st] += ms;
      g_cnt[p.st] += 1;
    }
    g_pool.push_back(p.a);
    g_pool.push_back(p.b);
  }
  g_pending.swap(keep);
}

void prof_get(int st, double* total_ms, long long* count) {
  std::lock_guard<std::mutex> l(g_mu);
  *total_ms = g_ms[st];
  *count = g_cnt[st];
}

}  // namespace fresco_hip
