// Stage timing with HIP events, recorded on the stream the stage's kernels
// are launched on.  Disabled by default; enabled through fresco_prof_enable.
#pragma once

#include <hip/hip_runtime.h>

namespace fresco_hip {

// ST_MERGE: the second kernel of a two-kernel spread (partial-tile sums)
enum Stage { ST_SORT = 0, ST_SPREAD, ST_FFT_FWD, ST_ZERO, ST_LOSS, ST_FFT_INV, ST_INTERP, ST_MERGE, ST_COUNT };

const char* stage_name(int st);
bool prof_enabled();             // any stage timed
bool prof_stage_on(int st);      // this stage timed
void prof_set(bool on);          // all stages / none
void prof_set_mask(unsigned m);  // the stages whose bit is set
unsigned prof_mask();
void prof_set_period(int k);     // time one launch in k of each stage
// whether this launch of stage st is timed (counts the stage's launches)
bool prof_take(int st);
void prof_reset();
void prof_begin(int st, hipStream_t s);
void prof_end(int st, hipStream_t s);
// Accumulate the elapsed time of every completed begin/end pair.  Call after
// the stream has been synchronised.
void prof_collect();
void prof_get(int st, double* total_ms, long long* count);

struct ProfScope {
  int st;
  hipStream_t s;
  bool on;
  ProfScope(int st_, hipStream_t s_) : st(st_), s(s_), on(prof_stage_on(st_) && prof_take(st_)) {
    if (on) prof_begin(st, s);
  }
  ~ProfScope() {
    if (on) prof_end(st, s);
  }
};

}  // namespace fresco_hip
