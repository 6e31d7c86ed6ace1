// Optional per-launch kernel timing with HIP events recorded on the launch
// stream (seis_profile_start / seis_profile_stop in include/seis.h).
#pragma once
#include <hip/hip_runtime.h>

namespace seis {

bool prof_enabled();
void prof_take_events(hipEvent_t* a, hipEvent_t* b);
void prof_record(const char* name, double alg_bytes, double cell_updates, hipEvent_t a, hipEvent_t b);

// Runs `launch` (which enqueues one kernel on `s`), bracketed by events when
// profiling is on.
template <class F>
inline void timed_launch(hipStream_t s, const char* name, double alg_bytes, double cell_updates, F&& launch) {
  if (!prof_enabled()) {
    launch();
    return;
  }
  hipEvent_t a, b;
  prof_take_events(&a, &b);
  (void)hipEventRecord(a, s);
  launch();
  (void)hipEventRecord(b, s);
  prof_record(name, alg_bytes, cell_updates, a, b);
}

}  // namespace seis
