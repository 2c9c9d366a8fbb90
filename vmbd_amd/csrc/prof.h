// Opt-in per-kernel timing with HIP events recorded on the launch stream.
// Disabled by default (no events are recorded); vmbd_prof_enable(1) turns it
// on, vmbd_prof_report() synchronises the recorded events and prints
// "name count total_us avg_us" per kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace vmbd {
bool prof_on();
void prof_begin(const char* name, hipStream_t st);
void prof_end(hipStream_t st);

struct ProfScope {
  hipStream_t st;
  bool on;
  ProfScope(const char* name, hipStream_t s) : st(s), on(prof_on()) {
    if (on) prof_begin(name, s);
  }
  ~ProfScope() {
    if (on) prof_end(st);
  }
};
}  // namespace vmbd
