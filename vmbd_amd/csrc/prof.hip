// Per-kernel event timing (see prof.h).
#include "prof.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "capi_util.h"

namespace vmbd {
namespace {
struct Rec {
  const char* name;
  hipEvent_t a, b;
};
std::mutex mu;
bool enabled = false;
std::vector<Rec> recs;
std::vector<hipEvent_t> pool;

hipEvent_t take() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
}  // namespace

bool prof_on() { return enabled; }

void prof_begin(const char* name, hipStream_t st) {
  std::lock_guard<std::mutex> lk(mu);
  Rec r{name, take(), take()};
  (void)hipEventRecord(r.a, st);
  recs.push_back(r);
}

void prof_end(hipStream_t st) {
  std::lock_guard<std::mutex> lk(mu);
  if (!recs.empty()) (void)hipEventRecord(recs.back().b, st);
}
}  // namespace vmbd

extern "C" {

int vmbd_prof_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(vmbd::mu);
  vmbd::enabled = on != 0;
  return VMBD_OK;
}

int vmbd_prof_reset(void) {
  std::lock_guard<std::mutex> lk(vmbd::mu);
  for (auto& r : vmbd::recs) {
    (void)hipEventSynchronize(r.b);
    vmbd::pool.push_back(r.a);
    vmbd::pool.push_back(r.b);
  }
  vmbd::recs.clear();
  return VMBD_OK;
}

// Writes "name count total_us avg_us\n" lines; returns bytes needed.
int vmbd_prof_report(char* buf, int32_t len) {
  std::lock_guard<std::mutex> lk(vmbd::mu);
  std::map<std::string, std::pair<long, double>> acc;
  for (auto& r : vmbd::recs) {
    if (hipEventSynchronize(r.b) != hipSuccess) return vmbd_fail(VMBD_E_HIP, "vmbd_prof_report: event sync");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) continue;
    auto& e = acc[r.name];
    e.first += 1;
    e.second += (double)ms * 1000.0;
  }
  std::string out;
  char line[256];
  for (auto& kv : acc) {
    snprintf(line, sizeof line, "%s %ld %.3f %.3f\n", kv.first.c_str(), kv.second.first, kv.second.second,
             kv.second.second / (double)kv.second.first);
    out += line;
  }
  if (buf && len > 0) {
    snprintf(buf, (size_t)len, "%s", out.c_str());
  }
  return (int)out.size() + 1;
}

}  // extern "C"
