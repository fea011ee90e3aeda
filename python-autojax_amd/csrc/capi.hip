// Library-wide C ABI helpers: error text, version.
#include <cstdarg>

#include "aj_common.h"

namespace aj {
static thread_local char g_err[1024] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }
}  // namespace aj

extern "C" const char* aj_last_error(void) { return aj::last_error(); }
extern "C" int aj_abi_version(void) { return 1; }

// ---------------------------------------------------------------------------
// Kernel timing: optional HIP-event pairs around selected launches.
// ---------------------------------------------------------------------------
#include <mutex>
#include <vector>

namespace aj {
namespace {
struct Slot {
  std::vector<hipEvent_t> start, stop;
  size_t used = 0;
};
std::mutex g_tm;
std::atomic<int> g_timing{0};  // read without the lock on every timed launch
Slot g_slots[AJ_TIMER_COUNT];
}  // namespace

void timer_begin(int id, hipStream_t st) {
  if (!g_timing || id < 0 || id >= AJ_TIMER_COUNT) return;
  std::lock_guard<std::mutex> lk(g_tm);
  Slot& s = g_slots[id];
  if (s.used == s.start.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    s.start.push_back(a);
    s.stop.push_back(b);
  }
  (void)hipEventRecord(s.start[s.used], st);
}

void timer_end(int id, hipStream_t st) {
  if (!g_timing || id < 0 || id >= AJ_TIMER_COUNT) return;
  std::lock_guard<std::mutex> lk(g_tm);
  Slot& s = g_slots[id];
  if (s.used < s.start.size()) {
    (void)hipEventRecord(s.stop[s.used], st);
    ++s.used;
  }
}
}  // namespace aj

extern "C" int aj_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(aj::g_tm);
  aj::g_timing.store(on ? 1 : 0);
  for (auto& s : aj::g_slots) s.used = 0;
  return AJ_OK;
}

extern "C" int aj_timing_read(int timer, double* total_ms, int64_t* count) {
  if (timer < 0 || timer >= AJ_TIMER_COUNT || !total_ms || !count) {
    aj::set_error("bad timer id %d", timer);
    return AJ_EINVAL;
  }
  std::lock_guard<std::mutex> lk(aj::g_tm);
  auto& s = aj::g_slots[timer];
  double tot = 0.0;
  for (size_t i = 0; i < s.used; ++i) {
    AJ_CHECK_HIP(hipEventSynchronize(s.stop[i]));
    float ms = 0.f;
    AJ_CHECK_HIP(hipEventElapsedTime(&ms, s.start[i], s.stop[i]));
    tot += ms;
  }
  *total_ms = tot;
  *count = (int64_t)s.used;
  return AJ_OK;
}
