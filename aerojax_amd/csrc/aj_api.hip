// Version and error reporting of the C ABI.
#include "aj_common.h"

#include <cstdio>
#include <vector>

namespace aj {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
  g_last_error = buf;
  return AJ_ERR_HIP;
}

struct ProfSlot {
  int tag;
  hipEvent_t a, b;
  double bytes;  // per launch
  int launches;  // launches between the two events
};
static std::vector<ProfSlot> g_slots;
static int g_used = 0;
static bool g_on = false;
static int g_repeat = 1;  // launches per tuned launch while profiling
static int g_open = -1;   // slot of the instrumented launch in progress

bool prof_on() { return g_on; }

int prof_begin_launch(int tag, hipStream_t st) {
  if (g_used >= (int)g_slots.size()) return -1;
  ProfSlot& s = g_slots[g_used];
  if (hipEventRecord(s.a, st) != hipSuccess) return -1;
  s.tag = tag;
  s.launches = 1;
  g_open = g_used;
  return g_used++;
}

void prof_end_launch(int slot, double bytes, hipStream_t st) {
  ProfSlot& s = g_slots[slot];
  s.bytes = bytes;
  if (hipEventRecord(s.b, st) != hipSuccess) s.tag = -1;
  g_open = -1;
}

int prof_repeat() { return (g_on && g_open >= 0) ? g_repeat : 1; }

void prof_count(int launches) {
  if (g_open >= 0) g_slots[g_open].launches = launches;
}

static void prof_release() {
  for (auto& s : g_slots) {
    (void)hipEventDestroy(s.a);
    (void)hipEventDestroy(s.b);
  }
  g_slots.clear();
  g_used = 0;
}

}  // namespace aj

extern "C" int aj_profile_begin(int max_records) {
  using namespace aj;
  if (max_records <= 0) return fail(AJ_ERR_INVALID, "aj_profile_begin: max_records must be > 0");
  prof_release();
  g_slots.resize(max_records);
  for (auto& s : g_slots) {
    AJ_HIP(hipEventCreate(&s.a));
    AJ_HIP(hipEventCreate(&s.b));
    s.tag = -1;
    s.bytes = 0;
  }
  g_used = 0;
  g_on = true;
  return AJ_OK;
}

extern "C" int aj_profile_end(void) {
  aj::g_on = false;
  aj::g_open = -1;
  return AJ_OK;
}

extern "C" int aj_profile_set_repeat(int repeat) {
  if (repeat < 1 || repeat > 64) return aj::fail(AJ_ERR_INVALID, "aj_profile_set_repeat: 1..64");
  aj::g_repeat = repeat;
  return AJ_OK;
}

extern "C" int aj_profile_read(int tag, int* count, double* total_ms, double* total_bytes) {
  using namespace aj;
  if (!count || !total_ms || !total_bytes) return fail(AJ_ERR_INVALID, "aj_profile_read: null");
  *count = 0;
  *total_ms = 0.0;
  *total_bytes = 0.0;
  for (int k = 0; k < g_used; ++k) {
    ProfSlot& s = g_slots[k];
    if (s.tag != tag) continue;
    AJ_HIP(hipEventSynchronize(s.b));
    float ms = 0.f;
    AJ_HIP(hipEventElapsedTime(&ms, s.a, s.b));
    *count += s.launches;
    *total_ms += ms;
    *total_bytes += s.bytes * s.launches;
  }
  return AJ_OK;
}

extern "C" const char* aj_version(void) { return "aerojax-amd 0.1.0 gfx950"; }

extern "C" const char* aj_last_error(void) { return aj::g_last_error.c_str(); }
