// Launch-shape autotuner.  A kernel with several compiled tile shapes or
// grid sizes asks tune_pick(key, n, launch) on its first launch for a given
// problem (key = kernel + shape); every candidate is launched once to warm
// up, then the candidates take turns for kTuneRounds rounds of kTuneReps
// launches between HIP events on the caller's stream, and the candidate
// with the fastest round is kept for the rest of the process (the last
// launch, of the winner, leaves the outputs in place: candidates are only
// kernels whose outputs depend on nothing they write).  Decisions can be
// read back (aj_tune_report), preset from such a report (aj_tune_load, so
// that a profiled run launches the shapes a timed run picked) and tuning
// switched off (aj_tune_enable(0) or AJ_TUNE=0: candidate 0 everywhere).
// Never tunes inside stream capture.
#include "aj_common.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace aj {

namespace {
constexpr int kTuneReps = 5;
constexpr int kTuneRounds = 3;
std::mutex g_tune_mu;
std::map<std::string, std::pair<int, std::vector<float>>> g_tune;  // key -> (pick, ms per cand)
int g_tune_on = -1;  // -1: read AJ_TUNE on first use

bool tune_enabled() {
  if (g_tune_on < 0) {
    const char* e = std::getenv("AJ_TUNE");
    g_tune_on = (e && e[0] == '0') ? 0 : 1;
  }
  return g_tune_on == 1;
}
}  // namespace

int tune_pick(const std::string& key, int ncand, const std::function<int(int, hipStream_t)>& launch,
              hipStream_t st, int* pick) {
  // while profiling with a repeat factor, a tunable kernel (its outputs do
  // not depend on what it writes) is launched that many times back to back
  // inside the instrumented launch's events
  auto run = [&](int c) -> int {
    const int reps = ncand >= 2 ? prof_repeat() : 1;
    for (int r = 0; r < reps; ++r)
      if (int rc = launch(c, st)) return rc;
    if (reps > 1) prof_count(reps);
    return AJ_OK;
  };
  // the decision is read under the lock and launched outside it: a launch
  // may tune a kernel of its own (nested keys)
  int known = -1;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end() && it->second.first < ncand) known = it->second.first;
  }
  if (known >= 0) {
    *pick = known;
    return run(known);
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(st, &cap);
  if (!tune_enabled() || ncand < 2 || cap != hipStreamCaptureStatusNone || prof_on()) {
    *pick = 0;
    return run(0);
  }
  hipEvent_t a, b;
  AJ_HIP(hipEventCreate(&a));
  AJ_HIP(hipEventCreate(&b));
  std::vector<float> ms(ncand, 1e30f);
  std::vector<char> ok(ncand, 0);
  for (int c = 0; c < ncand; ++c) {
    if (int rc = launch(c, st)) {  // warm-up; a failing candidate is skipped
      (void)rc;
      (void)hipGetLastError();
      continue;
    }
    ok[c] = 1;
  }
  for (int round = 0; round < kTuneRounds; ++round) {
    for (int c = 0; c < ncand; ++c) {
      if (!ok[c]) continue;
      AJ_HIP(hipEventRecord(a, st));
      for (int r = 0; r < kTuneReps; ++r)
        if (int rc = launch(c, st)) return rc;
      AJ_HIP(hipEventRecord(b, st));
      AJ_HIP(hipEventSynchronize(b));
      float t = 0.f;
      AJ_HIP(hipEventElapsedTime(&t, a, b));
      ms[c] = std::min(ms[c], t);
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  const int best = (int)(std::min_element(ms.begin(), ms.end()) - ms.begin());
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = {best, ms};
  }
  *pick = best;
  return launch(best, st);  // the winner's outputs are the call's result
}

}  // namespace aj

using namespace aj;

extern "C" int aj_tune_enable(int enabled) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_on = enabled ? 1 : 0;
  return AJ_OK;
}

extern "C" int aj_tune_reset(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  return AJ_OK;
}

extern "C" int aj_tune_load(const char* text, int* count) {
  if (!text) return fail(AJ_ERR_INVALID, "aj_tune_load: no text");
  if (count) *count = 0;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  int n = 0;
  for (const char* line = text; *line;) {
    const char* end = std::strchr(line, '\n');
    const std::string l(line, end ? (size_t)(end - line) : std::strlen(line));
    line = end ? end + 1 : line + l.size();
    const size_t arrow = l.find(" -> ");
    if (arrow == std::string::npos || arrow == 0) continue;
    char* stop = nullptr;
    const long pick = std::strtol(l.c_str() + arrow + 4, &stop, 10);
    if (stop == l.c_str() + arrow + 4 || pick < 0 || pick > 64)
      return fail(AJ_ERR_INVALID, "aj_tune_load: bad decision line");
    g_tune[l.substr(0, arrow)] = {(int)pick, {}};
    if (count) *count = ++n;
  }
  return AJ_OK;
}

extern "C" int aj_tune_report(char* buf, int len) {
  if (!buf || len < 1) return fail(AJ_ERR_INVALID, "aj_tune_report: no buffer");
  std::string s;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& kv : g_tune) {
      s += kv.first + " -> " + std::to_string(kv.second.first) + " [";
      for (size_t c = 0; c < kv.second.second.size(); ++c) {
        char t[32];
        std::snprintf(t, sizeof(t), "%s%.4f", c ? " " : "", kv.second.second[c] / kTuneReps);
        s += t;
      }
      s += " ms]\n";
    }
  }
  std::strncpy(buf, s.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return (int)std::min<size_t>(s.size(), (size_t)len - 1);
}
