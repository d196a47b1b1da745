// Library-wide ABI helpers: version, thread-local error message, launch checks,
// and the optional per-kernel event timer used by the benchmark.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "slx_common.h"

namespace slx {

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SLX_EHIP, "%s: %s", what, hipGetErrorString(e));
  return SLX_OK;
}

// ---- launch knobs ----
// (first-tier walk stack 24: on the 1M-vertex drape 21.07 ms per frame against
// 21.74 with 32 entries, whose deeper LDS stack costs occupancy; refit subtrees of
// 1024 leaves: two workgroups per CU overlap one's loads with the other's walk,
// 19.49 against 20.0 ms per frame with 2048)
static Knobs g_knobs = {256, 1, 24, 2048, 1024, 2048, 0, nullptr};
static thread_local const Knobs* t_knobs = nullptr;

Knobs& knob_defaults() { return g_knobs; }
const Knobs& knobs() { return t_knobs ? *t_knobs : g_knobs; }

KnobScope::KnobScope(const Knobs& k) : k_(k), prev_(t_knobs) { t_knobs = &k_; }
KnobScope::~KnobScope() { t_knobs = prev_; }

// ---- kernel timer: hipEvent pairs recorded on the launch stream ----
struct TimerRec {
  const char* name;
  hipEvent_t a, b;
};
static std::mutex g_tmu;
static std::vector<hipEvent_t> g_pool;
static std::vector<TimerRec> g_recs;
static size_t g_cap = 0;
static bool g_on = false;
static std::string g_only;        // non-empty: time only the launches of this name

bool timer_on() { return g_on; }

void timer_begin(const char* name, hipStream_t s, int* slot) {
  *slot = -1;
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  if (g_recs.size() >= g_cap) return;
  if (!g_only.empty() && g_only != name) return;
  size_t i = g_recs.size();
  TimerRec r{name, g_pool[2 * i], g_pool[2 * i + 1]};
  if (hipEventRecord(r.a, s) != hipSuccess) return;
  g_recs.push_back(r);
  *slot = (int)i;
}

void timer_end(int slot, hipStream_t s) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  (void)hipEventRecord(g_recs[(size_t)slot].b, s);
}

}  // namespace slx

extern "C" int slx_abi_version(void) { return SLX_ABI_VERSION; }
extern "C" const char* slx_last_error(void) { return slx::g_err; }

extern "C" int slx_timer_enable(int max_launches) {
  using namespace slx;
  std::lock_guard<std::mutex> lk(g_tmu);
  SLX_REQUIRE(max_launches >= 0, "slx_timer_enable: max_launches < 0");
  while (g_pool.size() < 2 * (size_t)max_launches) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return fail(SLX_EHIP, "slx_timer_enable: hipEventCreate failed");
    g_pool.push_back(e);
  }
  g_cap = (size_t)max_launches;
  g_recs.clear();
  g_on = max_launches > 0;
  return SLX_OK;
}

extern "C" int slx_timer_only(const char* name) {
  using namespace slx;
  std::lock_guard<std::mutex> lk(g_tmu);
  g_only = name ? name : "";
  return SLX_OK;
}

extern "C" int slx_timer_read(const char* name, double* total_ms, int* count) {
  using namespace slx;
  std::lock_guard<std::mutex> lk(g_tmu);
  SLX_REQUIRE(name && total_ms && count, "slx_timer_read: null pointer");
  double tot = 0.0;
  int n = 0;
  for (auto& r : g_recs) {
    if (strcmp(r.name, name) != 0) continue;
    if (hipEventSynchronize(r.b) != hipSuccess) return fail(SLX_EHIP, "slx_timer_read: sync failed");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return fail(SLX_EHIP, "slx_timer_read: elapsed failed");
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *count = n;
  return SLX_OK;
}

extern "C" int slx_timer_names(char* buf, int buf_len) {
  using namespace slx;
  std::lock_guard<std::mutex> lk(g_tmu);
  SLX_REQUIRE(buf && buf_len > 0, "slx_timer_names: bad buffer");
  std::vector<std::string> seen;
  std::string out;
  for (auto& r : g_recs) {
    bool dup = false;
    for (auto& s : seen) dup |= (s == r.name);
    if (dup) continue;
    seen.emplace_back(r.name);
    if (!out.empty()) out += ",";
    out += r.name;
  }
  snprintf(buf, (size_t)buf_len, "%s", out.c_str());
  return SLX_OK;
}
