// timing.hip -- per-kernel HIP-event timing registry (see timing.h).
#include <atomic>
#include <mutex>
#include <vector>

#include "common.h"
#include "timing.h"

namespace imp {

static std::atomic<int> g_on{0};
static std::mutex g_mu;
struct Rec {
  int kid;
  hipEvent_t a, b;
};
static std::vector<Rec> g_pending;
static double g_total_ms[K_COUNT];
static unsigned long long g_count[K_COUNT];

bool timing_on() { return g_on.load(std::memory_order_relaxed) != 0; }

hipEvent_t timing_start(hipStream_t s) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  hipEventRecord(e, s);
  return e;
}

void timing_stop(int kid, hipStream_t s, hipEvent_t start) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return;
  hipEventRecord(e, s);
  std::lock_guard<std::mutex> lk(g_mu);
  g_pending.push_back(Rec{kid, start, e});
}

static void drain() {
  for (auto& r : g_pending) {
    hipEventSynchronize(r.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, r.a, r.b);
    if (r.kid >= 0 && r.kid < K_COUNT) {
      g_total_ms[r.kid] += ms;
      g_count[r.kid] += 1;
    }
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  g_pending.clear();
}

}  // namespace imp

using namespace imp;

extern "C" void impeller_timing_enable(int on) { g_on.store(on ? 1 : 0); }

extern "C" void impeller_timing_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  drain();
  for (int i = 0; i < K_COUNT; i++) {
    g_total_ms[i] = 0;
    g_count[i] = 0;
  }
}

extern "C" int impeller_timing_read(int kid, double* total_ms, unsigned long long* count) {
  if (kid < 0 || kid >= K_COUNT || !total_ms || !count) return (int)hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(g_mu);
  drain();
  *total_ms = g_total_ms[kid];
  *count = g_count[kid];
  return 0;
}
