// timing.h -- optional per-kernel HIP-event timing (off by default).
// When enabled, each instrumented launch records an event pair on the stream
// it runs on; impeller_timing_read() synchronises them and accumulates the
// elapsed time per kernel id.
#pragma once
#include <hip/hip_runtime.h>

namespace imp {

enum KernelId {
  K_NTT_BITREV = 0,
  K_NTT_DIT = 1,
  K_MSM_DIGITS = 2,
  K_MSM_SCAN = 3,  // bucket bounds + task scan
  K_MSM_BUCKETS = 5,
  K_MSM_REDUCE = 6,
  K_MSM_SUM = 7,
  K_MSM_FINAL = 8,
  K_GATE = 9,
  K_MSM_SORT = 10,
  K_MSM_COMBINE = 11,
  K_COUNT = 16
};

bool timing_on();
hipEvent_t timing_start(hipStream_t s);
void timing_stop(int kid, hipStream_t s, hipEvent_t start);

// RAII helper: IMP_TIMED(kid, stream) { launch; }
struct TimedScope {
  int kid;
  hipStream_t s;
  hipEvent_t ev;
  TimedScope(int k, hipStream_t st) : kid(k), s(st), ev(nullptr) {
    if (timing_on()) ev = timing_start(st);
  }
  ~TimedScope() {
    if (ev) timing_stop(kid, s, ev);
  }
};

}  // namespace imp
