// Scoped HIP-event bracket around a kernel launch, active only while the
// kernel's name is armed with fv3_timing_arm.  `bytes` is the launch's
// algorithmic HBM traffic (each input read once, each output written once).
#pragma once

#include "common.hpp"

namespace fv3 {

class TimedLaunch {
 public:
  TimedLaunch(const char* name, hipStream_t s, double bytes);
  ~TimedLaunch();
  TimedLaunch(const TimedLaunch&) = delete;
  TimedLaunch& operator=(const TimedLaunch&) = delete;

 private:
  hipStream_t stream_;
  long slot_ = -1;
};

}  // namespace fv3
