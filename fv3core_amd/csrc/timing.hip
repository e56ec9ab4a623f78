// Launch timing of one named kernel with HIP events recorded on the stream
// the kernel is launched on.  A benchmark arms a kernel name, runs its timed
// region, then reads the count and summed duration of the armed launches;
// launches of other kernels record nothing.
#include <mutex>
#include <string>
#include <vector>

#include "timing.hpp"

namespace fv3 {
namespace {

struct Pair {
  hipEvent_t start, stop;
  double bytes;
};

std::mutex g_mu;
std::string g_armed;
bool g_any_armed = false;
bool g_held = false;
std::vector<Pair> g_pool;
size_t g_used = 0;

}  // namespace

TimedLaunch::TimedLaunch(const char* name, hipStream_t s, double bytes) : stream_(s) {
  if (!g_any_armed) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_armed != name) return;
  if (g_used == g_pool.size()) {
    Pair p;
    if (hipEventCreate(&p.start) != hipSuccess) return;
    if (hipEventCreate(&p.stop) != hipSuccess) {
      (void)hipEventDestroy(p.start);
      return;
    }
    g_pool.push_back(p);
  }
  slot_ = (long)g_used++;
  g_pool[slot_].bytes = bytes;
  (void)hipEventRecord(g_pool[slot_].start, stream_);
}

TimedLaunch::~TimedLaunch() {
  if (slot_ < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  (void)hipEventRecord(g_pool[slot_].stop, stream_);
}

}  // namespace fv3

extern "C" {

int32_t fv3_timing_arm(const char* kernel_name) {
  using namespace fv3;
  std::lock_guard<std::mutex> lk(g_mu);
  g_armed = kernel_name ? kernel_name : "";
  g_held = false;
  g_any_armed = !g_armed.empty();
  g_used = 0;
  return FV3_OK;
}

int32_t fv3_timing_hold(int32_t hold) {
  using namespace fv3;
  std::lock_guard<std::mutex> lk(g_mu);
  g_held = hold != 0;
  g_any_armed = !g_armed.empty() && !g_held;
  return FV3_OK;
}

int32_t fv3_timing_read(int64_t* count, double* total_ms, double* total_bytes) {
  using namespace fv3;
  FV3_REQUIRE(count && total_ms && total_bytes, FV3_E_ARG, "fv3_timing_read: NULL pointer");
  std::lock_guard<std::mutex> lk(g_mu);
  double sum = 0.0, bytes = 0.0;
  for (size_t n = 0; n < g_used; ++n) {
    if (hipEventSynchronize(g_pool[n].stop) != hipSuccess) {
      set_error("fv3_timing_read: event synchronize failed");
      return FV3_E_HIP;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_pool[n].start, g_pool[n].stop) != hipSuccess) {
      set_error("fv3_timing_read: elapsed time failed");
      return FV3_E_HIP;
    }
    sum += ms;
    bytes += g_pool[n].bytes;
  }
  *count = (int64_t)g_used;
  *total_ms = sum;
  *total_bytes = bytes;
  g_used = 0;
  return FV3_OK;
}

}  // extern "C"
