// Vertical remap of one scalar: RemapProfile (cs_profile) + MapSingle
// (fv3core/stencils/remap_profile.py:119-638, map_single.py:12-173).  The
// column kernel lives in remap_profile.hpp; this file holds the launcher and
// the C entry points.
#include "common.hpp"
#include "remap_profile.hpp"
#include "timing.hpp"

#include <mutex>

namespace fv3 {

constexpr int64_t kFusedRemapWaves = 1024;

template <typename T>
static int map_single_launch(const fv3_domain* dom, int kord, int iv, int i1, int i2, int j1,
                             int j2, T* q1, const T* pe1, const T* pe2, const T* qs, double qmin,
                             T* work, int64_t work_slab, void* stream) {
  const char* who = "fv3_map_single";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q1 && pe1 && pe2 && work, FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(kord <= 10, FV3_E_UNSUPPORTED, "%s: kord %d not implemented", who, kord);
  const int akord = kord < 0 ? -kord : kord;
  // The reference also accepts kord <= -11: for 10 < |kord| <= 16 it runs no
  // inner-level limiter and above 16 no constraints at all.  Those forms are
  // not built; they are rejected instead of running the kord-10 limiter.
  FV3_REQUIRE(akord <= 10, FV3_E_UNSUPPORTED,
              "%s: |kord| %d > 10 not implemented (only kord in -10..10)", who, akord);
  const int km = dom->nk;
  FV3_REQUIRE(km >= 6, FV3_E_ARG, "%s: need at least 6 layers (got %d)", who, km);
  FV3_REQUIRE(i1 >= 0 && j1 >= 0 && i2 < dom->ni && j2 < dom->nj && i2 >= i1 - 1 &&
                  j2 >= j1 - 1,
              FV3_E_ARG, "%s: box outside the storage", who);
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + dom->sk * (km + 1),
              FV3_E_ARG, "%s: workspace slab too small", who);
  const int nib = i2 - i1 + 1, njb = j2 - j1 + 1;
  if (nib <= 0 || njb <= 0 || dom->ntiles == 0) return FV3_OK;
  RemapArgs a{akord, iv, km, i1, j1, nib, njb, qmin, work_slab};
  Dom d = to_dom(dom);
  dim3 grid((unsigned)(((int64_t)nib * njb + 255) / 256), d.ntiles);
  // q1, pe1, pe2 read and q1 written once per cell-level of the box
  TimedLaunch timed("map_single_kernel", as_stream(stream),
                    4.0 * sizeof(T) * (double)nib * njb * km * d.ntiles);
  // Below one wave of columns per SIMD (1024 on MI355X) a column's dependent
  // chain sets the time and the four-sweep form, with more loads in flight,
  // is faster (one C96 tile: ~165 vs ~210 us); above it the fused form moves
  // fewer bytes (six C128 tiles: 315 vs 354 us).  Both give the same bits.
  const int64_t waves = ((int64_t)nib * njb + 63) / 64 * d.ntiles;
  if (waves >= kFusedRemapWaves)
    hipLaunchKernelGGL((map_single_kernel<T, true>), grid, dim3(256), 0, as_stream(stream), d, a,
                       q1, pe1, pe2, qs, work);
  else
    hipLaunchKernelGGL((map_single_kernel<T, false>), grid, dim3(256), 0, as_stream(stream), d, a,
                       q1, pe1, pe2, qs, work);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

// ---- scratch of the entry point that allocates nothing
// The six column slabs of map_single_kernel live in one file-scope device
// array; every GPU the library runs on holds its own copy until the process
// ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in): it uses the device address hipGetSymbolAddress gives
// for the current device, kept per device.
constexpr size_t kNoAllocScratchBytes = size_t(96) << 20;
constexpr int kRemapSlabs = 6;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static int scratch_on_current_device(void** out, const char* who) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  FV3_REQUIRE(e == hipSuccess, FV3_E_HIP, "%s: hipGetDevice failed: %s", who,
              hipGetErrorString(e));
  FV3_REQUIRE(dev >= 0 && dev < kMaxDevices, FV3_E_HIP, "%s: device %d out of range", who, dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    FV3_REQUIRE(e == hipSuccess && p != nullptr, FV3_E_HIP,
                "%s: hipGetSymbolAddress(scratch) failed on device %d: %s", who, dev,
                hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FV3_OK;
}

static int map_single_noalloc(const fv3_domain* dom, int kord, int iv, int i1, int i2, int j1,
                              int j2, double* q1, const double* pe1, const double* pe2,
                              const double* qs, double qmin, void* stream) {
  const char* who = "fv3_map_single_noalloc_f64";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(dom->st >= 0, FV3_E_ARG, "%s: negative tile stride", who);
  const int64_t tiles_before = dom->ntiles > 0 ? dom->ntiles - 1 : 0;
  // slab length in doubles, rounded to 256 bytes so every slab stays aligned
  const int64_t slab = (tiles_before * dom->st + dom->sk * (int64_t)(dom->nk + 1) + 31) / 32 * 32;
  const int64_t max_slab = (int64_t)(kNoAllocScratchBytes / sizeof(double) / kRemapSlabs);
  FV3_REQUIRE(slab <= max_slab, FV3_E_ARG,
              "%s: the columns need %lld bytes of scratch (6 slabs of %lld doubles), above the "
              "%zu-byte maximum",
              who, (long long)(slab * kRemapSlabs * (int64_t)sizeof(double)), (long long)slab,
              kNoAllocScratchBytes);
  void* p = nullptr;
  rc = scratch_on_current_device(&p, who);
  if (rc) return rc;
  return map_single_launch<double>(dom, kord, iv, i1, i2, j1, j2, q1, pe1, pe2, qs, qmin,
                                   static_cast<double*>(p), slab, stream);
}

}  // namespace fv3

extern "C" {

int32_t fv3_map_single_noalloc_f64(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                                   int32_t i2, int32_t j1, int32_t j2, double* q1,
                                   const double* pe1, const double* pe2, const double* qs,
                                   double qmin, void* stream) {
  return fv3::map_single_noalloc(d, kord, mode, i1, i2, j1, j2, q1, pe1, pe2, qs, qmin, stream);
}

int32_t fv3_map_single_f64(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                           int32_t i2, int32_t j1, int32_t j2, double* q1, const double* pe1,
                           const double* pe2, const double* qs, double qmin, double* work,
                           int64_t work_slab, void* stream) {
  return fv3::map_single_launch<double>(d, kord, mode, i1, i2, j1, j2, q1, pe1, pe2, qs, qmin,
                                        work, work_slab, stream);
}

int32_t fv3_map_single_f32(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                           int32_t i2, int32_t j1, int32_t j2, float* q1, const float* pe1,
                           const float* pe2, const float* qs, double qmin, float* work,
                           int64_t work_slab, void* stream) {
  return fv3::map_single_launch<float>(d, kord, mode, i1, i2, j1, j2, q1, pe1, pe2, qs, qmin,
                                       work, work_slab, stream);
}
}
