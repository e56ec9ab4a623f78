// Element-wise conversions between the FP32 storages of the mixed-precision
// acoustic step and the FP64 copies its FP64 solvers read (the Riemann
// solvers, the height updates on the C and D grids, the pressure halos).
// No counterpart in the reference, which computes in FP64 only.  A storage
// of the library is one dense block, so a conversion is a flat pass over
// its n elements: coalesced loads and stores, four elements per lane.
#include "common.hpp"

namespace fv3 {

namespace {

template <typename A, typename B>
__global__ void __launch_bounds__(256) cast_kernel(int64_t n, const A* __restrict__ a,
                                                   B* __restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * 256 * 4;
  for (int64_t base = ((int64_t)blockIdx.x * 256) * 4 + threadIdx.x; base < n; base += stride) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t i = base + (int64_t)u * 256;
      if (i < n) b[i] = (B)a[i];
    }
  }
}

template <typename A, typename B>
int cast_launch(int64_t n, const A* a, B* b, void* stream, const char* who) {
  FV3_REQUIRE(n >= 0, FV3_E_ARG, "%s: negative length", who);
  if (n == 0) return FV3_OK;
  FV3_REQUIRE(a && b, FV3_E_ARG, "%s: NULL pointer", who);
  int64_t blocks = (n + 1023) / 1024;
  if (blocks > 8192) blocks = 8192;  // grid-stride beyond 32 waves per CU
  hipLaunchKernelGGL((cast_kernel<A, B>), dim3((unsigned)blocks), dim3(256), 0,
                     as_stream(stream), n, a, b);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // namespace

}  // namespace fv3

extern "C" {

int32_t fv3_cast_f32_to_f64(int64_t n, const float* src, double* dst, void* stream) {
  return fv3::cast_launch<float, double>(n, src, dst, stream, "fv3_cast_f32_to_f64");
}

int32_t fv3_cast_f64_to_f32(int64_t n, const double* src, float* dst, void* stream) {
  return fv3::cast_launch<double, float>(n, src, dst, stream, "fv3_cast_f64_to_f32");
}
}
