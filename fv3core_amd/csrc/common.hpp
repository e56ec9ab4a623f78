// Shared helpers of the gfx950 kernels: error reporting across the C ABI,
// domain checks and small device utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/fv3core_amd.h"

namespace fv3 {

// Thread-local message of the last failed call (read by fv3_last_error).
void set_error(const char* fmt, ...);
void clear_error();

#define FV3_REQUIRE(cond, code, ...)      \
  do {                                    \
    if (!(cond)) {                        \
      ::fv3::set_error(__VA_ARGS__);      \
      return (code);                      \
    }                                     \
  } while (0)

#define FV3_HIP_CHECK_LAUNCH(name)                                               \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess) {                                                      \
      ::fv3::set_error("%s: launch failed: %s", name, hipGetErrorString(_e));    \
      return FV3_E_HIP;                                                          \
    }                                                                            \
  } while (0)

// Host-side validation shared by every entry point.
int check_domain(const fv3_domain* d, const char* who);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Kernel-side copy of the domain (plain values, passed by value).
struct Dom {
  int isc, iec, jsc, jec;
  int is, ie, js, je;  // edge anchors
  int ni, nj, nk, ntiles;
  int64_t sj, sk, st, sj2, st2;
};

inline Dom to_dom(const fv3_domain* d) {
  Dom o;
  o.isc = d->isc; o.iec = d->iec; o.jsc = d->jsc; o.jec = d->jec;
  o.is = d->i_start; o.ie = d->i_end; o.js = d->j_start; o.je = d->j_end;
  o.ni = d->ni; o.nj = d->nj; o.nk = d->nk; o.ntiles = d->ntiles;
  o.sj = d->sj; o.sk = d->sk; o.st = d->st; o.sj2 = d->sj2; o.st2 = d->st2;
  return o;
}

// Corner-copy destination/source pairs resolved to absolute indices
// (fv3core/utils/corners.py:305-540); dir 0 = x-differentiable, 1 = y.
struct CornerTable {
  int n;
  int dst_i[72], dst_j[72], src_i[72], src_j[72];
  unsigned char dir[72];
};
// which: bit 0 = x table, bit 1 = y table.
int build_corner_table(const fv3_domain* d, int which, CornerTable* tab);

// The reference's sign(a, b): |a| if b > 0 else -|a|
// (fv3core/stencils/basic_operations.py:33-40).
template <typename T>
__device__ __forceinline__ T sgn(T a, T b) {
  T aa = fabs(a);
  return b > T(0) ? aa : -aa;
}

// Number of XCDs on MI355X: consecutive workgroups are dispatched round-robin
// over them, each with its own L2.
constexpr int kNumXcd = 8;
constexpr int kSimdsPerGpu = 256 * 4;  // 256 CUs of 4 SIMDs

// Workgroup id -> logical unit so that each XCD runs a contiguous range of
// logical units (whole planes of one level, consecutive levels of one
// tile), keeping what neighbouring units share (halo rows, a tile's metric
// terms) in that XCD's L2.  A launch has a multiple of kNumXcd workgroups;
// ids mapped at or past `total` have no unit.
__device__ __forceinline__ int xcd_logical_block(int hw, int total) {
  const int per = (total + kNumXcd - 1) / kNumXcd;
  return (hw % kNumXcd) * per + hw / kNumXcd;
}

}  // namespace fv3
