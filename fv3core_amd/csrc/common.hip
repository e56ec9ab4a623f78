// Error reporting and argument checks shared by all entry points.
#include <cstdarg>

#include "common.hpp"

namespace fv3 {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

int check_domain(const fv3_domain* d, const char* who) {
  FV3_REQUIRE(d != nullptr, FV3_E_ARG, "%s: domain is NULL", who);
  FV3_REQUIRE(d->ni > 0 && d->nj > 0 && d->nk >= 0 && d->ntiles >= 0, FV3_E_ARG,
              "%s: bad extents ni=%d nj=%d nk=%d ntiles=%d", who, d->ni, d->nj, d->nk,
              d->ntiles);
  FV3_REQUIRE(d->isc >= 3 && d->jsc >= 3, FV3_E_ARG,
              "%s: compute domain must leave a 3-cell halo (isc=%d jsc=%d)", who, d->isc,
              d->jsc);
  FV3_REQUIRE(d->iec >= d->isc && d->jec >= d->jsc, FV3_E_ARG,
              "%s: empty compute domain", who);
  FV3_REQUIRE(d->iec + 4 < d->ni && d->jec + 4 < d->nj, FV3_E_ARG,
              "%s: storage (ni=%d nj=%d) too small for iec=%d jec=%d plus halo", who, d->ni,
              d->nj, d->iec, d->jec);
  FV3_REQUIRE(d->sj >= d->ni && d->sk >= d->sj * d->nj, FV3_E_ARG,
              "%s: 3-D strides overlap (sj=%lld sk=%lld)", who, (long long)d->sj,
              (long long)d->sk);
  FV3_REQUIRE(d->sj2 >= d->ni, FV3_E_ARG, "%s: IJ stride too small", who);
  return FV3_OK;
}

}  // namespace fv3

extern "C" {

const char* fv3_version(void) { return "fv3core_amd 0.1 gfx950"; }

const char* fv3_last_error(void) { return fv3::g_err; }

int32_t fv3_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}
}
