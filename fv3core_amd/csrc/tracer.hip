// Tracer sub-cycling helpers of TracerAdvection (tracer_2D_1L).
//
// Reference: fv3core/stencils/tracer_2d_1l.py:19-117: flux_compute (area
// fluxes from the accumulated Courant numbers), cmax_multiply_by_frac
// (divide the accumulated fluxes by the sub-cycle count), dp_fluxadjustment
// and q_adjust.  flux_compute and cmax_multiply_by_frac are fused into one
// pass over the storage; the transport itself is fv3_fvtp2d.
#include "common.hpp"
#include "kernels.hpp"

namespace fv3 {

template <typename T>
__global__ void __launch_bounds__(256)
    tracer_prep_kernel(Dom d, T frac, int do_scale, T* __restrict__ cx, T* __restrict__ cy,
                       T* __restrict__ mfx, T* __restrict__ mfy, T* __restrict__ xfx,
                       T* __restrict__ yfx, const T* __restrict__ dxa,
                       const T* __restrict__ dya, const T* __restrict__ dx,
                       const T* __restrict__ dy, const T* __restrict__ sin_sg1,
                       const T* __restrict__ sin_sg2, const T* __restrict__ sin_sg3,
                       const T* __restrict__ sin_sg4) {
  // domain: full + (1, 1) horizontally, all nk levels
  const int ni = d.iec + 5;  // isd .. ied + 1, with ied = iec + 3
  const int nj = d.jec + 5;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)ni * nj) return;
  const int i = (int)(n % ni), j = (int)(n / ni);
  if (i >= d.ni || j >= d.nj) return;
  const int k = blockIdx.y, t = blockIdx.z;
  const int64_t o = (int64_t)t * d.st + (int64_t)k * d.sk + (int64_t)j * d.sj + i;
  const int64_t o2 = (int64_t)t * d.st2 + (int64_t)j * d.sj2 + i;
  // flux_x on [isc, iec+1] x [jsc-3, jec+3]
  T xv = xfx[o];
  if (i >= d.isc && i <= d.iec + 1 && j >= d.jsc - 3 && j <= d.jec + 3) {
    const T c = cx[o];
    xv = c > T(0) ? c * dxa[o2 - 1] * dy[o2] * sin_sg3[o2 - 1] : c * dxa[o2] * dy[o2] * sin_sg1[o2];
  }
  T yv = yfx[o];
  if (i >= d.isc - 3 && i <= d.iec + 3 && j >= d.jsc && j <= d.jec + 1) {
    const T c = cy[o];
    yv = c > T(0) ? c * dya[o2 - d.sj2] * dx[o2] * sin_sg4[o2 - d.sj2]
                  : c * dya[o2] * dx[o2] * sin_sg2[o2];
  }
  if (do_scale) {
    cx[o] = cx[o] * frac;
    cy[o] = cy[o] * frac;
    mfx[o] = mfx[o] * frac;
    mfy[o] = mfy[o] * frac;
    xv = xv * frac;
    yv = yv * frac;
  }
  xfx[o] = xv;
  yfx[o] = yv;
}

// dp2 = dp1 + (mfx - mfx[i+1] + mfy - mfy[j+1]) * rarea   (mode 0)
// q = (q*dp1 + (fx - fx[i+1] + fy - fy[j+1]) * rarea) / dp2   (mode 1)
template <typename T>
__global__ void __launch_bounds__(256)
    tracer_adjust_kernel(Dom d, int mode, T* __restrict__ out, const T* __restrict__ dp1,
                         const T* __restrict__ fx, const T* __restrict__ fy,
                         const T* __restrict__ rarea, const T* __restrict__ dp2) {
  const int nci = d.iec - d.isc + 1, ncj = d.jec - d.jsc + 1;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)nci * ncj) return;
  const int i = d.isc + (int)(n % nci), j = d.jsc + (int)(n / nci);
  const int k = blockIdx.y, t = blockIdx.z;
  const int64_t o = (int64_t)t * d.st + (int64_t)k * d.sk + (int64_t)j * d.sj + i;
  const T ra = rarea[(int64_t)t * d.st2 + (int64_t)j * d.sj2 + i];
  const T div = (fx[o] - fx[o + 1] + fy[o] - fy[o + d.sj]) * ra;
  if (mode == 0) {
    out[o] = dp1[o] + div;
  } else {
    out[o] = (out[o] * dp1[o] + div) / dp2[o];
  }
}

// q_adjust of a batch of tracers: blockIdx.z = tile + ntiles * slot
struct TracerPtrs {
  double* q[FV3_MAX_TRACERS];
  const double* fx[FV3_MAX_TRACERS];
  const double* fy[FV3_MAX_TRACERS];
};

__global__ void __launch_bounds__(256)
    tracer_adjust_batch_kernel(Dom d, TracerPtrs p, const double* __restrict__ dp1,
                               const double* __restrict__ rarea,
                               const double* __restrict__ dp2) {
  const int nci = d.iec - d.isc + 1, ncj = d.jec - d.jsc + 1;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)nci * ncj) return;
  const int i = d.isc + (int)(n % nci), j = d.jsc + (int)(n / nci);
  const int k = blockIdx.y, t = (int)blockIdx.z % d.ntiles, b = (int)blockIdx.z / d.ntiles;
  double* q = p.q[0];
  const double *fx = p.fx[0], *fy = p.fy[0];
#pragma unroll
  for (int m = 1; m < FV3_MAX_TRACERS; ++m) {
    q = b == m ? p.q[m] : q;
    fx = b == m ? p.fx[m] : fx;
    fy = b == m ? p.fy[m] : fy;
  }
  const int64_t o = (int64_t)t * d.st + (int64_t)k * d.sk + (int64_t)j * d.sj + i;
  const double ra = rarea[(int64_t)t * d.st2 + (int64_t)j * d.sj2 + i];
  const double div = (fx[o] - fx[o + 1] + fy[o] - fy[o + d.sj]) * ra;
  q[o] = (q[o] * dp1[o] + div) / dp2[o];
}

}  // namespace fv3

extern "C" {

int32_t fv3_tracer_transport_f64(const fv3_domain* dom, int32_t hord, int32_t ntracers,
                                 double* const* q, double* const* fx, double* const* fy,
                                 const double* crx, const double* cry, const double* xfx,
                                 const double* yfx, const double* mfx, const double* mfy,
                                 const double* area, const double* dxa, const double* dya,
                                 const double* rarea, const double* dp1, const double* dp2,
                                 void* stream) {
  using namespace fv3;
  const char* who = "fv3_tracer_transport";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(ntracers >= 1 && ntracers <= FV3_MAX_TRACERS, FV3_E_ARG,
              "%s: %d tracers (1 .. %d)", who, ntracers, FV3_MAX_TRACERS);
  FV3_REQUIRE(q && fx && fy && crx && cry && xfx && yfx && mfx && mfy && area && dxa && dya &&
                  rarea && dp1 && dp2,
              FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  rc = fvtp2d_batch_corner_fill_f64(dom, hord, ntracers, q, fx, fy, crx, cry, xfx, yfx, mfx,
                                    mfy, area, dxa, dya, stream);
  if (rc) return rc;
  TracerPtrs p{};
  for (int b = 0; b < ntracers; ++b) {
    p.q[b] = q[b];
    p.fx[b] = fx[b];
    p.fy[b] = fy[b];
  }
  Dom d = to_dom(dom);
  const int64_t ncell = (int64_t)(d.iec - d.isc + 1) * (d.jec - d.jsc + 1);
  dim3 grid((unsigned)((ncell + 255) / 256), d.nk, d.ntiles * ntracers);
  hipLaunchKernelGGL(tracer_adjust_batch_kernel, grid, dim3(256), 0, as_stream(stream), d, p,
                     dp1, rarea, dp2);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_tracer_flux_prep_f64(const fv3_domain* dom, int32_t n_split, double* cxd,
                                 double* cyd, double* mfxd, double* mfyd, double* xfx,
                                 double* yfx, const double* dxa, const double* dya,
                                 const double* dx, const double* dy, const double* sin_sg1,
                                 const double* sin_sg2, const double* sin_sg3,
                                 const double* sin_sg4, void* stream) {
  using namespace fv3;
  const char* who = "fv3_tracer_flux_prep";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(cxd && cyd && mfxd && mfyd && xfx && yfx && dxa && dya && dx && dy && sin_sg1 &&
                  sin_sg2 && sin_sg3 && sin_sg4,
              FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(n_split >= 1, FV3_E_ARG, "%s: n_split must be >= 1", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const int64_t plane = (int64_t)(d.iec + 5) * (d.jec + 5);
  dim3 grid((unsigned)((plane + 255) / 256), d.nk, d.ntiles);
  // the reference multiplies by frac = 1.0 / n_split
  hipLaunchKernelGGL((tracer_prep_kernel<double>), grid, dim3(256), 0, as_stream(stream), d,
                     1.0 / (double)n_split, n_split > 1 ? 1 : 0, cxd, cyd, mfxd, mfyd, xfx, yfx,
                     dxa, dya, dx, dy, sin_sg1, sin_sg2, sin_sg3, sin_sg4);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_tracer_adjust_f64(const fv3_domain* dom, int32_t mode, double* out,
                              const double* dp1, const double* fx, const double* fy,
                              const double* rarea, const double* dp2, void* stream) {
  using namespace fv3;
  const char* who = "fv3_tracer_adjust";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(mode == 0 || mode == 1, FV3_E_ARG, "%s: mode must be 0 or 1", who);
  FV3_REQUIRE(out && dp1 && fx && fy && rarea && (mode == 0 || dp2), FV3_E_ARG,
              "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const int64_t ncell = (int64_t)(d.iec - d.isc + 1) * (d.jec - d.jsc + 1);
  dim3 grid((unsigned)((ncell + 255) / 256), d.nk, d.ntiles);
  hipLaunchKernelGGL((tracer_adjust_kernel<double>), grid, dim3(256), 0, as_stream(stream), d,
                     mode, out, dp1, fx, fy, rarea, dp2);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}
