// The dynamical-core step's wrap-up: AdjustNegativeTracerMixingRatio
// (fv3core/stencils/neg_adj3.py:279-372) and CubedToLatLon
// (fv3core/stencils/c2l_ord.py:74-141).
//
// neg_adj3 is four column stencils (fix_neg_water, two fillq calls,
// fix_water_vapor_down, fix_neg_cloud) that the reference runs one after the
// other over the compute box.  Every one of them only moves mass within a
// column, so one kernel with one lane per column does all of them in three
// walks over k:
//   A (top -> bottom): fix_neg_water at each level, the positive sums of
//     both fillq calls, the forward sweep of fix_water_vapor_down (its
//     upper_fix correction of the level above is applied as soon as the
//     level above is no longer read), and the forward sweep of
//     fix_neg_cloud with its bottom two levels;
//   B (bottom -> top): the first backward sweep of fillq for graupel and
//     rain, and the bottom-level vapor borrow of fix_water_vapor_down;
//   C (bottom -> top): the second backward sweep of fillq for both.
// Each statement keeps the reference's operand order, so the results equal
// the stencil sequence bit for bit.
//
// CubedToLatLon is one pointwise kernel on the compute box (4th order, with
// the 2nd-order tile-edge rows and columns) or on the box grown by one
// (2nd order); u and v arrive with their halos filled.
#include "common.hpp"
#include "kernels.hpp"
#include "stencil.hpp"

namespace fv3 {
namespace {

// global_constants.py and neg_adj3.py:315-320 (nonhydrostatic).
constexpr double kRvgas = 461.50;
constexpr double kCpAir = 1004.6;
constexpr double kRdgas = 287.05;
constexpr double kCvAir = kCpAir - kRdgas;
constexpr double kCvVap = 3.0 * kRvgas;
constexpr double kCIce = 1972.0;
constexpr double kCLiq = 4.1855e3;
constexpr double kHlv = 2.5e6;
constexpr double kHlf = 3.3358e5;
constexpr double kTice = 273.16;
constexpr double kDcIce = kCLiq - kCIce;
constexpr double kLi0 = kHlf - kDcIce * kTice;
constexpr double kD0Vap = kCvVap - kCLiq;
constexpr double kLv00 = kHlv - kD0Vap * kTice;

// `a if a < b else b`
__device__ __forceinline__ double lt_pick(double a, double b) { return a < b ? a : b; }

struct Water {
  double qv, ql, qr, qs, qi, qg, pt;
};

// fix_neg_water with fix_negative_ice and fix_negative_liq
// (neg_adj3.py:14-150) at one point.
__device__ __forceinline__ void fix_neg_water_point(Water& w) {
  const double q_liq = 0.0 > w.ql + w.qr ? 0.0 : w.ql + w.qr;
  const double q_sol = 0.0 > w.qi + w.qs ? 0.0 : w.qi + w.qs;
  const double cpm = (1.0 - (w.qv + q_liq + q_sol)) * kCvAir + w.qv * kCvVap +
                     q_liq * kCLiq + q_sol * kCIce;
  const double lcpk = (kLv00 + kD0Vap * w.pt) / cpm;
  const double icpk = (kLi0 + kDcIce * w.pt) / cpm;
  double dq;
  // fix_negative_ice
  {
    const double qsum = w.qi + w.qs;
    if (qsum > 0.0) {
      if (w.qi < 0.0) {
        w.qi = 0.0;
        w.qs = qsum;
      } else if (w.qs < 0.0) {
        w.qs = 0.0;
        w.qi = qsum;
      }
    } else {
      w.qi = 0.0;
      w.qs = 0.0;
      w.qg = w.qg + qsum;
    }
    if (w.qg < 0.0) {
      dq = lt_pick(w.qs, -w.qg);
      w.qs = w.qs - dq;
      w.qg = w.qg + dq;
      if (w.qg < 0.0) {
        dq = lt_pick(w.qi, -w.qg);
        w.qi = w.qi - dq;
        w.qg = w.qg + dq;
      }
    }
    // borrowing from rain takes the amount out of qliquid, as the stencil does
    if (w.qg < 0.0 && w.qr > 0.0) {
      dq = lt_pick(w.qr, -w.qg);
      w.qg = w.qg + dq;
      w.ql = w.ql - dq;
      w.pt = w.pt + dq * icpk;
    }
    if (w.qg < 0.0 && w.ql > 0.0) {
      dq = lt_pick(w.ql, -w.qg);
      w.qg = w.qg + dq;
      w.ql = w.ql - dq;
      w.pt = w.pt + dq * icpk;
    }
    if (w.qg < 0.0 && w.qv > 0.0) {
      dq = lt_pick(0.999 * w.qv, -w.qg);
      w.qg = w.qg + dq;
      w.qv = w.qv - dq;
      w.pt = w.pt + dq * (icpk + lcpk);
    }
  }
  // fix_negative_liq
  {
    const double qsum = w.ql + w.qr;
    const double pos_qg = 0.0 > w.qg ? 0.0 : w.qg;
    if (qsum > 0.0) {
      if (w.qr < 0.0) {
        w.qr = 0.0;
        w.ql = qsum;
      } else if (w.ql < 0.0) {
        w.ql = 0.0;
        w.qr = qsum;
      }
    } else {
      w.ql = 0.0;
      double qrain_tmp = qsum;
      dq = lt_pick(pos_qg, -qrain_tmp);
      qrain_tmp = qrain_tmp + dq;
      w.qg = w.qg - dq;
      w.pt = w.pt - dq * icpk;
      if (w.qr < 0.0) {
        dq = lt_pick(w.qi + w.qs, -qrain_tmp);
        qrain_tmp = qrain_tmp + dq;
        const double dq1 = lt_pick(dq, w.qs);
        w.qs = w.qs - dq1;
        w.qi = w.qi + dq1 - dq;
        w.pt = w.pt - dq * icpk;
      }
      w.qr = qrain_tmp;
      if (w.qr < 0.0 && w.qv > 0.0) {
        dq = lt_pick(0.999 * w.qv, -w.qr);
        w.qv = w.qv - dq;
        w.qr = w.qr + dq;
        w.pt = w.pt + dq * lcpk;
      }
    }
  }
}

struct NegAdjFields {
  double *qv, *ql, *qr, *qs, *qi, *qg, *qcld, *pt;
  const double* dp;
};

__global__ void __launch_bounds__(256) neg_adj3_kernel(Dom d, NegAdjFields f) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni;
  const int64_t c = at3(d, i, j, 0, (int)blockIdx.y);
  const int64_t sk = d.sk;
  const int nz = d.nk;
  double* __restrict__ qv = f.qv + c;
  double* __restrict__ ql = f.ql + c;
  double* __restrict__ qr = f.qr + c;
  double* __restrict__ qs = f.qs + c;
  double* __restrict__ qi = f.qi + c;
  double* __restrict__ qg = f.qg + c;
  double* __restrict__ qc = f.qcld + c;
  double* __restrict__ pt = f.pt + c;
  const double* __restrict__ dp = f.dp + c;

  // ---- walk A: top -> bottom
  double sum1g = 0.0, sum1r = 0.0;
  double qv_prev = 0.0, dp_prev = 0.0, lower_fix_prev = 0.0;
  double qc_prev = 0.0;  // fix_neg_cloud's forward value of the level above
  for (int k = 0; k < nz; ++k) {
    const int64_t o = (int64_t)k * sk;
    Water w{qv[o], ql[o], qr[o], qs[o], qi[o], qg[o], pt[o]};
    const double dpk = dp[o];
    const double dp_up = dp_prev;  // dp of the level above
    fix_neg_water_point(w);
    ql[o] = w.ql;
    qr[o] = w.qr;
    qs[o] = w.qs;
    qi[o] = w.qi;
    qg[o] = w.qg;
    pt[o] = w.pt;
    // fillq, forward sums (neg_adj3.py:105-107)
    if (w.qg > 0) sum1g = sum1g + w.qg * dpk;
    if (w.qr > 0) sum1r = sum1r + w.qr * dpk;
    // fix_water_vapor_down, interval(0, 1) / interval(1, 2) and the forward
    // sweep over interval(1, -1) (neg_adj3.py:197-220)
    double q = w.qv;
    if (k == 0) {
      if (q < 0.0) q = 0.0;
    } else {
      if (k == 1 && qv_prev < 0) q = q + qv_prev * dp_up / dpk;
      if (k < nz - 1) {
        double dq = qv_prev * dp_up;
        if (lower_fix_prev != 0) q += lower_fix_prev / dpk;
        double upper_fix = 0.0;
        if ((q < 0) && (qv_prev > 0)) {
          dq = lt_pick(dq, -q * dpk);
          upper_fix = dq;
          q += dq / dpk;
        }
        double lower_fix = 0.0;
        if (q < 0) {
          lower_fix = q * dpk;
          q = 0;
        }
        // interval(0, -2): the level above gives back what this level took
        const int64_t om = (int64_t)(k - 1) * sk;
        qv[om] = upper_fix != 0 ? qv_prev - upper_fix / dp_up : qv_prev;
        lower_fix_prev = lower_fix;
      } else {
        // bottom level: interval(-1, None) (its lower_fix is never written)
        qv[(int64_t)(k - 1) * sk] = qv_prev;
        if (lower_fix_prev > 0) q = q + 0.0 / dpk;
      }
    }
    qv_prev = q;
    // fix_neg_cloud, forward sweep over interval(1, -1) (neg_adj3.py:154-159)
    double qck = qc[o];
    if (k >= 1 && k < nz - 1) {
      if (qc_prev < 0.0) qck = qck + qc_prev * dp_up / dpk;
      qc_prev = qck;
      const double out = qck < 0.0 ? 0.0 : qck;  // the PARALLEL clamp over interval(1, -1)
      if (out != qc[o]) qc[o] = out;
    } else if (k == 0) {
      qc_prev = qck;
    }
    dp_prev = dpk;
  }
  const int kb = nz - 1;
  const int64_t ob = (int64_t)kb * sk;
  qv[ob] = qv_prev;
  // fix_neg_cloud, bottom two levels (neg_adj3.py:160-174)
  {
    const int64_t o2 = (int64_t)(kb - 1) * sk;
    const double dp2 = dp[o2], dpb = dp[ob];
    double q2 = qc[o2];
    const double qb0 = qc[ob];
    if (qb0 < 0.0 && q2 > 0) {
      const double dq = lt_pick(-q2 * dp2, qb0 * dpb);
      q2 = q2 - dq / dp2;
      qc[o2] = q2;
    }
    if (qb0 < 0 && q2 > 0.0) {
      const double dq = lt_pick(-qb0 * dpb, q2 * dp2);
      double qb = qb0 + dq / dpb;
      qb = 0.0 > qb ? 0.0 : qb;
      qc[ob] = qb;
    }
  }

  // ---- walk B: bottom -> top.  fillq's first backward sweep
  // (neg_adj3.py:108-113) and fix_water_vapor_down's bottom borrow
  // (neg_adj3.py:239-271, dp_bot = the bottom layer's dp).
  double sum2g = 0.0, sum2r = 0.0;
  const double dp_bot = dp[ob];
  double uf = qv[ob];
  for (int k = kb; k >= 0; --k) {
    const int64_t o = (int64_t)k * sk;
    const double dpk = dp[o];
    const double g = qg[o];
    if (g < 0.0 && sum1g >= 0) {
      const double dq = lt_pick(sum1g, -g * dpk);
      sum1g = sum1g - dq;
      sum2g = sum2g + dq;
      qg[o] = g + dq / dpk;
    }
    const double r = qr[o];
    if (r < 0.0 && sum1r >= 0) {
      const double dq = lt_pick(sum1r, -r * dpk);
      sum1r = sum1r - dq;
      sum2r = sum2r + dq;
      qr[o] = r + dq / dpk;
    }
    if (k < kb && uf < 0) {
      const double v = qv[o];
      if (v > 0) {
        double dq = v * dpk;
        if (dq >= -uf * dp_bot) dq = -uf * dp_bot;
        qv[o] = v - dq / dpk;
        uf = uf + dq / dp_bot;
      }
    }
  }
  qv[ob] = uf;

  // ---- walk C: bottom -> top.  fillq's second backward sweep
  // (neg_adj3.py:114-119).
  const bool cg = sum1g >= 1e-12 && sum2g > 0, cr = sum1r >= 1e-12 && sum2r > 0;
  if (!cg && !cr) return;
  for (int k = kb; k >= 0; --k) {
    const int64_t o = (int64_t)k * sk;
    const double dpk = dp[o];
    if (sum1g >= 1e-12 && sum2g > 0) {
      const double g = qg[o];
      if (g > 0.0) {
        const double dq = lt_pick(sum2g, g * dpk);
        sum2g = sum2g - dq;
        qg[o] = g - dq / dpk;
      }
    }
    if (sum1r >= 1e-12 && sum2r > 0) {
      const double r = qr[o];
      if (r > 0.0) {
        const double dq = lt_pick(sum2r, r * dpk);
        sum2r = sum2r - dq;
        qr[o] = r - dq / dpk;
      }
    }
  }
}

}  // namespace
}  // namespace fv3

extern "C" {

int32_t fv3_neg_adj3_f64(const fv3_domain* dom, double* qvapor, double* qliquid, double* qrain,
                         double* qsnow, double* qice, double* qgraupel, double* qcld, double* pt,
                         const double* delp, void* stream) {
  using namespace fv3;
  const char* who = "fv3_neg_adj3";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(qvapor && qliquid && qrain && qsnow && qice && qgraupel && qcld && pt && delp,
              FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(dom->nk >= 3, FV3_E_ARG, "%s: need at least 3 layers (got %d)", who, dom->nk);
  const int ni = dom->iec - dom->isc + 1, nj = dom->jec - dom->jsc + 1;
  if (ni <= 0 || nj <= 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  NegAdjFields f{qvapor, qliquid, qrain, qsnow, qice, qgraupel, qcld, pt, delp};
  dim3 grid((unsigned)((ni * nj + 255) / 256), (unsigned)d.ntiles);
  hipLaunchKernelGGL(neg_adj3_kernel, grid, dim3(256), 0, as_stream(stream), d, f);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_cubed_to_latlon_f64(const fv3_domain* dom, int32_t order, const double* u,
                                const double* v, const double* dx, const double* dy,
                                const double* a11, const double* a12, const double* a21,
                                const double* a22, double* ua, double* va, void* stream) {
  using namespace fv3;
  const char* who = "fv3_cubed_to_latlon";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(u && v && dx && dy && a11 && a12 && a21 && a22 && ua && va, FV3_E_ARG,
              "%s: NULL pointer", who);
  FV3_REQUIRE(order == 2 || order == 4, FV3_E_ARG, "%s: order must be 2 or 4 (got %d)", who,
              order);
  FV3_REQUIRE(dom->isc >= 2 && dom->jsc >= 2 && dom->iec + 3 <= dom->ni &&
                  dom->jec + 3 <= dom->nj,
              FV3_E_ARG, "%s: the stencil needs two halo points on each side", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const hipStream_t s = as_stream(stream);
  constexpr double c1 = 1.125, c2 = -0.125;  // c2l_ord.py:13-14
  if (order == 4) {
    const Box b{d.isc, d.iec, d.jsc, d.jec};
    launch_box(s, d, b, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      const int64_t o2 = at2(d, i, j, t);
      const int64_t sj = d.sj, sj2 = d.sj2;
      double utmp = c2 * (u[o - sj] + u[o + 2 * sj]) + c1 * (u[o] + u[o + sj]);
      double vtmp = c2 * (v[o - 1] + v[o + 2]) + c1 * (v[o] + v[o + 1]);
      const bool sn = j == d.js || j == d.je, we = i == d.is || i == d.ie;
      if (sn || we) {
        // 2nd-order rows and columns on the tile edges (c2l_ord.py:58-66)
        vtmp = 2.0 * ((v[o] * dy[o2]) + (v[o + 1] * dy[o2 + 1])) / (dy[o2] + dy[o2 + 1]);
        utmp = 2.0 * (u[o] * dx[o2] + u[o + sj] * dx[o2 + sj2]) / (dx[o2] + dx[o2 + sj2]);
      }
      ua[o] = a11[o2] * utmp + a12[o2] * vtmp;
      va[o] = a21[o2] * utmp + a22[o2] * vtmp;
    });
  } else {
    const Box b{d.isc - 1, d.iec + 1, d.jsc - 1, d.jec + 1};
    launch_box(s, d, b, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      const int64_t o2 = at2(d, i, j, t);
      const int64_t sj = d.sj, sj2 = d.sj2;
      const double wu = u[o] * dx[o2], wu1 = u[o + sj] * dx[o2 + sj2];
      const double wv = v[o] * dy[o2], wv1 = v[o + 1] * dy[o2 + 1];
      const double u1 = 2.0 * (wu + wu1) / (dx[o2] + dx[o2 + sj2]);
      const double v1 = 2.0 * (wv + wv1) / (dy[o2] + dy[o2 + 1]);
      ua[o] = a11[o2] * u1 + a12[o2] * v1;
      va[o] = a21[o2] * u1 + a22[o2] * v1;
    });
  }
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // extern "C"
