// Fast saturation adjustment: SatAdjust3d.
//
// Reference: fv3core/stencils/saturation_adjustment.py.  The saturation
// vapour-pressure "tables" (:43-168) are closed forms of the table index
// that the stencil evaluates at each lookup; here fv3_sat_adjust_tables
// evaluates the same closed forms once per index into four device tables
// (tablew, table2 and their differences desw, des2, for the indices -1 ..
// 2620 a lookup can reach), so a lookup is a load and every value is the one
// the closed form gives at that index.  The satadjust stencil (:554-896) is
// pointwise: one thread per cell of [isc, iec] x [jsc, jec] x [kmp, nz),
// its branches in the stencil's statement order, FP64 throughout.
#include "common.hpp"

namespace fv3 {

namespace {

constexpr double kGrav = 9.80665;
constexpr double kRdgas = 287.05;
constexpr double kRvgas = 461.50;
constexpr double kHlv = 2.5e6;
constexpr double kHlf = 3.3358e5;
constexpr double kCpVap = 4.0 * kRvgas;
constexpr double kCIce = 1972.0;
constexpr double kCLiq = 4.1855e3;
constexpr double kTice = 273.16;
constexpr double kRdg = -kRdgas / kGrav;
// saturation_adjustment.py:19-41
constexpr double kDcVap = kCpVap - kCLiq;
constexpr double kDcIce = kCLiq - kCIce;
constexpr double kLv0 = kHlv - kDcVap * kTice;
constexpr double kLi00 = kHlf - kDcIce * kTice;
constexpr double kLi2 = kLv0 + kLi00;
constexpr double kD2Ice = kDcVap + kDcIce;
constexpr double kE00 = 611.21;
constexpr double kTmin = kTice - 160.0;
constexpr double kDelt = 0.1;
constexpr double kTWfr = kTice - 40.0;
constexpr double kTice0 = kTice - 0.01;
constexpr double kLat2 = (kHlv + kHlf) * (kHlv + kHlf);
constexpr int kQsLength = 2621;
// table entries: index -1 .. kQsLength - 1 (entry e = index + 1)
constexpr int kTabN = kQsLength + 1;

__device__ __forceinline__ double q_table_oneline(double dhc, double lhc, double tem) {
  return kE00 * exp((dhc * log(tem / kTice) + (tem - kTice) / (tem * kTice) * lhc) / kRvgas);
}
__device__ __forceinline__ double table_vapor(double tem) {
  return q_table_oneline(kDcVap, kLv0, tem);
}
__device__ __forceinline__ double table_ice(double tem) {
  return q_table_oneline(kD2Ice, kLi2, tem);
}
__device__ __forceinline__ double tem_lower(double i) { return kTmin + kDelt * i; }
__device__ __forceinline__ double tem_upper(double i) { return 253.16 + kDelt * i; }

// qs_table2_fn (:95-129)
__device__ double qs_table2(double i) {
  const double tem0 = tem_lower(i);
  double table2 = i < 1600 ? table_ice(tem0) : table_vapor(tem0);
  if (i == 1599) {
    double table = table_ice(tem0);
    const double tu = tem_upper(i - 1400);
    table = (0.05 * (kTice - tu)) * table + (0.05 * (tu - 253.16)) * table_vapor(tu);
    const double m1 = table_ice(tem_lower(1598));
    const double p1 = table_vapor(tem_lower(1600));
    table2 = 0.25 * (m1 + 2.0 * table + p1);
  }
  if (i == 1600) {
    const double table = table_vapor(tem_upper(i - 1400));
    const double m1 = table_ice(tem_lower(1599));
    const double p1 = table_vapor(tem_lower(1601));
    table2 = 0.25 * (m1 + 2.0 * table + p1);
  }
  return table2;
}

// qs_tablew_fn (:132-135)
__device__ double qs_tablew(double i) { return table_vapor(tem_lower(i)); }

// des2_table / desw_table with des_end (:138-168)
__device__ double des_end(double t, double i, double des) {
  if (i == kQsLength - 1) des = fmax(0.0, t - qs_table2(i - 1));
  return des;
}

__global__ void __launch_bounds__(256) tables_kernel(double* tab) {
  const int e = (int)(blockIdx.x * 256 + threadIdx.x);
  if (e >= kTabN) return;
  const double i = (double)(e - 1);
  const double tw = qs_tablew(i), t2 = qs_table2(i);
  tab[e] = tw;
  tab[kTabN + e] = t2;
  tab[2 * kTabN + e] = des_end(tw, i, fmax(0.0, qs_tablew(i + 1) - tw));
  tab[3 * kTabN + e] = des_end(t2, i, fmax(0.0, qs_table2(i + 1) - t2));
}

struct Tables {
  const double *w, *t2, *dw, *d2;
  // entry of a table index (ap1 keeps indices in -1 .. 2620 for finite
  // temperatures; the clamp only guards non-finite input)
  __device__ __forceinline__ int at(double i) const {
    int e = (int)i + 1;
    return e < 0 ? 0 : (e >= kTabN ? kTabN - 1 : e);
  }
};

__device__ __forceinline__ double dimf(double a, double b) { return a - b > 0 ? a - b : 0.0; }

__device__ __forceinline__ double cvm_of(double mc_air, double qv, double c_vap, double q_liq,
                                         double q_sol) {
  return mc_air + qv * c_vap + q_liq * kCLiq + q_sol * kCIce;
}

// ap1_for_wqs2 (:454-457)
__device__ __forceinline__ double ap1_of(double ta) {
  const double ap1 = 10.0 * dimf(ta, kTmin) + 1.0;
  return fmin(ap1, (double)kQsLength) - 1;
}

// wqs2_fn_w (water) / wqs2_fn_2 (:503-522) with wqsat_and_dqdt (:487-494)
__device__ __forceinline__ void wqs2(const Tables& T, bool water, double ta, double den,
                                     double& wqsat, double& dqdt) {
  const double ap1 = ap1_of(ta);
  const double it = floor(ap1);
  const double it2 = floor(ap1 - 0.5);
  const double* tab = water ? T.w : T.t2;
  const double* des = water ? T.dw : T.d2;
  const double es = tab[T.at(it)] + (ap1 - it) * des[T.at(it)];
  const double denom = kRvgas * ta * den;
  wqsat = es / denom;
  const double d2 = des[T.at(it2)], dp1 = des[T.at(it2 + 1)];
  dqdt = 10.0 * (d2 + (ap1 - it2) * (dp1 - d2));
  dqdt = dqdt / denom;
}

// wqs1_fn_w / wqs1_fn_2 (:525-535)
__device__ __forceinline__ double wqs1(const Tables& T, bool water, double it, double ap1,
                                       double ta, double den) {
  const double* tab = water ? T.w : T.t2;
  const double* des = water ? T.dw : T.d2;
  return (tab[T.at(it)] + (ap1 - it) * des[T.at(it)]) / (kRvgas * ta * den);
}

struct SatFields {
  double *te, *qv, *ql, *qi, *qr, *qs, *qg, *qa;
  const double *area, *hs, *peln, *dp;
  const double* delz;
  double *q_con, *pt, *pkz, *cappa;
};

// hydrostatic form: delz[k] = delz[k - 1] on levels kmp+1 .. nz-1 (:612-615)
__global__ void __launch_bounds__(256)
    shift_delz_kernel(Dom d, int kmp, int nz, double* __restrict__ delz) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni, t = (int)blockIdx.y;
  const int64_t base = (int64_t)t * d.st + (int64_t)j * d.sj + i;
  for (int k = nz - 1; k > kmp; --k) delz[base + k * d.sk] = delz[base + (k - 1) * d.sk];
}

__global__ void __launch_bounds__(256)
    satadjust_kernel(Dom d, int kmp, fv3_satadj_params p, Tables T, SatFields f) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni;
  const int k = kmp + (int)blockIdx.y, t = (int)blockIdx.z;
  const int64_t o = (int64_t)t * d.st + (int64_t)k * d.sk + (int64_t)j * d.sj + i;
  const int64_t o2 = (int64_t)t * d.st2 + (int64_t)j * d.sj2 + i;
  const bool hydro = p.hydrostatic != 0;
  const double c_air = p.c_air, c_vap = p.c_vap, lv00 = p.lv00, d0_vap = p.d0_vap;

  double qv = f.qv[o], ql = f.ql[o], qi = f.qi[o], qr = f.qr[o], qs = f.qs[o], qg = f.qg[o];
  double pt = f.pt[o];
  const double dp = f.dp[o];
  double q_liq = ql + qr;
  double q_sol = qi + qs + qg;
  double qpz = q_liq + q_sol;
  double pt1 = pt / ((1.0 + p.zvir * qv) * (1.0 - qpz));
  const double t0 = pt1;
  qpz = qpz + qv;
  const double den = hydro ? dp / ((f.peln[o + d.sk] - f.peln[o]) * kRdgas * pt)
                           : -dp / (kGrav * f.delz[o]);
  const double mc_air = (1.0 - qpz) * c_air;
  double cvm = cvm_of(mc_air, qv, c_vap, q_liq, q_sol);
  double lhi = kLi00 + kDcIce * pt1;
  double icp2 = lhi / cvm;
  double lhl, lcp2;
  auto upd_i = [&]() {
    lhi = kLi00 + kDcIce * pt1;
    icp2 = lhi / cvm;
  };
  auto upd_li = [&]() {
    lhl = lv00 + d0_vap * pt1;
    lcp2 = lhl / cvm;
    upd_i();
  };
  double te0 = p.consv_te ? (hydro ? -c_air * t0 : -cvm * t0) : 0.0;
  if (qi < 0.0) {
    qs = qs + qi;
    qi = 0.0;
  }
  // melt_cloud_ice (:186-200)
  if (qi > 1.0e-8 && pt1 > kTice) {
    const double factmp = p.fac_imlt * (pt1 - kTice) / icp2;
    double sink = qi < factmp ? qi : factmp;
    qi = qi - sink;
    ql = ql + sink;
    q_liq = q_liq + sink;
    q_sol = q_sol - sink;
    cvm = cvm_of(mc_air, qv, c_vap, q_liq, q_sol);
    sink = -sink;
    pt1 = pt1 + sink * lhi / cvm;
  }
  upd_i();
  // fix_negative_snow (:209-219)
  if (qs < 0.0) {
    qg = qg + qs;
    qs = 0.0;
  } else if (qg < 0.0) {
    const double tmp = fmin(-qg, fmax(qs, 0.0));
    qg = qg + tmp;
    qs = qs - tmp;
  }
  // fix_negative_cloud_water (:222-233)
  if (ql < 0.0) {
    const double tmp = fmin(-ql, fmax(qr, 0.0));
    ql = ql + tmp;
    qr = qr - tmp;
  } else if (qr < 0.0) {
    const double tmp = fmin(-qr, fmax(ql, 0.0));
    ql = ql - tmp;
    qr = qr + tmp;
  }
  auto freeze = [&](double sink) {
    ql = ql - sink;
    qi = qi + sink;
    q_liq = q_liq - sink;
    q_sol = q_sol + sink;
    cvm = cvm_of(mc_air, qv, c_vap, q_liq, q_sol);
    pt1 = pt1 + sink * lhi / cvm;
  };
  // complete_freezing (:236-247)
  {
    const double dtmp = kTice - 48.0 - pt1;
    if (ql > 0.0 && dtmp > 0.0) freeze(fmin(ql, dtmp / icp2));
  }
  double wqsat, dq2dt;
  wqs2(T, true, pt1, den, wqsat, dq2dt);
  upd_li();
  double tcp3 = lcp2 + icp2 * fmin(1.0, dimf(kTice, pt1) / 48.0);
  // ql_evaporation (:437-441)
  auto evaporation = [&](double dq0) {
    const double factor = -fmin(1.0, p.fac_l2v * 10.0 * (1.0 - qv / wqsat));
    return -fmin(ql, factor * dq0);
  };
  // wqsat_correct (:444-451)
  auto correct = [&](double src) {
    qv = qv - src;
    ql = ql + src;
    q_liq = q_liq + src;
    cvm = cvm_of(mc_air, qv, c_vap, q_liq, q_sol);
    pt1 = pt1 + src * lhl / cvm;
  };
  {
    const double dq0 = (qv - wqsat) / (1.0 + tcp3 * dq2dt);
    const double src = dq0 > 0 ? fmin(p.sat_adj0 * dq0, fmax(p.ql_gen - ql, p.fac_v2l * dq0))
                               : evaporation(dq0);
    correct(src);
  }
  upd_li();
  tcp3 = lcp2 + icp2 * fmin(1.0, dimf(kTice, pt1) / 48.0);
  if (p.last_step) {
    wqs2(T, true, pt1, den, wqsat, dq2dt);
    const double dq0 = (qv - wqsat) / (1.0 + tcp3 * dq2dt);
    correct(dq0 > 0 ? dq0 : evaporation(dq0));
    upd_li();
  }
  // homogenous_freezing (:250-263)
  {
    const double dtmp = kTWfr - pt1;
    if (ql > 0.0 && dtmp > 0.0) freeze(fmin(fmin(ql, dtmp / icp2), ql * dtmp * 0.125));
  }
  upd_i();
  const double exptc = exp(0.66 * (kTice0 - pt1));
  // heterogeneous_freezing (:266-281)
  {
    const double tc = kTice0 - pt1;
    if (ql > 0.0 && tc > 0.0) {
      double sink = 3.3333e-10 * p.mdt * (exptc - 1.0) * den * (ql * ql);
      sink = fmin(ql, sink);
      sink = fmin(sink, tc / icp2);
      freeze(sink);
    }
  }
  upd_i();
  // make_graupel (:284-298)
  {
    const double dtmp = (kTice - 0.1) - pt1;
    if (qr > 1e-7 && dtmp > 0.0) {
      const double rainfac = (dtmp * 0.025) * (dtmp * 0.025);
      const double tmp = 1.0 < rainfac ? qr : rainfac * qr;
      const double sink = fmin(tmp, p.fac_r2g * dtmp / icp2);
      qr = qr - sink;
      qg = qg + sink;
      q_liq = q_liq - sink;
      q_sol = q_sol + sink;
      cvm = cvm_of(mc_air, qv, c_vap, q_liq, q_sol);
      pt1 = pt1 + sink * lhi / cvm;
    }
  }
  upd_i();
  // melt_snow (:301-321)
  {
    const double dtmp = pt1 - (kTice + 0.1);
    const double dimqs = dimf(p.qs_mlt, ql);
    if (qs > 1e-7 && dtmp > 0.0) {
      const double snowfac = (dtmp * 0.1) * (dtmp * 0.1);
      double tmp = 1.0 < snowfac ? qs : snowfac * qs;
      const double sink = fmin(tmp, p.fac_smlt * dtmp / icp2);
      tmp = fmin(sink, dimqs);
      qs = qs - sink;
      ql = ql + tmp;
      qr = qr + sink - tmp;
      q_liq = q_liq + sink;
      q_sol = q_sol - sink;
      cvm = cvm_of(mc_air, qv, c_vap, q_liq, q_sol);
      pt1 = pt1 - sink * lhi / cvm;
    }
  }
  // autoconversion_cloud_to_rain (:324-330)
  if (ql > p.ql0_max) {
    const double sink = p.fac_l2r * (ql - p.ql0_max);
    qr = qr + sink;
    ql = ql - sink;
  }
  double iqs2, dqsdt;
  wqs2(T, false, pt1, den, iqs2, dqsdt);
  const double expsubl = exp(0.875 * log(qi * den));
  upd_li();
  const double tcp2 = lcp2 + icp2;
  const double adj_fac = p.last_step ? 1.0 : p.sat_adj0;
  // sublimation (:333-396)
  {
    double src = 0.0;
    if (pt1 < p.t_sub) {
      src = dimf(qv, 1e-6);
    } else if (pt1 < kTice0) {
      const double dq = qv - iqs2;
      const double sink = adj_fac * dq / (1.0 + tcp2 * dqsdt);
      double pidep = 0.0;
      if (qi > 1.0e-8)
        pidep = p.sdt * dq * 349138.78 * expsubl /
                (iqs2 * den * kLat2 / (0.0243 * kRvgas * (pt1 * pt1)) + 4.42478e4);
      if (dq > 0.0) {
        const double tmp = kTice - pt1;
        const double qi_crt = p.qi_lim < 0.1 * tmp ? p.qi_gen * p.qi_lim / den
                                                   : p.qi_gen * 0.1 * tmp / den;
        const double maxtmp = qi_crt - qi > pidep ? qi_crt - qi : pidep;
        src = sink < maxtmp ? sink : maxtmp;
        src = src < tmp / tcp2 ? src : tmp / tcp2;
      } else {
        const double dimtmp = dimf(pt1, p.t_sub);
        pidep = 1.0 < dimtmp * 0.2 ? pidep : pidep * dimtmp * 0.2;
        src = pidep > sink ? pidep : sink;
        src = src > -qi ? src : -qi;
      }
    }
    qv = qv - src;
    qi = qi + src;
    q_sol = q_sol + src;
    cvm = cvm_of(mc_air, qv, c_vap, q_liq, q_sol);
    pt1 = pt1 + src * (lhl + lhi) / cvm;
  }
  // virtual temperature and cappa (:791-797)
  const double q_con = q_liq + q_sol;
  double tmp = 1.0 + p.zvir * qv;
  pt = pt1 * tmp * (1.0 - q_con);
  tmp *= kRdgas;
  const double cappa = tmp / (tmp + cvm);
  if (qg < 0) {
    const double mintmp = fmin(-qg, fmax(0.0, qi));
    qg = qg + mintmp;
    qi = qi - mintmp;
  }
  {
    const double qim = p.qi0_max / den;
    if (qi > qim) {
      const double sink = p.fac_i2s * (qi - qim);
      qi = qi - sink;
      qs = qs + sink;
    }
  }
  if (p.consv_te) te0 = hydro ? dp * (te0 + c_air * pt1) : dp * (te0 + cvm * pt1);
  cvm = mc_air + (qv + q_liq + q_sol) * c_vap;
  upd_li();
  if (p.do_qa && p.last_step) {
    // cloud fraction (:822-888)
    double qsol_r = p.rad_snow ? (p.rad_graupel ? qi + qs + qg : qi + qs) : qi;
    const double qliq_r = p.rad_rain ? ql + qr : ql;
    const double q_cond = qsol_r + qliq_r;
    const double tin = p.tintqs ? pt1 : pt1 - (lcp2 * q_cond + icp2 * qsol_r);
    const double ap1 = ap1_of(tin);
    const double it = floor(ap1);
    const double wqs1v = wqs1(T, true, it, ap1, tin, den);
    const double iqs1v = wqs1(T, false, it, ap1, tin, den);
    double qstar;
    if (tin < kTWfr) {
      qstar = iqs1v;
    } else if (tin >= kTice) {
      qstar = wqs1v;
    } else {
      const double rqi = q_cond > 1e-6 ? qsol_r / q_cond : (kTice - tin) / (kTice - kTWfr);
      qstar = rqi * iqs1v + (1.0 - rqi) * wqs1v;
    }
    const double mindw = fmin(1.0, fabs(f.hs[o2]) / (10.0 * kGrav));
    const double dw = p.dw_ocean + (p.dw_land - p.dw_ocean) * mindw;
    const double dbl_sqrt_area = dw * sqrt(sqrt(f.area[o2]) / 100.0e3);
    const double hvar = fmin(0.2, fmax(0.01, dbl_sqrt_area));
    const double rh = qpz / qstar;
    double qa;
    if (rh > 0.75 && qpz > 1.0e-8) {
      const double dq = hvar * qpz;
      const double q_plus = qpz + dq;
      const double q_minus = qpz - dq;
      if (p.icloud_f == 2) {
        if (qpz > qstar) {
          qa = 1.0;
        } else if (qstar < q_plus && q_cond > 1.0e-8) {
          const double r = (q_plus - qstar) / dq;
          qa = fmin(1.0, r * r);
        } else {
          qa = 0.0;
        }
      } else {
        if (qstar < q_minus) {
          qa = 1.0;
        } else {
          if (qstar < q_plus) {
            qa = p.icloud_f == 0 ? (q_plus - qstar) / (dq + dq)
                                 : (q_plus - qstar) / (2.0 * dq * (1.0 - q_cond));
          } else {
            qa = 0.0;
          }
          if (q_cond > 1.0e-8) qa = fmax(p.cld_min, qa);
          qa = fmin(1.0, qa);
        }
      }
    } else {
      qa = 0.0;
    }
    f.qa[o] = qa;
  }
  f.qv[o] = qv;
  f.ql[o] = ql;
  f.qi[o] = qi;
  f.qr[o] = qr;
  f.qs[o] = qs;
  f.qg[o] = qg;
  f.pt[o] = pt;
  f.q_con[o] = q_con;
  f.cappa[o] = cappa;
  if (p.consv_te) f.te[o] = te0;
  // compute_pkz_func (moist_cv.py:166-168)
  if (!hydro) f.pkz[o] = exp(cappa * log(kRdg * dp / f.delz[o] * pt));
}

}  // namespace

}  // namespace fv3

extern "C" {

int32_t fv3_sat_adjust_tables_f64(double* tables, void* stream) {
  using namespace fv3;
  FV3_REQUIRE(tables, FV3_E_ARG, "fv3_sat_adjust_tables: NULL pointer");
  hipLaunchKernelGGL(tables_kernel, dim3((kTabN + 255) / 256), dim3(256), 0, as_stream(stream),
                     tables);
  FV3_HIP_CHECK_LAUNCH("fv3_sat_adjust_tables");
  return FV3_OK;
}

int32_t fv3_sat_adjust_table_entries(void) { return fv3::kTabN; }

int32_t fv3_sat_adjust_f64(const fv3_domain* dom, int32_t kmp, const fv3_satadj_params* p,
                           const double* tables, double* te, double* qv, double* ql, double* qi,
                           double* qr, double* qs, double* qg, double* qa, const double* area,
                           const double* hs, const double* peln, const double* delp,
                           double* delz, double* q_con, double* pt, double* pkz, double* cappa,
                           void* stream) {
  using namespace fv3;
  const char* who = "fv3_sat_adjust";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(p && tables && te && qv && ql && qi && qr && qs && qg && qa && area && hs &&
                  delp && delz && q_con && pt && pkz && cappa && (!p->hydrostatic || peln),
              FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(kmp >= 0 && kmp <= dom->nk, FV3_E_ARG, "%s: kmp %d outside 0 .. %d", who, kmp,
              dom->nk);
  FV3_REQUIRE(p->icloud_f >= 0 && p->icloud_f <= 2, FV3_E_ARG, "%s: icloud_f must be 0..2",
              who);
  const int nk = dom->nk - kmp;
  if (nk <= 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const int ncol = (d.iec - d.isc + 1) * (d.jec - d.jsc + 1);
  if (p->hydrostatic) {
    hipLaunchKernelGGL(shift_delz_kernel, dim3((ncol + 255) / 256, d.ntiles), dim3(256), 0,
                       as_stream(stream), d, kmp, dom->nk, delz);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  Tables T{tables, tables + kTabN, tables + 2 * kTabN, tables + 3 * kTabN};
  SatFields f{te, qv, ql, qi, qr, qs, qg, qa, area, hs, peln, delp, delz, q_con, pt, pkz, cappa};
  hipLaunchKernelGGL(satadjust_kernel, dim3((ncol + 255) / 256, nk, d.ntiles), dim3(256), 0,
                     as_stream(stream), d, kmp, *p, T, f);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}
