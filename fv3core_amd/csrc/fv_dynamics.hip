// Pointwise stencils of the full dynamics step outside the acoustic loop
// and the remap: fv_setup + pt_adjust (the preamble) and set_omega.
//
// Reference: fv3core/stencils/moist_cv.py:197-228 (fv_setup with
// moist_phys, nwat = 6) and fv3core/stencils/fv_dynamics.py:32-39
// (pt_adjust, set_omega).  Both run on the compute domain, levels
// [0, d->nk).
#include "stencil.hpp"

extern "C" {

int32_t fv3_fv_setup_f64(const fv3_domain* dom, const double* const* q6, const double* delp,
                         const double* delz, double* pt, double* q_con, double* cvm, double* pkz,
                         double* cappa, double* dp1, void* stream) {
  using namespace fv3;
  const char* who = "fv3_fv_setup";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q6 && delp && delz && pt && q_con && cvm && pkz && cappa && dp1, FV3_E_ARG,
              "%s: NULL pointer", who);
  for (int n = 0; n < 6; ++n) FV3_REQUIRE(q6[n], FV3_E_ARG, "%s: NULL tracer %d", who, n);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const double* qv = q6[0];
  const double* ql = q6[1];
  const double* qr = q6[2];
  const double* qi = q6[3];
  const double* qs = q6[4];
  const double* qg = q6[5];
  const double rdgas = 287.05, rvgas = 461.50, grav = 9.80665, cp_air = 1004.6;
  const double cv_air = cp_air - rdgas, cv_vap = 3.0 * rvgas, c_ice = 1972.0, c_liq = 4.1855e3;
  const double zvir = rvgas / rdgas - 1.0;
  const double rdg = -rdgas / grav;
  const Box comp{d.isc, d.iec, d.jsc, d.jec};
  launch_box(as_stream(stream), d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    // moist_cv_nwat6_fn: (cvm, q_con)
    const double qliq = ql[o] + qr[o];
    const double qsol = qi[o] + qs[o] + qg[o];
    const double qc = qliq + qsol;
    const double v = qv[o];
    const double cv = (1.0 - (v + qc)) * cv_air + v * cv_vap + qliq * c_liq + qsol * c_ice;
    const double d1 = zvir * v;
    const double cap = rdgas / (rdgas + cv / (1.0 + d1));
    const double pz = exp(cap * log(rdg * delp[o] * pt[o] * (1.0 + d1) * (1.0 - qc) / delz[o]));
    cvm[o] = cv;
    q_con[o] = qc;
    dp1[o] = d1;
    cappa[o] = cap;
    pkz[o] = pz;
    // pt_adjust
    pt[o] = pt[o] * (1.0 + d1) * (1.0 - qc) / pz;
  });
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_set_omega_f64(const fv3_domain* dom, const double* delp, const double* delz,
                          const double* w, double* omga, void* stream) {
  using namespace fv3;
  const char* who = "fv3_set_omega";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(delp && delz && w && omga, FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const Box comp{d.isc, d.iec, d.jsc, d.jec};
  launch_box(as_stream(stream), d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    omga[o] = delp[o] / delz[o] * w[o];
  });
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // extern "C"
