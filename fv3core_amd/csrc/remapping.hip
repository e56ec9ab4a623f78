// Vertical remapping step: LagrangianToEulerian and fillz.
//
// Reference: fv3core/stencils/remapping.py:33-628 (init_pe,
// moist_cv_pt_pressure, pn2_pk_delp, pressures_mapu/v, update_ua,
// copy_from_below), moist_cv.py:8-194 and fillz.py:12-150.  The reference
// runs ~20 stencils and 14 MapSingle calls; here the pointwise and
// column-local stencils collapse into two column kernels that run before
// and after the remaps:
//   prep   init_pe + moist_cv_pt_pressure + pn2_pk_delp + pressures_mapu/v
//          (everything that reads the Lagrangian pe before it is replaced);
//   remaps pt, the tracers, w, delz (then fillz), u, v through map_single;
//   post   undo_delz_adjust_and_copy_peln + moist_pkz + update_ua +
//          copy_from_below + moist_pt_last_step / adjust_divide.
// One thread owns one column; lanes of a wave own neighbouring i, so every
// k-indexed access is coalesced across the wave.
#include "stencil.hpp"

namespace fv3 {
namespace {

constexpr double kRdgas = 287.05;
constexpr double kRvgas = 461.50;
constexpr double kGrav = 9.80665;
constexpr double kCpAir = 1004.6;
constexpr double kCvAir = kCpAir - kRdgas;
constexpr double kCvVap = 3.0 * kRvgas;
constexpr double kCIce = 1972.0;
constexpr double kCLiq = 4.1855e3;
constexpr double kTMin = 184.0;

// Work slabs of fv3_lagrangian_to_eulerian.
enum : int {
  kPe1 = 0, kPe2, kDp2, kPn2, kPe0u, kPe3u, kPe0v, kPe3v, kMapWork, kRemapSlabs = kMapWork + 7
};

struct Tracers {
  double* q[FV3_MAX_TRACERS];
};

// moist_cv_nwat6_fn + moist_cvm + set_cappa (moist_cv.py:8-41).
struct MoistCv {
  double cvm, gz, cappa;
};
__device__ __forceinline__ MoistCv moist_nwat6(double qv, double ql_, double qr, double qsn,
                                               double qi, double qg, double zvir) {
  const double ql = ql_ + qr;
  const double qs = qi + qsn + qg;
  const double gz = ql + qs;
  const double cvm = (1.0 - (qv + gz)) * kCvAir + qv * kCvVap + ql * kCLiq + qs * kCIce;
  const double cappa = kRdgas / (kRdgas + cvm / (1.0 + zvir * qv));
  return {cvm, gz, cappa};
}

// fix_tracer (fillz.py:12-94) on one column.  The reference's passes
// (forward borrow, upper-fix correction of the layer above, bottom borrow,
// then the rescale) touch level k-1 for the last time while level k is
// processed, so the correction of k-1 by upper_fix[k] is applied right
// there; the column sums run in the reference's order (levels 1..km-1).
__device__ void fix_tracer_column(double* q, const double* dp, int64_t sk, int km) {
  double qm = q[0];
  if (qm < 0.0) qm = 0.0;  // interval(0, 1); the (1, 2) fix sees qm >= 0
  double dpm = dp[0];
  double qk = q[sk], dpk = dp[sk];
  double lower = 0.0;
  int zfix = 0;
  double sum0 = 0.0, sum1 = 0.0;
  for (int k = 1; k <= km - 2; ++k) {
    const double qn = q[(int64_t)(k + 1) * sk], dpn = dp[(int64_t)(k + 1) * sk];
    if (lower != 0.0) qk = qk - lower / dpk;
    double upper = 0.0, low_k = 0.0;
    if (qk < 0.0) {
      zfix += 1;
      if (qm > 0.0) {
        const double cand = qm * dpm, lim = -(qk * dpk);
        const double dq = cand < lim ? cand : lim;
        qk = qk + dq / dpk;
        upper = dq;
      }
      if ((qk < 0.0) && (qn > 0.0)) {
        const double cand = qn * dpn, lim = -(qk * dpk);
        const double dq = cand < lim ? cand : lim;
        qk = qk + dq / dpk;
        low_k = dq;
      }
    }
    // PARALLEL interval(0, -1) on level k - 1
    if (upper != 0.0) qm = qm - upper / dpm;
    q[(int64_t)(k - 1) * sk] = qm;
    if (k - 1 >= 1) {
      const double dm = qm * dpm;
      sum0 += dm;
      sum1 += dm > 0.0 ? dm : 0.0;
    }
    qm = qk;
    dpm = dpk;
    qk = qn;
    dpk = dpn;
    lower = low_k;
  }
  // fix_bottom (level km-1) and the interval(-2, -1) correction of km-2
  if (lower != 0.0) qk = qk - lower / dpk;
  const double qup = qm * dpm, qly = -qk * dpk;
  const double dup = qup < qly ? qup : qly;
  double upb = 0.0;
  if ((qk < 0.0) && (qm > 0.0)) {
    zfix += 1;
    qk = qk + dup / dpk;
    upb = dup;
  }
  if (upb != 0.0) qm = qm - upb / dpm;
  q[(int64_t)(km - 2) * sk] = qm;
  q[(int64_t)(km - 1) * sk] = qk;
  {
    const double dm = qm * dpm;  // km - 2 >= 1
    sum0 += dm;
    sum1 += dm > 0.0 ? dm : 0.0;
    const double dmb = qk * dpk;
    sum0 += dmb;
    sum1 += dmb > 0.0 ? dmb : 0.0;
  }
  const double fac = sum0 > 0.0 ? sum0 / sum1 : 0.0;
  if (zfix > 0 && fac > 0.0) {
    for (int k = 1; k < km; ++k) {
      const int64_t o = (int64_t)k * sk;
      const double val = fac * (q[o] * dp[o]) / dp[o];
      q[o] = val > 0.0 ? val : 0.0;
    }
  }
}

__global__ void __launch_bounds__(256)
    fillz_kernel(Dom d, Tracers tr, const double* __restrict__ dp) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni;
  const int t = (int)blockIdx.y;
  const int64_t c = at3(d, i, j, 0, t);
  fix_tracer_column(tr.q[blockIdx.z] + c, dp + c, d.sk, d.nk);
}

struct RemapFields {
  fv3_remap_state s;
  fv3_remap_params p;
  double* w;  // work base
  int64_t slab;
};

__device__ __forceinline__ double* slab(const RemapFields& f, int which) {
  return f.w + (int64_t)which * f.slab;
}

// prep on [isc, iec+1] x [jsc, jec+1]: init_pe (remapping.py:33-40) and
// pressures_mapu (142-162) where i <= iec, pressures_mapv (165-179) where
// j <= jec, moist_cv_pt_pressure (56-125) and pn2_pk_delp (128-139) on the
// compute domain.
__global__ void __launch_bounds__(256) remap_prep_kernel(Dom d, RemapFields f) {
  const int ni = d.iec - d.isc + 2, nj = d.jec - d.jsc + 2;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni;
  const int t = (int)blockIdx.y;
  const int nz = d.nk;
  const int64_t sk = d.sk;
  const int64_t c = at3(d, i, j, 0, t);
  const fv3_remap_state& s = f.s;
  const double* pe = s.pe + c;
  const double* ak = s.ak;
  const double* bk = s.bk;
  const bool in_i = i <= d.iec, in_j = j <= d.jec;
  const double pe_bot = pe[nz * sk];
  if (in_i) {
    double* pe1 = slab(f, kPe1) + c;
    double* pe2 = slab(f, kPe2) + c;
    for (int k = 0; k <= nz; ++k) pe1[k * sk] = pe[k * sk];
    pe2[0] = f.p.ptop;
    pe2[nz * sk] = pe_bot;
    const double* pej = pe - d.sj;
    double* pe0 = slab(f, kPe0u) + c;
    double* pe3 = slab(f, kPe3u) + c;
    const double bsum = pej[nz * sk] + pe_bot;
    pe0[0] = pe[0];
    for (int k = 1; k <= nz; ++k) pe0[k * sk] = 0.5 * (pej[k * sk] + pe[k * sk]);
    for (int k = 0; k <= nz; ++k) pe3[k * sk] = ak[k] + (0.5 * bk[k]) * bsum;
  }
  if (in_j) {
    const double* pei = pe - 1;
    double* pe0 = slab(f, kPe0v) + c;
    double* pe3 = slab(f, kPe3v) + c;
    const double bsum = pei[nz * sk] + pe_bot;
    pe0[0] = pe[0];
    pe3[0] = ak[0];
    for (int k = 1; k <= nz; ++k) {
      pe0[k * sk] = 0.5 * (pei[k * sk] + pe[k * sk]);
      pe3[k * sk] = ak[k] + (0.5 * bk[k]) * bsum;
    }
  }
  if (!(in_i && in_j)) return;
  double* pt = s.pt + c;
  double* delp = s.delp + c;
  double* delz = s.delz + c;
  double* q_con = s.q_con + c;
  double* cappa = s.cappa + c;
  const double* qv = s.tracers[0] + c;
  const double* ql = s.tracers[1] + c;
  const double* qr = s.tracers[2] + c;
  const double* qi = s.tracers[3] + c;
  const double* qsn = s.tracers[4] + c;
  const double* qg = s.tracers[5] + c;
  const double rdg = -kRdgas / kGrav;
  for (int k = 0; k < nz; ++k) {
    const int64_t o = k * sk;
    const MoistCv m = moist_nwat6(qv[o], ql[o], qr[o], qsn[o], qi[o], qg[o], f.p.zvir);
    q_con[o] = m.gz;
    cappa[o] = m.cappa;
    const double p0 = pt[o], dz = delz[o], dp = delp[o];
    pt[o] = p0 * exp(m.cappa / (1.0 - m.cappa) * log(rdg * dp / dz * p0));
    delz[o] = -dz / dp;
  }
  s.ps[at2(d, i, j, t)] = pe_bot;
  double* pe2 = slab(f, kPe2) + c;
  double* dp2 = slab(f, kDp2) + c;
  double* pn2 = slab(f, kPn2) + c;
  double* pk = s.pk + c;
  for (int k = 1; k < nz; ++k) pe2[k * sk] = ak[k] + bk[k] * pe_bot;
  pn2[nz * sk] = s.peln[c + nz * sk];
  for (int k = nz - 1; k >= 0; --k) {
    const double v = pe2[(k + 1) * sk] - pe2[k * sk];
    dp2[k * sk] = v;
    delp[k * sk] = v;
  }
  for (int k = 0; k < nz; ++k) {
    const double l = log(pe2[k * sk]);
    pn2[k * sk] = l;
    pk[k * sk] = exp(f.p.akap * l);
  }
}

// moist_pt_last_step (moist_cv.py:139-162, dtmp = 0) on the last step,
// adjust_divide (pt / pkz) otherwise, for the column at offset c
__device__ __forceinline__ void final_pt(const RemapFields& f, int64_t c, int nz, int64_t sk) {
  const fv3_remap_state& s = f.s;
  double* pt = s.pt + c;
  const double* pkz = s.pkz + c;
  const double* qv = s.tracers[0] + c;
  const double* ql = s.tracers[1] + c;
  const double* qr = s.tracers[2] + c;
  const double* qi = s.tracers[3] + c;
  const double* qsn = s.tracers[4] + c;
  const double* qg = s.tracers[5] + c;
  if (f.p.last_step) {
    const double dtmp = 0.0;
    for (int k = 0; k <= nz; ++k) {
      const int64_t o = k * sk;
      const double gz = ql[o] + qr[o] + qi[o] + qsn[o] + qg[o];
      pt[o] = (pt[o] + dtmp * pkz[o]) / ((1.0 + f.p.zvir * qv[o]) * (1.0 - gz));
    }
  } else {
    for (int k = 0; k < nz; ++k) pt[k * sk] = pt[k * sk] / pkz[k * sk];
  }
}

// post on [isc, iec] x [jsc, jec+1]: undo_delz_adjust_and_copy_peln
// (remapping.py:43-53), moist_pkz (moist_cv.py:171-194), update_ua and
// copy_from_below (remapping.py:182-198), then moist_pt_last_step
// (moist_cv.py:139-162, dtmp = 0) or adjust_divide (pt / pkz).
__global__ void __launch_bounds__(256) remap_post_kernel(Dom d, RemapFields f) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 2;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni;
  const int t = (int)blockIdx.y;
  const int nz = d.nk;
  const int64_t sk = d.sk;
  const int64_t c = at3(d, i, j, 0, t);
  const fv3_remap_state& s = f.s;
  const double* pe2 = slab(f, kPe2) + c;
  double* ua = s.ua + c;
  if (j > d.jec) {
    // the extra row takes pe2 of the row below (remapping.py:190-193)
    const double* pe2m = pe2 - d.sj;
    for (int k = 0; k < nz - 1; ++k) ua[k * sk] = pe2m[(k + 1) * sk];
    ua[(nz - 1) * sk] = pe2[nz * sk];
    return;
  }
  double* delz = s.delz + c;
  const double* delp = s.delp + c;
  double* peln = s.peln + c;
  const double* pn2 = slab(f, kPn2) + c;
  double* pt = s.pt + c;
  double* pkz = s.pkz + c;
  double* q_con = s.q_con + c;
  double* cappa = s.cappa + c;
  double* pe = s.pe + c;
  const double* qv = s.tracers[0] + c;
  const double* ql = s.tracers[1] + c;
  const double* qr = s.tracers[2] + c;
  const double* qi = s.tracers[3] + c;
  const double* qsn = s.tracers[4] + c;
  const double* qg = s.tracers[5] + c;
  const double rdg = -kRdgas / kGrav;
  for (int k = 0; k <= nz; ++k) peln[k * sk] = pn2[k * sk];
  for (int k = 0; k < nz; ++k) {
    const int64_t o = k * sk;
    const double dz = -delz[o] * delp[o];
    delz[o] = dz;
    const MoistCv m = moist_nwat6(qv[o], ql[o], qr[o], qsn[o], qi[o], qg[o], f.p.zvir);
    q_con[o] = m.gz;
    cappa[o] = m.cappa;
    pkz[o] = exp(m.cappa * log(rdg * delp[o] / dz * pt[o]));
  }
  for (int k = 0; k < nz; ++k) ua[k * sk] = pe2[(k + 1) * sk];
  for (int k = 1; k < nz; ++k) pe[k * sk] = pe2[k * sk];
  if (!f.p.defer_pt) final_pt(f, c, nz, sk);
}

// the final pt step alone (after a saturation adjustment) on [isc, iec] x [jsc, jec]
__global__ void __launch_bounds__(256) remap_final_pt_kernel(Dom d, RemapFields f) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni;
  final_pt(f, at3(d, i, j, 0, (int)blockIdx.y), d.nk, d.sk);
}

}  // namespace
}  // namespace fv3

extern "C" {

int32_t fv3_fillz_f64(const fv3_domain* dom, int32_t ntracers, double* const* q, const double* dp,
                      void* stream) {
  using namespace fv3;
  const char* who = "fv3_fillz";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q && dp, FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(ntracers >= 0 && ntracers <= FV3_MAX_TRACERS, FV3_E_ARG,
              "%s: ntracers %d outside [0, %d]", who, ntracers, FV3_MAX_TRACERS);
  FV3_REQUIRE(dom->nk >= 3, FV3_E_ARG, "%s: need at least 3 layers (got %d)", who, dom->nk);
  Tracers tr{};
  for (int n = 0; n < ntracers; ++n) {
    FV3_REQUIRE(q[n], FV3_E_ARG, "%s: NULL tracer %d", who, n);
    tr.q[n] = q[n];
  }
  const int ni = dom->iec - dom->isc + 1, nj = dom->jec - dom->jsc + 1;
  if (ntracers == 0 || ni <= 0 || nj <= 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  dim3 grid((unsigned)((ni * nj + 255) / 256), (unsigned)d.ntiles, (unsigned)ntracers);
  hipLaunchKernelGGL(fillz_kernel, grid, dim3(256), 0, as_stream(stream), d, tr, dp);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_lagrangian_to_eulerian_f64(const fv3_domain* dom, const fv3_remap_params* p,
                                       const fv3_remap_state* s, double* work,
                                       int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_lagrangian_to_eulerian";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(p && s && work, FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(p->kord_tm < 0, FV3_E_UNSUPPORTED, "%s: kord_tm >= 0 (map ppm) is not supported",
              who);
  FV3_REQUIRE(p->nq >= 0 && p->nq <= FV3_MAX_TRACERS, FV3_E_ARG, "%s: nq %d outside [0, %d]",
              who, p->nq, FV3_MAX_TRACERS);
  FV3_REQUIRE(s->pt && s->delp && s->delz && s->peln && s->u && s->v && s->w && s->ua &&
                  s->cappa && s->q_con && s->pkz && s->pk && s->pe && s->ps && s->wsd && s->ak &&
                  s->bk,
              FV3_E_ARG, "%s: NULL field", who);
  const int nmoist = p->nq > 6 ? p->nq : 6;
  for (int n = 0; n < nmoist; ++n)
    FV3_REQUIRE(s->tracers[n], FV3_E_ARG, "%s: NULL tracer %d", who, n);
  const int nz = dom->nk;
  FV3_REQUIRE(nz >= 6, FV3_E_ARG, "%s: need at least 6 layers (got %d)", who, nz);
  FV3_REQUIRE(dom->isc >= 1 && dom->jsc >= 1 && dom->iec + 1 < dom->ni && dom->jec + 1 < dom->nj,
              FV3_E_ARG, "%s: the compute box needs one halo cell on each side", who);
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + dom->sk * (nz + 1),
              FV3_E_ARG, "%s: workspace slab too small", who);
  if (dom->ntiles == 0 || dom->iec < dom->isc || dom->jec < dom->jsc) return FV3_OK;
  Dom d = to_dom(dom);
  hipStream_t st = as_stream(stream);
  RemapFields f{*s, *p, work, work_slab};
  double* pe1 = work + kPe1 * work_slab;
  double* pe2 = work + kPe2 * work_slab;
  double* dp2 = work + kDp2 * work_slab;
  double* pn2 = work + kPn2 * work_slab;
  double* mw = work + kMapWork * work_slab;
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  {
    dim3 grid((unsigned)(((ni + 1) * (nj + 1) + 255) / 256), (unsigned)d.ntiles);
    hipLaunchKernelGGL(remap_prep_kernel, grid, dim3(256), 0, st, d, f);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  const int isc = d.isc, iec = d.iec, jsc = d.jsc, jec = d.jec;
  auto map = [&](int kord, int mode, int i2, int j2, double* q, const double* a, const double* b,
                 const double* qs, double qmin) {
    return fv3_map_single_f64(dom, kord, mode, isc, i2, jsc, j2, q, a, b, qs, qmin, mw, work_slab,
                              stream);
  };
  const int akord_tm = -p->kord_tm;
  const int akord_tr = p->kord_tr < 0 ? -p->kord_tr : p->kord_tr;
  rc = map(akord_tm, 1, iec, jec, s->pt, s->peln, pn2, nullptr, kTMin);
  if (rc) return rc;
  // MapNTracer forces kord 9 on the sixth tracer (mapn_tracer.py:41-42)
  for (int n = 0; n < p->nq; ++n) {
    rc = map(n == 5 ? 9 : akord_tr, 0, iec, jec, s->tracers[n], pe1, pe2, nullptr, 0.0);
    if (rc) return rc;
  }
  if (p->fill) {
    rc = fv3_fillz_f64(dom, p->nq, s->tracers, dp2, stream);
    if (rc) return rc;
  }
  rc = map(p->kord_wz, -2, iec, jec, s->w, pe1, pe2, s->wsd, 0.0);
  if (rc) return rc;
  rc = map(p->kord_wz, 1, iec, jec, s->delz, pe1, pe2, nullptr, 0.0);
  if (rc) return rc;
  {
    dim3 grid((unsigned)((ni * (nj + 1) + 255) / 256), (unsigned)d.ntiles);
    hipLaunchKernelGGL(remap_post_kernel, grid, dim3(256), 0, st, d, f);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  rc = map(p->kord_mt, -1, iec, jec + 1, s->u, work + kPe0u * work_slab,
           work + kPe3u * work_slab, nullptr, 0.0);
  if (rc) return rc;
  rc = map(p->kord_mt, -1, iec + 1, jec, s->v, work + kPe0v * work_slab,
           work + kPe3v * work_slab, nullptr, 0.0);
  if (rc) return rc;
  return FV3_OK;
}

int32_t fv3_remap_final_pt_f64(const fv3_domain* dom, const fv3_remap_params* p,
                               const fv3_remap_state* s, void* stream) {
  using namespace fv3;
  const char* who = "fv3_remap_final_pt";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(p && s && s->pt && s->pkz, FV3_E_ARG, "%s: NULL pointer", who);
  for (int n = 0; n < 6; ++n) FV3_REQUIRE(s->tracers[n], FV3_E_ARG, "%s: NULL tracer %d", who, n);
  if (dom->ntiles == 0 || dom->iec < dom->isc || dom->jec < dom->jsc) return FV3_OK;
  Dom d = to_dom(dom);
  RemapFields f{*s, *p, nullptr, 0};
  const int ncol = (d.iec - d.isc + 1) * (d.jec - d.jsc + 1);
  hipLaunchKernelGGL(remap_final_pt_kernel, dim3((unsigned)((ncol + 255) / 256), d.ntiles),
                     dim3(256), 0, as_stream(stream), d, f);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // extern "C"
