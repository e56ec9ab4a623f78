// D-grid step: FiniteVolumeFluxPrep (fxadv), DGridShallowWaterLagrangianDynamics
// (d_sw) with its kinetic energy (xtp_u / ytp_v) and DivergenceDamping.
//
// Reference: fv3core/stencils/fxadv.py:1-520, d_sw.py:1-1131, xtp_u.py,
// ytp_v.py, divergence_damping.py:1-490.  fv3_d_sw runs the reference's
// sequence from one host call; the transports reuse fv3_fvtp2d,
// fv3_delnflux and fv3_copy_corners_xy of this library, the B-grid
// interpolation launch_a2b_ord4.  Pointwise stencils that read only their
// own point of the fields another stencil writes are fused into one pass.
#include "kernels.hpp"
#include "ppm.hpp"

namespace fv3 {

namespace {

constexpr double kDconThreshold = 1e-5;

template <typename T>
__device__ __forceinline__ T contra(T v1, T v2, T cosa, T rsin2) {
  return (v1 - v2 * cosa) * rsin2;
}

// Typed views of fv3_dsw_params' columns and of the state structs (the
// same member order for every element type).
template <typename T>
struct DswCols {
  const T *damp_vt_deln, *damp_t_deln, *damp_w_nosg, *damp_v_nosg, *damp_w, *damp_vt, *ke_bg,
      *d_con_col, *d2_divg;
};

template <typename T>
DswCols<T> dsw_cols(const fv3_dsw_params* p) {
  DswCols<T> c;
  c.damp_vt_deln = static_cast<const T*>(p->damp_vt_deln);
  c.damp_t_deln = static_cast<const T*>(p->damp_t_deln);
  c.damp_w_nosg = static_cast<const T*>(p->damp_w_nosg);
  c.damp_v_nosg = static_cast<const T*>(p->damp_v_nosg);
  c.damp_w = static_cast<const T*>(p->damp_w);
  c.damp_vt = static_cast<const T*>(p->damp_vt);
  c.ke_bg = static_cast<const T*>(p->ke_bg);
  c.d_con_col = static_cast<const T*>(p->d_con_col);
  c.d2_divg = static_cast<const T*>(p->d2_divg);
  return c;
}

template <typename T>
struct DswState {
  T *delpc, *delp, *ptc, *pt, *u, *v, *w, *uc, *vc, *ua, *va, *divgd;
  T *mfx, *mfy, *cx, *cy, *crx, *cry, *xfx, *yfx, *q_con, *zh, *heat_source, *diss_est;
};
static_assert(sizeof(DswState<double>) == sizeof(fv3_dsw_state), "fv3_dsw_state layout");
static_assert(sizeof(DswState<float>) == sizeof(fv3_dsw_state_f32), "fv3_dsw_state_f32 layout");

template <typename T, typename S>
DswState<T> dsw_state(const S* st) {
  DswState<T> o;
  memcpy(&o, st, sizeof(o));
  return o;
}

// --------------------------------------------------------------------- fxadv
// main_uc_contra + uc_contra_y_edge and main_vc_contra (fxadv.py:11-58, 61-71)
// over the full domain.  Points the reference restores to their previous
// value are left untouched (no stencil downstream reads them).
template <typename T>
void launch_fxadv_main(hipStream_t s, const Dom& d, const Met<T>& m, const T* uc, const T* vc,
                       T* ucc, T* vcc) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj;
    if (i == d.is || i == d.ie + 1) {
      const T u = uc[o];
      ucc[o] = u > T(0) ? u / m.sin_sg3[o2 - 1] : u / m.sin_sg1[o2];
    } else if (i >= d.isc - 1 && i <= d.iec + 2 && !(j >= d.js - 1 && j <= d.js) &&
               !(j >= d.je && j <= d.je + 1)) {
      const T v = T(0.25) * (vc[o - 1] + vc[o] + vc[o - 1 + sj] + vc[o + sj]);
      ucc[o] = contra(uc[o], v, m.cosa_u[o2], m.rsin_u[o2]);
    }
    if (j >= d.jsc - 1 && j <= d.jec + 2 && j != d.js && j != d.je + 1) {
      const T u = T(0.25) * (uc[o - sj] + uc[o + 1 - sj] + uc[o] + uc[o + 1]);
      vcc[o] = contra(vc[o], u, m.cosa_v[o2], m.rsin_v[o2]);
    }
  });
}

// vc_contra_y_edge + vc_contra_x_edge (fxadv.py:102-153)
template <typename T>
void launch_fxadv_vc_edges(hipStream_t s, const Dom& d, const Met<T>& m, const T* vc,
                           const T* ucc, T* vcc) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj;
    if (j == d.js || j == d.je + 1) {
      const T v = vc[o];
      vcc[o] = v > T(0) ? v / m.sin_sg4[o2 - d.sj2] : v / m.sin_sg2[o2];
      return;
    }
    const bool col = (i >= d.is - 1 && i <= d.is) || (i >= d.ie && i <= d.ie + 1);
    if (!col || j < d.jsc || j > d.jec + 1) return;
    const bool restore = (j >= d.js && j <= d.js + 1) || (j >= d.je && j <= d.je + 1);
    if (restore) return;
    const T u = T(0.25) * (ucc[o - sj] + ucc[o + 1 - sj] + ucc[o] + ucc[o + 1]);
    vcc[o] = contra(vc[o], u, m.cosa_v[o2], T(1));
  });
}

// uc_contra_x_edge (fxadv.py:74-99)
template <typename T>
void launch_fxadv_uc_edges(hipStream_t s, const Dom& d, const Met<T>& m, const T* uc,
                           const T* vcc, T* ucc) {
  const Box bx{d.isc, d.iec + 1, d.jsc - 1, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const bool row = (j >= d.js - 1 && j <= d.js) || (j >= d.je && j <= d.je + 1);
    if (!row) return;
    const bool restore = (i >= d.is && i <= d.is + 1) || (i >= d.ie && i <= d.ie + 1);
    if (restore) return;
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const T v = T(0.25) * (vcc[o - 1] + vcc[o] + vcc[o - 1 + d.sj] + vcc[o + d.sj]);
    ucc[o] = contra(uc[o], v, m.cosa_u[o2], T(1));
  });
}

// uc_contra_corners + vc_contra_corners (fxadv.py:156-356): 16 points per
// level; none of them reads another one's new value, so all run at once.
template <typename T>
__global__ void __launch_bounds__(64)
    fxadv_corner_kernel(Dom d, Met<T> m, const T* __restrict__ uc, const T* __restrict__ vc,
                        T* ucc, T* vcc) {
  const int k = (int)(blockIdx.x * 64 + threadIdx.x);
  const int t = (int)blockIdx.y;
  if (k >= d.nk) return;
  const int64_t sj = d.sj, sj2 = d.sj2;
  T uval[8], vval[8];
  int64_t uo[8], vo[8];
  int nu = 0, nv = 0;
  // --- uc corners
  for (int r = 0; r < 8; ++r) {
    int i, j;
    switch (r) {
      case 0: i = d.is + 1; j = d.js - 1; break;
      case 1: i = d.is + 1; j = d.je; break;
      case 2: i = d.is + 1; j = d.js; break;
      case 3: i = d.is + 1; j = d.je + 1; break;
      case 4: i = d.ie; j = d.js - 1; break;
      case 5: i = d.ie; j = d.je; break;
      case 6: i = d.ie; j = d.js; break;
      default: i = d.ie; j = d.je + 1; break;
    }
    if (i < d.isc - 2 || i > d.iec + 3 || j < d.jsc - 2 || j > d.jec + 3) continue;
    const int64_t o = at3(d, i, j, k, t), o2 = at2(d, i, j, t);
    const T cu = m.cosa_u[o2];
    T cv, a, b, c, e, f, g, h;
    if (r < 2) {
      cv = m.cosa_v[o2 - 1];
      a = vcc[o - 1 + sj]; b = vcc[o + sj]; c = vcc[o]; e = vc[o - 1];
      f = ucc[o - 1]; g = ucc[o - 1 - sj]; h = ucc[o - sj];
    } else if (r < 4) {
      cv = m.cosa_v[o2 - 1 + sj2];
      a = vcc[o - 1]; b = vcc[o]; c = vcc[o + sj]; e = vc[o - 1 + sj];
      f = ucc[o - 1]; g = ucc[o - 1 + sj]; h = ucc[o + sj];
    } else if (r < 6) {
      cv = m.cosa_v[o2];
      a = vcc[o + sj]; b = vcc[o - 1 + sj]; c = vcc[o - 1]; e = vc[o];
      f = ucc[o + 1]; g = ucc[o + 1 - sj]; h = ucc[o - sj];
    } else {
      cv = m.cosa_v[o2 + sj2];
      a = vcc[o]; b = vcc[o - 1]; c = vcc[o - 1 + sj]; e = vc[o + sj];
      f = ucc[o + 1]; g = ucc[o + 1 + sj]; h = ucc[o + sj];
    }
    const T damp = T(1) / (T(1) - T(0.0625) * cu * cv);
    uval[nu] = (uc[o] - T(0.25) * cu * (a + b + c + e - T(0.25) * cv * (f + g + h))) * damp;
    uo[nu++] = o;
  }
  // --- vc corners (read the uc_contra values before this kernel's writes)
  for (int r = 0; r < 8; ++r) {
    int i, j;
    switch (r) {
      case 0: i = d.is - 1; j = d.js + 1; break;
      case 1: i = d.ie; j = d.js + 1; break;
      case 2: i = d.is; j = d.js + 1; break;
      case 3: i = d.ie + 1; j = d.js + 1; break;
      case 4: i = d.ie + 1; j = d.je; break;
      case 5: i = d.is; j = d.je; break;
      case 6: i = d.ie; j = d.je; break;
      default: i = d.is - 1; j = d.je; break;
    }
    if (i < d.isc - 2 || i > d.iec + 3 || j < d.jsc - 2 || j > d.jec + 3) continue;
    const int64_t o = at3(d, i, j, k, t), o2 = at2(d, i, j, t);
    const T cv = m.cosa_v[o2];
    T cu, a, b, c, e, f, g, h;
    if (r < 2) {
      cu = m.cosa_u[o2 - sj2];
      a = ucc[o + 1 - sj]; b = ucc[o + 1]; c = ucc[o]; e = uc[o - sj];
      f = vcc[o - sj]; g = vcc[o - 1 - sj]; h = vcc[o - 1];
    } else if (r < 4) {
      cu = m.cosa_u[o2 + 1 - sj2];
      a = ucc[o - sj]; b = ucc[o]; c = ucc[o + 1]; e = uc[o + 1 - sj];
      f = vcc[o - sj]; g = vcc[o + 1 - sj]; h = vcc[o + 1];
    } else if (r < 6) {
      cu = m.cosa_u[o2 + 1];
      a = ucc[o]; b = ucc[o - sj]; c = ucc[o + 1 - sj]; e = uc[o + 1];
      f = vcc[o + sj]; g = vcc[o + 1 + sj]; h = vcc[o + 1];
    } else {
      cu = m.cosa_u[o2];
      a = ucc[o + 1]; b = ucc[o + 1 - sj]; c = ucc[o - sj]; e = uc[o];
      f = vcc[o + sj]; g = vcc[o - 1 + sj]; h = vcc[o - 1];
    }
    const T damp = T(1) / (T(1) - T(0.0625) * cu * cv);
    vval[nv] = (vc[o] - T(0.25) * cv * (a + b + c + e - T(0.25) * cu * (f + g + h))) * damp;
    vo[nv++] = o;
  }
  for (int n = 0; n < nu; ++n) ucc[uo[n]] = uval[n];
  for (int n = 0; n < nv; ++n) vcc[vo[n]] = vval[n];
}

// fxadv_fluxes_stencil (fxadv.py:386-421)
template <typename T>
void launch_fxadv_fluxes(hipStream_t s, const Dom& d, const Met<T>& m, T dt, const T* ucc,
                         const T* vcc, T* crx, T* cry, T* xfx, T* yfx) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    if (i >= d.isc && i <= d.iec + 1) {
      const T prod = dt * ucc[o];
      if (prod > T(0)) {
        crx[o] = prod * m.rdxa[o2 - 1];
        xfx[o] = m.dy[o2] * prod * m.sin_sg3[o2 - 1];
      } else {
        crx[o] = prod * m.rdxa[o2];
        xfx[o] = m.dy[o2] * prod * m.sin_sg1[o2];
      }
    }
    if (j >= d.jsc && j <= d.jec + 1) {
      const T prod = dt * vcc[o];
      if (prod > T(0)) {
        cry[o] = prod * m.rdya[o2 - d.sj2];
        yfx[o] = m.dx[o2] * prod * m.sin_sg4[o2 - d.sj2];
      } else {
        cry[o] = prod * m.rdya[o2];
        yfx[o] = m.dx[o2] * prod * m.sin_sg2[o2];
      }
    }
  });
}

template <typename T>
int run_fxadv(hipStream_t s, const Dom& d, const Met<T>& m, T dt, const T* uc, const T* vc,
              T* crx, T* cry, T* xfx, T* yfx, T* ucc, T* vcc, const char* who) {
  launch_fxadv_main(s, d, m, uc, vc, ucc, vcc);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_fxadv_vc_edges(s, d, m, vc, ucc, vcc);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_fxadv_uc_edges(s, d, m, uc, vcc, ucc);
  FV3_HIP_CHECK_LAUNCH(who);
  if (has_edge(d.is) || has_edge(d.ie) || has_edge(d.js) || has_edge(d.je)) {
    dim3 grid((unsigned)((d.nk + 63) / 64), (unsigned)d.ntiles);
    hipLaunchKernelGGL((fxadv_corner_kernel<T>), grid, dim3(64), 0, s, d, m, uc, vc, ucc, vcc);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  launch_fxadv_fluxes(s, d, m, dt, ucc, vcc, crx, cry, xfx, yfx);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

// ------------------------------------------------------------- xtp_u / ytp_v
// bl, br of the cell at window position p (cells p-2..p+2 of q valid) for
// get_bl_br (xtp_u.py:8-47); `m` = absolute index of the cell along the
// direction, `dxr` the spacing row (dx for the hord < 8 face values, dxa for
// the hord = 8 edge formulas) with the window's origin.
template <typename T, int ORD>
__device__ __forceinline__ void xtp_blbr(const T* q, int p, int m, int is, int ie,
                                         const DxRow<T>& dxr, bool zero, T& bl, T& br) {
  if constexpr (ORD < 8) {
    bl = ppm_al(q, p, m, is, ie, dxr) - q[p];
    br = ppm_al(q, p + 1, m + 1, is, ie, dxr) - q[p];
  } else {
    const T dmm = ppm_dm(q[p - 2], q[p - 1], q[p]);
    const T dm0 = ppm_dm(q[p - 1], q[p], q[p + 1]);
    const T dmp = ppm_dm(q[p], q[p + 1], q[p + 2]);
    const T third = T(1.0 / 3.0);
    const T alL = T(0.5) * (q[p - 1] + q[p]) + third * (dmm - dm0);
    const T alR = T(0.5) * (q[p] + q[p + 1]) + third * (dm0 - dmp);
    const T xt = T(2) * dm0;
    bl = -sgn(fmin(fabs(xt), fabs(alL - q[p])), xt);
    br = sgn(fmin(fabs(xt), fabs(alR - q[p])), xt);
    const bool near_s = (m >= is - 1) && (m <= is + 1);
    const bool near_e = (m >= ie - 1) && (m <= ie + 1);
    if (near_s || near_e) {
      T dd[6];
      dxr.load(dd);
      const T s11 = T(FV3_PPM_S11), s14 = T(FV3_PPM_S14), s15 = T(FV3_PPM_S15);
      T xbl, xbr;
      if (m == ie + 1) {
        xbl = ppm_edge1(q, dd, p);
        xbr = s11 * (q[p + 1] - q[p]) - s14 * dmp + q[p];
      } else if (m == ie) {
        xbl = s15 * q[p] + s11 * q[p - 1] + s14 * dmm;
        xbr = ppm_edge0(q, dd, p);
      } else if (m == ie - 1) {
        xbl = alL;
        xbr = s15 * q[p + 1] + s11 * q[p] + s14 * dm0;
      } else if (m == is + 1) {
        xbl = s15 * q[p - 1] + s11 * q[p] - s14 * dm0;
        xbr = alR;
      } else if (m == is) {
        xbl = ppm_edge1(q, dd, p);
        xbr = s15 * q[p] + s11 * q[p + 1] - s14 * dmp;
      } else {
        xbl = s14 * dmm + s11 * (q[p - 1] - q[p]) + q[p];
        xbr = ppm_edge0(q, dd, p);
      }
      bl = xbl - q[p];
      br = xbr - q[p];
      // xtp_u constrains only the second cell from each edge (xtp_u.py:38-39)
      if (m == is + 1 || m == ie - 1) ppm_standard_constraint(bl, br);
    }
  }
  if (zero) {
    bl = T(0);
    br = T(0);
  }
}

// advect_u_along_x (xtp_u.py:50-91) at the face between cells f-1 and f.
// q6/dx6 hold cells f-3..f+2; zm/z0 mark cells f-1 / f in the zeroed corners.
template <typename T, int ORD>
__device__ __forceinline__ T xtp_face(const T (&q6)[6], const DxRow<T>& dxr, T ub, T cfl, int f,
                                      int is, int ie, bool zm, bool z0) {
  T blm, brm, bl0, br0;
  xtp_blbr<T, ORD>(q6, 2, f - 1, is, ie, dxr, zm, blm, brm);
  xtp_blbr<T, ORD>(q6, 3, f, is, ie, dxr, z0, bl0, br0);
  const T b0m = blm + brm, b00 = bl0 + br0;
  const T fx0 = cfl > T(0) ? (T(1) - cfl) * (brm - cfl * b0m) : (T(1) + cfl) * (bl0 + cfl * b00);
  T mask = T(1);
  if constexpr (ORD < 8) {
    bool sm, s0;
    if constexpr (ORD == 5) {
      sm = blm * brm < T(0);
      s0 = bl0 * br0 < T(0);
    } else {
      sm = (T(3) * fabs(b0m)) < fabs(blm - brm);
      s0 = (T(3) * fabs(b00)) < fabs(bl0 - br0);
    }
    mask = (sm || s0) ? T(1) : T(0);
  }
  return ub > T(0) ? q6[2] + fx0 * mask : q6[3] + fx0 * mask;
}

// kinetic_energy_update_part_1/2 (d_sw.py:196-237, interpolation 564-603,
// corner_ke 112-147) on [isc, iec+1]^2.
template <typename T, int ORD>
void launch_ke(hipStream_t s, const Dom& d, const Met<T>& m, T dt, const T* uc, const T* vc,
               const T* u, const T* v, const T* ucc, const T* vcc, T* ke) {
  const Box bx{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj, sj2 = d.sj2;
    // B-grid contravariant winds
    const T ub_cov = T(0.5) * (uc[o - sj] + uc[o]);
    const T vb_cov = T(0.5) * (vc[o - 1] + vc[o]);
    T ub = contra(ub_cov, vb_cov, m.cosa[o2], m.rsina[o2]);
    T vb = contra(vb_cov, ub_cov, m.cosa[o2], m.rsina[o2]);
    const bool rows = j == d.js || j == d.je + 1;
    const bool cols = i == d.is || i == d.ie + 1;
    if (rows) {
      ub = T(0.25) * (-ucc[o - 2 * sj] + T(3) * (ucc[o - sj] + ucc[o]) - ucc[o + sj]);
    }
    if (cols) {
      ub = T(0.5) * (ucc[o - sj] + ucc[o]);
      vb = T(0.25) * (-vcc[o - 2] + T(3) * (vcc[o - 1] + vcc[o]) - vcc[o + 1]);
    }
    if (rows) vb = T(0.5) * (vcc[o - 1] + vcc[o]);
    // advect_v_along_y: v along j with dy / dya
    T advv, advu;
    {
      T q6[6];
#pragma unroll
      for (int n = 0; n < 6; ++n) q6[n] = v[o + (int64_t)(n - 3) * sj];
      const bool dx_is_a = ORD == 8;
      DxRow<T> dr{(dx_is_a ? m.dya : m.dy) + o2 - 3 * sj2, sj2, 6};
      const T cfl = vb > T(0) ? vb * dt * m.rdy[o2 - sj2] : vb * dt * m.rdy[o2];
      // zeroed corners of ytp_v: i in {is, ie+1}, cells j in [js-1, js] or [je, je+1]
      const bool ci = i == d.is || i == d.ie + 1;
      auto zc = [&](int jj) {
        return ci && ((jj >= d.js - 1 && jj <= d.js) || (jj >= d.je && jj <= d.je + 1));
      };
      advv = xtp_face<T, ORD>(q6, dr, vb, cfl, j, d.js, d.je, zc(j - 1), zc(j));
    }
    {
      T q6[6];
#pragma unroll
      for (int n = 0; n < 6; ++n) q6[n] = u[o + n - 3];
      const bool dx_is_a = ORD == 8;
      DxRow<T> dr{(dx_is_a ? m.dxa : m.dx) + o2 - 3, 1, 6};
      const T cfl = ub > T(0) ? ub * dt * m.rdx[o2 - 1] : ub * dt * m.rdx[o2];
      const bool cj = j == d.js || j == d.je + 1;
      auto zc = [&](int ii) {
        return cj && ((ii >= d.is - 1 && ii <= d.is) || (ii >= d.ie && ii <= d.ie + 1));
      };
      advu = xtp_face<T, ORD>(q6, dr, ub, cfl, i, d.is, d.ie, zc(i - 1), zc(i));
    }
    T kev = T(0.5) * dt * (ub * advu + vb * advv);
    // all_corners_ke (d_sw.py:136-147)
    if (rows && cols) {
      int io1, jo1, io2;
      T vsign;
      if (i == d.is && j == d.js) {
        io1 = 0; jo1 = 0; io2 = -1; vsign = T(1);
      } else if (j == d.js) {
        io1 = -1; jo1 = 0; io2 = 0; vsign = T(-1);
      } else if (i == d.ie + 1) {
        io1 = -1; jo1 = -1; io2 = 0; vsign = T(1);
      } else {
        io1 = 0; jo1 = -1; io2 = -1; vsign = T(-1);
      }
      const T dt6 = dt / T(6);
      const T u0 = u[o], um = u[o - 1], v0 = v[o], vm = v[o - sj];
      const T ut0 = ucc[o], utm = ucc[o - sj], vt0 = vcc[o], vtm = vcc[o - 1];
      const T fi1 = T(io1 + 1), fi = T(io1), fj1 = T(jo1 + 1), fj = T(jo1);
      const T f21 = T(io2 + 1), f2 = T(io2);
      kev = dt6 * ((ut0 + utm) * (fi1 * u0 - (fi * um)) + (vt0 + vtm) * (fj1 * v0 - (fj * vm)) +
                   ((fj1 * ut0 - (fj * utm)) + vsign * (fi1 * vt0 - (fi * vtm))) *
                       (f21 * u0 - (f2 * um)));
    }
    ke[o] = kev;
  });
}

// ---------------------------------------------------------- pointwise d_sw
// flux_capacitor (full domain; the mass-flux temporaries are zero outside
// where the transport writes them) + heat_diss (compute domain).
template <typename T>
void launch_capacitor_heat(hipStream_t s, const Dom& d, const Met<T>& m, T dt, int nonhydro,
                           const DswCols<T>& cols, const DswState<T>& st, const T* fx,
                           const T* fy, const T* fx2, const T* fy2, T* dw, T* heat_s) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  T* cx = st.cx;
  T* cy = st.cy;
  T* mfx = st.mfx;
  T* mfy = st.mfy;
  const T* crx = st.crx;
  const T* cry = st.cry;
  const T* w = st.w;
  T* diss = st.diss_est;
  const T* damp_w = cols.damp_w;
  const T* ke_bg = cols.ke_bg;
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    cx[o] = cx[o] + crx[o];
    cy[o] = cy[o] + cry[o];
    if (i >= d.isc && i <= d.iec + 1 && j >= d.jsc && j <= d.jec) mfx[o] = mfx[o] + fx[o];
    if (i >= d.isc && i <= d.iec && j >= d.jsc && j <= d.jec + 1) mfy[o] = mfy[o] + fy[o];
    if (!nonhydro || i < d.isc || i > d.iec || j < d.jsc || j > d.jec) return;
    if (damp_w[k] > T(1e-5)) {
      const T dd8 = ke_bg[k] * fabs(dt);
      const T dwv = (fx2[o] - fx2[o + 1] + fy2[o] - fy2[o + d.sj]) * m.rarea[at2(d, i, j, t)];
      const T hs = dd8 - dwv * (w[o] + T(0.5) * dwv);
      dw[o] = dwv;
      heat_s[o] = hs;
      diss[o] = diss[o] + hs;
    }
  });
}

// flux_adjust (d_sw.py:50-64): q = q * delp + flux_increment(gx, gy) on the compute domain.
template <typename T>
void launch_flux_adjust(hipStream_t s, const Dom& d, const Met<T>& m, T* q, const T* delp,
                        const T* gx, const T* gy) {
  const Box comp{d.isc, d.iec, d.jsc, d.jec};
  launch_box(s, d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    q[o] = q[o] * delp[o] + (gx[o] - gx[o + 1] + gy[o] - gy[o + d.sj]) * m.rarea[at2(d, i, j, t)];
  });
}

// apply_pt_delp_fluxes (d_sw.py:149-180) on the compute domain.
template <typename T>
void launch_pt_delp(hipStream_t s, const Dom& d, const Met<T>& m, T* pt, T* delp, const T* gx,
                    const T* gy, const T* fx, const T* fy) {
  const Box comp{d.isc, d.iec, d.jsc, d.jec};
  launch_box(s, d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const T ra = m.rarea[at2(d, i, j, t)];
    const T dp = delp[o];
    const T ptn = pt[o] * dp + (gx[o] - gx[o + 1] + gy[o] - gy[o + d.sj]) * ra;
    const T dpn = dp + (fx[o] - fx[o + 1] + fy[o] - fy[o + d.sj]) * ra;
    pt[o] = ptn / dpn;
    delp[o] = dpn;
  });
}

// compute_vorticity (full domain) + adjust_w_and_qcon (compute domain).
template <typename T>
void launch_vorticity_adjust(hipStream_t s, const Dom& d, const Met<T>& m, const T* u,
                             const T* v, T* wk, T* w, const T* delp, const T* dw, T* q_con,
                             const T* damp_w) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const T dx = m.dx[o2], dy = m.dy[o2], ra = m.rarea[o2];
    wk[o] = (u[o] - u[o + d.sj] * m.dx[o2 + d.sj2] / dx) * (ra * dx) +
            (v[o + 1] * m.dy[o2 + 1] / dy - v[o]) * (ra * dy);
    if (i < d.isc || i > d.iec || j < d.jsc || j > d.jec) return;
    T wn = w[o] / delp[o];
    if (damp_w[k] > T(1e-5)) wn = wn + dw[o];
    w[o] = wn;
    q_con[o] = q_con[o] / delp[o];
  });
}

// vort_differencing (corners; levels with d_con > threshold) + compute_vort
// into a separate slab (full domain), so both read the old vorticity.
template <typename T>
void launch_vort_diff(hipStream_t s, const Dom& d, const Met<T>& m, const T* vort, const T* wk,
                      T* vxd, T* vyd, T* vort2, const T* dcon) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    if (dcon[k] > T(kDconThreshold)) {
      if (i >= d.isc && i <= d.iec && j >= d.jsc && j <= d.jec + 1) vxd[o] = vort[o] - vort[o + 1];
      if (i >= d.isc && i <= d.iec + 1 && j >= d.jsc && j <= d.jec) vyd[o] = vort[o] - vort[o + d.sj];
    }
    vort2[o] = wk[o] + m.f0[at2(d, i, j, t)];
  });
}

// u_and_v_from_ke (d_sw.py:244-266)
template <typename T>
void launch_uv_from_ke(hipStream_t s, const Dom& d, const Met<T>& m, const T* ke, const T* fx,
                       const T* fy, T* u, T* v) {
  const Box bx{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    if (i <= d.iec) u[o] = u[o] * m.dx[o2] + ke[o] - ke[o + 1] + fy[o];
    if (j <= d.jec) v[o] = v[o] * m.dy[o2] + ke[o] - ke[o + d.sj] - fx[o];
  });
}

// heat_source_from_vorticity_damping on the compute domain (the only part
// of the heat source that reaches the output; d_sw.py:353-434).
template <typename T>
void launch_heat_source(hipStream_t s, const Dom& d, const Met<T>& m, const T* vxd, const T* vyd,
                        const T* ut, const T* vt, const T* u, const T* v, const T* delp,
                        const T* heat_s, T* heat_total, const T* kef, int d_con_on) {
  if (!d_con_on) return;
  const Box comp{d.isc, d.iec, d.jsc, d.jec};
  launch_box(s, d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj, sj2 = d.sj2;
    T hs = heat_s[o];
    if (kef[k] > T(kDconThreshold)) {
      // ubt, gy at j and j+1; vbt, gx at i and i+1
      const T rdx0 = m.rdx[o2], rdx1 = m.rdx[o2 + sj2];
      const T rdy0 = m.rdy[o2], rdy1 = m.rdy[o2 + 1];
      const T ubt0 = (vxd[o] + vt[o]) * rdx0, ubt1 = (vxd[o + sj] + vt[o + sj]) * rdx1;
      const T fy0 = u[o] * rdx0, fy1 = u[o + sj] * rdx1;
      const T gy0 = fy0 * ubt0, gy1 = fy1 * ubt1;
      const T vbt0 = (vyd[o] - ut[o]) * rdy0, vbt1 = (vyd[o + 1] - ut[o + 1]) * rdy1;
      const T fx0 = v[o] * rdy0, fx1 = v[o + 1] * rdy1;
      const T gx0 = fx0 * vbt0, gx1 = fx1 * vbt1;
      const T u2 = fy0 + fy1, du2 = ubt0 + ubt1;
      const T v2 = fx0 + fx1, dv2 = vbt0 + vbt1;
      const T dampterm =
          m.rsin2[o2] * ((ubt0 * ubt0 + ubt1 * ubt1 + vbt0 * vbt0 + vbt1 * vbt1) +
                         T(2) * (gy0 + gy1 + gx0 + gx1) -
                         m.cosa_s[o2] * (u2 * dv2 + v2 * du2 + du2 * dv2));
      hs = delp[o] * (hs - T(0.25) * kef[k] * dampterm);
    }
    heat_total[o] = heat_total[o] + hs;
  });
}

// update_u_and_v (d_sw.py:437-460)
template <typename T>
void launch_update_uv(hipStream_t s, const Dom& d, const T* ut, const T* vt, T* u, T* v,
                      const T* damp_vt) {
  const Box bx{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    if (!(damp_vt[k] > T(1e-5))) return;
    const int64_t o = at3(d, i, j, k, t);
    if (i <= d.iec) u[o] = u[o] + vt[o];
    if (j <= d.jec) v[o] = v[o] - ut[o];
  });
}

// --------------------------------------------------------- divergence damping
template <typename T>
__device__ __forceinline__ T damp_tmp(T q, T da_min_c, T d2_bg, T dddmp) {
  const T mintmp = fmin(T(0.2), dddmp * fabs(q));
  return da_min_c * fmax(d2_bg, mintmp);
}

// get_delpc + damping (divergence_damping.py:31-107) on levels [0, nk_lo).
template <typename T>
void launch_div_low(hipStream_t s, const Dom& d, const Met<T>& m, int nk_lo, T dt, T da_min_c,
                    T dddmp, const T* d2_bg, const T* u, const T* v, const T* ua, const T* va,
                    const T* uc, const T* vc, T* delpc, T* ptc, T* vort, T* ke) {
  const Box bx{d.isc - 1, d.iec + 1, d.jsc - 1, d.jec + 1};
  launch_box(s, d, bx, 0, nk_lo, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj, sj2 = d.sj2;
    // u_contra * dyc at (i, j); v_contra * dxc at (i, j)
    auto u_dyc = [&](int64_t oo, int64_t oo2, int jj) {
      T uct;
      if (jj == d.js || jj == d.je + 1) {
        uct = vc[oo] > T(0) ? u[oo] * m.sin_sg4[oo2 - sj2] : u[oo] * m.sin_sg2[oo2];
      } else {
        const T vfa = T(0.5) * (va[oo - sj] + va[oo]);
        uct = contra(u[oo], vfa, m.cosa_v[oo2], m.sina_v[oo2]);
      }
      return uct * m.dyc[oo2];
    };
    auto v_dxc = [&](int64_t oo, int64_t oo2, int ii, int jj) {
      const bool zero_pt = (ii == d.is || ii == d.ie + 1) && (jj == d.je + 1 || jj == d.js - 1);
      if (zero_pt) return T(0);
      T vct;
      if (ii == d.is || ii == d.ie + 1) {
        vct = uc[oo] > T(0) ? v[oo] * m.sin_sg3[oo2 - 1] : v[oo] * m.sin_sg1[oo2];
      } else {
        const T ufa = T(0.5) * (ua[oo - 1] + ua[oo]);
        vct = contra(v[oo], ufa, m.cosa_u[oo2], m.sina_u[oo2]);
      }
      return vct * m.dxc[oo2];
    };
    const bool in_comp = i >= d.isc && j >= d.jsc;  // box ends at iec+1, jec+1
    if (j >= d.jsc) ptc[o] = u_dyc(o, o2, j);
    if (i >= d.isc && !in_comp) vort[o] = v_dxc(o, o2, i, j);
    if (!in_comp) return;
    const T vd0 = v_dxc(o, o2, i, j), vdm = v_dxc(o - sj, o2 - sj2, i, j - 1);
    const T ud0 = u_dyc(o, o2, j), udm = u_dyc(o - 1, o2 - 1, j);
    const T dpc = m.rarea_c[o2] * (vdm - vd0 + udm - ud0);
    delpc[o] = dpc;
    const T damp = damp_tmp(dpc * dt, da_min_c, d2_bg[k], dddmp);
    const T vv = damp * dpc;
    vort[o] = vv;
    ke[o] = ke[o] + vv;
  });
}

const int kOne[4] = {1, 1, 1, 1};

// B-grid and D-grid corner fills by the quarter-turn rule
// (corners.py:543-723, 993-1100): destination -> source in the 3x3 halo
// block of each tile corner.
void append_rotation_fill(FillTable* tab, const fv3_domain* dom, int kind, int fx, int fy,
                          double mysign) {
  // kind 0: B-grid x fill of field fx; 1: B-grid y fill; 2: D-grid vector (x = fx, y = fy)
  for (int c = 0; c < 4; ++c) {
    const int ai = (c == 0 || c == 2) ? dom->i_start : dom->i_end;
    const int aj = (c == 0 || c == 1) ? dom->j_start : dom->j_end;
    if (!has_edge(ai) || !has_edge(aj)) continue;
    const int sx = (c == 0 || c == 2) ? -1 : 1;
    const int sy = (c == 0 || c == 1) ? -1 : 1;
    // corner point in doubled coordinates (2 * storage position)
    const int cx2 = 2 * (sx < 0 ? ai : ai + 1), cy2 = 2 * (sy < 0 ? aj : aj + 1);
    const int ncomp = kind == 2 ? 2 : 1;
    for (int comp = 0; comp < ncomp; ++comp) {
      // staggers in doubled units: B-grid (0, 0); D-grid x (1, 0), y (0, 1)
      const int ox = (kind == 2 && comp == 0) ? 1 : 0;
      const int oy = (kind == 2 && comp == 1) ? 1 : 0;
      const bool keep_x = kind == 0 || (kind == 2 && comp == 0);
      for (int a = 1; a <= 3; ++a) {
        for (int b = 1; b <= 3; ++b) {
          const int px = ox ? sx * (2 * a - 1) : sx * 2 * a;
          const int py = oy ? sy * (2 * b - 1) : sy * 2 * b;
          int qx = py, qy = -px;  // (y, -x)
          const bool sign_ok = keep_x ? ((qx > 0) == (px > 0)) : ((qy > 0) == (py > 0));
          if (!sign_ok) {
            qx = -py;
            qy = px;
          }
          // the source field's stagger
          const int sox = (kind == 2 && comp == 1) ? 1 : 0;
          const int soy = (kind == 2 && comp == 0) ? 1 : 0;
          const int di = (cx2 + px - ox) / 2, dj = (cy2 + py - oy) / 2;
          const int si = (cx2 + qx - sox) / 2, sj = (cy2 + qy - soy) / 2;
          if (di < 0 || dj < 0 || si < 0 || sj < 0 || di >= dom->ni || si >= dom->ni ||
              dj >= dom->nj || sj >= dom->nj || tab->n >= kMaxFill)
            continue;
          const int e = tab->n++;
          tab->dst_i[e] = di;
          tab->dst_j[e] = dj;
          tab->src_i[e] = si;
          tab->src_j[e] = sj;
          if (kind == 2) {
            tab->dst_f[e] = (signed char)(comp == 0 ? fx : fy);
            tab->src_f[e] = (signed char)(comp == 0 ? fy : fx);
            tab->mult[e] = (signed char)((c == 0 || c == 3) ? (mysign < 0 ? -1 : 1) : 1);
          } else {
            tab->dst_f[e] = (signed char)fx;
            tab->src_f[e] = (signed char)fx;
            tab->mult[e] = 1;
          }
        }
      }
    }
  }
}

// The high-order part (levels [k0, nk)): nord passes of vc/uc from the
// divergence and the divergence from them, with the corner fills, then the
// Smagorinsky-like vorticity (from a2b_ord4 of wk) and the damping.
template <typename T>
int run_div_high(hipStream_t s, const fv3_domain* dom, const Met<T>& m, const fv3_dsw_params* p,
                 const DswCols<T>& cols, T dt, const T* wk, T* divg, T* vc, T* uc, T* delpc,
                 T* vort, T* ke, T* a2b_qx, T* a2b_edges, T* dg_last, const char* who) {
  const int k0 = p->div_k0;
  fv3_domain hd = *dom;
  hd.nk = dom->nk - k0;
  if (hd.nk <= 0) return FV3_OK;
  const int64_t ko = (int64_t)k0 * dom->sk;
  Dom d = to_dom(&hd);
  T* dg = divg + ko;
  T* vch = vc + ko;
  T* uch = uc + ko;
  T* dpc = delpc + ko;
  T* vo = vort + ko;
  T* keh = ke + ko;
  const T* wkh = wk + ko;
  const T* d2 = cols.d2_divg + k0;
  T* dgl = dg_last + ko;
  const int nord = p->div_nord;
  const bool corner = (has_edge(dom->i_start) || has_edge(dom->i_end)) &&
                      (has_edge(dom->j_start) || has_edge(dom->j_end));
  // In FP32 the product dd8 * (nord-fold Laplacian of divg_d) over- and
  // underflows on coarse grids (dd8 = (d4_bg da_min_c)^(nord+1) reaches 1e43
  // at C12, the Laplacian 1e-41): there each pass also multiplies by da_min_c
  // and the final coefficient divides that out, formed in FP64.  In FP64
  // the factor is 1, which leaves every result bit for bit as it was.
  const bool rescale = sizeof(T) < sizeof(double);
  const T pass_scale = rescale ? (T)p->da_min_c : T(1);
  // copy_computeplus (delpc = divg_d on the corners, before the passes
  // change divg_d) is done by the first pass's x differences, whose box
  // holds the corners; the corner fills before them write only halo points
  const Box comp{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  if (nord == 0) {
    launch_box(s, d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      dpc[o] = dg[o];
    });
    FV3_HIP_CHECK_LAUNCH(who);
  }
  FillFields<T> ff{};
  ff.f[0] = dg;
  ff.f[1] = vch;
  ff.f[2] = uch;
  // the last pass (nt = 0, no corner fills) is one launch: below
  for (int n = 0; n + 1 < nord; ++n) {
    const int nt = nord - (n + 1);
    const bool fillc = (n + 1 != nord) && corner;
    if (fillc) {
      FillTable tx;
      tx.n = 0;
      append_rotation_fill(&tx, &hd, 0, 0, 0, 1.0);
      launch_fill(s, d, tx, ff);
      FV3_HIP_CHECK_LAUNCH(who);
    }
    const Box bv{d.isc - nt - 1, d.iec + nt + 1, d.jsc - nt, d.jec + nt + 1};
    const bool first = n == 0;
    launch_box(s, d, bv, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      vch[o] = (dg[o + 1] - dg[o]) * m.divg_u[at2(d, i, j, t)];
      if (first && i >= d.isc && i <= d.iec + 1 && j >= d.jsc && j <= d.jec + 1) dpc[o] = dg[o];
    });
    FV3_HIP_CHECK_LAUNCH(who);
    if (fillc) {
      FillTable ty;
      ty.n = 0;
      append_rotation_fill(&ty, &hd, 1, 0, 0, 1.0);
      launch_fill(s, d, ty, ff);
      FV3_HIP_CHECK_LAUNCH(who);
    }
    const Box bu{d.isc - nt, d.iec + nt + 1, d.jsc - nt - 1, d.jec + nt + 1};
    launch_box(s, d, bu, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      uch[o] = (dg[o + d.sj] - dg[o]) * m.divg_v[at2(d, i, j, t)];
    });
    FV3_HIP_CHECK_LAUNCH(who);
    if (fillc) {
      FillTable td;
      td.n = 0;
      append_rotation_fill(&td, &hd, 2, 1, 2, -1.0);
      launch_fill(s, d, td, ff);
      FV3_HIP_CHECK_LAUNCH(who);
    }
    const Box br{d.isc - nt, d.iec + nt + 1, d.jsc - nt, d.jec + nt + 1};
    launch_box(s, d, br, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      const int64_t sj = d.sj;
      T val;
      const bool ci = i == d.is || i == d.ie + 1;
      if (ci && j == d.js) {
        val = vch[o - 1] - vch[o] - uch[o];
      } else if (ci && j == d.je + 1) {
        val = uch[o - sj] + vch[o - 1] - vch[o];
      } else {
        val = uch[o - sj] - uch[o] + vch[o - 1] - vch[o];
      }
      dg[o] = val * m.rarea_c[at2(d, i, j, t)] * pass_scale;
    });
    FV3_HIP_CHECK_LAUNCH(who);
  }
  // the last pass: vc on [isc-1, iec+1] x [jsc, jec+1], uc on
  // [isc, iec+1] x [jsc-1, jec+1] and the new divg_d on the corners, the
  // differences it reads formed again from divg_d where they lie (the same
  // expressions); the new divg_d goes to dg_last (its neighbours still read
  // the old one) and the final step below writes it back
  if (nord > 0) {
    const bool copy = nord == 1;
    const Box bl{d.isc - 1, d.iec + 1, d.jsc - 1, d.jec + 1};
    launch_box(s, d, bl, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      const int64_t sj = d.sj;
      auto vcx = [&](int64_t q, int ii, int jj) {
        return (dg[q + 1] - dg[q]) * m.divg_u[at2(d, ii, jj, t)];
      };
      auto ucy = [&](int64_t q, int ii, int jj) {
        return (dg[q + sj] - dg[q]) * m.divg_v[at2(d, ii, jj, t)];
      };
      const bool in_v = j >= d.jsc;   // i in [isc-1, iec+1], j in [jsc, jec+1]
      const bool in_u = i >= d.isc;   // i in [isc, iec+1], j in [jsc-1, jec+1]
      if (in_v) vch[o] = vcx(o, i, j);
      if (in_u) uch[o] = ucy(o, i, j);
      if (!(in_u && in_v)) return;
      if (copy) dpc[o] = dg[o];
      const T v_w = vcx(o - 1, i - 1, j), v_e = vcx(o, i, j);
      const T u_s = ucy(o - sj, i, j - 1), u_n = ucy(o, i, j);
      T val;
      const bool ci = i == d.is || i == d.ie + 1;
      if (ci && j == d.js) {
        val = v_w - v_e - u_n;
      } else if (ci && j == d.je + 1) {
        val = u_s + v_w - v_e;
      } else {
        val = u_s - u_n + v_w - v_e;
      }
      dgl[o] = val * m.rarea_c[at2(d, i, j, t)] * pass_scale;
    });
    FV3_HIP_CHECK_LAUNCH(who);
  }
  // _vorticity_calc
  const T dddmp = (T)p->dddmp;
  if (!(dddmp < T(1e-5))) {
    int rc = launch_a2b_ord4<T>(s, &hd, m, 0, hd.nk, wkh, vo, a2b_qx + ko, a2b_edges + ko, 0,
                                nullptr, who);
    if (rc) return rc;
  }
  const T adt = fabs(dt), da_min_c = (T)p->da_min_c;
  const T dd8 = rescale ? (T)(p->dd8 / pow(p->da_min_c, (double)nord)) : (T)p->dd8;
  const int zero_vort = dddmp < T(1e-5);
  launch_box(s, d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const T dp = dpc[o];
    T vv;
    if (zero_vort) {
      vv = T(0);
    } else {
      const T a = vo[o];
      vv = adt * sqrt(dp * dp + a * a);
    }
    const T damp = damp_tmp(vv, da_min_c, d2[k], dddmp);
    T g;
    if (nord > 0) {
      g = dgl[o];
      dg[o] = g;
    } else {
      g = dg[o];
    }
    const T nv = damp * dp + dd8 * g;
    vo[o] = nv;
    keh[o] = keh[o] + nv;
  });
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

template <typename T, int ORD>
void launch_ke_ord(hipStream_t s, const Dom& d, const Met<T>& m, T dt, const T* uc, const T* vc,
                   const T* u, const T* v, const T* ucc, const T* vcc, T* ke) {
  launch_ke<T, ORD>(s, d, m, dt, uc, vc, u, v, ucc, vcc, ke);
}

// The transports of d_sw for each storage type: fv_tp_2d with its corner
// fills, and the del-n damping of the fluxes.
inline int dsw_transport(const fv3_domain* d, int hord, double* q, double* qy, const double* crx,
                         const double* cry, const double* xfx, const double* yfx,
                         const double* xmf, const double* ymf, const double* area,
                         const double* dxa, const double* dya, double* fx, double* fy,
                         bool fp32_arithmetic, void* stream) {
  return fvtp2d_corner_fill_f64(d, hord, q, qy, crx, cry, xfx, yfx, xmf, ymf, area, dxa, dya, fx,
                                fy, fp32_arithmetic, stream);
}

inline int dsw_transport(const fv3_domain* d, int hord, float* q, float* qy, const float* crx,
                         const float* cry, const float* xfx, const float* yfx, const float* xmf,
                         const float* ymf, const float* area, const float* dxa, const float* dya,
                         float* fx, float* fy, bool, void* stream) {
  return fvtp2d_corner_fill_f32(d, hord, q, qy, crx, cry, xfx, yfx, xmf, ymf, area, dxa, dya, fx,
                                fy, stream);
}

inline int dsw_delnflux(const fv3_domain* d, int mode, const int* nord4, const double* q,
                        const double* mass, const double* damp, const double* del6_u,
                        const double* del6_v, const double* rarea, double* fx, double* fy,
                        void* stream) {
  return fv3_delnflux_f64(d, mode, nord4, q, mass, damp, del6_u, del6_v, rarea, fx, fy, stream);
}

inline int dsw_delnflux(const fv3_domain* d, int mode, const int* nord4, const float* q,
                        const float* mass, const float* damp, const float* del6_u,
                        const float* del6_v, const float* rarea, float* fx, float* fy,
                        void* stream) {
  return fv3_delnflux_f32(d, mode, nord4, q, mass, damp, del6_u, del6_v, rarea, fx, fy, stream);
}

int check_dsw_metrics(const fv3_metrics* g, const char* who) {
  FV3_REQUIRE(g != nullptr, FV3_E_ARG, "%s: NULL metrics", who);
  const void* need[] = {g->dx, g->dy, g->dxa, g->dya, g->dxc, g->dyc, g->rdx, g->rdy, g->rdxa,
                        g->rdya, g->area, g->rarea, g->rarea_c, g->sin_sg1, g->sin_sg2,
                        g->sin_sg3, g->sin_sg4, g->cosa, g->cosa_u, g->cosa_v, g->cosa_s,
                        g->sina_u, g->sina_v, g->rsin_u, g->rsin_v, g->rsina, g->rsin2, g->f0,
                        g->del6_u, g->del6_v, g->divg_u, g->divg_v};
  for (const void* q : need) FV3_REQUIRE(q != nullptr, FV3_E_ARG, "%s: a metric term is NULL", who);
  return check_a2b_metrics(g, who);
}

template <typename T, typename ST>
int run_d_sw(const fv3_domain* dom, const fv3_metrics* g, const fv3_dsw_params* p,
             const ST* state, T dt, T* work, int64_t work_slab, void* stream) {
  const char* who = "fv3_d_sw";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  rc = check_dsw_metrics(g, who);
  if (rc) return rc;
  FV3_REQUIRE(p && state && work, FV3_E_ARG, "%s: NULL argument", who);
  const DswState<T> st = dsw_state<T>(state);
  FV3_REQUIRE(st.delpc && st.delp && st.ptc && st.pt && st.u && st.v && st.w && st.uc &&
                  st.vc && st.ua && st.va && st.divgd && st.mfx && st.mfy && st.cx &&
                  st.cy && st.crx && st.cry && st.xfx && st.yfx && st.q_con &&
                  st.heat_source && st.diss_est,
              FV3_E_ARG, "%s: NULL state member", who);
  FV3_REQUIRE(p->damp_vt_deln && p->damp_t_deln && p->damp_w_nosg && p->damp_v_nosg &&
                  p->damp_w && p->damp_vt && p->ke_bg && p->d_con_col && p->d2_divg,
              FV3_E_ARG, "%s: NULL column", who);
  const DswCols<T> cols = dsw_cols<T>(p);
  const int hords[4] = {p->hord_dp, p->hord_tm, p->hord_mt, p->hord_vt};
  for (int h : hords)
    FV3_REQUIRE(h >= 5 && h <= 8, FV3_E_UNSUPPORTED, "%s: hord %d not in {5, 6, 7, 8}", who, h);
  FV3_REQUIRE(p->div_nord >= 0 && p->div_nord <= 3 && p->div_k0 >= 0, FV3_E_ARG,
              "%s: divergence-damping order must be 0..3", who);
  FV3_REQUIRE(dom->iec - dom->isc >= 4 && dom->jec - dom->jsc >= 4, FV3_E_ARG,
              "%s: compute domain must be at least 5 x 5", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  const int nk1 = dom->nk + 1;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)nk1 * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  hipStream_t s = as_stream(stream);
  const Met<T> m = to_met<T>(g);
  Dom d = to_dom(dom);
  fv3_domain d1 = *dom;  // transport levels: nk + 1
  d1.nk = nk1;
  T* W[22];
  for (int n = 0; n < 22; ++n) W[n] = work + (int64_t)n * work_slab;
  T *ucc = W[0], *vcc = W[1], *fx = W[2], *fy = W[3], *gx = W[4], *gy = W[5];
  T *fx2 = W[6], *fy2 = W[7], *dw = W[8], *heat_s = W[9], *wk = W[10], *vort = W[11];
  T *ke = W[12], *vxd = W[13], *vyd = W[14], *ut = W[15], *vt = W[16], *qy = W[17];
  T *a2x = W[18], *a2e = W[20], *vort2 = W[21];  // W[19]: the last divergence pass
  const T* area = m.area;
  const T* dxa = m.dxa;
  const T* dya = m.dya;
  const T* del6_u = m.del6_u;
  const T* del6_v = m.del6_v;
  const T* rarea = m.rarea;

  // transport of q: corner copies (x in place, y applied as the transport
  // reads q), fv_tp_2d on nk + 1 levels, then del-n damping of the fluxes on
  // nk levels
  auto transport = [&](T* q, int hord, const T* mfx_, const T* mfy_, const int* nord4,
                       const T* damp, const T* mass, T* ox, T* oy) -> int {
    int r = dsw_transport(&d1, hord, q, qy, st.crx, st.cry, st.xfx, st.yfx, mfx_, mfy_, area, dxa,
                          dya, ox, oy, p->transport_fp32 != 0, stream);
    if (r) return r;
    if (nord4) {
      r = dsw_delnflux(dom, mass ? FV3_DELN_MASS : FV3_DELN_ADD, nord4, q, mass, damp, del6_u,
                       del6_v, rarea, ox, oy, stream);
    }
    return r;
  };

  rc = run_fxadv<T>(s, d, m, dt, st.uc, st.vc, st.crx, st.cry, st.xfx, st.yfx, ucc, vcc, who);
  if (rc) return rc;
  rc = transport(st.delp, p->hord_dp, nullptr, nullptr, p->nord_v4, cols.damp_vt_deln, nullptr,
                 fx, fy);
  if (rc) return rc;
  const int nonhydro = !p->hydrostatic;
  if (nonhydro) {
    rc = dsw_delnflux(dom, FV3_DELN_STORE, p->nord_w4, st.w, nullptr, cols.damp_w_nosg, del6_u,
                      del6_v, rarea, fx2, fy2, stream);
    if (rc) return rc;
  }
  launch_capacitor_heat<T>(s, d, m, dt, nonhydro, cols, st, fx, fy, fx2, fy2, dw, heat_s);
  FV3_HIP_CHECK_LAUNCH(who);
  if (nonhydro) {
    rc = transport(st.w, p->hord_vt, fx, fy, nullptr, nullptr, nullptr, gx, gy);
    if (rc) return rc;
    launch_flux_adjust<T>(s, d, m, st.w, st.delp, gx, gy);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  rc = transport(st.q_con, p->hord_dp, fx, fy, p->nord_t4, cols.damp_t_deln, st.delp, gx, gy);
  if (rc) return rc;
  launch_flux_adjust<T>(s, d, m, st.q_con, st.delp, gx, gy);
  FV3_HIP_CHECK_LAUNCH(who);
  rc = transport(st.pt, p->hord_tm, fx, fy, p->nord_v4, cols.damp_vt_deln, st.delp, gx, gy);
  if (rc) return rc;
  launch_pt_delp<T>(s, d, m, st.pt, st.delp, gx, gy, fx, fy);
  FV3_HIP_CHECK_LAUNCH(who);
  switch (p->hord_mt) {
    case 5: launch_ke_ord<T, 5>(s, d, m, dt, st.uc, st.vc, st.u, st.v, ucc, vcc, ke); break;
    case 6: launch_ke_ord<T, 6>(s, d, m, dt, st.uc, st.vc, st.u, st.v, ucc, vcc, ke); break;
    case 7: launch_ke_ord<T, 7>(s, d, m, dt, st.uc, st.vc, st.u, st.v, ucc, vcc, ke); break;
    default: launch_ke_ord<T, 8>(s, d, m, dt, st.uc, st.vc, st.u, st.v, ucc, vcc, ke); break;
  }
  FV3_HIP_CHECK_LAUNCH(who);
  launch_vorticity_adjust<T>(s, d, m, st.u, st.v, wk, st.w, st.delp, dw, st.q_con, cols.damp_w);
  FV3_HIP_CHECK_LAUNCH(who);
  // divergence damping
  const int k_lo = p->div_k0 < dom->nk ? p->div_k0 : dom->nk;
  if (k_lo > 0) {
    launch_div_low<T>(s, d, m, k_lo, dt, (T)p->da_min_c, (T)p->dddmp, cols.d2_divg, st.u, st.v,
                      st.ua, st.va, st.uc, st.vc, st.delpc, st.ptc, vort, ke);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  rc = run_div_high<T>(s, dom, m, p, cols, dt, wk, st.divgd, st.vc, st.uc, st.delpc, vort, ke,
                       a2x, a2e, W[19], who);
  if (rc) return rc;
  launch_vort_diff<T>(s, d, m, vort, wk, vxd, vyd, vort2, cols.d_con_col);
  FV3_HIP_CHECK_LAUNCH(who);
  rc = transport(vort2, p->hord_vt, nullptr, nullptr, nullptr, nullptr, nullptr, fx, fy);
  if (rc) return rc;
  launch_uv_from_ke<T>(s, d, m, ke, fx, fy, st.u, st.v);
  FV3_HIP_CHECK_LAUNCH(who);
  rc = dsw_delnflux(dom, FV3_DELN_STORE, p->nord_v4, wk, nullptr, cols.damp_v_nosg, del6_u,
                    del6_v, rarea, ut, vt, stream);
  if (rc) return rc;
  launch_heat_source<T>(s, d, m, vxd, vyd, ut, vt, st.u, st.v, st.delp, heat_s, st.heat_source,
                        cols.d_con_col, p->d_con > kDconThreshold);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_update_uv<T>(s, d, ut, vt, st.u, st.v, cols.damp_vt);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // namespace

}  // namespace fv3

extern "C" {

int32_t fv3_d_sw_f64(const fv3_domain* dom, const fv3_metrics* g, const fv3_dsw_params* p,
                     const fv3_dsw_state* st, double dt, double* work, int64_t work_slab,
                     void* stream) {
  return fv3::run_d_sw<double>(dom, g, p, st, dt, work, work_slab, stream);
}

int32_t fv3_d_sw_f32(const fv3_domain* dom, const fv3_metrics* g, const fv3_dsw_params* p,
                     const fv3_dsw_state_f32* st, float dt, float* work, int64_t work_slab,
                     void* stream) {
  return fv3::run_d_sw<float>(dom, g, p, st, dt, work, work_slab, stream);
}
}
