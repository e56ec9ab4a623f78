// C-grid half step: DGrid2AGrid2CGridVectors (d2a2c_vect) and
// CGridShallowWaterDynamics (c_sw).
//
// Reference: fv3core/stencils/d2a2c_vect.py:1-651, fv3core/stencils/c_sw.py:
// 1-659.  Each launch is one pointwise pass over an index box; statements of
// the reference that only feed a neighbour's value (GT4Py temporaries read
// at an offset) are recomputed at that neighbour instead of being stored.
// Region statements become index tests against the tile-edge anchors, with
// the reference's statement order deciding which formula a point gets.
//
// Launch sequence of fv3_c_sw (workspace slabs in brackets):
//   1 d2a:    utmp, vtmp on the full domain [W0, W1]; ua, va
//   2 fill:   x corner fills of utmp (3 cells) and ua (2 cells)
//   3 uc:     uc, ut (ut_main + east_west_edges, then geoadjust_ut)
//   4 fill:   y corner fills of vtmp and va; x fills of delp, pt, w
//   5 vc:     vc, vt (north_south_edges + vt_main, then geoadjust_vt)
//   6 fluxx:  fx, fx1, fx2 [W2..W4]; divergence_corner -> divgd (nord > 0)
//   7 fill:   y corner fills of delp, pt, w
//   8 trans:  delpc, ptc, omga, ke [W5]
//   9 vort:   circulation_cgrid + absolute_vorticity [W6]
//  10 update: update_y_velocity (vc) and update_x_velocity (uc)
#include "stencil.hpp"

namespace fv3 {

namespace {

// Lagrange weights of a2b_ord4 (a2b_ord4.py:26-34) and the volume-conserving
// cubic weights of d2a2c_vect.py:12-15.
constexpr double kA1 = 9.0 / 16.0;
constexpr double kA2 = -1.0 / 16.0;
constexpr double kC1 = -2.0 / 14.0;
constexpr double kC2 = 11.0 / 14.0;
constexpr double kC3 = 5.0 / 14.0;
constexpr double kBigNumber = 1e30;

// Index bounds of DGrid2AGrid2CGridVectors.__init__ (d2a2c_vect.py:374-470).
struct D2A2CBounds {
  int is1, ie1, js1, je1, is2, ie2, js2, je2;
  int ifirst, ilast, jfirst, jlast;
  int avg_offset;
};

D2A2CBounds d2a2c_bounds(const fv3_domain* d) {
  const int nxc = d->iec - d->isc + 1, nyc = d->jec - d->jsc + 1;
  const int nx = d->iec + 1, ny = d->jec + 1;
  const int h = d->isc;  // n_halo
  const bool w = has_edge(d->i_start), e = has_edge(d->i_end);
  const bool s = has_edge(d->j_start), n = has_edge(d->j_end);
  int npt = 4;  // not nested
  if (npt > nxc - 1 || npt > nyc - 1) npt = 0;
  const int off = 2;
  D2A2CBounds b;
  b.js1 = s ? npt + off : d->jsc - 1;
  b.je1 = n ? ny - npt : d->jec + 1;
  b.is1 = w ? npt + off : d->isc - h;
  b.ie1 = e ? nx - npt : d->iec + h;
  b.is2 = w ? npt + off : d->isc - 1;
  b.ie2 = e ? nx - npt : d->iec + 1;
  b.js2 = s ? npt + off : d->jsc - h;
  b.je2 = n ? ny - npt : d->jec + h;
  b.ifirst = w ? d->isc + 2 : d->isc - 1;
  b.ilast = e ? d->iec - 1 : d->iec + 2;
  b.jfirst = s ? d->jsc + 2 : d->jsc - 1;
  b.jlast = n ? d->jec - 1 : d->jec + 2;
  b.avg_offset = npt == 0 ? -1 : 3;
  return b;
}

template <typename T>
__device__ __forceinline__ T contravariant(T v1, T v2, T cosa, T rsin2) {
  return (v1 - v2 * cosa) * rsin2;
}

// ---------------------------------------------------------------- d2a2c_vect
// set_tmps, lagrange_interpolation_y_p1/x_p1, avg_box, contravariant_components
template <typename T>
void launch_d2a(hipStream_t s, const Dom& d, const D2A2CBounds& b, const Met<T>& m,
                const T* u, const T* v, T* utmp, T* vtmp, T* ua, T* va) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int off = b.avg_offset;
    const bool band = j < d.js + off || j > d.je - off || i < d.is + off || i > d.ie - off;
    T ut, vt;
    if (band) {
      ut = T(0.5) * (u[o] + u[o + d.sj]);
      vt = T(0.5) * (v[o] + v[o + 1]);
    } else {
      ut = (i >= b.is1 && i <= b.ie1 && j >= b.js1 && j <= b.je1)
               ? T(kA2) * (u[o - d.sj] + u[o + 2 * d.sj]) + T(kA1) * (u[o] + u[o + d.sj])
               : T(kBigNumber);
      vt = (i >= b.is2 && i <= b.ie2 && j >= b.js2 && j <= b.je2)
               ? T(kA2) * (v[o - 1] + v[o + 2]) + T(kA1) * (v[o] + v[o + 1])
               : T(kBigNumber);
    }
    utmp[o] = ut;
    vtmp[o] = vt;
    if (i >= d.isc - 2 && i <= d.iec + 2 && j >= d.jsc - 2 && j <= d.jec + 2) {
      const int64_t o2 = at2(d, i, j, t);
      const T cs = m.cosa_s[o2], r2 = m.rsin2[o2];
      ua[o] = contravariant(ut, vt, cs, r2);
      va[o] = contravariant(vt, ut, cs, r2);
    }
  });
}

// ut_main + east_west_edges (+ geoadjust_ut when geo != 0), on
// [isc-1, iec+2] x [jsc-1, jec+1].
template <typename T>
void launch_uc(hipStream_t s, const Dom& d, const D2A2CBounds& b, const Met<T>& m, int geo,
               T dt2, const T* utmp, const T* ua, const T* v, T* uc, T* utc) {
  const Box bx{d.isc - 1, d.iec + 2, d.jsc - 1, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    T ucv, utv;
    if (i == d.is - 1 || i == d.ie) {
      ucv = T(kC1) * utmp[o - 2] + T(kC2) * utmp[o - 1] + T(kC3) * utmp[o];
      utv = contravariant(ucv, v[o], m.cosa_u[o2], m.rsin_u[o2]);
    } else if (i == d.is || i == d.ie + 1) {
      // edge_interpolate4_x (d2a2c_vect.py:254-260; n2 uses t1 as there)
      const T dm2 = m.dxa[o2 - 2], dm1 = m.dxa[o2 - 1], d0 = m.dxa[o2], dp1 = m.dxa[o2 + 1];
      const T t1 = dm2 + dm1, t2 = d0 + dp1;
      const T n1 = (t1 + dm1) * ua[o - 1] - dm1 * ua[o - 2];
      const T n2 = (t1 + d0) * ua[o] - d0 * ua[o + 1];
      utv = T(0.5) * (n1 / t1 + n2 / t2);
      ucv = utv > T(0) ? utv * m.sin_sg3[o2 - 1] : utv * m.sin_sg1[o2];
    } else if (i == d.is + 1 || i == d.ie + 2) {
      ucv = T(kC1) * utmp[o + 1] + T(kC2) * utmp[o] + T(kC3) * utmp[o - 1];
      utv = contravariant(ucv, v[o], m.cosa_u[o2], m.rsin_u[o2]);
    } else if (i >= b.ifirst && i <= b.ilast) {
      ucv = T(kA2) * (utmp[o - 2] + utmp[o + 1]) + T(kA1) * (utmp[o - 1] + utmp[o]);
      utv = contravariant(ucv, v[o], m.cosa_u[o2], m.rsin_u[o2]);
    } else {
      return;
    }
    uc[o] = ucv;
    if (geo) {
      utv = utv > T(0) ? dt2 * utv * m.dy[o2] * m.sin_sg3[o2 - 1]
                       : dt2 * utv * m.dy[o2] * m.sin_sg1[o2];
    }
    utc[o] = utv;
  });
}

// north_south_edges + vt_main (+ geoadjust_vt), on [isc-1, iec+1] x [jsc-1, jec+2].
template <typename T>
void launch_vc(hipStream_t s, const Dom& d, const D2A2CBounds& b, const Met<T>& m, int geo,
               T dt2, const T* vtmp, const T* va, const T* u, T* vc, T* vtc) {
  const Box bx{d.isc - 1, d.iec + 1, d.jsc - 1, d.jec + 2};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj, sj2 = d.sj2;
    T vcv, vtv;
    if (j >= b.jfirst && j <= b.jlast) {
      vcv = T(kA2) * (vtmp[o - 2 * sj] + vtmp[o + sj]) + T(kA1) * (vtmp[o - sj] + vtmp[o]);
      vtv = contravariant(vcv, u[o], m.cosa_v[o2], m.rsin_v[o2]);
    } else if (j == d.js - 1 || j == d.je) {
      vcv = T(kC1) * vtmp[o - 2 * sj] + T(kC2) * vtmp[o - sj] + T(kC3) * vtmp[o];
      vtv = contravariant(vcv, u[o], m.cosa_v[o2], m.rsin_v[o2]);
    } else if (j == d.js || j == d.je + 1) {
      const T dm2 = m.dya[o2 - 2 * sj2], dm1 = m.dya[o2 - sj2], d0 = m.dya[o2],
              dp1 = m.dya[o2 + sj2];
      const T t1 = dm2 + dm1, t2 = d0 + dp1;
      const T n1 = (t1 + dm1) * va[o - sj] - dm1 * va[o - 2 * sj];
      const T n2 = (t1 + d0) * va[o] - d0 * va[o + sj];
      vtv = T(0.5) * (n1 / t1 + n2 / t2);
      vcv = vtv > T(0) ? vtv * m.sin_sg4[o2 - sj2] : vtv * m.sin_sg2[o2];
    } else if (j == d.js + 1 || j == d.je + 2) {
      vcv = T(kC1) * vtmp[o + sj] + T(kC2) * vtmp[o] + T(kC3) * vtmp[o - sj];
      vtv = contravariant(vcv, u[o], m.cosa_v[o2], m.rsin_v[o2]);
    } else {
      vcv = T(kA2) * (vtmp[o - 2 * sj] + vtmp[o + sj]) + T(kA1) * (vtmp[o - sj] + vtmp[o]);
      vtv = contravariant(vcv, u[o], m.cosa_v[o2], m.rsin_v[o2]);
    }
    vc[o] = vcv;
    if (geo) {
      vtv = vtv > T(0) ? dt2 * vtv * m.dx[o2] * m.sin_sg4[o2 - sj2]
                       : dt2 * vtv * m.dx[o2] * m.sin_sg2[o2];
    }
    vtc[o] = vtv;
  });
}

const int kCornerMult[4] = {-1, 1, 1, -1};  // sw, se, nw, ne (d2a2c_vect.py:24-38)
const int kUnitMult[4] = {1, 1, 1, 1};

// The d2a2c part shared by fv3_d2a2c_vect and fv3_c_sw.  When extra_x is
// set, the x corner fills of delp, pt, w (fields 4, 5, 6 of ff) ride along
// with the y fills of vtmp and va.
template <typename T>
int run_d2a2c(hipStream_t s, const fv3_domain* dom, const Met<T>& m, int geo, T dt2, T* uc,
              T* vc, const T* u, const T* v, T* ua, T* va, T* utc, T* vtc, T* utmp, T* vtmp,
              T* delp, T* pt, T* w, const char* who) {
  Dom d = to_dom(dom);
  const D2A2CBounds b = d2a2c_bounds(dom);
  const Box edges{dom->isc - 3, dom->iec + 3, dom->jsc - 3, dom->jec + 3};
  FillFields<T> ff{};
  ff.f[0] = utmp;
  ff.f[1] = vtmp;
  ff.f[2] = ua;
  ff.f[3] = va;
  ff.f[4] = delp;
  ff.f[5] = pt;
  ff.f[6] = w;
  launch_d2a(s, d, b, m, u, v, utmp, vtmp, ua, va);
  FV3_HIP_CHECK_LAUNCH(who);
  FillTable fx;
  fx.n = 0;
  append_fill(&fx, dom, kFill3X, 0, 1, kCornerMult, edges);
  append_fill(&fx, dom, kFill2X, 2, 3, kCornerMult, edges);
  launch_fill(s, d, fx, ff);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_uc(s, d, b, m, geo, dt2, utmp, ua, v, uc, utc);
  FV3_HIP_CHECK_LAUNCH(who);
  FillTable fy;
  fy.n = 0;
  append_fill(&fy, dom, kFill3Y, 1, 0, kCornerMult, edges);
  append_fill(&fy, dom, kFill2Y, 3, 2, kCornerMult, edges);
  if (delp) {
    for (int f = 4; f <= 6; ++f) append_fill(&fy, dom, kFill2X, f, f, kUnitMult, edges);
  }
  launch_fill(s, d, fy, ff);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_vc(s, d, b, m, geo, dt2, vtmp, va, u, vc, vtc);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

// ----------------------------------------------------------------------- c_sw
// compute_nonhydro_fluxes_x on [isc-1, iec+2] x [jsc-1, jec+1] and, when
// divgd is given, divergence_corner on [isc, iec+1] x [jsc, jec+1].
template <typename T>
void launch_fluxx_div(hipStream_t s, const Dom& d, const Met<T>& m, const T* delp, const T* pt,
                      const T* w, const T* utc, T* fx, T* fx1, T* fx2, const T* u, const T* v,
                      const T* ua, const T* va, T* divg) {
  const Box bx{d.isc - 1, d.iec + 2, d.jsc - 1, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    {
      const T c = utc[o];
      const int64_t up = c > T(0) ? o - 1 : o;
      const T f1 = c * delp[up];
      fx1[o] = f1;
      fx[o] = f1 * pt[up];
      fx2[o] = f1 * w[up];
    }
    if (divg == nullptr || i < d.isc || i > d.iec + 1 || j < d.jsc) return;
    // divergence_corner (c_sw.py:172-289)
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj, sj2 = d.sj2;
    const T half = T(0.5), quarter = T(0.25);
    const T dxc = m.dxc[o2], dyc = m.dyc[o2];
    const T s1 = m.sin_sg1[o2], s2 = m.sin_sg2[o2];
    const T s3m = m.sin_sg3[o2 - 1], s4m = m.sin_sg4[o2 - sj2];
    const T ra = m.rarea_c[o2];
    const bool west = i == d.is || i == d.ie + 1;
    const bool south = j == d.js || j == d.je + 1;
    // uf, vf at the point; uf1 (i-1), vf1 (j-1) in the variants the regions use
    const T uf = (u[o] - quarter * (va[o - sj] + va[o]) * (m.cos_sg4[o2 - sj2] + m.cos_sg2[o2])) *
                 dyc * half * (s4m + s2);
    const T vf = (v[o] - quarter * (ua[o - 1] + ua[o]) * (m.cos_sg3[o2 - 1] + m.cos_sg1[o2])) *
                 dxc * half * (s3m + s1);
    T out;
    if (west && south) {
      const T uf1 = u[o - 1] * m.dyc[o2 - 1] * half * (m.sin_sg4[o2 - 1 - sj2] + m.sin_sg2[o2 - 1]);
      const T uf0 = u[o] * dyc * half * (s4m + s2);
      const T vf0 = v[o] * dxc * half * (s3m + s1);
      if (j == d.js) {
        out = (-vf0 + uf1 - uf0) * ra;
      } else {
        const T vf1 =
            v[o - sj] * m.dxc[o2 - sj2] * half * (m.sin_sg3[o2 - 1 - sj2] + m.sin_sg1[o2 - sj2]);
        out = (vf1 + uf1 - uf0) * ra;
      }
    } else if (south) {
      const T uf0 = u[o] * dyc * half * (s4m + s2);
      const T uf1 = u[o - 1] * m.dyc[o2 - 1] * half * (m.sin_sg4[o2 - 1 - sj2] + m.sin_sg2[o2 - 1]);
      const T vf1 = (v[o - sj] - quarter * (ua[o - 1 - sj] + ua[o - sj]) *
                                     (m.cos_sg3[o2 - 1 - sj2] + m.cos_sg1[o2 - sj2])) *
                    m.dxc[o2 - sj2] * half * (m.sin_sg3[o2 - 1 - sj2] + m.sin_sg1[o2 - sj2]);
      out = (vf1 - vf + uf1 - uf0) * ra;
    } else if (west) {
      const T vf0 = v[o] * dxc * half * (s3m + s1);
      const T vf1 =
          v[o - sj] * m.dxc[o2 - sj2] * half * (m.sin_sg3[o2 - 1 - sj2] + m.sin_sg1[o2 - sj2]);
      const T uf1 = (u[o - 1] - quarter * (va[o - 1 - sj] + va[o - 1]) *
                                    (m.cos_sg4[o2 - 1 - sj2] + m.cos_sg2[o2 - 1])) *
                    m.dyc[o2 - 1] * half * (m.sin_sg4[o2 - 1 - sj2] + m.sin_sg2[o2 - 1]);
      out = (vf1 - vf0 + uf1 - uf) * ra;
    } else {
      const T vfm = (v[o - sj] - quarter * (ua[o - 1 - sj] + ua[o - sj]) *
                                     (m.cos_sg3[o2 - 1 - sj2] + m.cos_sg1[o2 - sj2])) *
                    m.dxc[o2 - sj2] * half * (m.sin_sg3[o2 - 1 - sj2] + m.sin_sg1[o2 - sj2]);
      const T ufm = (u[o - 1] - quarter * (va[o - 1 - sj] + va[o - 1]) *
                                    (m.cos_sg4[o2 - 1 - sj2] + m.cos_sg2[o2 - 1])) *
                    m.dyc[o2 - 1] * half * (m.sin_sg4[o2 - 1 - sj2] + m.sin_sg2[o2 - 1]);
      out = (vfm - vf + ufm - uf) * ra;
    }
    divg[o] = out;
  });
}

// transportdelp_update_vorticity_and_kineticenergy on [isc-1, iec+1]^2, with
// initialize_delpc_ptc (zero) on the rest of the full domain.
template <typename T>
void launch_transport(hipStream_t s, const Dom& d, const Met<T>& m, T dt2, const T* delp,
                      const T* pt, const T* w, const T* utc, const T* vtc, const T* fx,
                      const T* fx1, const T* fx2, const T* ua, const T* va, const T* uc,
                      const T* vc, const T* u, const T* v, T* delpc, T* ptc, T* wc, T* ke) {
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    if (i < d.isc - 1 || i > d.iec + 1 || j < d.jsc - 1 || j > d.jec + 1) {
      delpc[o] = T(0);
      ptc[o] = T(0);
      return;
    }
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sj = d.sj;
    // y fluxes at j and j + 1
    const T c0 = vtc[o], c1 = vtc[o + sj];
    const int64_t up0 = c0 > T(0) ? o - sj : o;
    const int64_t up1 = c1 > T(0) ? o : o + sj;
    const T fy1_0 = c0 * delp[up0], fy1_1 = c1 * delp[up1];
    const T fy_0 = fy1_0 * pt[up0], fy_1 = fy1_1 * pt[up1];
    const T fy2_0 = fy1_0 * w[up0], fy2_1 = fy1_1 * w[up1];
    const T ra = m.rarea[o2];
    const T dp = delp[o];
    const T dpc = dp + (fx1[o] - fx1[o + 1] + fy1_0 - fy1_1) * ra;
    delpc[o] = dpc;
    ptc[o] = (pt[o] * dp + (fx[o] - fx[o + 1] + fy_0 - fy_1) * ra) / dpc;
    wc[o] = (w[o] * dp + (fx2[o] - fx2[o + 1] + fy2_0 - fy2_1) * ra) / dpc;
    // kinetic energy and vorticity on the A grid
    const T uav = ua[o], vav = va[o];
    T kev = uav > T(0) ? uc[o] : uc[o + 1];
    T vort = vav > T(0) ? vc[o] : vc[o + sj];
    if (j == d.js - 1 || j == d.je) {
      if (vav <= T(0)) vort = vort * m.sin_sg4[o2] + u[o + sj] * m.cos_sg4[o2];
    } else if (j == d.js || j == d.je + 1) {
      if (vav > T(0)) vort = vort * m.sin_sg2[o2] + u[o] * m.cos_sg2[o2];
    }
    if (i == d.ie || i == d.is - 1) {
      if (uav <= T(0)) kev = kev * m.sin_sg3[o2] + v[o + 1] * m.cos_sg3[o2];
    } else if (i == d.ie + 1 || i == d.is) {
      if (uav > T(0)) kev = kev * m.sin_sg1[o2] + v[o] * m.cos_sg1[o2];
    }
    ke[o] = T(0.5) * dt2 * (uav * kev + vav * vort);
  });
}

// circulation_cgrid + absolute_vorticity on [isc, iec+1]^2.
template <typename T>
void launch_vort(hipStream_t s, const Dom& d, const Met<T>& m, const T* uc, const T* vc,
                 T* vort) {
  const Box bx{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const T fx = m.dxc[o2] * uc[o];
    const T fy = m.dyc[o2] * vc[o];
    const T fx1 = m.dxc[o2 - d.sj2] * uc[o - d.sj];
    const T fy1 = m.dyc[o2 - 1] * vc[o - 1];
    const bool ej = j == d.js || j == d.je + 1;
    T vc_;
    if (i == d.is && ej) {
      vc_ = fx1 - fx + fy;
    } else if (i == d.ie + 1 && ej) {
      vc_ = fx1 - fx - fy1;
    } else {
      vc_ = fx1 - fx - fy1 + fy;
    }
    vort[o] = m.fC[o2] + m.rarea_c[o2] * vc_;
  });
}

// update_y_velocity (vc on [isc, iec] x [jsc, jec+1]) and update_x_velocity
// (uc on [isc, iec+1] x [jsc, jec]) in one pass over [isc, iec+1] x [jsc, jec+1].
template <typename T>
void launch_update(hipStream_t s, const Dom& d, const Met<T>& m, T dt2, const T* vort,
                   const T* ke, const T* u, const T* v, T* uc, T* vc) {
  const Box bx{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    if (i <= d.iec) {
      const T vcv = vc[o];
      T tmp = (j == d.js || j == d.je + 1) ? dt2 * u[o]
                                            : dt2 * (u[o] - vcv * m.cosa_v[o2]) / m.sina_v[o2];
      const T flux = tmp > T(0) ? vort[o] : vort[o + 1];
      vc[o] = vcv - tmp * flux + m.rdyc[o2] * (ke[o - d.sj] - ke[o]);
    }
    if (j <= d.jec) {
      const T ucv = uc[o];
      T tmp = (i == d.is || i == d.ie + 1) ? dt2 * v[o]
                                            : dt2 * (v[o] - ucv * m.cosa_u[o2]) / m.sina_u[o2];
      const T flux = tmp > T(0) ? vort[o] : vort[o + d.sj];
      uc[o] = ucv + tmp * flux + m.rdxc[o2] * (ke[o - 1] - ke[o]);
    }
  });
}

int check_csw_domain(const fv3_domain* dom, const char* who) {
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(dom->isc >= 3 && dom->jsc >= 3, FV3_E_ARG, "%s: the stencils need a 3-cell halo",
              who);
  FV3_REQUIRE(dom->iec - dom->isc + 1 >= 5 && dom->jec - dom->jsc + 1 >= 5, FV3_E_ARG,
              "%s: compute domain must be at least 5 x 5", who);
  return FV3_OK;
}

#define FV3_REQ_MET(g, name, who) \
  FV3_REQUIRE((g)->name != nullptr, FV3_E_ARG, "%s: metric " #name " is NULL", who)

int check_d2a2c_metrics(const fv3_metrics* g, const char* who) {
  FV3_REQUIRE(g != nullptr, FV3_E_ARG, "%s: NULL metrics", who);
  FV3_REQ_MET(g, cosa_s, who);
  FV3_REQ_MET(g, rsin2, who);
  FV3_REQ_MET(g, cosa_u, who);
  FV3_REQ_MET(g, cosa_v, who);
  FV3_REQ_MET(g, rsin_u, who);
  FV3_REQ_MET(g, rsin_v, who);
  FV3_REQ_MET(g, dxa, who);
  FV3_REQ_MET(g, dya, who);
  FV3_REQ_MET(g, sin_sg1, who);
  FV3_REQ_MET(g, sin_sg2, who);
  FV3_REQ_MET(g, sin_sg3, who);
  FV3_REQ_MET(g, sin_sg4, who);
  return FV3_OK;
}

int check_csw_metrics(const fv3_metrics* g, int nord, const char* who) {
  int rc = check_d2a2c_metrics(g, who);
  if (rc) return rc;
  FV3_REQ_MET(g, dx, who);
  FV3_REQ_MET(g, dy, who);
  FV3_REQ_MET(g, dxc, who);
  FV3_REQ_MET(g, dyc, who);
  FV3_REQ_MET(g, rdxc, who);
  FV3_REQ_MET(g, rdyc, who);
  FV3_REQ_MET(g, rarea, who);
  FV3_REQ_MET(g, rarea_c, who);
  FV3_REQ_MET(g, cos_sg1, who);
  FV3_REQ_MET(g, cos_sg2, who);
  FV3_REQ_MET(g, cos_sg3, who);
  FV3_REQ_MET(g, cos_sg4, who);
  FV3_REQ_MET(g, sina_u, who);
  FV3_REQ_MET(g, sina_v, who);
  FV3_REQ_MET(g, fC, who);
  (void)nord;
  return FV3_OK;
}

template <typename T>
int run_c_sw(const fv3_domain* dom, const fv3_metrics* g, int32_t nord, T dt2, T* delp, T* pt,
             const T* u, const T* v, T* w, T* uc, T* vc, T* ua, T* va, T* ut, T* vt, T* divgd,
             T* omga, T* delpc, T* ptc, T* work, int64_t work_slab, void* stream) {
  const char* who = "fv3_c_sw";
  int rc = check_csw_domain(dom, who);
  if (rc) return rc;
  rc = check_csw_metrics(g, nord, who);
  if (rc) return rc;
  FV3_REQUIRE(delp && pt && u && v && w && uc && vc && ua && va && ut && vt && omga && delpc &&
                  ptc && work && (nord <= 0 || divgd),
              FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)dom->nk * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  const Met<T> m = to_met<T>(g);
  hipStream_t s = as_stream(stream);
  T* utmp = work;
  T* vtmp = work + work_slab;
  T* fx = work + 2 * work_slab;
  T* fx1 = work + 3 * work_slab;
  T* fx2 = work + 4 * work_slab;
  T* ke = work + 5 * work_slab;
  T* vort = work + 6 * work_slab;
  rc = run_d2a2c<T>(s, dom, m, 1, dt2, uc, vc, u, v, ua, va, ut, vt, utmp, vtmp, delp, pt, w,
                    who);
  if (rc) return rc;
  Dom d = to_dom(dom);
  launch_fluxx_div<T>(s, d, m, delp, pt, w, ut, fx, fx1, fx2, u, v, ua, va,
                      nord > 0 ? divgd : nullptr);
  FV3_HIP_CHECK_LAUNCH(who);
  const Box edges{dom->isc - 3, dom->iec + 3, dom->jsc - 3, dom->jec + 3};
  FillTable fy;
  fy.n = 0;
  FillFields<T> ff{};
  ff.f[0] = delp;
  ff.f[1] = pt;
  ff.f[2] = w;
  for (int f = 0; f < 3; ++f) append_fill(&fy, dom, kFill2Y, f, f, kUnitMult, edges);
  launch_fill(s, d, fy, ff);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_transport<T>(s, d, m, dt2, delp, pt, w, ut, vt, fx, fx1, fx2, ua, va, uc, vc, u, v,
                      delpc, ptc, omga, ke);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_vort<T>(s, d, m, uc, vc, vort);
  FV3_HIP_CHECK_LAUNCH(who);
  launch_update<T>(s, d, m, dt2, vort, ke, u, v, uc, vc);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // namespace

}  // namespace fv3

extern "C" {

int32_t fv3_d2a2c_vect_f64(const fv3_domain* dom, const fv3_metrics* g, double* uc, double* vc,
                           const double* u, const double* v, double* ua, double* va,
                           double* utc, double* vtc, double* work, int64_t work_slab,
                           void* stream) {
  using namespace fv3;
  const char* who = "fv3_d2a2c_vect";
  int rc = check_csw_domain(dom, who);
  if (rc) return rc;
  rc = check_d2a2c_metrics(g, who);
  if (rc) return rc;
  FV3_REQUIRE(uc && vc && u && v && ua && va && utc && vtc && work, FV3_E_ARG,
              "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)dom->nk * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  const Met<double> m = to_met<double>(g);
  return run_d2a2c<double>(as_stream(stream), dom, m, 0, 0.0, uc, vc, u, v, ua, va, utc, vtc,
                           work, work + work_slab, nullptr, nullptr, nullptr, who);
}

int32_t fv3_c_sw_f64(const fv3_domain* dom, const fv3_metrics* g, int32_t nord, double dt2,
                     double* delp, double* pt, const double* u, const double* v, double* w,
                     double* uc, double* vc, double* ua, double* va, double* ut, double* vt,
                     double* divgd, double* omga, double* delpc, double* ptc, double* work,
                     int64_t work_slab, void* stream) {
  return fv3::run_c_sw<double>(dom, g, nord, dt2, delp, pt, u, v, w, uc, vc, ua, va, ut, vt,
                               divgd, omga, delpc, ptc, work, work_slab, stream);
}

int32_t fv3_c_sw_f32(const fv3_domain* dom, const fv3_metrics* g, int32_t nord, float dt2,
                     float* delp, float* pt, const float* u, const float* v, float* w, float* uc,
                     float* vc, float* ua, float* va, float* ut, float* vt, float* divgd,
                     float* omga, float* delpc, float* ptc, float* work, int64_t work_slab,
                     void* stream) {
  return fv3::run_c_sw<float>(dom, g, nord, dt2, delp, pt, u, v, w, uc, vc, ua, va, ut, vt,
                              divgd, omga, delpc, ptc, work, work_slab, stream);
}
}
