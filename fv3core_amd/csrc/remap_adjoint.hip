// Tangent-linear model and adjoint of MapSingle (map_scalar) with kord 9
// with respect to the remapped field, at the branch decisions of the forward
// (fv3core/stencils/map_single.py:69-173 with remap_profile.py:498-638; the
// reference has no derivative code).  The source and target interfaces
// pe1 / pe2, qs and qmin are held fixed.
//
// TLM: the forward column kernel itself (remap_profile.hpp) instantiated on
// Dual<1>, whose value part repeats the double arithmetic; the fields are
// packed into (value, perturbation) pairs and unpacked afterwards.
//
// Adjoint, one thread per column, after recomputing the forward edge values:
//   (4) lagrangian_contributions is linear in the layer means a4_1 and the
//       parabola coefficients a4_2..a4_4 with weights from pe1 / pe2 only:
//       its transpose re-walks the source layers and scatters the output
//       cotangents;
//   (3) set_interpolation_coefficients acts level by level on (a4_1 of the
//       level and its two neighbours each side, the two clamped edge values):
//       its local Jacobian comes from running the same statements on Dual<7>;
//   (2) apply_constraints clamps each edge value between the neighbouring
//       means: local Jacobian on Dual<3>;
//   (1) set_initial_vals is a tridiagonal solve for the edge values, linear
//       in a4_1 with coefficients from pe1: its transpose reverses the back
//       substitution and then the forward elimination.
#include "common.hpp"
#include "dual.hpp"
#include "remap_profile.hpp"
#include "timing.hpp"

namespace fv3 {
namespace {

// Edge value kk (1 <= kk <= km-1) after apply_constraints
// (remap_profile.py:210-280): qn clamped between the means of the layers
// on either side; gm1 / gp1 are the neighbouring mean differences that
// select the clamp (values only).
template <typename T>
__device__ __forceinline__ T edge_clamp(int kk, int km, int iv, T qn, T a0, T ak, double gm1,
                                        double gp1) {
  const T tmp = pick(a0 > ak, a0, ak);
  const T tmp2 = pick(a0 < ak, a0, ak);
  if (kk == 1 || kk == km - 1) {
    if (qn >= tmp) qn = tmp;
    if (qn <= tmp2) qn = tmp2;
  } else {
    if ((gm1 * gp1) > 0.0) {
      if (qn >= tmp) qn = tmp;
      if (qn <= tmp2) qn = tmp2;
    } else if (gm1 > 0.0) {
      if (qn <= tmp2) qn = tmp2;
    } else {
      if (qn >= tmp) qn = tmp;
      if (iv == 0) {
        if (qn < T(0)) qn = T(0);
      }
    }
  }
  return qn;
}

// set_interpolation_coefficients (remap_profile.py:283-495) for kord 9 at
// level k: a1 = a4_1(k), a2 / a3 its clamped edge values, am2..ap2 the means
// of levels k-2, k-1, k+1, k+2 (used at 2 <= k < km-2), em* the extremum
// flags of levels k-1, k, k+1.  Returns a4_2, a4_3, a4_4.
template <typename T>
__device__ __forceinline__ void level_coeffs9(int k, int km, int iv, double qmin, T a1, T a2, T a3,
                                              T am2, T am1, T ap1, T ap2, bool em_m1, bool em,
                                              bool em_p1, T& o2, T& o3, T& o4) {
  T a4 = T(0);
  if (k < 2) {
    if (iv == 0) {
      if (k == 0 && a2 < T(0)) a2 = T(0);
      a4 = T(3) * (T(2) * a1 - (a2 + a3));
    } else if (iv == -1) {
      if (k == 0 && a2 * a1 <= T(0)) a2 = T(0);
      a4 = T(3) * (T(2) * a1 - (a2 + a3));
    } else if (iv == 2) {
      if (k == 0) {
        a2 = a1;
        a3 = a1;
        a4 = T(0);
      } else {
        a4 = T(3) * (T(2) * a1 - (a2 + a3));
      }
    } else {
      a4 = T(3) * (T(2) * a1 - (a2 + a3));
    }
    if (k == 0) {
      if (iv != 2) posdef_iv1(a1, a2, a3, a4);
    } else {
      remap_constraint(a1, a2, a3, a4, em);
    }
  } else if (k < km - 2) {
    const T gk = a1 - am1, gp1 = ap1 - a1, gp2 = ap2 - ap1, gm1 = am1 - am2;
    const T pmp_1 = a1 - T(2) * gp1;
    const T lac_1 = pmp_1 + T(1.5) * gp2;
    const T pmp_2 = a1 + T(2) * gk;
    const T lac_2 = pmp_2 - T(1.5) * gm1;
    if ((em && em_m1) || (em && em_p1) || (em && (T(qmin) > T(0) && a1 < T(qmin)))) {
      a2 = a1;
      a3 = a1;
      a4 = T(0);
    } else {
      a4 = T(6) * a1 - T(3) * (a2 + a3);
      if (fabs(a4) > fabs(a2 - a3)) {
        T tmin = pick_min3(a1, pmp_1, lac_1);
        T tmax0 = pick(a2 > tmin, a2, tmin);
        T tmax = pick_max3(a1, pmp_1, lac_1);
        a2 = pick(tmax0 < tmax, tmax0, tmax);
        tmin = pick_min3(a1, pmp_2, lac_2);
        tmax0 = pick(a3 > tmin, a3, tmin);
        tmax = pick_max3(a1, pmp_2, lac_2);
        a3 = pick(tmax0 < tmax, tmax0, tmax);
        a4 = T(6) * a1 - T(3) * (a2 + a3);
      }
    }
    if (iv == 0) posdef_iv0(a1, a2, a3, a4);
  } else {
    if (k == km - 1) {
      if (iv == 0) {
        if (a3 < T(0)) a3 = T(0);
      }
      if (iv == -1) {
        if (a3 * a1 <= T(0)) a3 = T(0);
      }
    }
    a4 = T(3) * (T(2) * a1 - (a2 + a3));
    if (k == km - 2)
      remap_constraint(a1, a2, a3, a4, em);
    else
      posdef_iv1(a1, a2, a3, a4);
  }
  o2 = a2;
  o3 = a3;
  o4 = a4;
}

// workspace slabs of the adjoint: 0 edge values (km+1), 1 gam (km+1),
// 2..4 cotangents of a4_2..a4_4 from the overlap walk (km), 5 cotangent of
// a4_1 from the walk (km), 6 edge cotangents after the clamps and the
// forward substitution of the transposed solve (km+1), 7 cotangent of a4_1
// before the back substitution (km)
constexpr int kAdjSlabs = 8;
constexpr int KC = kRemapKC;

// The reverse sweep, one thread per column, in five sweeps:
//   F1/F2  edge values as the forward kernel (forward elimination, back
//          substitution), running values in registers;
//   R4     transpose of lagrangian_contributions: the overlap walk keeps the
//          current source layer's cotangents in registers and writes each
//          layer's once, when the walk leaves it;
//   R321   one upward sweep that, at step s, clamps edge s+1 and flags level
//          s (as the forward's apply_constraints), takes level s-1 back
//          through its Dual<7> coefficient Jacobian (3), takes the clamp of
//          edge s-3 back through its Dual<3> Jacobian (2) and runs the
//          forward substitution of the transposed tridiagonal solve on edge
//          s-3 (1); the cotangents of the means s-4 .. s+1 and of the edges
//          s-3 .. s ride in registers, and every sum is formed in the order
//          of the phase-by-phase statement (all of (3), level by level,
//          before (2), interface by interface);
//   R1     the back substitution of the transposed solve, downward.
__global__ void __launch_bounds__(256)
    map_single_adjoint_kernel(Dom d, RemapArgs a, const double* __restrict__ q1,
                              const double* __restrict__ pe1, const double* __restrict__ pe2,
                              const double* __restrict__ qs, const double* __restrict__ gout,
                              double* __restrict__ gin, double* __restrict__ work) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)a.ni_box * a.nj_box) return;
  const int i = a.i1 + (int)(n % a.ni_box);
  const int j = a.j1 + (int)(n / a.ni_box);
  const int t = blockIdx.y;
  const int64_t c = (int64_t)t * d.st + (int64_t)j * d.sj + i;
  const int64_t sk = d.sk;
  const int km = a.km, iv = a.iv;
  Col<const double> A1{q1 + c, sk}, P1{pe1 + c, sk}, P2{pe2 + c, sk}, GO{gout + c, sk};
  Col<double> GA{gin + c, sk};
  Col<double> EV{work + c, sk}, GAM{work + a.slab + c, sk}, G2{work + 2 * a.slab + c, sk},
      G3{work + 3 * a.slab + c, sk}, G4{work + 4 * a.slab + c, sk},
      GA4{work + 5 * a.slab + c, sk}, GEP{work + 6 * a.slab + c, sk},
      GAP{work + 7 * a.slab + c, sk};
  const double qsv = qs ? qs[(int64_t)t * d.st2 + (int64_t)j * d.sj2 + i] : 0.0;
  const int go = iv == -2 ? 1 : 0;

  // ---- F1: forward elimination (as the forward kernel)
  const double q_0 = A1[0], q_1 = A1[1];
  const double p_0 = P1[0], p_1 = P1[1], p_2 = P1[2];
  double ev_prev, gam_prev, ev_last;
  {
    double a_prev = q_1, p_k = p_2, dp_pp = p_1 - p_0, dp_prev = p_2 - p_1;
    if (iv == -2) {
      const double ev0 = 1.5 * q_0;
      EV[0] = ev0;
      GAM[1] = 0.5;
      {
        const double gr = dp_pp / dp_prev;
        const double bet = 2.0 + gr + gr - 0.5;
        ev_prev = (3.0 * (q_0 + q_1) - ev0) / bet;
        EV[1] = ev_prev;
      }
      gam_prev = 0.5;
      for (int k0 = 2; k0 < km; k0 += KC) {
        double qv[KC], pv[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
          const int k = min(k0 + u, km - 1);
          qv[u] = A1[k];
          pv[u] = P1[k + 1];
        }
#pragma unroll
        for (int u = 0; u < KC; ++u) {
          const int k = k0 + u;
          if (k < km) {
            const double dpk = pv[u] - p_k;
            const double ogr = dp_pp / dp_prev;
            const double obet = 2.0 + ogr + ogr - gam_prev;
            const double g = ogr / obet;
            GAM[k] = g;
            const double gr = dp_prev / dpk;
            double ev;
            if (k < km - 1) {
              const double bet = 2.0 + gr + gr - g;
              ev = (3.0 * (a_prev + qv[u]) - ev_prev) / bet;
            } else {
              ev = (3.0 * (a_prev + qv[u]) - gr * qsv - ev_prev) / (2.0 + gr + gr - g);
            }
            EV[k] = ev;
            ev_prev = ev;
            gam_prev = g;
            a_prev = qv[u];
            p_k = pv[u];
            dp_pp = dp_prev;
            dp_prev = dpk;
          }
        }
      }
      EV[km] = qsv;
      ev_last = ev_prev;
    } else {
      {
        const double gr = dp_prev / dp_pp;
        const double bet = gr * (gr + 0.5);
        ev_prev = ((gr + gr) * (gr + 1.0) * q_0 + q_1) / bet;
        gam_prev = (1.0 + gr * (gr + 1.5)) / bet;
        EV[0] = ev_prev;
        GAM[0] = gam_prev;
      }
      {
        const double d4 = dp_pp / dp_prev;
        const double bet = 2.0 + d4 + d4 - gam_prev;
        ev_prev = (3.0 * (q_0 + d4 * q_1) - ev_prev) / bet;
        gam_prev = d4 / bet;
        EV[1] = ev_prev;
        GAM[1] = gam_prev;
      }
      double a_pp = q_0;
      for (int k0 = 2; k0 < km; k0 += KC) {
        double qv[KC], pv[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
          const int k = min(k0 + u, km - 1);
          qv[u] = A1[k];
          pv[u] = P1[k + 1];
        }
#pragma unroll
        for (int u = 0; u < KC; ++u) {
          const int k = k0 + u;
          if (k < km) {
            const double dpk = pv[u] - p_k;
            const double d4 = dp_prev / dpk;
            const double bet = 2.0 + d4 + d4 - gam_prev;
            const double ev = (3.0 * (a_prev + d4 * qv[u]) - ev_prev) / bet;
            const double g = d4 / bet;
            EV[k] = ev;
            GAM[k] = g;
            ev_prev = ev;
            gam_prev = g;
            a_pp = a_prev;
            a_prev = qv[u];
            p_k = pv[u];
            dp_pp = dp_prev;
            dp_prev = dpk;
          }
        }
      }
      const double d4 = dp_pp / dp_prev;
      const double a_bot = 1.0 + d4 * (d4 + 1.5);
      ev_last = (2.0 * d4 * (d4 + 1.0) * a_prev + a_pp - a_bot * ev_prev) /
                (d4 * (d4 + 0.5) - a_bot * gam_prev);
      EV[km] = ev_last;
    }
  }

  // ---- F2: back substitution
  {
    double ev_next = ev_last;
    for (int k0 = km - 1 - go; k0 >= 0; k0 -= KC) {
      double ev[KC], g[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = max(k0 - u, 0);
        ev[u] = EV[k];
        g[u] = GAM[k + go];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 - u;
        if (k >= 0) {
          ev_next = ev[u] - g[u] * ev_next;
          EV[k] = ev_next;
        }
      }
    }
  }

  // ---- R4: transpose of lagrangian_contributions
  {
    int L = 0;
    double pL = p_0, pR = p_1, nR = p_2;
    double g2 = 0.0, g3 = 0.0, g4 = 0.0, ga4 = 0.0;  // cotangents of layer L
    auto leave = [&]() __attribute__((always_inline)) {  // L -> L + 1, for L < km - 1
      G2[L] = g2;
      G3[L] = g3;
      G4[L] = g4;
      GA4[L] = ga4;
      g2 = 0.0;
      g3 = 0.0;
      g4 = 0.0;
      ga4 = 0.0;
      L += 1;
      pL = pR;
      pR = nR;
      nR = P1[min(L + 2, km)];
    };
    double p2k = P2[0];
    for (int k0 = 0; k0 < km; k0 += KC) {
      double p2v[KC], gv[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        p2v[u] = P2[min(k0 + u + 1, km)];
        gv[u] = GO[min(k0 + u, km - 1)];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < km) {
          const double g = gv[u];
          const double p2n = p2v[u];
          const double dpl = pR - pL;
          const double pl = (p2k - pL) / dpl;
          if (p2n <= pR) {
            const double pr = (p2n - pL) / dpl;
            // out = b2 + 0.5 (b4 + b3 - b2)(pr + pl) - b4/3 (pr (pr + pl) + pl^2)
            g2 += g * (1.0 - 0.5 * (pr + pl));
            g3 += g * (0.5 * (pr + pl));
            g4 += g * (0.5 * (pr + pl) - (pr * (pr + pl) + pl * pl) / 3.0);
          } else {
            const double w = g / (p2n - p2k);
            const double wf = (pR - p2k) * w;
            g2 += wf * (1.0 - 0.5 * (1.0 + pl));
            g3 += wf * (0.5 * (1.0 + pl));
            g4 += wf * (0.5 * (1.0 + pl) - (1.0 + pl * (1.0 + pl)) / 3.0);
            if (L < km - 1) leave();
            while (L < km - 1 && pR < p2n) {
              ga4 += (pR - pL) * w;
              leave();
            }
            const double dpx = p2n - pL;
            const double esl = dpx / (pR - pL);
            const double u2 = dpx * w;
            // dpx (a2 + 0.5 esl (a3 - a2 + a4 (1 - 2/3 esl)))
            g2 += u2 * (1.0 - 0.5 * esl);
            g3 += u2 * (0.5 * esl);
            g4 += u2 * (0.5 * esl * (1.0 - (2.0 / 3.0) * esl));
          }
          p2k = p2n;
        }
      }
    }
    G2[L] = g2;
    G3[L] = g3;
    G4[L] = g4;
    GA4[L] = ga4;
    for (int k = L + 1; k < km; ++k) {
      G2[k] = 0.0;
      G3[k] = 0.0;
      G4[k] = 0.0;
      GA4[k] = 0.0;
    }
  }

  // ---- R321 (see above).  Windows at step s: a4_1 of levels s-5 .. s+2
  // (am5 .. ap2), edge values s-3 .. s+1 (w3 .. w0, wp1), clamped edges of
  // level s-1 (e2p, e3p), flags of levels s-2 .. s (fm2, fm1, f0),
  // cotangents of the means s-4 .. s+1 (ga0 .. ga5) and of the edges
  // s-3 .. s (ge0 .. ge3), and the substituted cotangent of edge s-4.
  {
    using D7 = Dual<7>;
    using D3 = Dual<3>;
    double am5 = 0.0, am4 = 0.0, am3 = 0.0, am2 = 0.0, am1 = 0.0, a0 = q_0, ap1 = q_1;
    double w3 = 0.0, w2 = 0.0, w1 = 0.0, w0 = EV[0];
    double qc_prev = w0;
    double e2p = 0.0, e3p = 0.0;
    bool fm2 = false, fm1 = false;
    double ga0 = 0.0, ga1 = 0.0, ga2 = 0.0, ga3 = 0.0, ga4 = 0.0;
    double ge0 = 0.0, ge1 = 0.0, ge2 = 0.0;
    double gpp = 0.0;
    // a step's loads are issued one step ahead: the step's own arithmetic
    // (the Dual<7> coefficient Jacobian) covers their latency
    struct StepIn {
      double a, ev, g2, g3, g4, ga, gm;
    };
    auto load_step = [&](int s, StepIn& x) __attribute__((always_inline)) {
      x.a = A1[min(s + 2, km - 1)];
      x.ev = EV[min(s + 1, km)];
      const int k = min(max(s - 1, 0), km - 1);
      x.g2 = G2[k];
      x.g3 = G3[k];
      x.g4 = G4[k];
      x.ga = GA4[min(s + 1, km - 1)];
      x.gm = GAM[min(max(s - 4 + go, 0), km)];
    };
    StepIn nx;
    load_step(0, nx);
    for (int s = 0; s <= km + 3; ++s) {
      {
        const StepIn cu = nx;
        load_step(s + 1, nx);
        {
          const double ap2 = cu.a;
          const double wp1 = cu.ev;
          double ga5 = s + 1 <= km - 1 ? cu.ga : 0.0;  // GA of level s+1, from the walk
          double ge3 = 0.0;
          // producer: clamped edge s+1 and flag of level s
          double e2 = 0.0, e3 = 0.0;
          bool f0 = false;
          if (s < km) {
            const double qk = qc_prev;
            double qn = wp1;
            if (s + 1 <= km - 1) {
              const double gm1 = (s + 1 >= 2) ? a0 - am1 : 0.0;
              const double gp1 = (s + 2 <= km - 1) ? ap2 - ap1 : 0.0;
              qn = edge_clamp<double>(s + 1, km, iv, qn, a0, ap1, gm1, gp1);
            }
            e2 = qk;
            e3 = qn;
            qc_prev = qn;
            if (s == 0 || s == km - 1) {
              f0 = (qk - a0) * (qn - a0) > 0.0;
            } else {
              f0 = (a0 - am1) * (ap1 - a0) < 0.0;
            }
          }
          // (3) level k = s - 1
          if (s >= 1 && s <= km) {
            const int k = s - 1;
            const bool mid = k >= 2 && k < km - 2;
            const D7 d1 = D7::seed(am1, 0);
            const D7 d2 = D7::seed(e2p, 1);
            const D7 d3 = D7::seed(e3p, 2);
            const D7 dm2 = pick(mid, D7::seed(am3, 3), D7(0.0));
            const D7 dm1 = pick(mid, D7::seed(am2, 4), D7(0.0));
            const D7 dp1 = pick(mid, D7::seed(a0, 5), D7(0.0));
            const D7 dp2 = pick(mid, D7::seed(ap1, 6), D7(0.0));
            D7 o2, o3, o4;
            level_coeffs9<D7>(k, km, iv, a.qmin, d1, d2, d3, dm2, dm1, dp1, dp2, k >= 1 && fm2,
                              fm1, k + 1 < km && f0, o2, o3, o4);
            const double g2 = cu.g2, g3 = cu.g3, g4 = cu.g4;
            double gat[7];
#pragma unroll
            for (int m = 0; m < 7; ++m) gat[m] = g2 * o2.d[m] + g3 * o3.d[m] + g4 * o4.d[m];
            ga3 += gat[0];
            ge2 += gat[1];
            ge3 += gat[2];
            if (mid) {
              ga1 += gat[3];
              ga2 += gat[4];
              ga4 += gat[5];
              ga5 += gat[6];
            }
          }
          // (2) and the forward substitution of (1), interface kk = s - 3
          const int kk = s - 3;
          if (kk >= 0 && kk <= km) {
            double gep = ge0;
            if (kk >= 1 && kk <= km - 1) {
              const double gm1 = (kk >= 2) ? am4 - am5 : 0.0;
              const double gp1 = (kk + 1 <= km - 1) ? am2 - am3 : 0.0;
              const D3 e = edge_clamp<D3>(kk, km, iv, D3::seed(w3, 0), D3::seed(am4, 1),
                                          D3::seed(am3, 2), gm1, gp1);
              gep = ge0 * e.d[0];
              ga0 += ge0 * e.d[1];
              ga1 += ge0 * e.d[2];
            }
            if (kk >= 1 && !(go && kk == km)) gep = gep - cu.gm * gpp;
            GEP[kk] = gep;
            gpp = gep;
          }
          if (s - 4 >= 0 && s - 4 <= km - 1) GAP[s - 4] = ga0;
          // shift the windows
          am5 = am4;
          am4 = am3;
          am3 = am2;
          am2 = am1;
          am1 = a0;
          a0 = ap1;
          ap1 = ap2;
          w3 = w2;
          w2 = w1;
          w1 = w0;
          w0 = wp1;
          e2p = e2;
          e3p = e3;
          fm2 = fm1;
          fm1 = f0;
          ga0 = ga1;
          ga1 = ga2;
          ga2 = ga3;
          ga3 = ga4;
          ga4 = ga5;
          ge0 = ge1;
          ge1 = ge2;
          ge2 = ge3;
        }
      }
    }
  }

  // ---- R1: back substitution of the transposed edge-value solve
  {
    // bottom terms (iv != -2) and the downward sweep k = km-1 .. 1; the
    // cotangent of edge k is GEP[k] minus the h of step k+1 (a_bot h at the
    // bottom), that of mean k gathers the bottom terms, 3 h of step k+1 and
    // its own step's term, in the phase-by-phase order
    double h_next = 0.0, hb = 0.0, d4b = 0.0, a_botb = 0.0;
    double pk1 = P1[km], pk0 = P1[km - 1];  // pe1 at interfaces k+1, k (k = km-1)
    if (!go) {
      const double pkm2 = P1[km - 2];
      d4b = (pk0 - pkm2) / (pk1 - pk0);
      a_botb = 1.0 + d4b * (d4b + 1.5);
      const double den = d4b * (d4b + 0.5) - a_botb * GAM[km - 1];
      hb = GEP[km] / den;
    }
    double ga_k1 = 0.0;  // GA of level 1, written after the top terms
    for (int k0 = km - 1; k0 >= 1; k0 -= KC) {
      double gev[KC], gav[KC], gmv[KC], pv[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = max(k0 - u, 1);
        gev[u] = GEP[k];
        gav[u] = GAP[k];
        gmv[u] = GAM[k - 1 + go];
        pv[u] = P1[k - 1];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 - u;
        if (k >= 1) {
          double ga = gav[u];
          double ge = gev[u];
          if (!go) {
            if (k == km - 1) ga += 2.0 * d4b * (d4b + 1.0) * hb;
            if (k == km - 2) ga += hb;
          }
          if (k < km - 1) {
            ga += 3.0 * h_next;
            ge -= h_next;
          } else if (!go) {
            ge -= a_botb * hb;
          }
          const double pkm = pv[u];  // pe1 at interface k-1
          const double c4 = (pk0 - pkm) / (pk1 - pk0);
          const double bet = 2.0 + c4 + c4 - gmv[u];
          const double h = ge / bet;
          ga += go ? 3.0 * h : 3.0 * c4 * h;
          if (k >= 2)
            GA[k] = ga;
          else
            ga_k1 = ga;
          h_next = h;
          pk1 = pk0;
          pk0 = pkm;
        }
      }
    }
    // k = 0: pk1 = pe1[1], pk0 = pe1[0]
    double ga0 = GAP[0] + 3.0 * h_next;
    const double ge0 = GEP[0] - h_next;
    if (go) {
      ga0 += 1.5 * ge0;
      GA[1] = ga_k1;
    } else {
      const double gr = (P1[2] - pk1) / (pk1 - pk0);
      const double bet = gr * (gr + 0.5);
      const double h = ge0 / bet;
      ga0 += (gr + gr) * (gr + 1.0) * h;
      GA[1] = ga_k1 + h;
    }
    GA[0] = ga0;
  }
}

// (value, perturbation) packing for the Dual<1> TLM
__global__ void __launch_bounds__(256)
    pack_kernel(int64_t n, const double* __restrict__ v, const double* __restrict__ dv,
                Dual<1>* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  Dual<1> x(v[e]);
  x.d[0] = dv ? dv[e] : 0.0;
  out[e] = x;
}

__global__ void __launch_bounds__(256)
    unpack_kernel(int64_t n, const Dual<1>* __restrict__ in, double* __restrict__ v,
                  double* __restrict__ dv) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  v[e] = in[e].v;
  dv[e] = in[e].d[0];
}

int check_remap_args(const fv3_domain* dom, const char* who, int kord, int i1, int i2, int j1,
                     int j2, int64_t work_slab) {
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(kord == 9 || kord == -9, FV3_E_UNSUPPORTED,
              "%s: only kord 9 has a derivative here (got %d)", who, kord);
  FV3_REQUIRE(dom->nk >= 6, FV3_E_ARG, "%s: need at least 6 layers (got %d)", who, dom->nk);
  FV3_REQUIRE(i1 >= 0 && j1 >= 0 && i2 < dom->ni && j2 < dom->nj && i2 >= i1 - 1 &&
                  j2 >= j1 - 1,
              FV3_E_ARG, "%s: box outside the storage", who);
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + dom->sk * (dom->nk + 1),
              FV3_E_ARG, "%s: workspace slab too small", who);
  return FV3_OK;
}

}  // namespace
}  // namespace fv3

extern "C" {

int32_t fv3_map_single_tlm_workspace_slabs(void) { return 22; }
int32_t fv3_map_single_adjoint_workspace_slabs(void) { return fv3::kAdjSlabs; }

int32_t fv3_map_single_tlm_f64(const fv3_domain* dom, int32_t kord, int32_t mode, int32_t i1,
                               int32_t i2, int32_t j1, int32_t j2, double* q1, double* dq1,
                               const double* pe1, const double* pe2, const double* qs,
                               double qmin, double* work, int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_map_single_tlm";
  int rc = check_remap_args(dom, who, kord, i1, i2, j1, j2, work_slab);
  if (rc) return rc;
  FV3_REQUIRE(q1 && dq1 && pe1 && pe2 && work, FV3_E_ARG, "%s: NULL pointer", who);
  const int nib = i2 - i1 + 1, njb = j2 - j1 + 1;
  if (nib <= 0 || njb <= 0 || dom->ntiles == 0) return FV3_OK;
  hipStream_t s = as_stream(stream);
  // Dual<1> arrays: each slab of Dual elements spans two double slabs
  Dual<1>* W = reinterpret_cast<Dual<1>*>(work);
  Dual<1>* kwork = W;                 // 7 kernel slabs (14 double slabs)
  Dual<1>* dq = W + 7 * work_slab;    // q1 (2 double slabs)
  Dual<1>* dp1 = W + 8 * work_slab;   // pe1
  Dual<1>* dp2 = W + 9 * work_slab;   // pe2
  Dual<1>* dqs = W + 10 * work_slab;  // qs (IJ, fits in one slab)
  // exact extents the kernel addresses: q1 on km layers, pe1 / pe2 on km + 1
  // interfaces, qs on one IJ plane per tile
  const int64_t tail = (int64_t)(dom->nj - 1) * dom->sj + dom->ni;
  const int64_t base = (int64_t)(dom->ntiles - 1) * dom->st;
  const int64_t nq = base + (int64_t)(dom->nk - 1) * dom->sk + tail;
  const int64_t npe = base + (int64_t)dom->nk * dom->sk + tail;
  const int64_t n2 = (int64_t)(dom->ntiles - 1) * dom->st2 + (int64_t)(dom->nj - 1) * dom->sj2 +
                     dom->ni;
  FV3_REQUIRE(npe <= work_slab && n2 <= work_slab, FV3_E_ARG,
              "%s: the fields do not fit a workspace slab", who);
  auto blocks = [](int64_t n) { return dim3((unsigned)((n + 255) / 256)); };
  hipLaunchKernelGGL(pack_kernel, blocks(nq), dim3(256), 0, s, nq, q1, dq1, dq);
  hipLaunchKernelGGL(pack_kernel, blocks(npe), dim3(256), 0, s, npe, pe1, nullptr, dp1);
  hipLaunchKernelGGL(pack_kernel, blocks(npe), dim3(256), 0, s, npe, pe2, nullptr, dp2);
  if (qs) {
    hipLaunchKernelGGL(pack_kernel, blocks(n2), dim3(256), 0, s, n2, qs, nullptr, dqs);
  }
  RemapArgs a{9, mode, dom->nk, i1, j1, nib, njb, qmin, work_slab};
  Dom d = to_dom(dom);
  dim3 grid((unsigned)(((int64_t)nib * njb + 255) / 256), d.ntiles);
  hipLaunchKernelGGL((map_single_kernel<Dual<1>, false>), grid, dim3(256), 0, s, d, a, dq, dp1, dp2,
                     qs ? dqs : nullptr, kwork);
  hipLaunchKernelGGL(unpack_kernel, blocks(nq), dim3(256), 0, s, nq, dq, q1, dq1);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_map_single_adjoint_f64(const fv3_domain* dom, int32_t kord, int32_t mode, int32_t i1,
                                   int32_t i2, int32_t j1, int32_t j2, const double* q1,
                                   const double* pe1, const double* pe2, const double* qs,
                                   double qmin, const double* g_out, double* g_q1,
                                   double* work, int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_map_single_adjoint";
  int rc = check_remap_args(dom, who, kord, i1, i2, j1, j2, work_slab);
  if (rc) return rc;
  FV3_REQUIRE(q1 && pe1 && pe2 && g_out && g_q1 && work, FV3_E_ARG, "%s: NULL pointer", who);
  const int nib = i2 - i1 + 1, njb = j2 - j1 + 1;
  if (nib <= 0 || njb <= 0 || dom->ntiles == 0) return FV3_OK;
  RemapArgs a{9, mode, dom->nk, i1, j1, nib, njb, qmin, work_slab};
  Dom d = to_dom(dom);
  dim3 grid((unsigned)(((int64_t)nib * njb + 255) / 256), d.ntiles);
  hipLaunchKernelGGL(map_single_adjoint_kernel, grid, dim3(256), 0, as_stream(stream), d, a, q1,
                     pe1, pe2, qs, g_out, g_q1, work);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}
