// RemapProfile (cs_profile) + MapSingle column kernel, shared by the
// forward entry points (remap.hip) and the derivative code
// (remap_adjoint.hip), which instantiates it on derivative-carrying numbers.
//
// Reference: fv3core/stencils/remap_profile.py:119-638 (three stencils:
// tridiagonal edge values, monotonicity constraints, kord 9/10 limiters) and
// map_single.py:12-173 (copy, set_dp, lagrangian_contributions with a
// data-dependent walk over source layers).  One thread owns one column and
// runs four sweeps back to back; lanes of a wave own neighbouring i, so
// every k-indexed access is coalesced across the wave.  What one sweep hands
// to a later one (edge values, tridiagonal factors, the limited
// coefficients) lives in a caller-provided workspace of field-shaped slabs
// (the reference's component-owned temporaries); what a sweep carries from
// level to level stays in registers.
#pragma once

#include "common.hpp"

namespace fv3 {

template <typename T>
struct Col {
  T* p;
  int64_t sk;
  __device__ __forceinline__ T& operator[](int k) const { return p[(int64_t)k * sk]; }
};

template <typename T>
__device__ __forceinline__ void posdef_iv1(T a1, T& a2, T& a3, T& a4) {
  const T da1 = a3 - a2;
  const T da2 = da1 * da1;
  const T a6da = a4 * da1;
  if (((a1 - a2) * (a1 - a3)) >= T(0)) {
    a2 = a1;
    a3 = a1;
    a4 = T(0);
  } else if (a6da < T(-1) * da2) {
    a4 = T(3) * (a2 - a1);
    a3 = a2 - a4;
  } else if (a6da > da2) {
    a4 = T(3) * (a3 - a1);
    a2 = a3 - a4;
  }
}

template <typename T>
__device__ __forceinline__ void remap_constraint(T a1, T& a2, T& a3, T& a4, bool extm) {
  const T da1 = a3 - a2;
  const T da2 = da1 * da1;
  const T a6da = a4 * da1;
  if (extm) {
    a2 = a1;
    a3 = a1;
    a4 = T(0);
  } else if (a6da < -da2) {
    a4 = T(3) * (a2 - a1);
    a3 = a2 - a4;
  } else if (a6da > da2) {
    a4 = T(3) * (a3 - a1);
    a2 = a3 - a4;
  }
}

template <typename T>
__device__ __forceinline__ void posdef_iv0(T a1, T& a2, T& a3, T& a4) {
  if (a1 <= T(0)) {
    a2 = a1;
    a3 = a1;
    a4 = T(0);
  } else {
    const T d = a3 - a2;
    if (fabs(d) < -a4 && (a1 + T(0.25) * (d * d) / a4 + a4 * T(1.0 / 12.0)) < T(0)) {
      if ((a1 < a3) && (a1 < a2)) {
        a3 = a1;
        a2 = a1;
        a4 = T(0);
      } else if (a3 > a2) {
        a4 = T(3) * (a2 - a1);
        a3 = a2 - a4;
      } else {
        a4 = T(3) * (a3 - a1);
        a2 = a3 - a4;
      }
    }
  }
}

// c ? a : b on stencil numbers.  Derivative-carrying number types overload it
// with a component-wise select: a conditional expression on two structures
// is lowered through memory.
template <typename T>
__device__ __forceinline__ T pick(bool c, T a, T b) {
  return c ? a : b;
}

template <typename T>
__device__ __forceinline__ T pick_min3(T a, T p, T l) {
  return pick((a < p) && (a < l), a, pick(p < l, p, l));
}
template <typename T>
__device__ __forceinline__ T pick_max3(T a, T p, T l) {
  return pick((a > p) && (a > l), a, pick(p > l, p, l));
}

struct RemapArgs {
  int kord, iv, km;
  int i1, j1, ni_box, nj_box;
  double qmin;
  int64_t slab;  // elements between workspace slabs
};

// Levels per chunk of a column sweep: the chunk's loads, which never depend
// on the values the sweep carries from level to level, are issued together
// and the recurrence then runs on registers, so a chunk waits for memory
// once instead of once per level.
constexpr int kRemapKC = 8;

template <typename T>
struct RemapCoef {
  T a1, a2, a3, a4;
};

// One step s of apply_constraints (remap_profile.py:210-280, edge at
// interface s+1) and set_interpolation_coefficients (remap_profile.py:283-495,
// level s-1) over a register window of the column: a4_1 of levels s-3 .. s+2,
// the clamped edges of level s-1 and the flags (extm | ext5 << 1 | ext6 << 2)
// of levels s-2 .. s.  Steps run for s = 0 .. km in order; both forms of
// map_single_kernel drive the same steps, so they give the same bits.
template <typename T>
struct RemapLimiter {
  int kord, iv, km;
  double qmin;
  T am3, am2, am1, a0, ap1;
  T qc_prev;  // clamped edge at interface s (level 0's top edge is never clamped)
  T e2p, e3p;
  int fm2, fm1;

  __device__ __forceinline__ RemapLimiter(int kord_, int iv_, int km_, double qmin_, T q0, T q1,
                                          T ev0)
      : kord(kord_), iv(iv_), km(km_), qmin(qmin_), am3(T(0)), am2(T(0)), am1(T(0)), a0(q0),
        ap1(q1), qc_prev(ev0), e2p(T(0)), e3p(T(0)), fm2(0), fm1(0) {}

  // ap2: a4_1 of level s+2 (clamped to the last level); wp1: the edge value
  // at interface s+1 from the back substitution.  For s >= 1 stores level
  // s-1's a4_1 .. a4_4 in o and returns true.
  __device__ __forceinline__ bool step(int s, T ap2, T wp1, RemapCoef<T>& o) {
    bool produced = false;
    T e2 = T(0), e3 = T(0);
    int f0 = 0;
    if (s < km) {
      const T qk = qc_prev;
      T qn = wp1;
      if (s + 1 <= km - 1) {
        const int kk = s + 1;
        const T tmp = pick(a0 > ap1, a0, ap1);
        const T tmp2 = pick(a0 < ap1, a0, ap1);
        if (kk == 1 || kk == km - 1) {
          if (qn >= tmp) qn = tmp;
          if (qn <= tmp2) qn = tmp2;
        } else {
          const T gm1 = a0 - am1, gp1 = ap2 - ap1;
          if ((gm1 * gp1) > T(0)) {
            if (qn >= tmp) qn = tmp;
            if (qn <= tmp2) qn = tmp2;
          } else if (gm1 > T(0)) {
            if (qn <= tmp2) qn = tmp2;
          } else {
            if (qn >= tmp) qn = tmp;
            if (iv == 0) {
              if (qn < T(0)) qn = T(0);
            }
          }
        }
      }
      e2 = qk;
      e3 = qn;
      qc_prev = qn;
      bool extm;
      if (s == 0 || s == km - 1) {
        extm = (qk - a0) * (qn - a0) > T(0);
      } else {
        extm = (a0 - am1) * (ap1 - a0) < T(0);
      }
      bool e5 = false, e6 = false;
      if (kord > 9) {
        const T x0 = T(2) * a0 - (qk + qn);
        const T x1 = fabs(qk - qn);
        const T a4 = T(3) * x0;
        e5 = fabs(x0) > x1;
        e6 = fabs(a4) > x1;
      }
      f0 = (extm ? 1 : 0) | (e5 ? 2 : 0) | (e6 ? 4 : 0);
    }
    if (s >= 1) {
      const int k = s - 1;
      const T a1 = am1;
      T a2 = e2p, a3 = e3p, a4;
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
          remap_constraint(a1, a2, a3, a4, (fm1 & 1) != 0);
        }
      } else if (k < km - 2) {
        const T gk = am1 - am2, gp1 = a0 - am1, gp2 = ap1 - a0, gm1 = am2 - am3;
        if (kord < 9) {
          const T pmp_1 = a1 - gp1;
          const T lac_1 = pmp_1 + T(1.5) * gp2;
          T tmin = pick_min3(a1, pmp_1, lac_1);
          T tmax0 = pick(a2 > tmin, a2, tmin);
          T tmax = pick_max3(a1, pmp_1, lac_1);
          a2 = pick(tmax0 < tmax, tmax0, tmax);
          const T pmp_2 = a1 + T(2) * gp1;
          const T lac_2 = pmp_2 - T(1.5) * gm1;
          tmin = pick_min3(a1, pmp_2, lac_2);
          tmax0 = pick(a3 > tmin, a3, tmin);
          tmax = pick_max3(a1, pmp_2, lac_2);
          a3 = pick(tmax0 < tmax, tmax0, tmax);
          a4 = T(3) * (T(2) * a1 - (a2 + a3));
        } else if (kord == 9) {
          const T pmp_1 = a1 - T(2) * gp1;
          const T lac_1 = pmp_1 + T(1.5) * gp2;
          const T pmp_2 = a1 + T(2) * gk;
          const T lac_2 = pmp_2 - T(1.5) * gm1;
          const bool em = (fm1 & 1) != 0;
          if ((em && (fm2 & 1)) || (em && (f0 & 1)) ||
              (em && (T(qmin) > T(0) && a1 < T(qmin)))) {
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
        } else {  // kord == 10
          const T pmp_1 = a1 - T(2) * gp1;
          const T lac_1 = pmp_1 + T(1.5) * gp2;
          const T pmp_2 = a1 + T(2) * gk;
          const T lac_2 = pmp_2 - T(1.5) * gm1;
          const T tmin2 = pick_min3(a1, pmp_1, lac_1);
          const T tmax2 = pick_max3(a1, pmp_1, lac_1);
          const T tmp2 = pick(a2 > tmin2, a2, tmin2);
          T tmin3 = pick(a1 < pmp_2, a1, pmp_2);
          tmin3 = pick(lac_2 < tmin3, lac_2, tmin3);
          T tmax3 = pick(a1 > pmp_2, a1, pmp_2);
          tmax3 = pick(lac_2 > tmax3, lac_2, tmax3);
          const T tmp3 = pick(a3 > tmin3, a3, tmin3);
          const bool e5 = (fm1 & 2) != 0, e6 = (fm1 & 4) != 0;
          const bool e5n = ((fm2 | f0) & 2) != 0;
          const bool e6n = ((fm2 | f0) & 4) != 0;
          if (e5) {
            if (e5n) {
              a2 = a1;
              a3 = a1;
            } else if (e6n) {
              a2 = pick(tmp2 < tmax2, tmp2, tmax2);
              a3 = pick(tmp3 < tmax3, tmp3, tmax3);
            }
          } else if (e6) {
            if (e5n) {
              a2 = pick(tmp2 < tmax2, tmp2, tmax2);
              a3 = pick(tmp3 < tmax3, tmp3, tmax3);
            }
          }
          a4 = T(3) * (T(2) * a1 - (a2 + a3));
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
          remap_constraint(a1, a2, a3, a4, (fm1 & 1) != 0);
        else
          posdef_iv1(a1, a2, a3, a4);
      }
      o = RemapCoef<T>{a1, a2, a3, a4};
      produced = true;
    }
    am3 = am2;
    am2 = am1;
    am1 = a0;
    a0 = ap1;
    ap1 = ap2;
    fm2 = fm1;
    fm1 = f0;
    e2p = e2;
    e3p = e3;
    return produced;
  }
};

// slabs: 0 edge values (km+1), 1 gam (km+1), 2 a4_2, 3 a4_3, 4 a4_4 (not
//        with FUSED), 5 copy of a4_1 (read after the walk has started
//        overwriting q1)
//
// Sweeps: (1) upward forward elimination of the edge-value system with the
// running values in registers; (2) downward back substitution; (3) one
// upward sweep that clamps the edge values and forms the interpolation
// coefficients, with a register window over the layer means, the clamped
// edges and the extremum flags of the neighbouring levels; (4) the overlap
// walk, which prefetches the next source layer.  FUSED runs (3) as a
// producer one source layer ahead of the walk instead, so the coefficients
// never go through memory: fewer bytes, but a longer dependent chain per
// column, which pays only when the launch has waves enough to cover it.  Every value is
// formed by the same operations in the same order as the stencils.
template <typename T, bool FUSED>
__global__ void __launch_bounds__(256)
    map_single_kernel(Dom d, RemapArgs a, T* __restrict__ q1, const T* __restrict__ pe1,
                      const T* __restrict__ pe2, const T* __restrict__ qs,
                      T* __restrict__ work) {
  constexpr int KC = kRemapKC;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)a.ni_box * a.nj_box) return;
  const int i = a.i1 + (int)(n % a.ni_box);
  const int j = a.j1 + (int)(n / a.ni_box);
  const int t = blockIdx.y;
  const int64_t c = (int64_t)t * d.st + (int64_t)j * d.sj + i;
  const int64_t sk = d.sk;
  const int km = a.km;
  const int kord = a.kord, iv = a.iv;
  Col<T> Q{q1 + c, sk};
  Col<const T> P1{pe1 + c, sk}, P2{pe2 + c, sk};
  Col<T> EV{work + c, sk}, GAM{work + a.slab + c, sk}, A2{work + 2 * a.slab + c, sk},
      A3{work + 3 * a.slab + c, sk}, A4{work + 4 * a.slab + c, sk},
      A1{work + 5 * a.slab + c, sk};
  const T qsv = qs ? qs[(int64_t)t * d.st2 + (int64_t)j * d.sj2 + i] : T(0);

  // ---- (1) set_initial_vals: forward elimination (EV = q, GAM = gam)
  const T q_0 = Q[0], q_1 = Q[1];
  const T p_0 = P1[0], p_1 = P1[1], p_2 = P1[2];
  A1[0] = q_0;
  A1[1] = q_1;
  T ev_prev, gam_prev;  // EV and GAM of level k-1
  T a_prev = q_1;       // a4_1 of level k-1
  T p_k = p_2;          // pe1 at interface k
  T dp_pp = p_1 - p_0;  // dp(k-2)
  T dp_prev = p_2 - p_1;  // dp(k-1)
  T ev_last;              // the value the back substitution starts from
  if (iv == -2) {
    const T ev0 = T(1.5) * q_0;
    EV[0] = ev0;
    GAM[1] = T(0.5);
    {
      const T gr = dp_pp / dp_prev;
      const T bet = T(2) + gr + gr - T(0.5);
      ev_prev = (T(3) * (q_0 + q_1) - ev0) / bet;
      EV[1] = ev_prev;
    }
    gam_prev = T(0.5);
    for (int k0 = 2; k0 < km; k0 += KC) {
      T qv[KC], pv[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = min(k0 + u, km - 1);
        qv[u] = Q[k];
        pv[u] = P1[k + 1];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < km) {
          const T dpk = pv[u] - p_k;
          const T ogr = dp_pp / dp_prev;
          const T obet = T(2) + ogr + ogr - gam_prev;
          const T g = ogr / obet;
          GAM[k] = g;
          const T gr = dp_prev / dpk;
          T ev;
          if (k < km - 1) {
            const T bet = T(2) + gr + gr - g;
            ev = (T(3) * (a_prev + qv[u]) - ev_prev) / bet;
          } else {
            ev = (T(3) * (a_prev + qv[u]) - gr * qsv - ev_prev) / (T(2) + gr + gr - g);
          }
          EV[k] = ev;
          A1[k] = qv[u];
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
    ev_last = ev_prev;  // EV[km-1], final as it stands
  } else {
    {
      const T gr = dp_prev / dp_pp;
      const T bet = gr * (gr + T(0.5));
      ev_prev = ((gr + gr) * (gr + T(1)) * q_0 + q_1) / bet;
      gam_prev = (T(1) + gr * (gr + T(1.5))) / bet;
      EV[0] = ev_prev;
      GAM[0] = gam_prev;
    }
    // level 1 uses dp(0) / dp(1), the pair already held
    {
      const T d4 = dp_pp / dp_prev;
      const T bet = T(2) + d4 + d4 - gam_prev;
      ev_prev = (T(3) * (q_0 + d4 * q_1) - ev_prev) / bet;
      gam_prev = d4 / bet;
      EV[1] = ev_prev;
      GAM[1] = gam_prev;
    }
    T a_pp = q_0;  // a4_1 of level k-2
    for (int k0 = 2; k0 < km; k0 += KC) {
      T qv[KC], pv[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = min(k0 + u, km - 1);
        qv[u] = Q[k];
        pv[u] = P1[k + 1];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < km) {
          const T dpk = pv[u] - p_k;
          const T d4 = dp_prev / dpk;
          const T bet = T(2) + d4 + d4 - gam_prev;
          const T ev = (T(3) * (a_prev + d4 * qv[u]) - ev_prev) / bet;
          const T g = d4 / bet;
          EV[k] = ev;
          GAM[k] = g;
          A1[k] = qv[u];
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
    {
      const T d4 = dp_pp / dp_prev;
      const T a_bot = T(1) + d4 * (d4 + T(1.5));
      ev_last = (T(2) * d4 * (d4 + T(1)) * a_prev + a_pp - a_bot * ev_prev) /
                (d4 * (d4 + T(0.5)) - a_bot * gam_prev);
      EV[km] = ev_last;
    }
  }

  // ---- (2) back substitution: EV[k] -= gam * EV[k+1], with gam of level
  // k+1 (iv = -2, from k = km-2) or of level k (from k = km-1)
  {
    const int go = iv == -2 ? 1 : 0;
    T ev_next = ev_last;
    for (int k0 = km - 1 - go; k0 >= 0; k0 -= KC) {
      T ev[KC], g[KC];
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

  if constexpr (!FUSED) {
  // ---- (3) one upward sweep of limiter steps; step s writes level s-1's
  // a4_2 .. a4_4
  {
    RemapLimiter<T> lim(kord, iv, km, a.qmin, q_0, q_1, EV[0]);
    for (int s0 = 0; s0 <= km; s0 += KC) {
      T av[KC], evv[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        av[u] = Q[min(s0 + u + 2, km - 1)];
        evv[u] = EV[min(s0 + u + 1, km)];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int s = s0 + u;
        RemapCoef<T> o;
        if (s <= km && lim.step(s, av[u], evv[u], o)) {
          A2[s - 1] = o.a2;
          A3[s - 1] = o.a3;
          A4[s - 1] = o.a4;
        }
      }
    }
  }

  // ---- (4) lagrangian_contributions (map_single.py:19-66): the overlap walk
  // over source layers L, with layer L's values in registers and layer
  // L+1's loads issued as soon as L moves
  {
    int L = 0;
    T pL = p_0, pR = p_1;
    T c1 = A1[0], c2 = A2[0], c3 = A3[0], c4 = A4[0];
    T nR = p_2, n1 = A1[1], n2 = A2[1], n3 = A3[1], n4 = A4[1];
    auto advance = [&]() __attribute__((always_inline)) {  // L -> L + 1, for L < km - 1
      L += 1;
      pL = pR;
      pR = nR;
      c1 = n1;
      c2 = n2;
      c3 = n3;
      c4 = n4;
      const int m = L + 1 < km ? L + 1 : km - 1;
      nR = P1[m + 1];
      n1 = A1[m];
      n2 = A2[m];
      n3 = A3[m];
      n4 = A4[m];
    };
    T p2k = P2[0];
    for (int k0 = 0; k0 < km; k0 += KC) {
      T p2v[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) p2v[u] = P2[min(k0 + u + 1, km)];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < km) {
          const T p2n = p2v[u];
          const T dpl = pR - pL;
          const T pl = (p2k - pL) / dpl;
          T out;
          if (p2n <= pR) {
            const T pr = (p2n - pL) / dpl;
            out = c2 + T(0.5) * (c4 + c3 - c2) * (pr + pl) -
                  (c4 * T(1) / T(3)) * (pr * (pr + pl) + pl * pl);
          } else {
            T qsum = (pR - p2k) * (c2 + T(0.5) * (c4 + c3 - c2) * (T(1) + pl) -
                                   (c4 * T(1) / T(3)) * (T(1) + pl * (T(1) + pl)));
            // valid columns (pe1 bottom at or below pe2 bottom) never pass the
            // last layer; the bound only keeps a malformed column inside its
            // storage
            if (L < km - 1) advance();
            while (L < km - 1 && pR < p2n) {
              qsum += (pR - pL) * c1;
              advance();
            }
            const T dpx = p2n - pL;
            const T esl = dpx / (pR - pL);
            qsum += dpx * (c2 + T(0.5) * esl * (c3 - c2 + c4 * (T(1) - T(2.0 / 3.0) * esl)));
            out = qsum / (p2n - p2k);
          }
          Q[k] = out;
          p2k = p2n;
        }
      }
    }
  }
  } else {
  // ---- (3) + (4): the limiter steps run as a producer one source layer
  // ahead of the overlap walk that consumes them, so the limited coefficients
  // never go through memory; a step's two loads are issued one step ahead.
  {
    RemapLimiter<T> lim(kord, iv, km, a.qmin, q_0, q_1, EV[0]);
    int s = 0;
    T nx_a = A1[2], nx_e = EV[1];
    // producer step s; for s >= 1 it returns level s-1's a4_1 .. a4_4
    auto produce = [&]() __attribute__((always_inline)) -> RemapCoef<T> {
      RemapCoef<T> o{T(0), T(0), T(0), T(0)};
      const T ap2 = nx_a, wp1 = nx_e;
      nx_a = A1[min(s + 3, km - 1)];
      nx_e = EV[min(s + 2, km)];
      lim.step(s, ap2, wp1, o);
      ++s;
      return o;
    };
    // overlap walk (map_single.py:19-66) over source layers L, with layer L's
    // coefficients in registers and layer L+1's produced as soon as L moves
    produce();  // step 0: no level yet
    T c1, c2, c3, c4, n1, n2, n3, n4;
    {
      const RemapCoef<T> a = produce();  // level 0
      c1 = a.a1;
      c2 = a.a2;
      c3 = a.a3;
      c4 = a.a4;
      const RemapCoef<T> b = produce();  // level 1
      n1 = b.a1;
      n2 = b.a2;
      n3 = b.a3;
      n4 = b.a4;
    }
    int L = 0;
    T pL = p_0, pR = p_1, nR = p_2;
    auto advance = [&]() __attribute__((always_inline)) {  // L -> L + 1, for L < km - 1
      L += 1;
      pL = pR;
      pR = nR;
      c1 = n1;
      c2 = n2;
      c3 = n3;
      c4 = n4;
      if (L + 1 < km) {
        nR = P1[L + 2];
        const RemapCoef<T> b = produce();  // level L + 1
        n1 = b.a1;
        n2 = b.a2;
        n3 = b.a3;
        n4 = b.a4;
      }
    };
    T p2k = P2[0];
    for (int k0 = 0; k0 < km; k0 += KC) {
      T p2v[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) p2v[u] = P2[min(k0 + u + 1, km)];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < km) {
          const T p2n = p2v[u];
          const T dpl = pR - pL;
          const T pl = (p2k - pL) / dpl;
          T out;
          if (p2n <= pR) {
            const T pr = (p2n - pL) / dpl;
            out = c2 + T(0.5) * (c4 + c3 - c2) * (pr + pl) -
                  (c4 * T(1) / T(3)) * (pr * (pr + pl) + pl * pl);
          } else {
            T qsum = (pR - p2k) * (c2 + T(0.5) * (c4 + c3 - c2) * (T(1) + pl) -
                                   (c4 * T(1) / T(3)) * (T(1) + pl * (T(1) + pl)));
            // valid columns (pe1 bottom at or below pe2 bottom) never pass the
            // last layer; the bound only keeps a malformed column inside its
            // storage
            if (L < km - 1) advance();
            while (L < km - 1 && pR < p2n) {
              qsum += (pR - pL) * c1;
              advance();
            }
            const T dpx = p2n - pL;
            const T esl = dpx / (pR - pL);
            qsum += dpx * (c2 + T(0.5) * esl * (c3 - c2 + c4 * (T(1) - T(2.0 / 3.0) * esl)));
            out = qsum / (p2n - p2k);
          }
          Q[k] = out;
          p2k = p2n;
        }
      }
    }
  }
  }
}

}  // namespace fv3
