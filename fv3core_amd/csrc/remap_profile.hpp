// RemapProfile (cs_profile) + MapSingle column kernel, shared by the
// forward entry points (remap.hip) and the derivative code
// (remap_adjoint.hip), which instantiates it on derivative-carrying numbers.
//
// Reference: fv3core/stencils/remap_profile.py:119-638 (three stencils:
// tridiagonal edge values, monotonicity constraints, kord 9/10 limiters) and
// map_single.py:12-173 (copy, set_dp, lagrangian_contributions with a
// data-dependent walk over source layers).  One thread owns one column and
// runs the five passes back to back; lanes of a wave own neighbouring i, so
// every k-indexed access is coalesced across the wave.  Column intermediates
// live in a caller-provided workspace of field-shaped slabs (the reference's
// component-owned temporaries).
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

template <typename T>
__device__ __forceinline__ T pick_min3(T a, T p, T l) {
  return ((a < p) && (a < l)) ? a : (p < l ? p : l);
}
template <typename T>
__device__ __forceinline__ T pick_max3(T a, T p, T l) {
  return ((a > p) && (a > l)) ? a : (p > l ? p : l);
}

struct RemapArgs {
  int kord, iv, km;
  int i1, j1, ni_box, nj_box;
  double qmin;
  int64_t slab;  // elements between workspace slabs
};

// slabs: 0 edge values q (km+1), 1 gam (km+1), 2 a4_2, 3 a4_3, 4 a4_4,
//        5 copy of a4_1, 6 flags (extm | ext5 << 1 | ext6 << 2)
template <typename T>
__global__ void __launch_bounds__(256)
    map_single_kernel(Dom d, RemapArgs a, T* __restrict__ q1, const T* __restrict__ pe1,
                      const T* __restrict__ pe2, const T* __restrict__ qs,
                      T* __restrict__ work) {
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
  Col<T> W0{work + c, sk}, W1{work + a.slab + c, sk}, A2{work + 2 * a.slab + c, sk},
      A3{work + 3 * a.slab + c, sk}, A4{work + 4 * a.slab + c, sk},
      A1{work + 5 * a.slab + c, sk}, FL{work + 6 * a.slab + c, sk};
  const T qsv = qs ? qs[(int64_t)t * d.st2 + (int64_t)j * d.sj2 + i] : T(0);

  for (int k = 0; k < km; ++k) A1[k] = Q[k];
  auto dp = [&](int k) -> T { return P1[k + 1] - P1[k]; };

  // ---- set_initial_vals: forward elimination (W0 = q, W1 = gam)
  if (iv == -2) {
    W0[0] = T(1.5) * A1[0];
    W1[1] = T(0.5);
    {
      const T gr = dp(0) / dp(1);
      const T bet = T(2) + gr + gr - W1[1];
      W0[1] = (T(3) * (A1[0] + A1[1]) - W0[0]) / bet;
    }
    T gr_prev = T(0);
    for (int k = 2; k < km; ++k) {
      const T ogr = dp(k - 2) / dp(k - 1);
      const T obet = T(2) + ogr + ogr - W1[k - 1];
      const T g = ogr / obet;
      W1[k] = g;
      const T gr = dp(k - 1) / dp(k);
      if (k < km - 1) {
        const T bet = T(2) + gr + gr - g;
        W0[k] = (T(3) * (A1[k - 1] + A1[k]) - W0[k - 1]) / bet;
      } else {
        W0[k] = (T(3) * (A1[k - 1] + A1[k]) - gr * qsv - W0[k - 1]) / (T(2) + gr + gr - g);
      }
      gr_prev = gr;
    }
    (void)gr_prev;
    W0[km] = qsv;
    for (int k = km - 2; k >= 0; --k) W0[k] = W0[k] - W1[k + 1] * W0[k + 1];
  } else {
    {
      const T gr = dp(1) / dp(0);
      const T bet = gr * (gr + T(0.5));
      W0[0] = ((gr + gr) * (gr + T(1)) * A1[0] + A1[1]) / bet;
      W1[0] = (T(1) + gr * (gr + T(1.5))) / bet;
    }
    for (int k = 1; k < km; ++k) {
      const T d4 = dp(k - 1) / dp(k);
      const T bet = T(2) + d4 + d4 - W1[k - 1];
      W0[k] = (T(3) * (A1[k - 1] + d4 * A1[k]) - W0[k - 1]) / bet;
      W1[k] = d4 / bet;
    }
    {
      const T d4 = dp(km - 2) / dp(km - 1);
      const T a_bot = T(1) + d4 * (d4 + T(1.5));
      W0[km] = (T(2) * d4 * (d4 + T(1)) * A1[km - 1] + A1[km - 2] - a_bot * W0[km - 1]) /
               (d4 * (d4 + T(0.5)) - a_bot * W1[km - 1]);
    }
    for (int k = km - 1; k >= 0; --k) W0[k] = W0[k] - W1[k] * W0[k + 1];
  }

  // ---- apply_constraints: clamp edge values, set a4_2/a4_3, flags
  // gam_c(k) = a4_1[k] - a4_1[k-1] for k >= 1
  auto gc = [&](int k) -> T { return A1[k] - A1[k - 1]; };
  T qc_prev = W0[0];  // level 0 edge is never clamped
  for (int k = 0; k < km; ++k) {
    T qk = (k == 0) ? W0[0] : qc_prev;
    T qn = W0[k + 1];
    if (k + 1 <= km - 1) {
      const int kk = k + 1;
      const T a0 = A1[kk - 1], ak = A1[kk];
      const T tmp = a0 > ak ? a0 : ak;
      const T tmp2 = a0 < ak ? a0 : ak;
      if (kk == 1 || kk == km - 1) {
        if (qn >= tmp) qn = tmp;
        if (qn <= tmp2) qn = tmp2;
      } else {
        const T gm1 = gc(kk - 1), gp1 = gc(kk + 1);
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
    A2[k] = qk;
    A3[k] = qn;
    qc_prev = qn;
    const T a1 = A1[k];
    bool extm;
    if (k == 0 || k == km - 1) {
      extm = (qk - a1) * (qn - a1) > T(0);
    } else {
      extm = gc(k) * gc(k + 1) < T(0);
    }
    bool e5 = false, e6 = false;
    if (kord > 9) {
      const T x0 = T(2) * a1 - (qk + qn);
      const T x1 = fabs(qk - qn);
      const T a4 = T(3) * x0;
      A4[k] = a4;
      e5 = fabs(x0) > x1;
      e6 = fabs(a4) > x1;
    }
    FL[k] = T((extm ? 1 : 0) | (e5 ? 2 : 0) | (e6 ? 4 : 0));
  }

  // ---- set_interpolation_coefficients
  auto flag = [&](int k, int bit) -> bool { return ((int)FL[k] & bit) != 0; };
  for (int k = 0; k < 2; ++k) {
    const T a1 = A1[k];
    T a2 = A2[k], a3 = A3[k], a4 = A4[k];
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
      remap_constraint(a1, a2, a3, a4, flag(1, 1));
    }
    A2[k] = a2;
    A3[k] = a3;
    A4[k] = a4;
  }
  for (int k = 2; k < km - 2; ++k) {
    const T a1 = A1[k];
    T a2 = A2[k], a3 = A3[k], a4 = A4[k];
    const T gk = gc(k), gp1 = gc(k + 1), gp2 = gc(k + 2), gm1 = gc(k - 1);
    if (kord < 9) {
      const T pmp_1 = a1 - gp1;
      const T lac_1 = pmp_1 + T(1.5) * gp2;
      T tmin = pick_min3(a1, pmp_1, lac_1);
      T tmax0 = a2 > tmin ? a2 : tmin;
      T tmax = pick_max3(a1, pmp_1, lac_1);
      a2 = tmax0 < tmax ? tmax0 : tmax;
      const T pmp_2 = a1 + T(2) * gp1;
      const T lac_2 = pmp_2 - T(1.5) * gm1;
      tmin = pick_min3(a1, pmp_2, lac_2);
      tmax0 = a3 > tmin ? a3 : tmin;
      tmax = pick_max3(a1, pmp_2, lac_2);
      a3 = tmax0 < tmax ? tmax0 : tmax;
      a4 = T(3) * (T(2) * a1 - (a2 + a3));
    } else if (kord == 9) {
      const T pmp_1 = a1 - T(2) * gp1;
      const T lac_1 = pmp_1 + T(1.5) * gp2;
      const T pmp_2 = a1 + T(2) * gk;
      const T lac_2 = pmp_2 - T(1.5) * gm1;
      const bool em = flag(k, 1);
      if ((em && flag(k - 1, 1)) || (em && flag(k + 1, 1)) ||
          (em && (T(a.qmin) > T(0) && a1 < T(a.qmin)))) {
        a2 = a1;
        a3 = a1;
        a4 = T(0);
      } else {
        a4 = T(6) * a1 - T(3) * (a2 + a3);
        if (fabs(a4) > fabs(a2 - a3)) {
          T tmin = pick_min3(a1, pmp_1, lac_1);
          T tmax0 = a2 > tmin ? a2 : tmin;
          T tmax = pick_max3(a1, pmp_1, lac_1);
          a2 = tmax0 < tmax ? tmax0 : tmax;
          tmin = pick_min3(a1, pmp_2, lac_2);
          tmax0 = a3 > tmin ? a3 : tmin;
          tmax = pick_max3(a1, pmp_2, lac_2);
          a3 = tmax0 < tmax ? tmax0 : tmax;
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
      const T tmp2 = a2 > tmin2 ? a2 : tmin2;
      T tmin3 = a1 < pmp_2 ? a1 : pmp_2;
      tmin3 = lac_2 < tmin3 ? lac_2 : tmin3;
      T tmax3 = a1 > pmp_2 ? a1 : pmp_2;
      tmax3 = lac_2 > tmax3 ? lac_2 : tmax3;
      const T tmp3 = a3 > tmin3 ? a3 : tmin3;
      const bool e5 = flag(k, 2), e6 = flag(k, 4);
      const bool e5n = flag(k - 1, 2) || flag(k + 1, 2);
      const bool e6n = flag(k - 1, 4) || flag(k + 1, 4);
      if (e5) {
        if (e5n) {
          a2 = a1;
          a3 = a1;
        } else if (e6n) {
          a2 = tmp2 < tmax2 ? tmp2 : tmax2;
          a3 = tmp3 < tmax3 ? tmp3 : tmax3;
        }
      } else if (e6) {
        if (e5n) {
          a2 = tmp2 < tmax2 ? tmp2 : tmax2;
          a3 = tmp3 < tmax3 ? tmp3 : tmax3;
        }
      }
      a4 = T(3) * (T(2) * a1 - (a2 + a3));
    }
    if (iv == 0) posdef_iv0(a1, a2, a3, a4);
    A2[k] = a2;
    A3[k] = a3;
    A4[k] = a4;
  }
  for (int k = km - 2; k < km; ++k) {
    const T a1 = A1[k];
    T a2 = A2[k], a3 = A3[k], a4;
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
      remap_constraint(a1, a2, a3, a4, flag(k, 1));
    else
      posdef_iv1(a1, a2, a3, a4);
    A2[k] = a2;
    A3[k] = a3;
    A4[k] = a4;
  }

  // ---- lagrangian_contributions (map_single.py:19-66)
  int L = 0;
  for (int k = 0; k < km; ++k) {
    const T p2k = P2[k], p2n = P2[k + 1];
    const T pL = P1[L];
    const T dpl = P1[L + 1] - pL;
    const T pl = (p2k - pL) / dpl;
    T out;
    if (p2n <= P1[L + 1]) {
      const T pr = (p2n - pL) / dpl;
      const T b2 = A2[L], b3 = A3[L], b4 = A4[L];
      out = b2 + T(0.5) * (b4 + b3 - b2) * (pr + pl) - (b4 * T(1) / T(3)) * (pr * (pr + pl) + pl * pl);
    } else {
      const T b2 = A2[L], b3 = A3[L], b4 = A4[L];
      T qsum = (P1[L + 1] - p2k) * (b2 + T(0.5) * (b4 + b3 - b2) * (T(1) + pl) -
                                   (b4 * T(1) / T(3)) * (T(1) + pl * (T(1) + pl)));
      L += 1;
      // valid columns (pe1 bottom at or below pe2 bottom) never pass the last
      // layer; the bound only keeps a malformed column inside its storage
      if (L > km - 1) L = km - 1;
      while (L < km - 1 && P1[L + 1] < p2n) {
        qsum += (P1[L + 1] - P1[L]) * A1[L];
        L += 1;
      }
      const T dpx = p2n - P1[L];
      const T esl = dpx / (P1[L + 1] - P1[L]);
      qsum += dpx * (A2[L] + T(0.5) * esl *
                                 (A3[L] - A2[L] + A4[L] * (T(1) - T(2.0 / 3.0) * esl)));
      out = qsum / (p2n - p2k);
    }
    Q[k] = out;
  }
}

}  // namespace fv3
