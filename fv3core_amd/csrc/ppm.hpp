// One-dimensional PPM face evaluation shared by the x and y sweeps.
//
// The reference computes the PPM flux in a stencil per direction
// (fv3core/stencils/xppm.py:18-280; yppm.py is the same code with i and j
// exchanged).  Here one device function evaluates the mean of q advected
// through one interface from the six cells q[i-3..i+2] around it, so the
// fused transport kernel can call it on LDS rows (x) or LDS columns (y).
//
// Indices: `i` is the absolute storage index of the interface (between cells
// i-1 and i); `is`/`ie` are the tile-edge anchors of that direction
// (i_start/i_end, or far-away sentinels when the tile has no edge there).
#pragma once

#include "common.hpp"

namespace fv3 {

// fv3core/stencils/ppm.py:8-18
#define FV3_PPM_C1 (-2.0 / 14.0)
#define FV3_PPM_C2 (11.0 / 14.0)
#define FV3_PPM_C3 (5.0 / 14.0)
#define FV3_PPM_P1 (7.0 / 12.0)
#define FV3_PPM_P2 (-1.0 / 12.0)
#define FV3_PPM_S11 (11.0 / 14.0)
#define FV3_PPM_S14 (4.0 / 7.0)
#define FV3_PPM_S15 (3.0 / 14.0)

// Monotone slope of cell p from its neighbours (dm_iord8plus, xppm.py:84-89).
template <typename T>
__device__ __forceinline__ T ppm_dm(T qm, T q0, T qp) {
  T xt = T(0.25) * (qp - qm);
  T dqr = fmax(fmax(q0, qm), qp) - q0;
  T dql = q0 - fmin(fmin(q0, qm), qp);
  return sgn(fmin(fmin(fabs(xt), dqr), dql), xt);
}

// pert_ppm_standard_constraint_fcn (ppm.py:21-34)
template <typename T>
__device__ __forceinline__ void ppm_standard_constraint(T& al, T& ar) {
  if (al * ar < T(0)) {
    T da1 = al - ar;
    T da2 = da1 * da1;
    T a6da = T(3) * (al + ar) * da1;
    if (a6da < -da2) {
      ar = T(-2) * al;
    } else if (a6da > da2) {
      al = T(-2) * ar;
    }
  } else {
    al = T(0);
    ar = T(0);
  }
}

// xt_dxa_edge_0 (xppm.py:105-113, 125-133) at cell p of q[], dx[] (absolute
// offsets p-1..p+2 must be valid).
template <typename T>
__device__ __forceinline__ T ppm_edge0(const T* q, const T* dx, int p) {
  T xt = T(0.5) * (((T(2) * dx[p] + dx[p - 1]) * q[p] - dx[p] * q[p - 1]) / (dx[p - 1] + dx[p]) +
                   ((T(2) * dx[p + 1] + dx[p + 2]) * q[p + 1] - dx[p + 1] * q[p + 2]) /
                       (dx[p + 1] + dx[p + 2]));
  T minq = fmin(fmin(fmin(q[p - 1], q[p]), q[p + 1]), q[p + 2]);
  T maxq = fmax(fmax(fmax(q[p - 1], q[p]), q[p + 1]), q[p + 2]);
  return fmin(fmax(xt, minq), maxq);
}

// xt_dxa_edge_1 (xppm.py:115-122, 136-144)
template <typename T>
__device__ __forceinline__ T ppm_edge1(const T* q, const T* dx, int p) {
  T xt = T(0.5) * (((T(2) * dx[p - 1] + dx[p - 2]) * q[p - 1] - dx[p - 1] * q[p - 2]) /
                       (dx[p - 2] + dx[p - 1]) +
                   ((T(2) * dx[p] + dx[p + 1]) * q[p] - dx[p] * q[p + 1]) / (dx[p] + dx[p + 1]));
  T minq = fmin(fmin(fmin(q[p - 2], q[p - 1]), q[p]), q[p + 1]);
  T maxq = fmax(fmax(fmax(q[p - 2], q[p - 1]), q[p]), q[p + 1]);
  return fmin(fmax(xt, minq), maxq);
}

// Loads dx[i-3..i+2] of the interface's row/column only when an edge formula
// needs it.  S is the stored type of the widths, T the type the PPM
// arithmetic runs in (the same, or a derivative-carrying type built from S).
template <typename T, typename S = T>
struct DxRow {
  const S* base;  // points at dx of cell i-3
  int64_t stride;
  int count = 6;  // entries that may be read (the rest are filled with 1)
  __device__ __forceinline__ void load(T (&d)[6]) const {
#pragma unroll
    for (int n = 0; n < 6; ++n) d[n] = n < count ? T(base[n * stride]) : T(1);
  }
};

// Edge perturbations bl, br of cell m = absolute index, located at q[p]
// (hord = 8: compute_blbr_ord8plus + bl_br_edges, xppm.py:92-102, 184-266).
// q[p-2..p+2] and dx[p-2..p+2] must be valid.
template <typename T, typename S>
__device__ __forceinline__ void ppm_blbr8(const T* q, int p, int m, int is, int ie,
                                          const DxRow<T, S>& dxr, T& bl, T& br) {
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
    T d[6];
    dxr.load(d);
    // d[] holds dx of cells (interface-3 .. interface+2); cell m sits at
    // d[p] because q[] and d[] share the same origin.
    const T s11 = T(FV3_PPM_S11), s14 = T(FV3_PPM_S14), s15 = T(FV3_PPM_S15);
    T xbl, xbr;
    // Regions are applied in the reference's order, so the last match wins;
    // test them in reverse and take the first match.
    if (m == ie + 1) {
      xbl = ppm_edge1(q, d, p);
      xbr = s11 * (q[p + 1] - q[p]) - s14 * dmp + q[p];
    } else if (m == ie) {
      xbl = s15 * q[p] + s11 * q[p - 1] + s14 * dmm;
      xbr = ppm_edge0(q, d, p);
    } else if (m == ie - 1) {
      xbl = alL;
      xbr = s15 * q[p + 1] + s11 * q[p] + s14 * dm0;
    } else if (m == is + 1) {
      xbl = s15 * q[p - 1] + s11 * q[p] - s14 * dm0;
      xbr = alR;
    } else if (m == is) {
      xbl = ppm_edge1(q, d, p);
      xbr = s15 * q[p] + s11 * q[p + 1] - s14 * dmp;
    } else {  // m == is - 1
      xbl = s14 * dmm + s11 * (q[p - 1] - q[p]) + q[p];
      xbr = ppm_edge0(q, d, p);
    }
    bl = xbl - q[p];
    br = xbr - q[p];
    ppm_standard_constraint(bl, br);
  }
}


// ppm_blbr8 on a 5-cell window w[0..4] centred on the cell (w[2]); dxr points
// at dx of the window's first cell.
template <typename T, typename S>
__device__ __forceinline__ void ppm_blbr8_w(const T (&w)[5], int m, int is, int ie,
                                            const DxRow<T, S>& dxr, T& bl, T& br) {
  T q[6] = {w[0], w[1], w[2], w[3], w[4], w[4]};
  // ppm_blbr8 reads q[p-2..p+2] and loads dx[0..5] relative to its own
  // origin; with p = 2 the window's first cell is the origin
  ppm_blbr8(q, 2, m, is, ie, dxr, bl, br);
}

// Interface value al at face f whose left cell is q[p-1] (hord < 8,
// compute_al, xppm.py:147-181).  q[p-2..p+1] valid.
template <typename T, typename S>
__device__ __forceinline__ T ppm_al(const T* q, int p, int f, int is, int ie,
                                    const DxRow<T, S>& dxr) {
  if (f == is + 1 || f == ie + 2) {
    return T(FV3_PPM_C3) * q[p - 1] + T(FV3_PPM_C2) * q[p] + T(FV3_PPM_C1) * q[p + 1];
  }
  if (f == is || f == ie + 1) {
    T d[6];
    dxr.load(d);
    return T(0.5) * (((T(2) * d[p - 1] + d[p - 2]) * q[p - 1] - d[p - 1] * q[p - 2]) /
                         (d[p - 2] + d[p - 1]) +
                     ((T(2) * d[p] + d[p + 1]) * q[p] - d[p] * q[p + 1]) / (d[p] + d[p + 1]));
  }
  if (f == is - 1 || f == ie) {
    return T(FV3_PPM_C1) * q[p - 2] + T(FV3_PPM_C2) * q[p - 1] + T(FV3_PPM_C3) * q[p];
  }
  return T(FV3_PPM_P1) * (q[p - 1] + q[p]) + T(FV3_PPM_P2) * (q[p - 2] + q[p + 1]);
}

// Mean of q advected through interface i (compute_x_flux, xppm.py:269-280).
// q6[n] = q of cell i-3+n; c = Courant number on the interface.
// ORD is the reference's iord/jord (5, 6, 7 or 8).
template <typename T, int ORD, typename S>
__device__ __forceinline__ T ppm_face(const T (&q6)[6], T c, int i, int is, int ie,
                                      const DxRow<T, S>& dxr) {
  if constexpr (ORD == 8) {
    T bl, br;
    if (c > T(0)) {
      ppm_blbr8(q6, 2, i - 1, is, ie, dxr, bl, br);
      T b0 = bl + br;
      T fx1 = (T(1) - c) * (br - c * b0);
      return q6[2] + fx1;
    } else {
      ppm_blbr8(q6, 3, i, is, ie, dxr, bl, br);
      T b0 = bl + br;
      T fx1 = (T(1) + c) * (bl + c * b0);
      return q6[3] + fx1;
    }
  } else {
    // al at faces i-1, i, i+1
    const T alm = ppm_al(q6, 2, i - 1, is, ie, dxr);
    const T al0 = ppm_al(q6, 3, i, is, ie, dxr);
    const T alp = ppm_al(q6, 4, i + 1, is, ie, dxr);
    // cell i-1
    const T blm = alm - q6[2];
    const T brm = al0 - q6[2];
    const T b0m = blm + brm;
    // cell i
    const T bl0 = al0 - q6[3];
    const T br0 = alp - q6[3];
    const T b00 = bl0 + br0;
    bool smt_m, smt_0;
    if constexpr (ORD == 5) {
      smt_m = blm * brm < T(0);
      smt_0 = bl0 * br0 < T(0);
    } else {
      smt_m = (T(3) * fabs(b0m)) < fabs(blm - brm);
      smt_0 = (T(3) * fabs(b00)) < fabs(bl0 - br0);
    }
    const T mask = (smt_m || smt_0) ? T(1) : T(0);
    if (c > T(0)) {
      T fx1 = (T(1) - c) * (brm - c * b0m);
      return q6[2] + fx1 * mask;
    } else {
      T fx1 = (T(1) + c) * (bl0 + c * b00);
      return q6[3] + fx1 * mask;
    }
  }
}


// Branch-free form of ppm_face: the upwind cell's 5-point window is selected
// first (lanes of a wave may differ in the sign of c), then one straight-line
// evaluation follows.  Arithmetic and rounding are those of ppm_face.  With
// EDGES = false the tile-edge regions are not tested (blocks far from any
// tile edge).
template <typename T, int ORD, bool EDGES, typename S>
__device__ __forceinline__ T ppm_face_sel(const T (&q6)[6], T c, int i, int is, int ie,
                                          const DxRow<T, S>& dxr) {
  const bool pos = c > T(0);
  if constexpr (ORD == 8) {
    T w[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) w[m] = pos ? q6[m] : q6[m + 1];
    T bl, br;
    if constexpr (EDGES) {
      DxRow<T, S> dw{dxr.base + (pos ? 0 : dxr.stride), dxr.stride, 5};
      ppm_blbr8_w(w, pos ? i - 1 : i, is, ie, dw, bl, br);
    } else {
      const T dmm = ppm_dm(w[0], w[1], w[2]);
      const T dm0 = ppm_dm(w[1], w[2], w[3]);
      const T dmp = ppm_dm(w[2], w[3], w[4]);
      const T third = T(1.0 / 3.0);
      const T alL = T(0.5) * (w[1] + w[2]) + third * (dmm - dm0);
      const T alR = T(0.5) * (w[2] + w[3]) + third * (dm0 - dmp);
      const T xt = T(2) * dm0;
      bl = -sgn(fmin(fabs(xt), fabs(alL - w[2])), xt);
      br = sgn(fmin(fabs(xt), fabs(alR - w[2])), xt);
    }
    const T b0 = bl + br;
    const T fpos = (T(1) - c) * (br - c * b0);
    const T fneg = (T(1) + c) * (bl + c * b0);
    return w[2] + (pos ? fpos : fneg);
  } else {
    T alm, al0, alp;
    if constexpr (EDGES) {
      alm = ppm_al(q6, 2, i - 1, is, ie, dxr);
      al0 = ppm_al(q6, 3, i, is, ie, dxr);
      alp = ppm_al(q6, 4, i + 1, is, ie, dxr);
    } else {
      alm = T(FV3_PPM_P1) * (q6[1] + q6[2]) + T(FV3_PPM_P2) * (q6[0] + q6[3]);
      al0 = T(FV3_PPM_P1) * (q6[2] + q6[3]) + T(FV3_PPM_P2) * (q6[1] + q6[4]);
      alp = T(FV3_PPM_P1) * (q6[3] + q6[4]) + T(FV3_PPM_P2) * (q6[2] + q6[5]);
    }
    const T blm = alm - q6[2];
    const T brm = al0 - q6[2];
    const T b0m = blm + brm;
    const T bl0 = al0 - q6[3];
    const T br0 = alp - q6[3];
    const T b00 = bl0 + br0;
    bool smt_m, smt_0;
    if constexpr (ORD == 5) {
      smt_m = blm * brm < T(0);
      smt_0 = bl0 * br0 < T(0);
    } else {
      smt_m = (T(3) * fabs(b0m)) < fabs(blm - brm);
      smt_0 = (T(3) * fabs(b00)) < fabs(bl0 - br0);
    }
    const T mask = (smt_m || smt_0) ? T(1) : T(0);
    const T fpos = q6[2] + (T(1) - c) * (brm - c * b0m) * mask;
    const T fneg = q6[3] + (T(1) + c) * (bl0 + c * b00) * mask;
    return pos ? fpos : fneg;
  }
}

}  // namespace fv3
