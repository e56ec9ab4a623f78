// Fused PPM transport (fv_tp_2d) and the CopyCornersXY corner copies.
//
// Reference: fv3core/stencils/fvtp2d.py:151-362 runs eight stencils over
// full-size HBM temporaries (y_ppm inner, q_i, x_ppm outer, x_ppm inner,
// q_j, y_ppm outer, final_fluxes).  Here one kernel evaluates all of them in
// registers (the wave-strip march below) and writes only the two flux
// fields.  HBM traffic is the compulsory reads of q, the Courant numbers and
// the area/mass fluxes plus the two flux writes.
#include "common.hpp"
#include "timing.hpp"
#include "ppm.hpp"

namespace fv3 {

// Wave-strip march.  One wave (64 lanes) owns a strip of 64 columns of one
// level of one tile: 58 output interfaces in i plus a 3-cell halo on each
// side, lane L <-> column i0 - 3 + L.  It marches up the rows of a segment of
// SEG output rows (MTJ or MTJ_TALL, chosen per launch; SEG + 6 march steps,
// a multiple of the unroll factor 3).  Stencils across i read the neighbouring lanes' values
// with DPP lane shifts; stencils along j keep a sliding window of rows in
// registers.  Each PPM face value al and each cell's edge perturbations
// (bl, br and the smoothness flag, or the monotone slope dm for hord 8) are
// computed once and shared by the two interfaces that use them, with the
// reference's expressions, so results equal the per-interface form to the
// last bit.  No LDS and no barriers: waves run independently.
constexpr int MW = 64;        // lanes = strip columns
constexpr int MTI = MW - 6;   // output interfaces per strip in i
constexpr int MTJ = 33;       // output rows per segment in j (default)
constexpr int MTJ_TALL = 51;  // the taller segment of large batches of narrow tiles

// Whole-wave shifts by one lane with DPP: wave_shr:1 (lane L reads lane
// L - 1) and wave_shl:1 (lane L reads lane L + 1); the lane without a source
// receives 0.  They are register moves on the VALU, without the LDS round
// trip of a general lane permute.
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

template <int CTRL>
__device__ __forceinline__ int dpp32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_shift(T v) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp32<CTRL>((int)(unsigned)u);
    const unsigned hi = (unsigned)dpp32<CTRL>((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, dpp32<CTRL>(__builtin_bit_cast(int, v)));
  } else {
    return (T)dpp32<CTRL>((int)v);
  }
}
template <typename T>
__device__ __forceinline__ T lane_dn(T v, int d) {  // value of lane + d
  for (int n = 0; n < d; ++n) v = dpp_shift<kWaveShl1>(v);
  return v;
}
template <typename T>
__device__ __forceinline__ T lane_up(T v, int d) {  // value of lane - d
  for (int n = 0; n < d; ++n) v = dpp_shift<kWaveShr1>(v);
  return v;
}


// Face value al at face f (hord < 8, compute_al, xppm.py:147-181) from the
// cells f-2, f-1, f, f+1 (qa..qd) and their widths (da..dd, read only at the
// tile-edge faces).  With EDGES every candidate is evaluated and the match is
// selected without branches (the reference's regions are applied in its
// order, so the last selection below is its first test); lanes or steps that
// need none of them keep the interior value.
template <typename T, bool EDGES>
__device__ __forceinline__ T al_face(T qa, T qb, T qc, T qd, int f, int is, int ie, T da, T db,
                                     T dc, T dd) {
  T al = T(FV3_PPM_P1) * (qb + qc) + T(FV3_PPM_P2) * (qa + qd);
  if constexpr (EDGES) {
    const T c123 = T(FV3_PPM_C1) * qa + T(FV3_PPM_C2) * qb + T(FV3_PPM_C3) * qc;
    const T ediv = T(0.5) * (((T(2) * db + da) * qb - db * qa) / (da + db) +
                             ((T(2) * dc + dd) * qc - dc * qd) / (dc + dd));
    const T c321 = T(FV3_PPM_C3) * qb + T(FV3_PPM_C2) * qc + T(FV3_PPM_C1) * qd;
    al = (f == is - 1 || f == ie) ? c123 : al;
    al = (f == is || f == ie + 1) ? ediv : al;
    al = (f == is + 1 || f == ie + 2) ? c321 : al;
  }
  return al;
}

// Edge perturbations of a cell from its two face values (hord < 8) and the
// smoothness test of hord 5 (bl br < 0) or 6/7 (3|b0| < |bl - br|).
template <typename T, int ORD>
__device__ __forceinline__ void cell_lo(T alL, T alR, T q, T& bl, T& br, bool& smt) {
  bl = alL - q;
  br = alR - q;
  const T b0 = bl + br;
  if constexpr (ORD == 5) {
    smt = bl * br < T(0);
  } else {
    smt = (T(3) * fabs(b0)) < fabs(bl - br);
  }
}

// Mean advected through face f from cells f-1 (m) and f (0), hord < 8.
template <typename T>
__device__ __forceinline__ T face_lo(T c, T qm, T blm, T brm, bool smm, T q0, T bl0, T br0,
                                     bool sm0) {
  const T b0m = blm + brm;
  const T b00 = bl0 + br0;
  const T mask = (smm || sm0) ? T(1) : T(0);
  const T fpos = qm + (T(1) - c) * (brm - c * b0m) * mask;
  const T fneg = q0 + (T(1) + c) * (bl0 + c * b00) * mask;
  return c > T(0) ? fpos : fneg;
}

// hord 8 edge perturbations of cell m from cells m-2..m+2 and the slopes of
// cells m-1, m, m+1 (compute_blbr_ord8plus away from the tile edges).
template <typename T>
__device__ __forceinline__ void cell8(T w1, T w2, T w3, T dmm, T dm0, T dmp, T& bl, T& br) {
  const T third = T(1.0 / 3.0);
  const T alL = T(0.5) * (w1 + w2) + third * (dmm - dm0);
  const T alR = T(0.5) * (w2 + w3) + third * (dm0 - dmp);
  const T xt = T(2) * dm0;
  bl = -sgn(fmin(fabs(xt), fabs(alL - w2)), xt);
  br = sgn(fmin(fabs(xt), fabs(alR - w2)), xt);
}

// Mean advected through face f, hord 8: the upwind cell's perturbations.
template <typename T>
__device__ __forceinline__ T face8(T c, T qm, T blm, T brm, T q0, T bl0, T br0) {
  const bool pos = c > T(0);
  const T bl = pos ? blm : bl0, br = pos ? brm : br0, qu = pos ? qm : q0;
  const T b0 = bl + br;
  const T fpos = (T(1) - c) * (br - c * b0);
  const T fneg = (T(1) + c) * (bl + c * b0);
  return qu + (pos ? fpos : fneg);
}

// PPM along i on one row held across the wave: lane L has cell i (= face i
// on its left) with q, the Courant number c of face i and (tile-edge waves)
// the width dxv of cell i; dxrow points at dx of column 0 of the row.
template <typename T, int ORD, bool EDGES, typename S>
__device__ __forceinline__ T ppm_row(T q, T c, int i, int is, int ie, T dxv,
                                     const S* __restrict__ dxrow) {
  const T qm1 = lane_up(q, 1), qm2 = lane_up(q, 2), qp1 = lane_dn(q, 1);
  if constexpr (ORD < 8) {
    T dm2 = T(0), dm1 = T(0), dp1 = T(0);
    if constexpr (EDGES) {
      dm2 = lane_up(dxv, 2);
      dm1 = lane_up(dxv, 1);
      dp1 = lane_dn(dxv, 1);
    }
    const T al = al_face<T, EDGES>(qm2, qm1, q, qp1, i, is, ie, dm2, dm1, dxv, dp1);
    const T alp = lane_dn(al, 1);
    T bl, br;
    bool sm;
    cell_lo<T, ORD>(al, alp, q, bl, br, sm);
    const T blm = lane_up(bl, 1), brm = lane_up(br, 1);
    const bool smm = lane_up((int)sm, 1) != 0;
    return face_lo(c, qm1, blm, brm, smm, q, bl, br, sm);
  } else {
    const T qp2 = lane_dn(q, 2);
    T bl, br;
    if constexpr (EDGES) {
      const T w[5] = {qm2, qm1, q, qp1, qp2};
      DxRow<T, S> dw{dxrow + (i - 2), 1, 5};
      ppm_blbr8_w(w, i, is, ie, dw, bl, br);
    } else {
      const T dm = ppm_dm(qm1, q, qp1);
      cell8(qm1, q, qp1, lane_up(dm, 1), dm, lane_dn(dm, 1), bl, br);
    }
    return face8(c, qm1, lane_up(bl, 1), lane_up(br, 1), q, bl, br);
  }
}

// PPM along j in one column: cells are pushed bottom to top; after cell c is
// pushed, face c - 2 (between cells c - 3 and c - 2) can be evaluated.
template <typename T, int ORD, bool EDGES>
struct ColumnPPM {
  T q[5] = {};    // cells c-4 .. c
  T dw[5] = {};   // their widths (tile-edge waves, hord < 8)
  T al1 = 0, al2 = 0;            // faces c-2, c-1 (hord < 8)
  T dm0 = 0, dm1 = 0, dm2 = 0;   // slopes of cells c-3, c-2, c-1 (hord 8)
  T bl0 = 0, br0 = 0, bl1 = 0, br1 = 0;  // cells c-3, c-2
  bool sm0 = false, sm1 = false;

  template <typename S>
  __device__ __forceinline__ void push(T qc, T dc, int c, int is, int ie,
                                       const S* __restrict__ dycol, int64_t sj2) {
#pragma unroll
    for (int n = 0; n < 4; ++n) q[n] = q[n + 1];
    q[4] = qc;
    bl0 = bl1;
    br0 = br1;
    sm0 = sm1;
    if constexpr (ORD < 8) {
      if constexpr (EDGES) {
#pragma unroll
        for (int n = 0; n < 4; ++n) dw[n] = dw[n + 1];
        dw[4] = dc;
      }
      al1 = al2;
      al2 = al_face<T, EDGES>(q[1], q[2], q[3], q[4], c - 1, is, ie, dw[1], dw[2], dw[3], dw[4]);
      cell_lo<T, ORD>(al1, al2, q[2], bl1, br1, sm1);
    } else {
      dm0 = dm1;
      dm1 = dm2;
      dm2 = ppm_dm(q[2], q[3], q[4]);
      if constexpr (EDGES) {
        const T w[5] = {q[0], q[1], q[2], q[3], q[4]};
        DxRow<T, S> dwr{dycol + (int64_t)(c - 4) * sj2, sj2, 5};
        ppm_blbr8_w(w, c - 2, is, ie, dwr, bl1, br1);
      } else {
        cell8(q[1], q[2], q[3], dm0, dm1, dm2, bl1, br1);
      }
    }
  }
  // mean advected through face c - 2
  __device__ __forceinline__ T face(T cc) const {
    if constexpr (ORD < 8) {
      return face_lo(cc, q[1], bl0, br0, sm0, q[2], bl1, br1, sm1);
    } else {
      return face8(cc, q[1], bl0, br0, q[2], bl1, br1);
    }
  }
};

// The y-direction corner fill applied as q_y is read.  Each present corner c
// (0 = (is, js), 1 = (ie, js), 2 = (is, je), 3 = (ie, je)) fills the 3 x 3
// halo cells beyond it from source (si, sj) = (ai[c] + bi[c] i + ci[c] j,
// aj[c] + bj[c] i + cj[c] j), a rotation fitted to the reference's table on
// the host.  Sources are never corner cells.  present = 0: q_y is read as it
// is.  (Coefficients are picked with constant indices so that the table stays
// in the kernel arguments.)
// Most scalars one transport launch carries (the tracers of one call).
constexpr int kMaxScalarBatch = FV3_MAX_TRACERS;

struct YFill {
  int present;  // bit c: corner c is filled
  int ai[4], bi[4], ci[4], aj[4], bj[4], cj[4];
};

__device__ __forceinline__ int sel4(int c, int v0, int v1, int v2, int v3) {
  return c == 0 ? v0 : c == 1 ? v1 : c == 2 ? v2 : v3;
}

// Scalars transported by one launch with the same Courant numbers, area and
// mass fluxes (the tracers of TracerAdvection): each slot's storage and flux
// outputs.  The plain call is a batch of one.
template <typename S>
struct ScalarBatch {
  int n;
  const S* qx[kMaxScalarBatch];
  const S* qy[kMaxScalarBatch];
  S* fx[kMaxScalarBatch];
  S* fy[kMaxScalarBatch];
};

// a[b] for a wave-uniform slot b, as a chain of selects (no indexed access
// to the kernel arguments)
template <typename P>
__device__ __forceinline__ P pick_slot(const P (&a)[kMaxScalarBatch], int b) {
  P r = a[0];
#pragma unroll
  for (int m = 1; m < kMaxScalarBatch; ++m) r = (b == m) ? a[m] : r;
  return r;
}

// S is the storage type of every field, T the type the march computes in
// (the same, or float for FP64 storage in the mixed-precision form).
template <typename S, typename T, int ORD, bool EX, bool EY, int SEG>
__device__ __forceinline__ void fvtp2d_strip(
    const Dom& d, const YFill& yf, int i0, int j0, const S* __restrict__ qx,
    const S* __restrict__ qy,
    const S* __restrict__ cx, const S* __restrict__ cy, const S* __restrict__ ax,
    const S* __restrict__ ay, const S* __restrict__ ux, const S* __restrict__ uy,
    const S* __restrict__ ar, const S* __restrict__ dx, const S* __restrict__ dy,
    S* __restrict__ fx, S* __restrict__ fy) {
  const int lane = threadIdx.x;
  const int i = i0 - 3 + lane;  // column of this lane
  const int ni = d.ni, nj = d.nj;
  const int64_t sj = d.sj, sj2 = d.sj2;
  const bool out_lane = lane >= 3 && lane < 3 + MTI;
  // Loads read rows and columns clamped into the storage, unconditionally.
  // A lane whose column lies outside the storage, or a step whose row does,
  // feeds only interfaces beyond the compute domain (every flux kept reads
  // cells inside the storage), so what such loads return is never used.
  const int ic = min(max(i, 0), ni - 1);
  const uint32_t icb = (uint32_t)ic * (uint32_t)sizeof(S);  // byte offset of the column
  // The column offset is passed through an empty asm at each use so that the
  // compiler keeps (uniform row address) + (32-bit lane offset), which maps
  // to the scalar-base form of the global loads and stores, instead of
  // hoisting 64-bit per-lane pointers and adding the row in vector registers.
  auto lane_off = [&]() {
    uint32_t o = icb;
    asm volatile("" : "+v"(o));
    return o;
  };
  auto ld = [&](const S* __restrict__ p, int j, int64_t stride) -> T {
    const int jc = __builtin_amdgcn_readfirstlane(min(max(j, 0), nj - 1));
    const char* __restrict__ row = (const char*)(p + (int64_t)jc * stride);  // wave-uniform
    return T(*(const S*)(row + lane_off()));
  };
  // Flux stores are buffer stores: lanes or steps that own no output give
  // an offset beyond the buffer's record range, which the hardware drops,
  // so the march body has no branches around its stores.
  const __amdgpu_buffer_rsrc_t rfx = __builtin_amdgcn_make_buffer_rsrc(fx, 0, 0x7fffffff, 0x00020000);
  const __amdgpu_buffer_rsrc_t rfy = __builtin_amdgcn_make_buffer_rsrc(fy, 0, 0x7fffffff, 0x00020000);
  auto st = [&](__amdgpu_buffer_rsrc_t rs, int j, bool keep, T v) {
    const uint32_t off = keep ? icb : 0x80000000u;
    const int rowb = __builtin_amdgcn_readfirstlane((int)((int64_t)min(max(j, 0), nj - 1) * sj *
                                                          (int64_t)sizeof(S)));
    const S sv = S(v);
    if constexpr (sizeof(S) == 8) {
      typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, sv), rs, off, rowb, 0);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, sv), rs, off, rowb, 0);
    }
  };
  ColumnPPM<T, ORD, EY> yin, yout;
  T sxm[4] = {};     // x-inner means of rows r-3 .. r at face i
  T sym_prev = T(0); // y-inner mean of face r-3
  // tile-edge width rows are clamped into the storage the same way
  auto jrow = [&](int j) { return (int64_t)min(max(j, 0), nj - 1) * sj2; };
  const S* dycol = dy + ic;
  // every value a march step uses is loaded by the step before it, so the
  // loads of one step are in flight while the previous step computes
  struct StepIn {
    T qx, qy, cx, ax, ar, dx, dy;  // row r
    T cyf, ayf;                    // row f = r - 2
    T ag, cxg, dxg;                // row g = r - 3
    T uxg, uyf;                    // the mass fluxes at the outputs
  };
  auto load_step = [&](int r, StepIn& v) {
    const int f = r - 2, g = r - 3;
    v.qx = ld(qx, r, sj);
    v.qy = ld(qy, r, sj);
    if constexpr (EX && EY) {
      // corner strips only: the y corner fill of q_y (the row test is
      // wave-uniform, the column test per lane)
      const bool south = r >= d.js - 3 && r < d.js, north = r > d.je && r <= d.je + 3;
      if (yf.present && (south || north)) {
        const bool west = i >= d.is - 3 && i < d.is, east = i > d.ie && i <= d.ie + 3;
        const int c = (north ? 2 : 0) + (east ? 1 : 0);
        if ((west || east) && ((yf.present >> c) & 1)) {
          const int si = sel4(c, yf.ai[0], yf.ai[1], yf.ai[2], yf.ai[3]) +
                         sel4(c, yf.bi[0], yf.bi[1], yf.bi[2], yf.bi[3]) * i +
                         sel4(c, yf.ci[0], yf.ci[1], yf.ci[2], yf.ci[3]) * r;
          const int sjj = sel4(c, yf.aj[0], yf.aj[1], yf.aj[2], yf.aj[3]) +
                          sel4(c, yf.bj[0], yf.bj[1], yf.bj[2], yf.bj[3]) * i +
                          sel4(c, yf.cj[0], yf.cj[1], yf.cj[2], yf.cj[3]) * r;
          v.qy = T(qy[(int64_t)sjj * sj + si]);
        }
      }
    }
    v.cx = ld(cx, r, sj);
    v.ax = ld(ax, r, sj);
    v.ar = ld(ar, r, sj2);
    v.dx = EX ? ld(dx, r, sj2) : T(0);
    v.dy = EY ? ld(dy, r, sj2) : T(0);
    v.cyf = ld(cy, f, sj);
    v.ayf = ld(ay, f, sj);
    v.ag = ld(ar, g, sj2);
    v.cxg = ld(cx, g, sj);
    v.dxg = EX ? ld(dx, g, sj2) : T(0);
    v.uxg = ld(ux, g, sj);
    v.uyf = ld(uy, f, sj);
  };
  StepIn nx;
  load_step(j0 - 3, nx);
  T ay_g = T(0);  // ay of row g = r - 3 (row f of the previous step)
#pragma unroll 3
  for (int s = 0; s < SEG + 6; ++s) {
    const int r = j0 - 3 + s;
    const StepIn cu = nx;
    load_step(r + 1, nx);
    const int f = r - 2, g = r - 3;
    const bool wfx = out_lane && g >= j0 && g < j0 + SEG && i <= d.iec + 1 && g <= d.jec;
    const bool wfy = out_lane && f >= j0 && f < j0 + SEG && i <= d.iec && f <= d.jec + 1;
    // x-inner PPM of q_x on row r (faces i), then q_j of row r
    // (fvtp2d.py:50-60)
    const T xm = ppm_row<T, ORD, EX>(cu.qx, cu.cx, i, d.is, d.ie, cu.dx, dx + jrow(r));
    const T sxr = xm;
    const T f1 = lane_dn(cu.ax, 1);
    const T fx1_0 = cu.ax * sxr;
    const T fx1_1 = f1 * lane_dn(sxr, 1);
    const T qjr = (cu.qx * cu.ar + fx1_0 - fx1_1) / (cu.ar + cu.ax - f1);
#pragma unroll
    for (int n = 0; n < 3; ++n) sxm[n] = sxm[n + 1];
    sxm[3] = sxr;
    // y-inner PPM of q_y: face f = r - 2
    yin.push(cu.qy, cu.dy, r, d.js, d.je, dycol, sj2);
    const T symf = yin.face(cu.cyf);
    // q_i of row g = r - 3 (fvtp2d.py:33-47)
    const T qig = (yin.q[1] * cu.ag + ay_g * sym_prev - cu.ayf * symf) /
                  (cu.ag + ay_g - cu.ayf);
    sym_prev = symf;
    ay_g = cu.ayf;
    // x-outer on q_i of row g: q_x_flux (fvtp2d.py:63-92)
    const T outx = ppm_row<T, ORD, EX>(qig, cu.cxg, i, d.is, d.ie, cu.dxg, dx + jrow(g));
    st(rfx, g, wfx, T(0.5) * (outx + sxm[0]) * cu.uxg);
    // y-outer PPM of q_j: face f, q_y_flux
    yout.push(qjr, cu.dy, r, d.js, d.je, dycol, sj2);
    const T outy = yout.face(cu.cyf);
    st(rfy, f, wfy, T(0.5) * (outy + symf) * cu.uyf);
  }
}

template <typename S, typename T, int ORD, int SEG>
__global__ void __launch_bounds__(MW)
    fvtp2d_kernel(Dom d, ScalarBatch<S> sb, const S* __restrict__ crx,
                  const S* __restrict__ cry, const S* __restrict__ xfx,
                  const S* __restrict__ yfx, const S* __restrict__ xmf,
                  const S* __restrict__ ymf, const S* __restrict__ area,
                  const S* __restrict__ dxa, const S* __restrict__ dya, int nbi, int nbj,
                  YFill yf) {
  const int per_plane = nbi * nbj;
  const int total = per_plane * d.nk * d.ntiles * sb.n;
  const int lb = xcd_logical_block(blockIdx.x, total);
  if (lb >= total) return;
  const int plane = lb / per_plane;
  const int within = lb % per_plane;
  const int bi = within % nbi, bj = within / nbi;
  const int k = plane % d.nk, tb = plane / d.nk;
  const int t = tb % d.ntiles, b = tb / d.ntiles;
  const S* __restrict__ q_x = pick_slot(sb.qx, b);
  const S* __restrict__ q_y = pick_slot(sb.qy, b);
  S* __restrict__ fx = pick_slot(sb.fx, b);
  S* __restrict__ fy = pick_slot(sb.fy, b);
  const int i0 = d.isc + bi * MTI;
  const int j0 = d.jsc + bj * SEG;
  const int64_t off3 = (int64_t)t * d.st + (int64_t)k * d.sk;
  const int64_t off2 = (int64_t)t * d.st2;
  const S* ux = (xmf ? xmf : xfx) + off3;
  const S* uy = (ymf ? ymf : yfx) + off3;
  // the tile-edge formulas of a direction are needed when the strip's
  // stencil reach (cells i0-4 .. i0+MTI+4, rows j0-4 .. j0+SEG+4) touches
  // an edge region of that direction
  auto near = [](int lo, int hi, int a) { return hi >= a - 2 && lo <= a + 3; };
  const bool ex = near(i0 - 4, i0 + MTI + 4, d.is) || near(i0 - 4, i0 + MTI + 4, d.ie);
  const bool ey = near(j0 - 4, j0 + SEG + 4, d.js) || near(j0 - 4, j0 + SEG + 4, d.je);
#define FV3_TP_STRIP(EX_, EY_)                                                                  \
  fvtp2d_strip<S, T, ORD, EX_, EY_, SEG>(d, yf, i0, j0, q_x + off3, q_y + off3, crx + off3, \
                                 cry + off3,                                                  \
                                 xfx + off3, yfx + off3, ux, uy, area + off2, dxa + off2,     \
                                 dya + off2, fx + off3, fy + off3)
  if (ex && ey) {
    FV3_TP_STRIP(true, true);
  } else if (ex) {
    FV3_TP_STRIP(true, false);
  } else if (ey) {
    FV3_TP_STRIP(false, true);
  } else {
    FV3_TP_STRIP(false, false);
  }
#undef FV3_TP_STRIP
}

// T: storage type; C: arithmetic type of the kernel.  sb: the scalars
// transported with the shared winds and fluxes (one for the plain call).
template <typename T, typename C = T>
static int fvtp2d_batch_launch(const fv3_domain* dom, int hord, const ScalarBatch<T>& sb,
                               const T* crx, const T* cry, const T* xfx, const T* yfx,
                               const T* xmf, const T* ymf, const T* area, const T* dxa,
                               const T* dya, void* stream, const YFill* yfill = nullptr) {
  const char* who = "fv3_fvtp2d";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(sb.n >= 1 && sb.n <= kMaxScalarBatch, FV3_E_ARG,
              "%s: %d scalars in one launch (1 .. %d)", who, sb.n, kMaxScalarBatch);
  bool have = crx && cry && xfx && yfx && area && dxa && dya;
  for (int b = 0; b < sb.n; ++b) have = have && sb.qx[b] && sb.qy[b] && sb.fx[b] && sb.fy[b];
  FV3_REQUIRE(have, FV3_E_ARG, "%s: required pointer is NULL", who);
  FV3_REQUIRE(hord == 5 || hord == 6 || hord == 7 || hord == 8, FV3_E_UNSUPPORTED,
              "%s: hord=%d not supported (5, 6, 7, 8)", who, hord);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(dom->sj < (1 << 30) && dom->sj2 < (1 << 30), FV3_E_ARG, "%s: row stride too large",
              who);
  // flux stores address a level by a 32-bit byte offset
  FV3_REQUIRE((int64_t)dom->nj * dom->sj * (int64_t)sizeof(T) < ((int64_t)1 << 31), FV3_E_ARG,
              "%s: level too large", who);
  Dom d = to_dom(dom);
  const int nfi = d.iec + 2 - d.isc, nfj = d.jec + 2 - d.jsc;
  const int nbi = (nfi + MTI - 1) / MTI;
  // Segment height: each segment marches its rows plus 6 halo rows, so
  // taller segments take fewer steps per output row, but fewer waves.  The
  // taller segment is taken when it needs fewer march steps and the launch
  // still has four waves per SIMD (large batches of narrow tiles, such as
  // 192 C48 tiles); otherwise 33 rows (six C96 tiles: 97 rows in three
  // segments, the fastest height of a sweep over 13-49 rows).
  const int64_t planes = (int64_t)nbi * d.nk * d.ntiles * sb.n;
  int mtj = MTJ;
  {
    const int nb_s = (nfj + MTJ - 1) / MTJ, nb_t = (nfj + MTJ_TALL - 1) / MTJ_TALL;
    if ((int64_t)nb_t * (MTJ_TALL + 6) < (int64_t)nb_s * (MTJ + 6) &&
        planes * nb_t >= 4 * kSimdsPerGpu)
      mtj = MTJ_TALL;
  }
  const int nbj = (nfj + mtj - 1) / mtj;
  const int64_t total = (int64_t)nbi * nbj * d.nk * d.ntiles * sb.n;
  FV3_REQUIRE(total < (int64_t)1 << 30, FV3_E_ARG, "%s: grid too large", who);
  const int64_t launched = (total + kNumXcd - 1) / kNumXcd * kNumXcd;
  dim3 grid((unsigned)launched), block(MW);
  hipStream_t s = as_stream(stream);
  YFill yf{};
  if (yfill) yf = *yfill;
  // algorithmic bytes: q (and its y-corner copy when it is a separate
  // array), crx, cry, both area fluxes and the optional mass fluxes read
  // once, both flux fields written once, over the compute cells
  // (per scalar: q, its y-corner copy when that is a separate array, both
  // fluxes; shared: crx, cry, both area fluxes, the optional mass fluxes)
  double nfld = 4.0 + (xmf ? 1.0 : 0.0) + (ymf ? 1.0 : 0.0);
  for (int b = 0; b < sb.n; ++b) nfld += 3.0 + (sb.qy[b] != sb.qx[b] ? 1.0 : 0.0);
  TimedLaunch timed("fvtp2d_kernel", s,
                    nfld * sizeof(T) * (double)(d.iec - d.isc + 1) * (d.jec - d.jsc + 1) *
                        d.nk * d.ntiles);
  auto go = [&](auto seg) {
    constexpr int SEG = decltype(seg)::value;
    switch (hord) {
      case 5:
        hipLaunchKernelGGL((fvtp2d_kernel<T, C, 5, SEG>), grid, block, 0, s, d, sb, crx, cry, xfx,
                           yfx, xmf, ymf, area, dxa, dya, nbi, nbj, yf);
        break;
      case 6:
        hipLaunchKernelGGL((fvtp2d_kernel<T, C, 6, SEG>), grid, block, 0, s, d, sb, crx, cry, xfx,
                           yfx, xmf, ymf, area, dxa, dya, nbi, nbj, yf);
        break;
      case 7:
        hipLaunchKernelGGL((fvtp2d_kernel<T, C, 7, SEG>), grid, block, 0, s, d, sb, crx, cry, xfx,
                           yfx, xmf, ymf, area, dxa, dya, nbi, nbj, yf);
        break;
      default:
        hipLaunchKernelGGL((fvtp2d_kernel<T, C, 8, SEG>), grid, block, 0, s, d, sb, crx, cry, xfx,
                           yfx, xmf, ymf, area, dxa, dya, nbi, nbj, yf);
        break;
    }
  };
  if (mtj == MTJ_TALL)
    go(std::integral_constant<int, MTJ_TALL>{});
  else
    go(std::integral_constant<int, MTJ>{});
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

template <typename T, typename C = T>
static int fvtp2d_launch(const fv3_domain* dom, int hord, const T* q_x, const T* q_y,
                         const T* crx, const T* cry, const T* xfx, const T* yfx,
                         const T* xmf, const T* ymf, const T* area, const T* dxa,
                         const T* dya, T* fx, T* fy, void* stream,
                         const YFill* yfill = nullptr) {
  ScalarBatch<T> sb{};
  sb.n = 1;
  sb.qx[0] = q_x;
  sb.qy[0] = q_y;
  sb.fx[0] = fx;
  sb.fy[0] = fy;
  return fvtp2d_batch_launch<T, C>(dom, hord, sb, crx, cry, xfx, yfx, xmf, ymf, area, dxa, dya,
                                   stream, yfill);
}

// ---------------------------------------------------------------------------
// Stand-alone 1-D PPM (XPiecewiseParabolic / YPiecewiseParabolic, xppm.py:283-348):
// one thread per interface of an arbitrary origin/extent.
// ---------------------------------------------------------------------------

template <typename T, int ORD, int DIR>
__global__ void ppm_flux_kernel(Dom d, int oi, int oj, int ei, int ej, const T* __restrict__ q,
                                const T* __restrict__ c, const T* __restrict__ dxa,
                                T* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)ei * ej) return;
  const int i = oi + (int)(n % ei), j = oj + (int)(n / ei);
  const int64_t off3 = (int64_t)blockIdx.z * d.st + (int64_t)blockIdx.y * d.sk;
  const int64_t off2 = (int64_t)blockIdx.z * d.st2;
  const T* qq = q + off3;
  T q6[6];
  if (DIR == 0) {
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int ii = i - 3 + m;
      q6[m] = (ii >= 0 && ii < d.ni) ? qq[(int64_t)j * d.sj + ii] : T(0);
    }
    if (i - 3 < 0 || i + 2 >= d.ni) return;
    DxRow<T> dr{dxa + off2 + (int64_t)j * d.sj2 + (i - 3), 1};
    out[off3 + (int64_t)j * d.sj + i] =
        ppm_face<T, ORD>(q6, c[off3 + (int64_t)j * d.sj + i], i, d.is, d.ie, dr);
  } else {
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int jj = j - 3 + m;
      q6[m] = (jj >= 0 && jj < d.nj) ? qq[(int64_t)jj * d.sj + i] : T(0);
    }
    if (j - 3 < 0 || j + 2 >= d.nj) return;
    DxRow<T> dr{dxa + off2 + (int64_t)(j - 3) * d.sj2 + i, d.sj2};
    out[off3 + (int64_t)j * d.sj + i] =
        ppm_face<T, ORD>(q6, c[off3 + (int64_t)j * d.sj + i], j, d.js, d.je, dr);
  }
}

template <typename T, int DIR>
static void ppm_flux_dispatch(int ord, dim3 g, hipStream_t s, const Dom& d, int oi, int oj,
                              int ei, int ej, const T* q, const T* c, const T* dxa, T* out) {
  switch (ord) {
    case 5:
      hipLaunchKernelGGL((ppm_flux_kernel<T, 5, DIR>), g, dim3(256), 0, s, d, oi, oj, ei, ej, q,
                         c, dxa, out);
      break;
    case 6:
      hipLaunchKernelGGL((ppm_flux_kernel<T, 6, DIR>), g, dim3(256), 0, s, d, oi, oj, ei, ej, q,
                         c, dxa, out);
      break;
    case 7:
      hipLaunchKernelGGL((ppm_flux_kernel<T, 7, DIR>), g, dim3(256), 0, s, d, oi, oj, ei, ej, q,
                         c, dxa, out);
      break;
    default:
      hipLaunchKernelGGL((ppm_flux_kernel<T, 8, DIR>), g, dim3(256), 0, s, d, oi, oj, ei, ej, q,
                         c, dxa, out);
      break;
  }
}

template <typename T>
static int ppm_flux_launch(const fv3_domain* dom, int direction, int ord, int oi, int oj, int ei,
                           int ej, const T* q, const T* c, const T* dxa, T* out, void* stream) {
  const char* who = "fv3_ppm_flux";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q && c && dxa && out, FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(direction == 0 || direction == 1, FV3_E_ARG, "%s: direction must be 0 or 1", who);
  FV3_REQUIRE(ord == 5 || ord == 6 || ord == 7 || ord == 8, FV3_E_UNSUPPORTED,
              "%s: order %d not supported (5, 6, 7, 8)", who, ord);
  FV3_REQUIRE(ei >= 0 && ej >= 0 && oi >= 0 && oj >= 0 && oi + ei <= dom->ni &&
                  oj + ej <= dom->nj,
              FV3_E_ARG, "%s: origin/extent outside the storage", who);
  FV3_REQUIRE(direction == 1 || (oi >= 3 && oi + ei + 2 <= dom->ni), FV3_E_ARG,
              "%s: x interfaces need 3 cells on each side inside the storage", who);
  FV3_REQUIRE(direction == 0 || (oj >= 3 && oj + ej + 2 <= dom->nj), FV3_E_ARG,
              "%s: y interfaces need 3 cells on each side inside the storage", who);
  if ((int64_t)ei * ej == 0 || dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  dim3 g((unsigned)(((int64_t)ei * ej + 255) / 256), d.nk, d.ntiles);
  hipStream_t s = as_stream(stream);
  if (direction == 0)
    ppm_flux_dispatch<T, 0>(ord, g, s, d, oi, oj, ei, ej, q, c, dxa, out);
  else
    ppm_flux_dispatch<T, 1>(ord, g, s, d, oi, oj, ei, ej, q, c, dxa, out);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

// ---------------------------------------------------------------------------
// CopyCornersXY: q_x/q_y = q, then the 36 x- and 36 y-corner cells of each
// level are replaced by rotated neighbours.  The destination -> source table
// (corners.py:425-540) is resolved to absolute indices on the host.
// ---------------------------------------------------------------------------


template <typename T>
__global__ void copy_full_kernel(Dom d, const T* __restrict__ q, T* __restrict__ a,
                                 T* __restrict__ b) {
  const int64_t plane = (int64_t)d.ni * d.nj;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= plane) return;
  const int i = (int)(n % d.ni), j = (int)(n / d.ni);
  const int64_t o = (int64_t)blockIdx.z * d.st + (int64_t)blockIdx.y * d.sk + (int64_t)j * d.sj + i;
  const T v = q[o];
  if (a) a[o] = v;
  if (b) b[o] = v;
}

template <typename T>
__global__ void corners_kernel(Dom d, CornerTable tab, const T* __restrict__ q, T* qx,
                               T* __restrict__ qy) {
  const int e = threadIdx.x;
  if (e >= tab.n) return;
  const int64_t base = (int64_t)blockIdx.z * d.st + (int64_t)blockIdx.y * d.sk;
  const T v = q[base + (int64_t)tab.src_j[e] * d.sj + tab.src_i[e]];
  T* out = tab.dir[e] ? qy : qx;
  out[base + (int64_t)tab.dst_j[e] * d.sj + tab.dst_i[e]] = v;
}

// (di, dj) source offsets of the x table rows; destination anchors as
// (start=0/end=1, offset) for i then j, two regions per row.
static const int kXRows[18][10] = {
    {0, -3, 0, -3, 1, 3, 0, -3, 0, 5},   {0, -2, 0, -3, 1, 3, 0, -2, -1, 4},
    {0, -1, 0, -3, 1, 3, 0, -1, -2, 3},  {0, -3, 0, -2, 1, 2, 0, -3, 1, 4},
    {0, -2, 0, -2, 1, 2, 0, -2, 0, 3},   {0, -1, 0, -2, 1, 2, 0, -1, -1, 2},
    {0, -3, 0, -1, 1, 1, 0, -3, 2, 3},   {0, -2, 0, -1, 1, 1, 0, -2, 1, 2},
    {0, -1, 0, -1, 1, 1, 0, -1, 0, 1},   {0, -3, 1, 1, 1, 1, 1, 3, 2, -3},
    {0, -2, 1, 1, 1, 1, 1, 2, 1, -2},    {0, -1, 1, 1, 1, 1, 1, 1, 0, -1},
    {0, -3, 1, 2, 1, 2, 1, 3, 1, -4},    {0, -2, 1, 2, 1, 2, 1, 2, 0, -3},
    {0, -1, 1, 2, 1, 2, 1, 1, -1, -2},   {0, -3, 1, 3, 1, 3, 1, 3, 0, -5},
    {0, -2, 1, 3, 1, 3, 1, 2, -1, -4},   {0, -1, 1, 3, 1, 3, 1, 1, -2, -3},
};
static const int kYRows[18][10] = {
    {0, -3, 0, -3, 0, -3, 1, 3, 5, 0},   {0, -2, 0, -3, 0, -3, 1, 2, 4, 1},
    {0, -1, 0, -3, 0, -3, 1, 1, 3, 2},   {0, -3, 0, -2, 0, -2, 1, 3, 4, -1},
    {0, -2, 0, -2, 0, -2, 1, 2, 3, 0},   {0, -1, 0, -2, 0, -2, 1, 1, 2, 1},
    {0, -3, 0, -1, 0, -1, 1, 3, 3, -2},  {0, -2, 0, -1, 0, -1, 1, 2, 2, -1},
    {0, -1, 0, -1, 0, -1, 1, 1, 1, 0},   {1, 1, 0, -3, 1, 3, 1, 1, -3, 2},
    {1, 2, 0, -3, 1, 3, 1, 2, -4, 1},    {1, 3, 0, -3, 1, 3, 1, 3, -5, 0},
    {1, 1, 0, -2, 1, 2, 1, 1, -2, 1},    {1, 2, 0, -2, 1, 2, 1, 2, -3, 0},
    {1, 3, 0, -2, 1, 2, 1, 3, -4, -1},   {1, 1, 0, -1, 1, 1, 1, 1, -1, 0},
    {1, 2, 0, -1, 1, 1, 1, 2, -2, -1},   {1, 3, 0, -1, 1, 1, 1, 3, -3, -2},
};

// Fills the table for the edges this domain owns; corner regions need both
// anchors of the region (e.g. i_start and j_start) to exist.
int build_corner_table(const fv3_domain* d, int which, CornerTable* tab) {
  const bool have_is = d->i_start > FV3_NO_EDGE_LO / 2 && d->i_start < FV3_NO_EDGE_HI / 2;
  const bool have_ie = d->i_end > FV3_NO_EDGE_LO / 2 && d->i_end < FV3_NO_EDGE_HI / 2;
  const bool have_js = d->j_start > FV3_NO_EDGE_LO / 2 && d->j_start < FV3_NO_EDGE_HI / 2;
  const bool have_je = d->j_end > FV3_NO_EDGE_LO / 2 && d->j_end < FV3_NO_EDGE_HI / 2;
  tab->n = 0;
  for (int dir = 0; dir < 2; ++dir) {
    if (!(which & (1 << dir))) continue;
    const int(*rows)[10] = dir == 0 ? kXRows : kYRows;
    for (int r = 0; r < 18; ++r) {
      for (int reg = 0; reg < 2; ++reg) {
        const int* a = rows[r] + reg * 4;
        const bool ok_i = a[0] == 0 ? have_is : have_ie;
        const bool ok_j = a[2] == 0 ? have_js : have_je;
        if (!ok_i || !ok_j) continue;
        const int di = (a[0] == 0 ? d->i_start : d->i_end) + a[1];
        const int dj = (a[2] == 0 ? d->j_start : d->j_end) + a[3];
        const int si = di + rows[r][8], sj = dj + rows[r][9];
        if (di < 0 || di >= d->ni || dj < 0 || dj >= d->nj || si < 0 || si >= d->ni ||
            sj < 0 || sj >= d->nj)
          continue;
        const int e = tab->n++;
        tab->dst_i[e] = di;
        tab->dst_j[e] = dj;
        tab->src_i[e] = si;
        tab->src_j[e] = sj;
        tab->dir[e] = (unsigned char)dir;
      }
    }
  }
  return FV3_OK;
}

template <typename T>
static int corners_launch(const fv3_domain* dom, const T* q, T* q_x, T* q_y, void* stream) {
  const char* who = "fv3_copy_corners_xy";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q && q_x && q_y, FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(q_y != q, FV3_E_ARG, "%s: q_y must not alias q", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  hipStream_t s = as_stream(stream);
  const int64_t plane = (int64_t)d.ni * d.nj;
  dim3 g1((unsigned)((plane + 255) / 256), d.nk, d.ntiles);
  hipLaunchKernelGGL((copy_full_kernel<T>), g1, dim3(256), 0, s, d, q,
                     q_x == q ? (T*)nullptr : q_x, q_y);
  FV3_HIP_CHECK_LAUNCH(who);
  CornerTable tab;
  build_corner_table(dom, 3, &tab);
  if (tab.n > 0) {
    hipLaunchKernelGGL((corners_kernel<T>), dim3(1, d.nk, d.ntiles), dim3(128), 0, s, d, tab,
                       q, q_x, q_y);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  return FV3_OK;
}

// fv_tp_2d of q preceded by the reference's corner copies: copy_corners_x in
// place on q (as CopyCornersXY with q_x = q) and copy_corners_y applied as the
// march reads q, instead of into a full copy of q.  Every y destination is
// also an x destination and no source is a corner cell, so q after its x
// fill, read with the y fill, is exactly the copy the two-step form builds.
// When the two destination sets differ, the copy is built in qy_scratch.
// The x corner table of the domain and, when every y destination is also an
// x destination and each corner's 3 x 3 y square is an affine map of its
// destinations, the fit of those maps (yf); returns false otherwise.
static bool fit_corner_fill(const fv3_domain* dom, CornerTable* xt, YFill* yf_out) {
  CornerTable tab;
  build_corner_table(dom, 3, &tab);
  xt->n = 0;
  int ny = 0;
  int ydi[72], ydj[72], ysi[72], ysj[72];
  for (int e = 0; e < tab.n; ++e) {
    if (tab.dir[e] == 0) {
      const int m = xt->n++;
      xt->dst_i[m] = tab.dst_i[e];
      xt->dst_j[m] = tab.dst_j[e];
      xt->src_i[m] = tab.src_i[e];
      xt->src_j[m] = tab.src_j[e];
      xt->dir[m] = 0;
    } else {
      ydi[ny] = tab.dst_i[e];
      ydj[ny] = tab.dst_j[e];
      ysi[ny] = tab.src_i[e];
      ysj[ny] = tab.src_j[e];
      ++ny;
    }
  }
  // every y destination must be an x destination and vice versa, and each
  // corner's 9 y entries an exact affine map of the destination
  bool same = xt->n == ny;
  for (int a = 0; same && a < ny; ++a) {
    bool hit = false;
    for (int b = 0; b < xt->n; ++b)
      hit = hit || (xt->dst_i[b] == ydi[a] && xt->dst_j[b] == ydj[a]);
    same = hit;
  }
  YFill yf{};
  const int ci_[4] = {dom->i_start, dom->i_end, dom->i_start, dom->i_end};
  const int cj_[4] = {dom->j_start, dom->j_start, dom->j_end, dom->j_end};
  int counted = 0;
  for (int c = 0; same && c < 4; ++c) {
    // entries of corner c: dst beyond both of its anchors
    int idx[9], n9 = 0;
    for (int e = 0; e < ny; ++e) {
      const int di = ydi[e] - ci_[c], dj = ydj[e] - cj_[c];
      const bool in_i = (c & 1) ? (di >= 1 && di <= 3) : (di >= -3 && di <= -1);
      const bool in_j = (c & 2) ? (dj >= 1 && dj <= 3) : (dj >= -3 && dj <= -1);
      if (in_i && in_j) {
        if (n9 == 9) {
          same = false;
          break;
        }
        idx[n9++] = e;
      }
    }
    if (!same || n9 == 0) continue;
    if (n9 != 9) {
      same = false;
      break;
    }
    // fit s = a + b i + c j from three cells of the 3 x 3 square
    int e0 = -1, e1 = -1, e2 = -1;
    for (int m = 0; m < 9; ++m) {
      const int e = idx[m];
      if (e0 < 0) e0 = e;
      else if (e1 < 0 && ydj[e] == ydj[e0] && ydi[e] != ydi[e0]) e1 = e;
      else if (e2 < 0 && ydi[e] == ydi[e0] && ydj[e] != ydj[e0]) e2 = e;
    }
    if (e1 < 0 || e2 < 0) {
      same = false;
      break;
    }
    const int dxi = ydi[e1] - ydi[e0], dyj = ydj[e2] - ydj[e0];
    const int bi = (ysi[e1] - ysi[e0]) / dxi, bj = (ysj[e1] - ysj[e0]) / dxi;
    const int cI = (ysi[e2] - ysi[e0]) / dyj, cJ = (ysj[e2] - ysj[e0]) / dyj;
    const int ai = ysi[e0] - bi * ydi[e0] - cI * ydj[e0];
    const int aj = ysj[e0] - bj * ydi[e0] - cJ * ydj[e0];
    for (int m = 0; m < 9; ++m) {
      const int e = idx[m];
      same = same && ai + bi * ydi[e] + cI * ydj[e] == ysi[e] &&
             aj + bj * ydi[e] + cJ * ydj[e] == ysj[e];
    }
    yf.present |= 1 << c;
    yf.ai[c] = ai;
    yf.bi[c] = bi;
    yf.ci[c] = cI;
    yf.aj[c] = aj;
    yf.bj[c] = bj;
    yf.cj[c] = cJ;
    counted += 9;
  }
  *yf_out = yf;
  return same && counted == ny;
}

// x corner fill in place of every scalar of a batch: blockIdx.z is
// tile + ntiles * slot.
template <typename T>
__global__ void corners_x_batch_kernel(Dom d, CornerTable tab, ScalarBatch<T> sb) {
  const int e = threadIdx.x;
  if (e >= tab.n) return;
  const int t = (int)blockIdx.z % d.ntiles, b = (int)blockIdx.z / d.ntiles;
  T* q = const_cast<T*>(pick_slot(sb.qx, b));
  const int64_t base = (int64_t)t * d.st + (int64_t)blockIdx.y * d.sk;
  q[base + (int64_t)tab.dst_j[e] * d.sj + tab.dst_i[e]] =
      q[base + (int64_t)tab.src_j[e] * d.sj + tab.src_i[e]];
}

// fv_tp_2d of q preceded by the reference's corner copies: copy_corners_x in
// place on q (as CopyCornersXY with q_x = q) and copy_corners_y applied as the
// march reads q, instead of into a full copy of q.  Every y destination is
// also an x destination and no source is a corner cell, so q after its x
// fill, read with the y fill, is exactly the copy the two-step form builds.
// When the two destination sets differ, the copy is built in qy_scratch.
template <typename T, typename C>
static int fvtp2d_corner_fill_launch(const fv3_domain* dom, int hord, T* q, T* qy_scratch,
                                     const T* crx, const T* cry, const T* xfx, const T* yfx,
                                     const T* xmf, const T* ymf, const T* area, const T* dxa,
                                     const T* dya, T* fx, T* fy, void* stream) {
  const char* who = "fv3_fvtp2d";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q && qy_scratch, FV3_E_ARG, "%s: required pointer is NULL", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  CornerTable xt;
  YFill yf{};
  if (!fit_corner_fill(dom, &xt, &yf)) {
    rc = corners_launch<T>(dom, q, q, qy_scratch, stream);
    if (rc) return rc;
    return fvtp2d_launch<T, C>(dom, hord, q, qy_scratch, crx, cry, xfx, yfx, xmf, ymf, area,
                               dxa, dya, fx, fy, stream);
  }
  if (xt.n > 0) {
    Dom d = to_dom(dom);
    hipLaunchKernelGGL((corners_kernel<T>), dim3(1, d.nk, d.ntiles), dim3(128), 0,
                       as_stream(stream), d, xt, q, q, q);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  return fvtp2d_launch<T, C>(dom, hord, q, q, crx, cry, xfx, yfx, xmf, ymf, area, dxa, dya, fx,
                             fy, stream, &yf);
}

int fvtp2d_batch_corner_fill_f64(const fv3_domain* dom, int hord, int n, double* const* q,
                                 double* const* fx, double* const* fy, const double* crx,
                                 const double* cry, const double* xfx, const double* yfx,
                                 const double* xmf, const double* ymf, const double* area,
                                 const double* dxa, const double* dya, void* stream) {
  const char* who = "fv3_fvtp2d";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(n >= 1 && n <= kMaxScalarBatch, FV3_E_ARG, "%s: %d scalars (1 .. %d)", who, n,
              kMaxScalarBatch);
  FV3_REQUIRE(q && fx && fy, FV3_E_ARG, "%s: NULL pointer table", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  CornerTable xt;
  YFill yf{};
  FV3_REQUIRE(fit_corner_fill(dom, &xt, &yf), FV3_E_UNSUPPORTED,
              "%s: the batched transport needs the x and y corner sets of a whole cube tile "
              "(or no corners); transport the scalars one by one", who);
  ScalarBatch<double> sb{};
  sb.n = n;
  for (int b = 0; b < n; ++b) {
    FV3_REQUIRE(q[b] && fx[b] && fy[b], FV3_E_ARG, "%s: NULL pointer in slot %d", who, b);
    sb.qx[b] = sb.qy[b] = q[b];
    sb.fx[b] = fx[b];
    sb.fy[b] = fy[b];
  }
  if (xt.n > 0) {
    Dom d = to_dom(dom);
    hipLaunchKernelGGL((corners_x_batch_kernel<double>), dim3(1, d.nk, d.ntiles * n), dim3(128),
                       0, as_stream(stream), d, xt, sb);
    FV3_HIP_CHECK_LAUNCH(who);
  }
  return fvtp2d_batch_launch<double, double>(dom, hord, sb, crx, cry, xfx, yfx, xmf, ymf, area,
                                             dxa, dya, stream, &yf);
}

int fvtp2d_corner_fill_f64(const fv3_domain* dom, int hord, double* q, double* qy_scratch,
                           const double* crx, const double* cry, const double* xfx,
                           const double* yfx, const double* xmf, const double* ymf,
                           const double* area, const double* dxa, const double* dya, double* fx,
                           double* fy, bool fp32_arithmetic, void* stream) {
  if (fp32_arithmetic)
    return fvtp2d_corner_fill_launch<double, float>(dom, hord, q, qy_scratch, crx, cry, xfx, yfx,
                                                    xmf, ymf, area, dxa, dya, fx, fy, stream);
  return fvtp2d_corner_fill_launch<double, double>(dom, hord, q, qy_scratch, crx, cry, xfx, yfx,
                                                   xmf, ymf, area, dxa, dya, fx, fy, stream);
}

int fvtp2d_corner_fill_f32(const fv3_domain* dom, int hord, float* q, float* qy_scratch,
                           const float* crx, const float* cry, const float* xfx,
                           const float* yfx, const float* xmf, const float* ymf,
                           const float* area, const float* dxa, const float* dya, float* fx,
                           float* fy, void* stream) {
  return fvtp2d_corner_fill_launch<float, float>(dom, hord, q, qy_scratch, crx, cry, xfx, yfx,
                                                 xmf, ymf, area, dxa, dya, fx, fy, stream);
}

}  // namespace fv3

extern "C" {

int32_t fv3_fvtp2d_f64(const fv3_domain* d, int32_t hord, const double* q_x, const double* q_y,
                       const double* crx, const double* cry, const double* x_area_flux,
                       const double* y_area_flux, const double* x_mass_flux,
                       const double* y_mass_flux, const double* area, const double* dxa,
                       const double* dya, double* q_x_flux, double* q_y_flux, void* stream) {
  return fv3::fvtp2d_launch<double>(d, hord, q_x, q_y, crx, cry, x_area_flux, y_area_flux,
                                    x_mass_flux, y_mass_flux, area, dxa, dya, q_x_flux,
                                    q_y_flux, stream);
}

int32_t fv3_fvtp2d_corner_fill_f64(const fv3_domain* d, int32_t hord, double* q,
                                   double* qy_scratch, const double* crx, const double* cry,
                                   const double* x_area_flux, const double* y_area_flux,
                                   const double* x_mass_flux, const double* y_mass_flux,
                                   const double* area, const double* dxa, const double* dya,
                                   double* q_x_flux, double* q_y_flux, void* stream) {
  return fv3::fvtp2d_corner_fill_f64(d, hord, q, qy_scratch, crx, cry, x_area_flux, y_area_flux,
                                     x_mass_flux, y_mass_flux, area, dxa, dya, q_x_flux,
                                     q_y_flux, false, stream);
}

int32_t fv3_fvtp2d_f32(const fv3_domain* d, int32_t hord, const float* q_x, const float* q_y,
                       const float* crx, const float* cry, const float* x_area_flux,
                       const float* y_area_flux, const float* x_mass_flux,
                       const float* y_mass_flux, const float* area, const float* dxa,
                       const float* dya, float* q_x_flux, float* q_y_flux, void* stream) {
  return fv3::fvtp2d_launch<float>(d, hord, q_x, q_y, crx, cry, x_area_flux, y_area_flux,
                                   x_mass_flux, y_mass_flux, area, dxa, dya, q_x_flux,
                                   q_y_flux, stream);
}

int32_t fv3_fvtp2d_mixed_f64(const fv3_domain* d, int32_t hord, const double* q_x,
                             const double* q_y, const double* crx, const double* cry,
                             const double* x_area_flux, const double* y_area_flux,
                             const double* x_mass_flux, const double* y_mass_flux,
                             const double* area, const double* dxa, const double* dya,
                             double* q_x_flux, double* q_y_flux, void* stream) {
  return fv3::fvtp2d_launch<double, float>(d, hord, q_x, q_y, crx, cry, x_area_flux,
                                           y_area_flux, x_mass_flux, y_mass_flux, area, dxa,
                                           dya, q_x_flux, q_y_flux, stream);
}

int32_t fv3_ppm_flux_f64(const fv3_domain* d, int32_t direction, int32_t ord, int32_t oi,
                         int32_t oj, int32_t ei, int32_t ej, const double* q, const double* c,
                         const double* dxa, double* out, void* stream) {
  return fv3::ppm_flux_launch<double>(d, direction, ord, oi, oj, ei, ej, q, c, dxa, out, stream);
}

int32_t fv3_ppm_flux_f32(const fv3_domain* d, int32_t direction, int32_t ord, int32_t oi,
                         int32_t oj, int32_t ei, int32_t ej, const float* q, const float* c,
                         const float* dxa, float* out, void* stream) {
  return fv3::ppm_flux_launch<float>(d, direction, ord, oi, oj, ei, ej, q, c, dxa, out, stream);
}

int32_t fv3_copy_corners_xy_f64(const fv3_domain* d, const double* q, double* q_x, double* q_y,
                                void* stream) {
  return fv3::corners_launch<double>(d, q, q_x, q_y, stream);
}

int32_t fv3_copy_corners_xy_f32(const fv3_domain* d, const float* q, float* q_x, float* q_y,
                                void* stream) {
  return fv3::corners_launch<float>(d, q, q_x, q_y, stream);
}
}
