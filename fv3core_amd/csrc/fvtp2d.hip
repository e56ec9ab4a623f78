// Fused PPM transport (fv_tp_2d) and the CopyCornersXY corner copies.
//
// Reference: fv3core/stencils/fvtp2d.py:151-362 runs eight stencils over
// full-size HBM temporaries (y_ppm inner, q_i, x_ppm outer, x_ppm inner,
// q_j, y_ppm outer, final_fluxes).  Here one workgroup owns a TI x TJ block
// of interfaces on one level of one tile: it stages q (x- and
// y-differentiable copies) with the 3-cell halo in LDS, evaluates the inner
// sweeps, the cross-term updates and the outer sweeps in LDS, and writes only
// the two flux fields.  HBM traffic is the compulsory reads of q, the Courant
// numbers and the area/mass fluxes plus the two flux writes.
#include "common.hpp"
#include "timing.hpp"
#include "ppm.hpp"

namespace fv3 {

constexpr int TR_TI = 16;   // interfaces per block in i
constexpr int TR_TJ = 24;   // interfaces per block in j
constexpr int TR_NT = 256;  // threads per block (4 waves of 64)

template <typename T>
__device__ __forceinline__ T ld2(const T* __restrict__ p, int i, int j, int64_t sj, int ni,
                                 int nj) {
  return (i >= 0 && i < ni && j >= 0 && j < nj) ? p[(int64_t)j * sj + i] : T(0);
}

// Number of XCDs on MI355X: consecutive workgroups are dispatched round-robin
// over them, each with its own L2.
constexpr int kNumXcd = 8;

// Workgroup id -> logical block so that each XCD runs a contiguous range of
// logical blocks (whole planes of one level), keeping the halo rows that
// neighbouring blocks share in that XCD's L2.
__device__ __forceinline__ int xcd_logical_block(int hw, int total) {
  const int per = (total + kNumXcd - 1) / kNumXcd;
  return (hw % kNumXcd) * per + hw / kNumXcd;
}

template <typename T, int ORD, bool EDGES>
__device__ __forceinline__ void fvtp2d_block(
    const Dom& d, int i0, int j0, bool same_q, const T* __restrict__ qx,
    const T* __restrict__ qy, const T* __restrict__ cx, const T* __restrict__ cy,
    const T* __restrict__ ax, const T* __restrict__ ay, const T* __restrict__ ux,
    const T* __restrict__ uy, const T* __restrict__ ar, const T* __restrict__ dx,
    const T* __restrict__ dy, T* __restrict__ fx, T* __restrict__ fy,
    T (*__restrict__ sqx)[TR_TI + 6], T (*__restrict__ sqy_own)[TR_TI + 6],
    T (*__restrict__ sym)[TR_TI + 5], T (*__restrict__ sxm)[TR_TI + 1],
    T (*__restrict__ sady)[TR_TI + 5], T (*__restrict__ sadx)[TR_TI]) {
  constexpr int TI = TR_TI, TJ = TR_TJ;
  constexpr int QW = TI + 6, QH = TJ + 6;
  const int tid = threadIdx.x;
  const int ni = d.ni, nj = d.nj;
  const int sj = (int)d.sj, sj2 = (int)d.sj2;
  T (*sqy)[QW] = same_q ? sqx : sqy_own;
  auto ld = [&](const T* __restrict__ p, int i, int j, int stride) -> T {
    return (i >= 0 && i < ni && j >= 0 && j < nj) ? p[j * stride + i] : T(0);
  };

  // 1. stage q_x (and q_y unless it is the same storage) with a 3-cell halo
  for (int n = tid; n < QW * QH; n += TR_NT) {
    const int li = n % QW, lj = n / QW;
    const int i = i0 - 3 + li, j = j0 - 3 + lj;
    sqx[lj][li] = ld(qx, i, j, sj);
    if (!same_q) sqy_own[lj][li] = ld(qy, i, j, sj);
  }
  __syncthreads();

  // 2a. x-inner PPM on q_x: faces i in [i0, i0+TI], rows j in [j0-3, j0+TJ+2)
  for (int n = tid; n < (TI + 1) * (TJ + 5); n += TR_NT) {
    const int li = n % (TI + 1), lj = n / (TI + 1);
    const int i = i0 + li, j = j0 - 3 + lj;
    T q6[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) q6[m] = sqx[lj][li + m];
    const T c = ld(cx, i, j, sj);
    DxRow<T> dr{dx + (int64_t)j * sj2 + (i - 3), 1};
    sxm[lj][li] = (j < nj) ? ppm_face_sel<T, ORD, EDGES>(q6, c, i, d.is, d.ie, dr) : T(0);
  }
  // 2b. y-inner PPM on q_y: faces j in [j0, j0+TJ], columns i in [i0-3, i0+TI+2)
  for (int n = tid; n < (TI + 5) * (TJ + 1); n += TR_NT) {
    const int li = n % (TI + 5), lj = n / (TI + 5);
    const int i = i0 - 3 + li, j = j0 + lj;
    T q6[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) q6[m] = sqy[lj + m][li];
    const T c = ld(cy, i, j, sj);
    DxRow<T> dr{dy + (int64_t)(j - 3) * sj2 + i, sj2};
    sym[lj][li] = (i < ni) ? ppm_face_sel<T, ORD, EDGES>(q6, c, j, d.js, d.je, dr) : T(0);
  }
  __syncthreads();

  // 3a. q_j (fvtp2d.py:50-60): cells i in [i0, i0+TI), rows [j0-3, j0+TJ+2)
  for (int n = tid; n < TI * (TJ + 5); n += TR_NT) {
    const int li = n % TI, lj = n / TI;
    const int i = i0 + li, j = j0 - 3 + lj;
    const T a = ld(ar, i, j, sj2);
    const T f0 = ld(ax, i, j, sj);
    const T f1 = ld(ax, i + 1, j, sj);
    const T fx1_0 = f0 * sxm[lj][li];
    const T fx1_1 = f1 * sxm[lj][li + 1];
    const T area_with_x_flux = a + f0 - f1;
    sadx[lj][li] = (sqx[lj][li + 3] * a + fx1_0 - fx1_1) / area_with_x_flux;
  }
  // 3b. q_i (fvtp2d.py:33-47): cells i in [i0-3, i0+TI+2), rows [j0, j0+TJ)
  for (int n = tid; n < (TI + 5) * TJ; n += TR_NT) {
    const int li = n % (TI + 5), lj = n / (TI + 5);
    const int i = i0 - 3 + li, j = j0 + lj;
    const T a = ld(ar, i, j, sj2);
    const T g0 = ld(ay, i, j, sj);
    const T g1 = ld(ay, i, j + 1, sj);
    const T fyy0 = g0 * sym[lj][li];
    const T fyy1 = g1 * sym[lj + 1][li];
    sady[lj][li] = (sqy[lj + 3][li] * a + fyy0 - fyy1) / (a + g0 - g1);
  }
  __syncthreads();

  // 4. outer sweeps and final_fluxes (fvtp2d.py:63-92)
  for (int n = tid; n < TI * TJ; n += TR_NT) {
    const int li = n % TI, lj = n / TI;
    const int i = i0 + li, j = j0 + lj;
    const int o = j * sj + i;
    if (i <= d.iec + 1 && j <= d.jec) {
      T q6[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) q6[m] = sady[lj][li + m];
      DxRow<T> dr{dx + (int64_t)j * sj2 + (i - 3), 1};
      const T outer = ppm_face_sel<T, ORD, EDGES>(q6, cx[o], i, d.is, d.ie, dr);
      fx[o] = T(0.5) * (outer + sxm[lj + 3][li]) * ux[o];
    }
    if (i <= d.iec && j <= d.jec + 1) {
      T q6[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) q6[m] = sadx[lj + m][li];
      DxRow<T> dr{dy + (int64_t)(j - 3) * sj2 + i, sj2};
      const T outer = ppm_face_sel<T, ORD, EDGES>(q6, cy[o], j, d.js, d.je, dr);
      fy[o] = T(0.5) * (outer + sym[lj][li + 3]) * uy[o];
    }
  }
}

template <typename T, int ORD>
__global__ void __launch_bounds__(TR_NT)
    fvtp2d_kernel(Dom d, const T* __restrict__ q_x, const T* __restrict__ q_y,
                  const T* __restrict__ crx, const T* __restrict__ cry,
                  const T* __restrict__ xfx, const T* __restrict__ yfx,
                  const T* __restrict__ xmf, const T* __restrict__ ymf,
                  const T* __restrict__ area, const T* __restrict__ dxa,
                  const T* __restrict__ dya, T* __restrict__ fx, T* __restrict__ fy,
                  int nbi, int nbj) {
  constexpr int TI = TR_TI, TJ = TR_TJ;
  constexpr int QW = TI + 6, QH = TJ + 6;
  __shared__ T sqx[QH][QW];
  __shared__ T sqy[QH][QW];
  __shared__ T sym[TJ + 1][TI + 5];   // y-inner means, faces (i0-3.., j0..)
  __shared__ T sxm[TJ + 5][TI + 1];   // x-inner means, faces (i0.., j0-3..)
  __shared__ T sady[TJ][TI + 5];      // q advected in y, cells (i0-3.., j0..)
  __shared__ T sadx[TJ + 5][TI];      // q advected in x, cells (i0.., j0-3..)

  const int per_plane = nbi * nbj;
  const int total = per_plane * d.nk * d.ntiles;
  const int lb = xcd_logical_block(blockIdx.x, total);
  if (lb >= total) return;
  const int plane = lb / per_plane;
  const int within = lb % per_plane;
  const int bi = within % nbi, bj = within / nbi;
  const int k = plane % d.nk, t = plane / d.nk;
  const int i0 = d.isc + bi * TI;
  const int j0 = d.jsc + bj * TJ;
  const int64_t off3 = (int64_t)t * d.st + (int64_t)k * d.sk;
  const int64_t off2 = (int64_t)t * d.st2;
  const bool same_q = (q_x == q_y);
  const T* qx = q_x + off3;
  const T* qy = q_y + off3;
  const T* cx = crx + off3;
  const T* cy = cry + off3;
  const T* ax = xfx + off3;
  const T* ay = yfx + off3;
  const T* ux = (xmf ? xmf : xfx) + off3;
  const T* uy = (ymf ? ymf : yfx) + off3;
  const T* ar = area + off2;
  const T* dx = dxa + off2;
  const T* dy = dya + off2;
  T* fxp = fx + off3;
  T* fyp = fy + off3;
  // a block needs the tile-edge formulas when its stencil reach (cells
  // i0-4 .. i0+TI+4) touches the edge regions (anchor -2 .. +3)
  auto near = [](int lo, int hi, int a) { return hi >= a - 2 && lo <= a + 3; };
  const bool edges = near(i0 - 4, i0 + TI + 4, d.is) || near(i0 - 4, i0 + TI + 4, d.ie) ||
                     near(j0 - 4, j0 + TJ + 4, d.js) || near(j0 - 4, j0 + TJ + 4, d.je);
  if (edges) {
    fvtp2d_block<T, ORD, true>(d, i0, j0, same_q, qx, qy, cx, cy, ax, ay, ux, uy, ar, dx, dy,
                               fxp, fyp, sqx, sqy, sym, sxm, sady, sadx);
  } else {
    fvtp2d_block<T, ORD, false>(d, i0, j0, same_q, qx, qy, cx, cy, ax, ay, ux, uy, ar, dx, dy,
                                fxp, fyp, sqx, sqy, sym, sxm, sady, sadx);
  }
}

template <typename T>
static int fvtp2d_launch(const fv3_domain* dom, int hord, const T* q_x, const T* q_y,
                         const T* crx, const T* cry, const T* xfx, const T* yfx,
                         const T* xmf, const T* ymf, const T* area, const T* dxa,
                         const T* dya, T* fx, T* fy, void* stream) {
  const char* who = "fv3_fvtp2d";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q_x && q_y && crx && cry && xfx && yfx && area && dxa && dya && fx && fy,
              FV3_E_ARG, "%s: required pointer is NULL", who);
  FV3_REQUIRE(hord == 5 || hord == 6 || hord == 7 || hord == 8, FV3_E_UNSUPPORTED,
              "%s: hord=%d not supported (5, 6, 7, 8)", who, hord);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(dom->sj < (1 << 30) && dom->sj2 < (1 << 30), FV3_E_ARG, "%s: row stride too large",
              who);
  Dom d = to_dom(dom);
  const int nfi = d.iec + 2 - d.isc, nfj = d.jec + 2 - d.jsc;
  const int nbi = (nfi + TR_TI - 1) / TR_TI, nbj = (nfj + TR_TJ - 1) / TR_TJ;
  const int64_t total = (int64_t)nbi * nbj * d.nk * d.ntiles;
  FV3_REQUIRE(total < (int64_t)1 << 30, FV3_E_ARG, "%s: grid too large", who);
  const int64_t launched = (total + kNumXcd - 1) / kNumXcd * kNumXcd;
  dim3 grid((unsigned)launched), block(TR_NT);
  hipStream_t s = as_stream(stream);
  // algorithmic bytes: q (and its y-corner copy when it is a separate
  // array), crx, cry, both area fluxes and the optional mass fluxes read
  // once, both flux fields written once, over the compute cells
  const double nfld = 7.0 + (q_y != q_x ? 1.0 : 0.0) + (xmf ? 1.0 : 0.0) + (ymf ? 1.0 : 0.0);
  TimedLaunch timed("fvtp2d_kernel", s,
                    nfld * sizeof(T) * (double)(d.iec - d.isc + 1) * (d.jec - d.jsc + 1) *
                        d.nk * d.ntiles);
  switch (hord) {
    case 5:
      hipLaunchKernelGGL((fvtp2d_kernel<T, 5>), grid, block, 0, s, d, q_x, q_y, crx, cry,
                         xfx, yfx, xmf, ymf, area, dxa, dya, fx, fy, nbi, nbj);
      break;
    case 6:
      hipLaunchKernelGGL((fvtp2d_kernel<T, 6>), grid, block, 0, s, d, q_x, q_y, crx, cry,
                         xfx, yfx, xmf, ymf, area, dxa, dya, fx, fy, nbi, nbj);
      break;
    case 7:
      hipLaunchKernelGGL((fvtp2d_kernel<T, 7>), grid, block, 0, s, d, q_x, q_y, crx, cry,
                         xfx, yfx, xmf, ymf, area, dxa, dya, fx, fy, nbi, nbj);
      break;
    default:
      hipLaunchKernelGGL((fvtp2d_kernel<T, 8>), grid, block, 0, s, d, q_x, q_y, crx, cry,
                         xfx, yfx, xmf, ymf, area, dxa, dya, fx, fy, nbi, nbj);
      break;
  }
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
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

int32_t fv3_fvtp2d_f32(const fv3_domain* d, int32_t hord, const float* q_x, const float* q_y,
                       const float* crx, const float* cry, const float* x_area_flux,
                       const float* y_area_flux, const float* x_mass_flux,
                       const float* y_mass_flux, const float* area, const float* dxa,
                       const float* dya, float* q_x_flux, float* q_y_flux, void* stream) {
  return fv3::fvtp2d_launch<float>(d, hord, q_x, q_y, crx, cry, x_area_flux, y_area_flux,
                                   x_mass_flux, y_mass_flux, area, dxa, dya, q_x_flux,
                                   q_y_flux, stream);
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
