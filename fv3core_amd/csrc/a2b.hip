// AGrid2BGridFourthOrder (a2b_ord4): cell-centre values to cell corners.
//
// Reference: fv3core/stencils/a2b_ord4.py:1-695.  Four launches:
//   0 weights: the corner extrapolation weights from great-circle
//             distances on the A/B-grid longitudes and latitudes (12 per
//             tile, grid only);
//   1 edges:  the four corner extrapolations and the tile-edge lines
//             qout_x_edge / qout_y_edge, into qout and the edge temporary;
//   2 means + interp: ppm_volume_mean_x on [isc, iec+1] x [jsc-2, jec+2]
//             and ppm_volume_mean_y on [isc-2, iec+2] x [jsc, jec+1], each
//             formed in LDS only where a block's corners read it, then
//             a2b_interpolation on the corners off the tile-edge lines;
//   3 copy:   the optional copy back into qin.
#include "kernels.hpp"
#include "tile_config.hpp"
#include "timing.hpp"

namespace fv3 {

namespace {

constexpr double kA1 = 9.0 / 16.0;
constexpr double kA2 = -1.0 / 16.0;
constexpr double kB1 = 7.0 / 12.0;
constexpr double kB2 = -1.0 / 12.0;
constexpr double kCc1 = 2.0 / 3.0;
constexpr double kCc2 = -1.0 / 6.0;

// Runs f(i, j, k, t) on the perimeter of box b only (its first and last
// rows and columns), for levels [k0, k0 + nk) of every tile.
template <typename F>
__global__ void __launch_bounds__(256) perimeter_kernel(Box b, int k0, F f) {
  const int nbi = b.i1 - b.i0 + 1, nbj = b.j1 - b.j0 + 1;
  const int per = 2 * nbi + 2 * (nbj - 2);
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= per) return;
  int i, j;
  if (n < nbi) {
    i = b.i0 + n;
    j = b.j0;
  } else if (n < 2 * nbi) {
    i = b.i0 + n - nbi;
    j = b.j1;
  } else {
    const int m = n - 2 * nbi;
    i = m < nbj - 2 ? b.i0 : b.i1;
    j = b.j0 + 1 + (m < nbj - 2 ? m : m - (nbj - 2));
  }
  f(i, j, k0 + (int)blockIdx.y, (int)blockIdx.z);
}

template <typename F>
inline void launch_perimeter(hipStream_t s, const Dom& d, Box b, int k0, int nk, F f) {
  const int nbi = b.i1 - b.i0 + 1, nbj = b.j1 - b.j0 + 1;
  if (nbi <= 0 || nbj < 3 || nk <= 0 || d.ntiles <= 0) return;
  const int per = 2 * nbi + 2 * (nbj - 2);
  dim3 grid((unsigned)((per + 255) / 256), (unsigned)nk, (unsigned)d.ntiles);
  hipLaunchKernelGGL((perimeter_kernel<F>), grid, dim3(256), 0, s, b, k0, f);
}

// great_circle_dist (a2b_ord4.py:37-41)
template <typename T>
__device__ __forceinline__ T gc_dist(T p1a, T p1b, T p2a, T p2b) {
  const T sb = sin((p1b - p2b) / T(2));
  const T sa = sin((p1a - p2a) / T(2));
  const T tb = sb * sb;
  const T ta = sa * sa;
  return asin(sqrt(tb + cos(p1b) * cos(p2b) * ta)) * T(2);
}

// The three extrapolations of each cube corner (a2b_ord4.py:617-658), as
// centre offsets (di1, dj1, di2, dj2) from the corner point; corners in the
// order (isc, jsc), (iec+1, jsc), (iec+1, jec+1), (isc, jec+1).
__constant__ int kCornerTaps[4][3][4] = {
    {{0, 0, 1, 1}, {-1, 0, -2, 1}, {0, -1, 1, -2}},
    {{-1, 0, -2, 1}, {-1, -1, -2, -2}, {0, 0, 1, 1}},
    {{-1, -1, -2, -2}, {0, -1, 1, -2}, {-1, 0, -2, 1}},
    {{0, -1, 1, -2}, {-1, -1, -2, -2}, {0, 0, 1, 1}},
};

// Weight x1 / (x2 - x1) of extrapolation e (corner e / 3, tap e % 3) of
// tile blockIdx.x: extrap_corner is qa + x1 / (x2 - x1) (qa - qb) with
// great-circle distances that depend on the grid only, so they are formed
// once per call here instead of once per level in the perimeter launch.
//
// The same launch forms, per tile and row j in [jsc-2, jec+2], the cell-length
// ratios of the one-sided x volume means (qx_edge_west's g_in = dxa(p+1) /
// dxa(p) and g_ou = dxa(p-2) / dxa(p-1) at p = is, then at p = ie+1), which
// depend on the grid only; the means read them instead of dividing again on
// every level (the same divisions, so the same bits).
template <typename T>
__global__ void __launch_bounds__(64) corner_weights_kernel(Met<T> m, Dom d, T* w, T* xr,
                                                            int nrow) {
  const int e = threadIdx.x;
  const int t = blockIdx.x;
  for (int r = e; r < nrow; r += 64) {
    const T* dx = m.dxa + (int64_t)t * d.st2 + (int64_t)(d.jsc - 2 + r) * d.sj2;
    T* out = xr + ((int64_t)t * nrow + r) * 4;
    if (d.is >= 2 && d.is + 1 < d.ni) {
      out[0] = dx[d.is + 1] / dx[d.is];
      out[1] = dx[d.is - 2] / dx[d.is - 1];
    }
    if (d.ie >= 1 && d.ie + 2 < d.ni) {
      out[2] = dx[d.ie + 2] / dx[d.ie + 1];
      out[3] = dx[d.ie - 1] / dx[d.ie];
    }
  }
  if (e >= 12) return;
  const int corner = e / 3, tap = e % 3;
  const int ic = (corner == 1 || corner == 2) ? d.iec + 1 : d.isc;
  const int jc = (corner >= 2) ? d.jec + 1 : d.jsc;
  const int64_t o2 = at2(d, ic, jc, t);
  const int* tp = kCornerTaps[corner][tap];
  const int64_t a1 = o2 + tp[0] + (int64_t)tp[1] * d.sj2, a2 = o2 + tp[2] + (int64_t)tp[3] * d.sj2;
  const T p0a = m.bgrid1[o2], p0b = m.bgrid2[o2];
  const T x1 = gc_dist(m.agrid1[a1], m.agrid2[a1], p0a, p0b);
  const T x2 = gc_dist(m.agrid1[a2], m.agrid2[a2], p0a, p0b);
  w[t * 12 + e] = x1 / (x2 - x1);
}

template <typename T>
__device__ __forceinline__ T qx_edge_west(const T* q, const T* dx) {
  // q[], dx[] centred at the point (index 0 = the point)
  const T g_in = dx[1] / dx[0];
  const T g_ou = dx[-2] / dx[-1];
  return T(0.5) * (((T(2) + g_in) * q[0] - q[1]) / (T(1) + g_in) +
                   ((T(2) + g_ou) * q[-1] - q[-2]) / (T(1) + g_ou));
}

template <typename T>
__device__ __forceinline__ T qx_edge_west2(const T* q, const T* dx) {
  const T g_in = dx[0] / dx[-1];
  const T g_ou = dx[-3] / dx[-2];
  const T left = T(0.5) * (((T(2) + g_in) * q[-1] - q[0]) / (T(1) + g_in) +
                           ((T(2) + g_ou) * q[-2] - q[-3]) / (T(1) + g_ou));
  const T right = T(kB2) * (q[-1] + q[2]) + T(kB1) * (q[0] + q[1]);
  return (T(3) * (g_in * q[-1] + q[0]) - (g_in * left + right)) / (T(2) + T(2) * g_in);
}

template <typename T>
__device__ __forceinline__ T qx_edge_east(const T* q, const T* dx) {
  const T g_in = dx[-2] / dx[-1];
  const T g_ou = dx[1] / dx[0];
  return T(0.5) * (((T(2) + g_in) * q[-1] - q[-2]) / (T(1) + g_in) +
                   ((T(2) + g_ou) * q[0] - q[1]) / (T(1) + g_ou));
}

template <typename T>
__device__ __forceinline__ T qx_edge_east2(const T* q, const T* dx) {
  const T g_in = dx[-1] / dx[0];
  const T g_ou = dx[2] / dx[1];
  const T right = T(0.5) * (((T(2) + g_in) * q[0] - q[-1]) / (T(1) + g_in) +
                            ((T(2) + g_ou) * q[1] - q[2]) / (T(1) + g_ou));
  const T left = T(kB2) * (q[-3] + q[0]) + T(kB1) * (q[-2] + q[-1]);
  return (T(3) * (q[-1] + g_in * q[0]) - (g_in * right + left)) / (T(2) + T(2) * g_in);
}

// ppm_volume_mean_x / _y at a point: q7/dx7 hold cells p-3..p+3 along the
// direction; a = absolute index of the point along it, s/e the anchors.
template <typename T>
__device__ __forceinline__ T volume_mean(const T (&q7)[7], const T (&dx7)[7], int a, int s,
                                         int e) {
  const T* q = q7 + 3;
  const T* dx = dx7 + 3;
  // regions in the reference's order: start, start+1, end+1, end (last wins)
  if (a == e) return qx_edge_east2(q, dx);
  if (a == e + 1) return qx_edge_east(q, dx);
  if (a == s + 1) return qx_edge_west2(q, dx);
  if (a == s) return qx_edge_west(q, dx);
  return T(kB2) * (q[-2] + q[1]) + T(kB1) * (q[-1] + q[0]);
}

// Launches 2 and 3 fused: one workgroup owns a TI x TJ block of corners of
// one level.  It stages q over the block plus a 3-cell halo in LDS, forms
// the x volume means on the rows j-3 .. j+2 and the y volume means on the
// columns i-3 .. i+2 of its corners there, then interpolates.  Lane x of
// every wave owns LDS column x; wave y owns rows y, y+8, ...  (The metric
// terms are read from global memory: staging them in LDS as well, once per
// block over several levels, measured slower through the lower occupancy.)
constexpr int AB_W = 64;           // LDS columns = TI + 6
constexpr int AB_TI = AB_W - 6;    // corners per block in i
constexpr int AB_TJ = FV3_AB_TJ;    // corners per block in j
constexpr int AB_H = AB_TJ + 6;    // q rows: j0-3 .. j0+TJ+2
constexpr int AB_NY = FV3_AB_NY;    // waves per block

struct A2BArgs {
  int ci0, cj0, nci, ncj;  // corner box (interpolation off the edge lines)
  int nbi;                 // blocks along i
  int k0;
  int nrow;                // rows of the edge ratios (jsc-2 .. jec+2)
};

template <typename T>
__global__ void __launch_bounds__(AB_W * AB_NY)
    a2b_interp_kernel(Dom d, Met<T> m, A2BArgs a, const T* __restrict__ qin,
                      const T* __restrict__ edges, const T* __restrict__ xr,
                      T* __restrict__ qout) {
  __shared__ T sq[AB_H][AB_W];
  __shared__ T sx[AB_H - 1][AB_W];  // x means on rows j0-3 .. j0+TJ+1
  __shared__ T sy[AB_TJ][AB_W];     // y means on rows j0 .. j0+TJ-1
  const int li = threadIdx.x, ty = threadIdx.y;
  const int bi = blockIdx.x % a.nbi, bj = blockIdx.x / a.nbi;
  const int k = a.k0 + (int)blockIdx.y, t = blockIdx.z;
  const int i0 = a.ci0 + bi * AB_TI, j0 = a.cj0 + bj * AB_TJ;
  const int i = i0 - 3 + li;  // absolute column of this lane
  const int jb = j0 - 3;      // absolute row of LDS row 0
  const bool col_ok = i >= 0 && i < d.ni;
  const int64_t off3 = (int64_t)t * d.st + (int64_t)k * d.sk;
  const int64_t off2 = (int64_t)t * d.st2;
  for (int r = ty; r < AB_H; r += AB_NY) {
    const int j = jb + r;
    sq[r][li] = (col_ok && j >= 0 && j < d.nj) ? qin[off3 + (int64_t)j * d.sj + i] : T(0);
  }
  __syncthreads();
  // x means where launch 2 formed them: i in [isc, iec+1], j in [jsc-2, jec+2].
  // The lanes of one wave span the whole row, so the one-sided forms of the
  // tile-edge interfaces diverge within it.  They share one instruction
  // stream here: the outer form at is equals the one at ie+1 (the same two
  // halves, added in the other order), and the inner form at is+1 or ie is
  // the outer form of its neighbour combined with interior means, mirrored
  // about the interface for the east side.  Each value keeps the operations
  // of volume_mean, so the same bits; a wave runs 6 divisions a row instead
  // of one set per form present in it.
  if (li >= 3 && li < AB_W - 3 && i >= d.isc && i <= d.iec + 1) {
    // form: 0 interior, 1 outer, 2 inner west, 3 inner east, in volume_mean's
    // order of precedence; sh: the outer form's point, relative to i
    int form = 0, sh = 0;
    if (i == d.ie) {
      form = 3;
      sh = 1;
    } else if (i == d.ie + 1) {
      form = 1;
    } else if (i == d.is + 1) {
      form = 2;
      sh = -1;
    } else if (i == d.is) {
      form = 1;
    }
    const bool west = form == 2;
    for (int r = ty; r < AB_H - 1; r += AB_NY) {
      const int j = jb + r;
      if (j < d.jsc - 2 || j > d.jec + 2) continue;
      const T* q = &sq[r][li];
      T v = T(kB2) * (q[-2] + q[1]) + T(kB1) * (q[-1] + q[0]);
      if (form != 0) {
        // the outer form at is or ie+1 (qx_edge_west there) from its ratios
        const T* gr = xr + ((int64_t)t * a.nrow + (j - (d.jsc - 2))) * 4;
        const bool wside = i <= d.is + 1;
        const T g_in = wside ? gr[0] : gr[2], g_ou = wside ? gr[1] : gr[3];
        const T* qb = q + sh;
        const T outer = T(0.5) * (((T(2) + g_in) * qb[0] - qb[1]) / (T(1) + g_in) +
                                  ((T(2) + g_ou) * qb[-1] - qb[-2]) / (T(1) + g_ou));
        if (form == 1) {
          v = outer;
        } else {
          // qx_edge_west2 at is+1 (its left = outer) or qx_edge_east2 at ie
          // (its right = outer), with q[k] -> q[-1-k] on the east side; its
          // ratio dxa(is+1) / dxa(is) or dxa(ie-1) / dxa(ie) is the outer
          // form's g_in or g_ou
          auto qm = [&](int k) { return west ? q[k] : q[-1 - k]; };
          const T g = west ? g_in : g_ou;
          const T inner = T(kB2) * (qm(-1) + qm(2)) + T(kB1) * (qm(0) + qm(1));
          v = (T(3) * (g * qm(-1) + qm(0)) - (g * outer + inner)) / (T(2) + T(2) * g);
        }
      }
      sx[r][li] = v;
    }
  }
  // y means where launch 2 formed them: i in [isc-2, iec+2], j in [jsc, jec+1]
  if (li < AB_W - 1 && i >= d.isc - 2 && i <= d.iec + 2) {
    for (int r = ty; r < AB_TJ; r += AB_NY) {
      const int j = j0 + r;
      if (j >= d.jsc && j <= d.jec + 1) {
        // the row is uniform across the wave: only the rows on the tile-edge
        // interfaces read the cell lengths and take the one-sided forms
        T v = T(kB2) * (sq[r + 1][li] + sq[r + 4][li]) + T(kB1) * (sq[r + 2][li] + sq[r + 3][li]);
        if (j == d.js || j == d.js + 1 || j == d.je || j == d.je + 1) {
          const int64_t o2 = off2 + (int64_t)j * d.sj2 + i;
          T q7[7], dy7[7];
#pragma unroll
          for (int n = 0; n < 7; ++n) {
            q7[n] = sq[r + n][li];
            dy7[n] = m.dya[o2 + (int64_t)(n - 3) * d.sj2];
          }
          v = volume_mean(q7, dy7, j, d.js, d.je);
        }
        sy[r][li] = v;
      }
    }
  }
  __syncthreads();
  // a2b_interpolation on the block's corners (qx rows and qy columns -3 .. +2)
  if (li < 3 || li >= 3 + AB_TI || i >= a.ci0 + a.nci) return;
  const int64_t sj = d.sj;
  for (int r = ty; r < AB_TJ; r += AB_NY) {
    const int j = j0 + r;
    if (j >= a.cj0 + a.ncj) break;
    const int x = r + 3;  // sx row of j
    const int64_t o = off3 + (int64_t)j * sj + i;
    T qxx = T(kA2) * (sx[x - 2][li] + sx[x + 1][li]) + T(kA1) * (sx[x - 1][li] + sx[x][li]);
    T qyy = T(kA2) * (sy[r][li - 2] + sy[r][li + 1]) + T(kA1) * (sy[r][li - 1] + sy[r][li]);
    if (j == d.js + 1) {
      const T up = T(kA2) * (sx[x - 1][li] + sx[x + 2][li]) + T(kA1) * (sx[x][li] + sx[x + 1][li]);
      qxx = T(kCc1) * (sx[x - 1][li] + sx[x][li]) + T(kCc2) * (edges[o - sj] + up);
    }
    if (j == d.je) {
      const T lo = T(kA2) * (sx[x - 3][li] + sx[x][li]) + T(kA1) * (sx[x - 2][li] + sx[x - 1][li]);
      qxx = T(kCc1) * (sx[x - 1][li] + sx[x][li]) + T(kCc2) * (edges[o + sj] + lo);
    }
    if (i == d.is + 1) {
      const T rr = T(kA2) * (sy[r][li - 1] + sy[r][li + 2]) + T(kA1) * (sy[r][li] + sy[r][li + 1]);
      qyy = T(kCc1) * (sy[r][li - 1] + sy[r][li]) + T(kCc2) * (edges[o - 1] + rr);
    }
    if (i == d.ie) {
      const T l = T(kA2) * (sy[r][li - 3] + sy[r][li]) + T(kA1) * (sy[r][li - 2] + sy[r][li - 1]);
      qyy = T(kCc1) * (sy[r][li - 1] + sy[r][li]) + T(kCc2) * (edges[o + 1] + l);
    }
    qout[o] = T(0.5) * (qxx + qyy);
  }
}

}  // namespace

template <typename T>
int launch_a2b_ord4(hipStream_t s, const fv3_domain* dom, const Met<T>& m, int k0, int nk,
                    const T* qin, T* qout, T* qx, T* edges, int replace, T* qin_w,
                    const char* who) {
  if (nk <= 0) return FV3_OK;
  Dom d = to_dom(dom);
  const bool w = has_edge(dom->i_start), e = has_edge(dom->i_end);
  const bool so = has_edge(dom->j_start), no = has_edge(dom->j_end);
  const int js2 = so ? d.jsc + 1 : d.jsc, je1 = no ? d.jec : d.jec + 1;
  const int is2 = w ? d.isc + 1 : d.isc, ie1 = e ? d.iec : d.iec + 1;
  // 0: corner extrapolation weights, 12 per tile, into the head of the qx
  // slab, and the x volume means' edge ratios after them (nothing else
  // writes that slab during the call)
  T* const cw = qx;
  T* const xr = qx + 12 * (int64_t)d.ntiles;
  const int nrow = d.jec - d.jsc + 5;
  hipLaunchKernelGGL((corner_weights_kernel<T>), dim3((unsigned)d.ntiles), dim3(64), 0, s, m, d,
                     cw, xr, nrow);
  FV3_HIP_CHECK_LAUNCH(who);
  // 1: corners and edge lines: only the perimeter of [isc, iec+1]^2 can hold them
  const Box bx{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_perimeter(s, d, bx, k0, nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    T val;
    const bool ci = i == d.isc || i == d.iec + 1, cj = j == d.jsc || j == d.jec + 1;
    if (ci && cj) {
      // _sw_corner at (isc, jsc), _nw_corner at (iec+1, jsc), _ne_corner at
      // (iec+1, jec+1), _se_corner at (isc, jec+1) (a2b_ord4.py:617-658)
      const int corner = j == d.jsc ? (i == d.isc ? 0 : 1) : (i == d.iec + 1 ? 2 : 3);
      T ex[3];
#pragma unroll
      for (int tap = 0; tap < 3; ++tap) {
        const int* tp = kCornerTaps[corner][tap];
        const T qa = qin[o + tp[0] + (int64_t)tp[1] * d.sj];
        const T qb = qin[o + tp[2] + (int64_t)tp[3] * d.sj];
        ex[tap] = qa + cw[t * 12 + corner * 3 + tap] * (qa - qb);
      }
      val = (ex[0] + ex[1] + ex[2]) * (T(1) / T(3));
    } else if (ci && ((i == d.isc && w) || (i == d.iec + 1 && e)) && j >= js2 && j <= je1) {
      // qout_x_edge (a2b_ord4.py:313-323): q2 at j and j - 1
      const T* dxa = m.dxa;
      const T q2 = (qin[o - 1] * dxa[o2] + qin[o] * dxa[o2 - 1]) / (dxa[o2 - 1] + dxa[o2]);
      const int64_t om = o - d.sj, om2 = o2 - d.sj2;
      const T q2m = (qin[om - 1] * dxa[om2] + qin[om] * dxa[om2 - 1]) / (dxa[om2 - 1] + dxa[om2]);
      const T wgt = (i == d.isc ? m.edge_w : m.edge_e)[o2];
      val = wgt * q2m + (T(1) - wgt) * q2;
    } else if (cj && ((j == d.jsc && so) || (j == d.jec + 1 && no)) && i >= is2 && i <= ie1) {
      const T* dya = m.dya;
      const int64_t sj = d.sj, sj2 = d.sj2;
      const T q1 = (qin[o - sj] * dya[o2] + qin[o] * dya[o2 - sj2]) / (dya[o2 - sj2] + dya[o2]);
      const T q1m = (qin[o - 1 - sj] * dya[o2 - 1] + qin[o - 1] * dya[o2 - 1 - sj2]) /
                    (dya[o2 - 1 - sj2] + dya[o2 - 1]);
      const T wgt = (j == d.jsc ? m.edge_s : m.edge_n)[o2];
      val = wgt * q1m + (T(1) - wgt) * q1;
    } else {
      return;
    }
    qout[o] = val;
    edges[o] = val;
  });
  FV3_HIP_CHECK_LAUNCH(who);
  // 2 + 3: volume means and interpolation off the tile-edge lines
  {
    A2BArgs a;
    a.ci0 = w ? d.isc + 1 : d.isc;
    a.cj0 = so ? d.jsc + 1 : d.jsc;
    a.nci = (e ? d.iec : d.iec + 1) - a.ci0 + 1;
    a.ncj = (no ? d.jec : d.jec + 1) - a.cj0 + 1;
    a.k0 = k0;
    a.nrow = nrow;
    if (a.nci > 0 && a.ncj > 0) {
      a.nbi = (a.nci + AB_TI - 1) / AB_TI;
      const int nbj = (a.ncj + AB_TJ - 1) / AB_TJ;
      dim3 grid((unsigned)(a.nbi * nbj), (unsigned)nk, (unsigned)d.ntiles);
      // q read, qout written (per corner-level)
      TimedLaunch timed("a2b_interp_kernel", s,
                        2.0 * sizeof(T) * (double)a.nci * a.ncj * nk * d.ntiles);
      hipLaunchKernelGGL((a2b_interp_kernel<T>), grid, dim3(AB_W, AB_NY), 0, s, d, m, a, qin,
                         edges, xr, qout);
      FV3_HIP_CHECK_LAUNCH(who);
    }
  }
  if (replace && qin_w) {
    launch_box(s, d, bx, k0, nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      qin_w[o] = qout[o];
    });
    FV3_HIP_CHECK_LAUNCH(who);
  }
  return FV3_OK;
}

template int launch_a2b_ord4<double>(hipStream_t, const fv3_domain*, const Met<double>&, int, int,
                                     const double*, double*, double*, double*, int, double*,
                                     const char*);
template int launch_a2b_ord4<float>(hipStream_t, const fv3_domain*, const Met<float>&, int, int,
                                    const float*, float*, float*, float*, int, float*,
                                    const char*);

int check_a2b_metrics(const fv3_metrics* g, const char* who) {
  FV3_REQUIRE(g && g->dxa && g->dya && g->agrid1 && g->agrid2 && g->bgrid1 && g->bgrid2 &&
                  g->edge_w && g->edge_e && g->edge_s && g->edge_n,
              FV3_E_ARG, "%s: a2b_ord4 needs dxa, dya, agrid1/2, bgrid1/2 and edge_w/e/s/n",
              who);
  return FV3_OK;
}

}  // namespace fv3

extern "C" int32_t fv3_a2b_ord4_f64(const fv3_domain* dom, const fv3_metrics* g, int32_t replace,
                                    double* qin, double* qout, double* work, int64_t work_slab,
                                    void* stream) {
  using namespace fv3;
  const char* who = "fv3_a2b_ord4";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  rc = check_a2b_metrics(g, who);
  if (rc) return rc;
  FV3_REQUIRE(qin && qout && work, FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)dom->nk * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  const Met<double> m = to_met<double>(g);
  return launch_a2b_ord4<double>(as_stream(stream), dom, m, 0, dom->nk, qin, qout, work,
                                 work + 2 * work_slab, replace, qin, who);
}
