// The acoustic step's remaining stencils: geopotential on the C grid
// (updatedzc), height on the D grid (updatedzd), the nonhydrostatic
// pressure gradient (nh_p_grad), the halo pressures (pk3_halo, pe_halo), the
// C-grid pressure gradient (p_grad_c), Rayleigh damping (ray_fast),
// del2cubed hyperdiffusion, the heating adjustment of pt and the small
// column stencils of dyn_core.
//
// Reference: fv3core/stencils/updatedzc.py, updatedzd.py, nh_p_grad.py,
// pk3_halo.py, pe_halo.py, ray_fast.py, del2cubed.py, temperature_adjust.py
// and dyn_core.py:47-145.  Vertical recurrences run one column per lane.
#include "kernels.hpp"

namespace fv3 {

namespace {

constexpr double kDzMin = 2.0;

// Value of the 2-cell corner-filled copy (fill_corners_2cells_x/_y,
// corners.py:127-300) of q at (i, j): the rotated source for the 8
// destinations of a tile's corners, q itself elsewhere.
// The same as an offset: q at (i, j) through the fill is q[o + fill2_off].
__device__ __forceinline__ int64_t fill2_off(const Dom& d, int i, int j, bool ydir) {
  const bool w = i == d.is - 1 || i == d.is - 2, e = i == d.ie + 1 || i == d.ie + 2;
  const bool s = j == d.js - 1 || j == d.js - 2, n = j == d.je + 1 || j == d.je + 2;
  if (!((w || e) && (s || n))) return 0;
  int di = 0, dj = 0;
  if (!ydir) {
    // destinations (is-1, js-1), (is-2, js-1) ... at rows js-1 / je+1
    if (j != d.js - 1 && j != d.je + 1) return 0;
    const int a = w ? d.is - 1 - i : i - d.ie - 1;  // 0 or 1 cells beyond the corner column
    const int sy = s ? 1 : -1;
    di = w ? a : -a;
    dj = sy * (1 + a);
  } else {
    if (i != d.is - 1 && i != d.ie + 1) return 0;
    const int b = s ? d.js - 1 - j : j - d.je - 1;
    const int sx = w ? 1 : -1;
    di = sx * (1 + b);
    dj = s ? b : -b;
  }
  return di + (int64_t)dj * d.sj;
}


// Levels per chunk of a column sweep: a chunk's loads are issued together
// before the recurrence runs over them, so a column waits for memory once
// per chunk rather than once per level (the sweeps are latency-bound: a
// cube of C96 tiles has under one wave of columns per SIMD).
constexpr int KC = 8;

// h[k] = max(h[k], h[k+1] + DZ_MIN) for k = nk-2 .. 0 (BACKWARD), given
// h[nk-1] = bottom.
template <typename T>
__device__ __forceinline__ void floor_column_up(T* __restrict__ h, int64_t sk, int nk,
                                                T bottom) {
  T above = bottom;
  for (int k0 = nk - 2; k0 >= 0; k0 -= KC) {
    T v[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u)
      if (k0 - u >= 0) v[u] = h[(int64_t)(k0 - u) * sk];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      if (k0 - u >= 0) {
        const T kp1 = above + T(kDzMin);
        const T w = v[u] > kp1 ? v[u] : kp1;
        h[(int64_t)(k0 - u) * sk] = w;
        above = w;
      }
    }
  }
}

// ----------------------------------------------------------------- updatedzc
// One column per lane on [isc-1, iec+1]^2, k in chunks of KZ levels: the
// interface winds at level k come from the layer winds k-1 and k, so the
// sweep carries the two layers above the chunk.  The corner fill of gz is
// a per-column offset, formed once.
constexpr int KZ = 4;

// U is the storage type of ut and vt: FP64, or FP32 in the mixed-precision
// step (widened exactly on load, so the sums are the FP64 kernel's).
template <typename T, typename U>
__global__ void __launch_bounds__(256)
    update_dz_c_kernel(Dom d, const T* __restrict__ dp0, const T* __restrict__ zs,
                       const T* __restrict__ area, const U* __restrict__ ut,
                       const U* __restrict__ vt, const T* __restrict__ gz, T* __restrict__ gzn,
                       T* __restrict__ ws, T rdt) {
  const int ni = d.iec - d.isc + 3, nj = d.jec - d.jsc + 3;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc - 1 + n % ni, j = d.jsc - 1 + n / ni, t = (int)blockIdx.y;
  const int64_t o2 = at2(d, i, j, t);
  const int nk = d.nk;
  const int64_t sk = d.sk, sj = d.sj;
  const T a = area[o2];
  const int64_t oc = at3(d, i, j, 0, t);
  const int64_t gxm = -1 + fill2_off(d, i - 1, j, false), gx0 = fill2_off(d, i, j, false);
  const int64_t gxp = 1 + fill2_off(d, i + 1, j, false);
  const int64_t gym = -sj + fill2_off(d, i, j - 1, true), gy0 = fill2_off(d, i, j, true);
  const int64_t gyp = sj + fill2_off(d, i, j + 1, true);
  // interface value at level k from the layer values vk (k), vm1 (k-1),
  // vm2 (k-2) and, at the top, v1 (layer 1)
  auto iface = [&](T vk, T vm1, T vm2, T v1, int k) -> T {
    if (k == 0) {
      const T ratio = dp0[0] / (dp0[0] + dp0[1]);
      return vk + (vk - v1) * ratio;
    }
    if (k == nk - 1) {
      const T ratio = dp0[k - 1] / (dp0[k - 2] + dp0[k - 1]);
      return vm1 + (vm1 - vm2) * ratio;
    }
    const T ir = T(1) / (dp0[k - 1] + dp0[k]);
    return (dp0[k] * vm1 + dp0[k - 1] * vk) * ir;
  };
  T u0m1 = T(0), u1m1 = T(0), v0m1 = T(0), v1m1 = T(0);  // layer k-1
  T u0m2 = T(0), u1m2 = T(0), v0m2 = T(0), v1m2 = T(0);  // layer k-2
  T last = T(0);
  for (int k0 = 0; k0 < nk; k0 += KZ) {
    T u0[KZ], u1[KZ], v0[KZ], v1[KZ], gx[KZ][3], gy[KZ][3];
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      const int64_t ol = oc + (int64_t)min(k0 + u, nk - 2) * sk;  // layer
      const int64_t og = oc + (int64_t)min(k0 + u, nk - 1) * sk;  // interface
      u0[u] = T(ut[ol]);
      u1[u] = T(ut[ol + 1]);
      v0[u] = T(vt[ol]);
      v1[u] = T(vt[ol + sj]);
      gx[u][0] = gz[og + gxm];
      gx[u][1] = gz[og + gx0];
      gx[u][2] = gz[og + gxp];
      gy[u][0] = gz[og + gym];
      gy[u][1] = gz[og + gy0];
      gy[u][2] = gz[og + gyp];
    }
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        const T xfx0 = iface(u0[u], u0m1, u0m2, u0[1], k);
        const T xfx1 = iface(u1[u], u1m1, u1m2, u1[1], k);
        const T yfx0 = iface(v0[u], v0m1, v0m2, v0[1], k);
        const T yfx1 = iface(v1[u], v1m1, v1m2, v1[1], k);
        const T fx0 = xfx0 * (xfx0 > T(0) ? gx[u][0] : gx[u][1]);
        const T fx1 = xfx1 * (xfx1 > T(0) ? gx[u][1] : gx[u][2]);
        const T fy0 = yfx0 * (yfx0 > T(0) ? gy[u][0] : gy[u][1]);
        const T fy1 = yfx1 * (yfx1 > T(0) ? gy[u][1] : gy[u][2]);
        const T v = (gy[u][1] * a + fx0 - fx1 + fy0 - fy1) / (a + xfx0 - xfx1 + yfx0 - yfx1);
        gzn[oc + (int64_t)k * sk] = v;
        last = v;
        u0m2 = u0m1;
        u1m2 = u1m1;
        v0m2 = v0m1;
        v1m2 = v1m1;
        u0m1 = u0[u];
        u1m1 = u1[u];
        v0m1 = v0[u];
        v1m1 = v1[u];
      }
    }
  }
  ws[o2] = (zs[o2] - last) * rdt;
  floor_column_up(gzn + oc, sk, nk, last);
}

// ----------------------------------------------------------------- updatedzd
// cubic_spline_interpolation_from_layer_center_to_interfaces for up to four
// fields, one column per lane over the full domain.
// U is the storage type of the layer values (FP64, or FP32 in the
// mixed-precision step, widened exactly on load).
template <typename T, typename U>
__global__ void __launch_bounds__(256)
    spline_kernel(Dom d, const T* __restrict__ gk, const T* __restrict__ beta,
                  const T* __restrict__ gamma, const U* q0, const U* q1, const U* q2,
                  const U* q3, T* o0, T* o1, T* o2_, T* o3) {
  const int ni = d.iec - d.isc + 7, nj = d.jec - d.jsc + 7;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc - 3 + n % ni, j = d.jsc - 3 + n / ni, t = (int)blockIdx.y;
  const int f = (int)blockIdx.z;
  const U* qc = f == 0 ? q0 : f == 1 ? q1 : f == 2 ? q2 : q3;
  T* qi = f == 0 ? o0 : f == 1 ? o1 : f == 2 ? o2_ : o3;
  const int nk = d.nk;
  const int64_t base = at3(d, i, j, 0, t), sk = d.sk;
  T prev;
  {
    const T xt1 = T(2) * gk[0] * (gk[0] + T(1));
    prev = (xt1 * T(qc[base]) + T(qc[base + sk])) / beta[0];
    qi[base] = prev;
  }
  T q_prev = T(qc[base]);  // qc[k - 1]
  for (int k0 = 1; k0 < nk - 1; k0 += KC) {
    T qv[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u)
      if (k0 + u < nk - 1) qv[u] = T(qc[base + (k0 + u) * sk]);
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk - 1) {
        prev = (T(3) * (q_prev + gk[k] * qv[u]) - prev) / beta[k];
        qi[base + k * sk] = prev;
        q_prev = qv[u];
      }
    }
  }
  {
    const int k = nk - 1;
    const T g = gk[k - 1];
    const T a_bot = T(1) + g * (g + T(1.5));
    const T xt1 = T(2) * g * (g + T(1));
    const T xt2 = g * (g + T(0.5)) - a_bot * gamma[k - 1];
    prev = (xt1 * T(qc[base + (k - 1) * sk]) + T(qc[base + (k - 2) * sk]) - a_bot * prev) / xt2;
    qi[base + k * sk] = prev;
  }
  for (int k0 = nk - 2; k0 >= 0; k0 -= KC) {
    T qv[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u)
      if (k0 - u >= 0) qv[u] = qi[base + (k0 - u) * sk];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        const T v = qv[u] - gamma[k] * prev;
        qi[base + k * sk] = v;
        prev = v;
      }
    }
  }
}

// apply_height_fluxes (updatedzd.py:63-118), one column per lane on the
// compute domain.
template <typename T>
__global__ void __launch_bounds__(256)
    apply_height_kernel(Dom d, const T* __restrict__ area, const T* __restrict__ zs,
                        T* __restrict__ h, const T* __restrict__ fx, const T* __restrict__ fy,
                        const T* __restrict__ xfx, const T* __restrict__ yfx,
                        const T* __restrict__ gxd, const T* __restrict__ gyd, T* __restrict__ ws,
                        T dt) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni, t = (int)blockIdx.y;
  const int64_t o2 = at2(d, i, j, t);
  const T a = area[o2];
  const int nk = d.nk;
  const int64_t sj = d.sj;
  const int64_t sk = d.sk;
  const int64_t oc = at3(d, i, j, 0, t);
  for (int k0 = 0; k0 < nk; k0 += KZ) {
    T xa[KZ], xb[KZ], ya[KZ], yb[KZ], hv[KZ], fa[KZ], fb[KZ], ga[KZ], gb[KZ];
    T xg[KZ], xh[KZ], yg[KZ], yh[KZ];
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      const int64_t o = oc + (int64_t)min(k0 + u, nk - 1) * sk;
      xa[u] = xfx[o];
      xb[u] = xfx[o + 1];
      ya[u] = yfx[o];
      yb[u] = yfx[o + sj];
      hv[u] = h[o];
      fa[u] = fx[o];
      fb[u] = fx[o + 1];
      ga[u] = fy[o];
      gb[u] = fy[o + sj];
      xg[u] = gxd[o];
      xh[u] = gxd[o + 1];
      yg[u] = gyd[o];
      yh[u] = gyd[o + sj];
    }
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        const T after = (a + xa[u] - xb[u]) + (a + ya[u] - yb[u]) - a;
        const T adv = (hv[u] * a + fa[u] - fb[u] + ga[u] - gb[u]) / after;
        h[oc + (int64_t)k * sk] = adv + (xg[u] - xh[u] + yg[u] - yh[u]) / a;
      }
    }
  }
  const T bottom = h[at3(d, i, j, nk - 1, t)];
  ws[o2] = (zs[o2] - bottom) / dt;
  floor_column_up(h + at3(d, i, j, 0, t), d.sk, nk, bottom);
}

// ------------------------------------------------------------------- rings
// pk3_halo (width 2, pk3 = pe^akap) and pe_halo (width 1, pe) column sums on
// the ring of halo columns around the compute box.  The ring holds few
// columns, so a thread owns one (ring column, level) pair and re-forms its
// level's pe = ptop + dp[0] + ... + dp[k-1] itself, in the order of the
// column sum, instead of one thread walking each column with a pow per level.
// Lanes of a wave share k and own neighbouring ring columns.
template <typename T>
__global__ void __launch_bounds__(256)
    ring_kernel(Dom d, int width, int mode, T ptop, T akap, const T* __restrict__ delp,
                T* __restrict__ out) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  const int wi = ni + 2 * width;                   // strip length along i
  const int nstrip = width * wi;                   // cells of the south / north strip
  const int nside = width * nj;                    // cells of the west / east strip
  const int nring = 2 * nstrip + 2 * nside;
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= (int64_t)nring * d.nk) return;
  const int k = (int)(n / nring);
  int c = (int)(n % nring);
  int i, j;
  if (c < 2 * nstrip) {
    const int north = c >= nstrip;
    c -= north * nstrip;
    i = d.isc - width + c % wi;
    j = north ? d.jec + 1 + c / wi : d.jsc - width + c / wi;
  } else {
    c -= 2 * nstrip;
    const int east = c >= nside;
    c -= east * nside;
    i = east ? d.iec + 1 + c % width : d.isc - width + c % width;
    j = d.jsc + c / width;
  }
  const int t = (int)blockIdx.y;
  const int64_t base = at3(d, i, j, 0, t), sk = d.sk;
  if (k == 0) {
    if (mode == 1) out[base] = ptop;
    return;
  }
  T pe = ptop;
  for (int m = 0; m < k; ++m) pe = pe + delp[base + m * sk];
  out[base + k * sk] = mode == 0 ? pow(pe, akap) : pe;
}

// ------------------------------------------------------------------ ray_fast
// RayleighDamping (ray_fast.py:25-127) with the per-level rf, masks and
// column mass dm prepared on the host; one column per lane on [isc, iec+1]^2.
template <typename T>
__global__ void __launch_bounds__(256)
    ray_fast_kernel(Dom d, int nonhydro, const T* __restrict__ rf, const T* __restrict__ dp,
                    const T* __restrict__ dm, const int* __restrict__ use_rf,
                    const int* __restrict__ use_n, T* __restrict__ u, T* __restrict__ v,
                    T* __restrict__ w) {
  const int ni = d.iec - d.isc + 2, nj = d.jec - d.jsc + 2;
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= ni * nj) return;
  const int i = d.isc + n % ni, j = d.jsc + n / ni, t = (int)blockIdx.y;
  const int64_t base = at3(d, i, j, 0, t), sk = d.sk;
  const int nk = d.nk;
  auto damp = [&](T* q) {
    T dmdir = T(0);
    for (int k = 0; k < nk; ++k) {
      if (use_rf[k]) {
        const int64_t o = base + k * sk;
        dmdir = dmdir + (T(1) - rf[k]) * dp[k] * q[o];
        q[o] = q[o] * rf[k];
      }
    }
    for (int k = 0; k < nk; ++k) {
      if (use_n[k]) {
        const int64_t o = base + k * sk;
        q[o] = q[o] + dmdir / dm[k];
      }
    }
  };
  if (i <= d.iec) damp(u);
  if (j <= d.jec) damp(v);
  if (nonhydro && i <= d.iec && j <= d.jec) {
    for (int k = 0; k < nk; ++k)
      if (use_rf[k]) w[base + k * sk] = w[base + k * sk] * rf[k];
  }
}

}  // namespace

inline int copy_corners_xy(const fv3_domain* d, const double* q, double* qx, double* qy,
                           void* stream) {
  return fv3_copy_corners_xy_f64(d, q, qx, qy, stream);
}

inline int copy_corners_xy(const fv3_domain* d, const float* q, float* qx, float* qy,
                           void* stream) {
  return fv3_copy_corners_xy_f32(d, q, qx, qy, stream);
}

// p_grad_c_stencil (dyn_core.py:101-145) on [isc, iec+1]^2; the winds and
// delpc in storage type S (FP32 in the mixed form), the pressure and
// geopotential terms and the arithmetic in FP64.
template <typename S>
int p_grad_c(const fv3_domain* dom, const double* rdxc, const double* rdyc, int hydrostatic,
             S* uc, S* vc, const S* delpc, const double* pkc, const double* gz, double dt2,
             void* stream) {
  const char* who = "fv3_p_grad_c";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(rdxc && rdyc && uc && vc && pkc && gz && (hydrostatic || delpc), FV3_E_ARG,
              "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const Box bx{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_box(as_stream(stream), d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sk = d.sk, sj = d.sj;
    auto wk = [&](int64_t q) -> double {
      return hydrostatic ? pkc[q + sk] - pkc[q] : (double)delpc[q];
    };
    const double w0 = wk(o);
    {
      const int64_t q = o - 1;
      uc[o] = (S)((double)uc[o] + dt2 * rdxc[o2] / (wk(q) + w0) *
                                      ((gz[q + sk] - gz[o]) * (pkc[o + sk] - pkc[q]) +
                                       (gz[q] - gz[o + sk]) * (pkc[q + sk] - pkc[o])));
    }
    {
      const int64_t q = o - sj;
      vc[o] = (S)((double)vc[o] + dt2 * rdyc[o2] / (wk(q) + w0) *
                                      ((gz[q + sk] - gz[o]) * (pkc[o + sk] - pkc[q]) +
                                       (gz[q] - gz[o + sk]) * (pkc[q + sk] - pkc[o])));
    }
  });
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

template <typename T>
int ray_fast(const fv3_domain* dom, int hydrostatic, const T* rf, const T* dp, const T* dm,
             const int32_t* use_rf, const int32_t* use_nudge, T* u, T* v, T* w, void* stream) {
  const char* who = "fv3_ray_fast";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(rf && dp && dm && use_rf && use_nudge && u && v && (hydrostatic || w), FV3_E_ARG,
              "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const int nc = (d.iec - d.isc + 2) * (d.jec - d.jsc + 2);
  hipLaunchKernelGGL((ray_fast_kernel<T>), dim3((nc + 255) / 256, d.ntiles), dim3(256), 0,
                     as_stream(stream), d, hydrostatic ? 0 : 1, rf, dp, dm, use_rf, use_nudge, u,
                     v, w);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

template <typename T>
int del2cubed(const fv3_domain* dom, const T* del6_u, const T* del6_v, const T* rarea, int nmax,
              T cd, T* qdel, T* work, int64_t work_slab, void* stream) {
  const char* who = "fv3_del2cubed";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(del6_u && del6_v && rarea && qdel && work, FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)dom->nk * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  Dom d = to_dom(dom);
  hipStream_t s = as_stream(stream);
  T* q = work;
  T* qy = work + work_slab;
  const int ntimes = nmax < 3 ? nmax : 3;
  const Box full{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
  for (int n = 0; n < ntimes; ++n) {
    const int nt = ntimes - (n + 1);
    // corner_fill (del2cubed.py:30-62): q = qdel with three-cell averages
    // at the 12 points around the tile corners
    launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      const int64_t sj = d.sj;
      const T third = T(1.0) / T(3.0);
      T val = qdel[o];
      const bool w = i == d.is || i == d.is - 1, e = i == d.ie || i == d.ie + 1;
      const bool so = j == d.js || j == d.js - 1, no = j == d.je || j == d.je + 1;
      if ((w || e) && (so || no)) {
        const int ci = w ? d.is : d.ie, cj = so ? d.js : d.je;
        const int ox = w ? -1 : 1, oy = so ? -1 : 1;  // outward directions
        if (i == ci && j == cj) {
          val = (qdel[o] + qdel[o + ox] + qdel[o + oy * sj]) * third;
        } else if (i == ci + ox && j == cj) {
          val = (qdel[o - ox] + qdel[o] + qdel[o - ox + oy * sj]) * third;
        } else if (i == ci && j == cj + oy) {
          val = (qdel[o - oy * sj] + qdel[o + ox - oy * sj] + qdel[o]) * third;
        }
      }
      q[o] = val;
    });
    FV3_HIP_CHECK_LAUNCH(who);
    const T* qxp = q;
    const T* qyp = q;
    if (nt > 0) {
      rc = copy_corners_xy(dom, q, q, qy, stream);
      if (rc) return rc;
      qyp = qy;
    }
    // copy to qdel and the flux update on the compute domain + nt halo
    launch_box(s, d, full, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      T val = qyp[o];
      if (i >= d.isc - nt && i <= d.iec + nt && j >= d.jsc - nt && j <= d.jec + nt) {
        const int64_t o2 = at2(d, i, j, t);
        const int64_t sj = d.sj, sj2 = d.sj2;
        const T fx0 = del6_v[o2] * (qxp[o - 1] - qxp[o]);
        const T fx1 = del6_v[o2 + 1] * (qxp[o] - qxp[o + 1]);
        const T fy0 = del6_u[o2] * (qyp[o - sj] - qyp[o]);
        const T fy1 = del6_u[o2 + sj2] * (qyp[o] - qyp[o + sj]);
        val = val + cd * rarea[o2] * (fx0 - fx1 + fy0 - fy1);
      }
      qdel[o] = val;
    });
    FV3_HIP_CHECK_LAUNCH(who);
  }
  return FV3_OK;
}

// compute_pkz_tempadjust (temperature_adjust.py:8-44): delp, cappa,
// heat_source and pt in storage type S, delz and pkz and the arithmetic FP64.
template <typename S>
int pkz_tempadjust(const fv3_domain* dom, const S* delp, const double* delz, const S* cappa,
                   const S* heat_source, S* pt, double* pkz, double delt_time_factor,
                   void* stream) {
  const char* who = "fv3_pkz_tempadjust";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(delp && delz && cappa && heat_source && pt && pkz, FV3_E_ARG, "%s: NULL pointer",
              who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const Box comp{d.isc, d.iec, d.jsc, d.jec};
  const double rdg = -287.05 / 9.80665;
  const double cv_air = 1004.6 - 287.05;
  launch_box(as_stream(stream), d, comp, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const double c = cappa[o];
    const double dp = delp[o], ptv = pt[o];
    const double pk = pow(rdg * dp / delz[o] * ptv, c / (1.0 - c));
    pkz[o] = pk;
    const double dtmp = (double)heat_source[o] / (cv_air * dp);
    const double f = k == 0 ? delt_time_factor * 0.1
                            : (k == 1 ? delt_time_factor * 0.5 : delt_time_factor);
    const double deltmin = sgn(fmin(f, fabs(dtmp)), dtmp);
    pt[o] = (S)(ptv + deltmin / pk);
  });
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

// NonHydrostaticPressureGradient.__call__ (nh_p_grad.py:147-206): the
// winds in storage type S (FP32 in the mixed form), the pressures, heights,
// delp and the arithmetic FP64.
template <typename S>
int nh_p_grad(const fv3_domain* dom, const fv3_metrics* g, S* u, S* v, double* pp, double* gz,
              double* pk3, const double* delp, double dt, double ptk, double* work,
              int64_t work_slab, void* stream) {
  const char* who = "fv3_nh_p_grad";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  rc = check_a2b_metrics(g, who);
  if (rc) return rc;
  FV3_REQUIRE(g->rdx && g->rdy, FV3_E_ARG, "%s: rdx/rdy are NULL", who);
  FV3_REQUIRE(u && v && pp && gz && pk3 && delp && work, FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  const int nz = dom->nk;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)(nz + 1) * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  hipStream_t s = as_stream(stream);
  const Met<double> m = to_met<double>(g);
  // slabs: B-grid pp, delp's B-grid values (wk1), the a2b scratch (corner
  // weights, edges), B-grid pk3 and gz
  double *ppb = work, *wk1 = work + work_slab, *qx = work + 2 * work_slab;
  double *pk3b = work + 3 * work_slab, *edges = work + 4 * work_slab;
  double* gzb = work + 5 * work_slab;
  // a2b_k1 on pp and pk3 (levels 1..nz), a2b_kbuffer on gz (0..nz), into
  // the slabs; the u/v update below writes them back into pp, pk3 and gz
  // (the reference's replace), so no separate copies are launched
  rc = launch_a2b_ord4<double>(s, dom, m, 1, nz, pp, ppb, qx, edges, 0, nullptr, who);
  if (rc) return rc;
  rc = launch_a2b_ord4<double>(s, dom, m, 1, nz, pk3, pk3b, qx, edges, 0, nullptr, who);
  if (rc) return rc;
  rc = launch_a2b_ord4<double>(s, dom, m, 0, nz + 1, gz, gzb, qx, edges, 0, nullptr, who);
  if (rc) return rc;
  rc = launch_a2b_ord4<double>(s, dom, m, 0, nz, delp, wk1, qx, edges, 0, nullptr, who);
  if (rc) return rc;
  Dom d = to_dom(dom);
  // set_k0_and_calc_wk, calc_u and calc_v on [isc, iec+1]^2 in one pass:
  // level 0 of pp and pk3 is 0 and ptk, wk = pk3(k+1) - pk3(k) is formed
  // where it is read (the same subtraction), and every point writes its
  // B-grid pp, pk3 and gz back (the top thread also level nz)
  const Box bc{d.isc, d.iec + 1, d.jsc, d.jec + 1};
  launch_box(s, d, bc, 0, nz, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const int64_t sk = d.sk;
    auto pk3_at = [&](int64_t p, int kk) { return kk == 0 ? ptk : pk3b[p]; };
    auto pp_at = [&](int64_t p, int kk) { return kk == 0 ? 0.0 : ppb[p]; };
    const double pk3_o = pk3_at(o, k), pk3_ok = pk3b[o + sk];
    const double wk_o = pk3_ok - pk3_o;
    if (i <= d.iec) {
      const int64_t q = o + 1;
      const double pk3_q = pk3_at(q, k), pk3_qk = pk3b[q + sk];
      const double wk_q = pk3_qk - pk3_q;
      const double du = dt / (wk_o + wk_q) *
                        ((gzb[o + sk] - gzb[q]) * (pk3_qk - pk3_o) +
                         (gzb[o] - gzb[q + sk]) * (pk3_ok - pk3_q));
      u[o] = (S)(((double)u[o] + du +
              dt / (wk1[o] + wk1[q]) *
                  ((gzb[o + sk] - gzb[q]) * (ppb[q + sk] - pp_at(o, k)) +
                   (gzb[o] - gzb[q + sk]) * (ppb[o + sk] - pp_at(q, k)))) *
             m.rdx[o2]);
    }
    if (j <= d.jec) {
      const int64_t q = o + d.sj;
      const double pk3_q = pk3_at(q, k), pk3_qk = pk3b[q + sk];
      const double wk_q = pk3_qk - pk3_q;
      const double dv = dt / (wk_o + wk_q) *
                        ((gzb[o + sk] - gzb[q]) * (pk3_qk - pk3_o) +
                         (gzb[o] - gzb[q + sk]) * (pk3_ok - pk3_q));
      v[o] = (S)(((double)v[o] + dv +
              dt / (wk1[o] + wk1[q]) *
                  ((gzb[o + sk] - gzb[q]) * (ppb[q + sk] - pp_at(o, k)) +
                   (gzb[o] - gzb[q + sk]) * (ppb[o + sk] - pp_at(q, k)))) *
             m.rdy[o2]);
    }
    pp[o] = pp_at(o, k);
    pk3[o] = pk3_o;
    gz[o] = gzb[o];
    if (k == nz - 1) {
      pp[o + sk] = ppb[o + sk];
      pk3[o + sk] = pk3_ok;
      gz[o + sk] = gzb[o + sk];
    }
  });
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

template <typename T>
int dyncore_column(const fv3_domain* dom, int op, double scalar, const T* a, const T* b, T* out,
                   void* stream) {
  const char* who = "fv3_dyncore_column";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(out && (op == 4 || a), FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(op >= 0 && op <= 4, FV3_E_ARG, "%s: unknown op %d", who, op);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  hipStream_t s = as_stream(stream);
  if (op == 0) {
    // set_gz (dyn_core.py:80-85): out = gz on nk levels, a = zs (IJ), b = delz
    FV3_REQUIRE(b, FV3_E_ARG, "%s: NULL delz", who);
    const Box comp{d.isc, d.iec, d.jsc, d.jec};
    launch_box(s, d, comp, 0, 1, [=] __device__(int i, int j, int, int t) {
      int64_t o = at3(d, i, j, d.nk - 1, t);
      T g = a[at2(d, i, j, t)];
      out[o] = g;
      for (int k = d.nk - 2; k >= 0; --k) {
        o -= d.sk;
        g = g - b[o];
        out[o] = g;
      }
    });
  } else if (op == 1) {
    // set_pem (dyn_core.py:88-93): out = pem, a = delp, scalar = ptop
    const Box bx{d.isc - 1, d.iec + 1, d.jsc - 1, d.jec + 1};
    launch_box(s, d, bx, 0, 1, [=] __device__(int i, int j, int, int t) {
      int64_t o = at3(d, i, j, 0, t);
      T p = (T)scalar;
      out[o] = p;
      for (int k = 1; k < d.nk; ++k) {
        o += d.sk;
        p = p + a[o];
        out[o] = p;
      }
    });
  } else if (op == 2) {
    // compute_geopotential (dyn_core.py:96-98) on compute + 2 halo: out = a * GRAV
    const Box bx{d.isc - 2, d.iec + 2, d.jsc - 2, d.jec + 2};
    launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      out[o] = a[o] * T(9.80665);
    });
  } else if (op == 3) {
    // copy_defn on the full domain
    const Box bx{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
    launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      const int64_t o = at3(d, i, j, k, t);
      out[o] = a[o];
    });
  } else {
    // fill the full domain with a value
    const Box bx{d.isc - 3, d.iec + 3, d.jsc - 3, d.jec + 3};
    launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
      out[at3(d, i, j, k, t)] = (T)scalar;
    });
  }
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

// (ring column, level) pairs of a ring of the given width around the box
static int64_t ring_threads(const Dom& d, int width) {
  const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
  return (int64_t)(2 * width * (ni + 2 * width) + 2 * width * nj) * d.nk;
}

int check_metrics_list(const fv3_metrics* g, const void* const* ptrs, int n, const char* who) {
  FV3_REQUIRE(g != nullptr, FV3_E_ARG, "%s: NULL metrics", who);
  for (int a = 0; a < n; ++a)
    FV3_REQUIRE(ptrs[a] != nullptr, FV3_E_ARG, "%s: a metric term is NULL", who);
  return FV3_OK;
}

// UpdateGeopotentialHeightOnCGrid (updatedzc.py) with the layer winds ut, vt
// of storage type U.
template <typename U>
int update_dz_c(const fv3_domain* dom, const double* area, const double* zs,
                const double* dp_ref, const U* ut, const U* vt, double* gz, double* ws,
                double dt, double* work, int64_t work_slab, void* stream) {
  const char* who = "fv3_update_dz_c";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(area && zs && dp_ref && ut && vt && gz && ws && work, FV3_E_ARG,
              "%s: NULL pointer", who);
  FV3_REQUIRE(dom->nk >= 3, FV3_E_ARG, "%s: needs at least 3 interface levels", who);
  if (dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)dom->nk * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  Dom d = to_dom(dom);
  hipStream_t s = as_stream(stream);
  const int ncol = (d.iec - d.isc + 3) * (d.jec - d.jsc + 3);
  hipLaunchKernelGGL((update_dz_c_kernel<double, U>), dim3((ncol + 255) / 256, d.ntiles),
                     dim3(256), 0, s, d, dp_ref, zs, area, ut, vt, gz, work, ws, 1.0 / dt);
  FV3_HIP_CHECK_LAUNCH(who);
  const Box bx{d.isc - 1, d.iec + 1, d.jsc - 1, d.jec + 1};
  const double* src = work;
  launch_box(s, d, bx, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    gz[o] = src[o];
  });
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

// UpdateHeightOnDGrid (updatedzd.py) with the Courant numbers and area
// fluxes of storage type U.
template <typename U>
int update_dz_d(const fv3_domain* dom, const fv3_metrics* g, int32_t hord_tm,
                const int32_t nord4[4], const double* damp, const double* gk,
                const double* beta, const double* gamma, const double* zs, double* height,
                const U* crx, const U* cry, const U* xfx, const U* yfx, double* ws, double dt,
                double* work, int64_t work_slab, void* stream) {
  const char* who = "fv3_update_dz_d";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  const void* need[] = {g ? g->area : nullptr, g ? g->dxa : nullptr, g ? g->dya : nullptr,
                        g ? g->rarea : nullptr, g ? g->del6_u : nullptr,
                        g ? g->del6_v : nullptr};
  rc = check_metrics_list(g, need, 6, who);
  if (rc) return rc;
  FV3_REQUIRE(nord4 && damp && gk && beta && gamma && zs && height && crx && cry && xfx && yfx &&
                  ws && work,
              FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(dom->nk >= 3, FV3_E_ARG, "%s: needs at least 3 interface levels", who);
  if (dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)dom->nk * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  Dom d = to_dom(dom);
  hipStream_t s = as_stream(stream);
  double* W[9];
  for (int n = 0; n < 9; ++n) W[n] = work + (int64_t)n * work_slab;
  double *crx_i = W[0], *xfx_i = W[1], *cry_i = W[2], *yfx_i = W[3], *fx = W[4], *fy = W[5];
  double *gxd = W[6], *gyd = W[7], *qy = W[8];
  const int ncol = (d.iec - d.isc + 7) * (d.jec - d.jsc + 7);
  hipLaunchKernelGGL((spline_kernel<double, U>), dim3((ncol + 255) / 256, d.ntiles, 4),
                     dim3(256), 0, s, d, gk, beta, gamma, crx, xfx, cry, yfx, crx_i, xfx_i,
                     cry_i, yfx_i);
  FV3_HIP_CHECK_LAUNCH(who);
  const double* area = static_cast<const double*>(g->area);
  rc = fvtp2d_corner_fill_f64(dom, hord_tm, height, qy, crx_i, cry_i, xfx_i, yfx_i, nullptr,
                              nullptr, area, static_cast<const double*>(g->dxa),
                              static_cast<const double*>(g->dya), fx, fy, false, stream);
  if (rc) return rc;
  rc = fv3_delnflux_f64(dom, FV3_DELN_STORE, nord4, height, nullptr, damp,
                        static_cast<const double*>(g->del6_u), static_cast<const double*>(g->del6_v),
                        static_cast<const double*>(g->rarea), gxd, gyd, stream);
  if (rc) return rc;
  const int nc = (d.iec - d.isc + 1) * (d.jec - d.jsc + 1);
  hipLaunchKernelGGL((apply_height_kernel<double>), dim3((nc + 255) / 256, d.ntiles), dim3(256), 0,
                     s, d, area, zs, height, fx, fy, xfx_i, yfx_i, gxd, gyd, ws, dt);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // namespace fv3

extern "C" {

int32_t fv3_update_dz_c_f64(const fv3_domain* dom, const double* area, const double* zs,
                            const double* dp_ref, const double* ut, const double* vt,
                            double* gz, double* ws, double dt, double* work, int64_t work_slab,
                            void* stream) {
  return fv3::update_dz_c<double>(dom, area, zs, dp_ref, ut, vt, gz, ws, dt, work, work_slab,
                                  stream);
}

int32_t fv3_update_dz_c_mixed(const fv3_domain* dom, const double* area, const double* zs,
                              const double* dp_ref, const float* ut, const float* vt,
                              double* gz, double* ws, double dt, double* work,
                              int64_t work_slab, void* stream) {
  return fv3::update_dz_c<float>(dom, area, zs, dp_ref, ut, vt, gz, ws, dt, work, work_slab,
                                 stream);
}

int32_t fv3_update_dz_d_f64(const fv3_domain* dom, const fv3_metrics* g, int32_t hord_tm,
                            const int32_t nord4[4], const double* damp, const double* gk,
                            const double* beta, const double* gamma, const double* zs,
                            double* height, const double* crx, const double* cry,
                            const double* xfx, const double* yfx, double* ws, double dt,
                            double* work, int64_t work_slab, void* stream) {
  return fv3::update_dz_d<double>(dom, g, hord_tm, nord4, damp, gk, beta, gamma, zs, height, crx,
                                  cry, xfx, yfx, ws, dt, work, work_slab, stream);
}

int32_t fv3_update_dz_d_mixed(const fv3_domain* dom, const fv3_metrics* g, int32_t hord_tm,
                              const int32_t nord4[4], const double* damp, const double* gk,
                              const double* beta, const double* gamma, const double* zs,
                              double* height, const float* crx, const float* cry,
                              const float* xfx, const float* yfx, double* ws, double dt,
                              double* work, int64_t work_slab, void* stream) {
  return fv3::update_dz_d<float>(dom, g, hord_tm, nord4, damp, gk, beta, gamma, zs, height, crx,
                                 cry, xfx, yfx, ws, dt, work, work_slab, stream);
}

int32_t fv3_nh_p_grad_f64(const fv3_domain* dom, const fv3_metrics* g, double* u, double* v,
                          double* pp, double* gz, double* pk3, const double* delp, double dt,
                          double ptk, double* work, int64_t work_slab, void* stream) {
  return fv3::nh_p_grad<double>(dom, g, u, v, pp, gz, pk3, delp, dt, ptk, work, work_slab,
                                stream);
}

int32_t fv3_nh_p_grad_mixed(const fv3_domain* dom, const fv3_metrics* g, float* u, float* v,
                            double* pp, double* gz, double* pk3, const double* delp, double dt,
                            double ptk, double* work, int64_t work_slab, void* stream) {
  return fv3::nh_p_grad<float>(dom, g, u, v, pp, gz, pk3, delp, dt, ptk, work, work_slab,
                               stream);
}

int32_t fv3_pk3_halo_f64(const fv3_domain* dom, double* pk3, const double* delp, double ptop,
                         double akap, void* stream) {
  using namespace fv3;
  const char* who = "fv3_pk3_halo";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(pk3 && delp, FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const int64_t nc = ring_threads(d, 2);
  hipLaunchKernelGGL((ring_kernel<double>), dim3((unsigned)((nc + 255) / 256), d.ntiles),
                     dim3(256), 0, as_stream(stream), d, 2, 0, ptop, akap, delp, pk3);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_pe_halo_f64(const fv3_domain* dom, double* pe, const double* delp, double ptop,
                        void* stream) {
  using namespace fv3;
  const char* who = "fv3_pe_halo";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(pe && delp, FV3_E_ARG, "%s: NULL pointer", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const int64_t nc = ring_threads(d, 1);
  hipLaunchKernelGGL((ring_kernel<double>), dim3((unsigned)((nc + 255) / 256), d.ntiles),
                     dim3(256), 0, as_stream(stream), d, 1, 1, ptop, 0.0, delp, pe);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_p_grad_c_f64(const fv3_domain* dom, const double* rdxc, const double* rdyc,
                         int32_t hydrostatic, double* uc, double* vc, const double* delpc,
                         const double* pkc, const double* gz, double dt2, void* stream) {
  return fv3::p_grad_c<double>(dom, rdxc, rdyc, hydrostatic, uc, vc, delpc, pkc, gz, dt2, stream);
}

int32_t fv3_p_grad_c_mixed(const fv3_domain* dom, const double* rdxc, const double* rdyc,
                           int32_t hydrostatic, float* uc, float* vc, const float* delpc,
                           const double* pkc, const double* gz, double dt2, void* stream) {
  return fv3::p_grad_c<float>(dom, rdxc, rdyc, hydrostatic, uc, vc, delpc, pkc, gz, dt2, stream);
}

int32_t fv3_ray_fast_f64(const fv3_domain* dom, int32_t hydrostatic, const double* rf,
                         const double* dp, const double* dm, const int32_t* use_rf,
                         const int32_t* use_nudge, double* u, double* v, double* w,
                         void* stream) {
  return fv3::ray_fast<double>(dom, hydrostatic, rf, dp, dm, use_rf, use_nudge, u, v, w, stream);
}

int32_t fv3_ray_fast_f32(const fv3_domain* dom, int32_t hydrostatic, const float* rf,
                         const float* dp, const float* dm, const int32_t* use_rf,
                         const int32_t* use_nudge, float* u, float* v, float* w, void* stream) {
  return fv3::ray_fast<float>(dom, hydrostatic, rf, dp, dm, use_rf, use_nudge, u, v, w, stream);
}

int32_t fv3_del2cubed_f64(const fv3_domain* dom, const double* del6_u, const double* del6_v,
                          const double* rarea, int32_t nmax, double cd, double* qdel,
                          double* work, int64_t work_slab, void* stream) {
  return fv3::del2cubed<double>(dom, del6_u, del6_v, rarea, nmax, cd, qdel, work, work_slab,
                                stream);
}

int32_t fv3_del2cubed_f32(const fv3_domain* dom, const float* del6_u, const float* del6_v,
                          const float* rarea, int32_t nmax, float cd, float* qdel, float* work,
                          int64_t work_slab, void* stream) {
  return fv3::del2cubed<float>(dom, del6_u, del6_v, rarea, nmax, cd, qdel, work, work_slab,
                               stream);
}

int32_t fv3_pkz_tempadjust_f64(const fv3_domain* dom, const double* delp, const double* delz,
                               const double* cappa, const double* heat_source, double* pt,
                               double* pkz, double delt_time_factor, void* stream) {
  return fv3::pkz_tempadjust<double>(dom, delp, delz, cappa, heat_source, pt, pkz,
                                     delt_time_factor, stream);
}

int32_t fv3_pkz_tempadjust_mixed(const fv3_domain* dom, const float* delp, const double* delz,
                                 const float* cappa, const float* heat_source, float* pt,
                                 double* pkz, double delt_time_factor, void* stream) {
  return fv3::pkz_tempadjust<float>(dom, delp, delz, cappa, heat_source, pt, pkz,
                                    delt_time_factor, stream);
}

int32_t fv3_dyncore_column_f64(const fv3_domain* dom, int32_t op, double scalar, const double* a,
                               const double* b, double* out, void* stream) {
  return fv3::dyncore_column<double>(dom, op, scalar, a, b, out, stream);
}

int32_t fv3_dyncore_column_f32(const fv3_domain* dom, int32_t op, double scalar, const float* a,
                               const float* b, float* out, void* stream) {
  return fv3::dyncore_column<float>(dom, op, scalar, a, b, out, stream);
}
}
