// Fused del-n diffusive fluxes (DelnFlux / DelnFluxNoSG).
//
// Reference: fv3core/stencils/delnflux.py:901-1207 runs, per call,
// 5 + 5*nmax stencils (d2 preamble, x/y corner copies, fx/fy, and per extra
// order a divergence d2 plus corner copies and fluxes) on full-size HBM
// temporaries.  Here one workgroup owns a TI x TJ block of interfaces on one
// level; d2 and the intermediate fluxes live in LDS over the block plus a
// 3-cell halo, and the corner copies are done in place in LDS.  Levels whose
// band has nord == 0 restrict the del-2 flux to the compute interfaces, as
// the reference's regions do.
#include "common.hpp"
#include "tile_config.hpp"
#include "timing.hpp"

namespace fv3 {


// A workgroup is 4 waves; lane x of every wave owns LDS column x (the block
// plus its 3-cell halo is exactly one wave wide), and wave y owns rows y,
// y+4, ... so no index is ever divided.
constexpr int DN_W = 64;            // LDS columns = TI + 6
constexpr int DN_TI = DN_W - 6;     // interfaces per block in i
// Block shapes (interface rows, waves): the default, and a taller one for
// FP32 launches where it computes fewer LDS rows (the 6 halo rows are paid
// once per block); in FP64 its LDS tile halves the blocks a CU holds.
constexpr int DN_TJ = FV3_DN_TJ;
constexpr int DN_NY = FV3_DN_NY;
constexpr int DN_TJ_TALL = FV3_DN_TJ_TALL;
constexpr int DN_NY_TALL = FV3_DN_NY_TALL;

struct Box {
  int i0, i1, j0, j1;  // inclusive absolute bounds
  __device__ __forceinline__ bool in(int i, int j) const {
    return i >= i0 && i <= i1 && j >= j0 && j <= j1;
  }
};

template <typename T, int TJ_, int NY_>
__global__ void __launch_bounds__(DN_W * NY_)
    delnflux_kernel(Dom d, int mode, int4 nord4, int nmax, CornerTable tab,
                    const T* __restrict__ q, const T* __restrict__ mass,
                    const T* __restrict__ damp, const T* __restrict__ del6_u,
                    const T* __restrict__ del6_v, const T* __restrict__ rarea,
                    T* __restrict__ fx, T* __restrict__ fy, int nbi) {
  constexpr int W = DN_W, TI = DN_TI, TJ = TJ_, H = TJ_ + 6;  // LDS rows H
  constexpr int DN_NY = NY_, DN_NT = DN_W * NY_;
  constexpr int NR = (H + DN_NY - 1) / DN_NY;  // LDS rows per lane
  __shared__ T sd2[H][W];
  __shared__ T sfx[H][W];
  __shared__ T sfy[H][W];

  const int li = threadIdx.x;  // LDS column of this lane
  const int ty = threadIdx.y;  // this wave owns LDS rows ty, ty + 4, ...
  const int tid = ty * W + li;
  const int bi = blockIdx.x % nbi, bj = blockIdx.x / nbi;
  const int k = blockIdx.y, t = blockIdx.z;
  const int i0 = d.isc + bi * TI, j0 = d.jsc + bj * TJ;
  const int ib = i0 - 3, jb = j0 - 3;  // absolute index of LDS [0][0]
  const int i = ib + li;               // absolute column of this lane
  const bool col_ok = i >= 0 && i < d.ni;
  const int64_t off3 = (int64_t)t * d.st + (int64_t)k * d.sk;
  const int64_t off2 = (int64_t)t * d.st2;
  const int band = k < 3 ? k : 3;
  const int nord_k = band == 0 ? nord4.x : band == 1 ? nord4.y : band == 2 ? nord4.z : nord4.w;
  const bool pos = nord_k > 0;
  const bool use_damp = (mode == FV3_DELN_ADD || mode == FV3_DELN_STORE);
  const T dk = damp[k];
  const int nj = d.nj;

  // Every global value the passes need at this lane's cells is loaded up
  // front, so the passes below wait only on LDS and barriers.
  T qv[NR], dv[NR], du[NR], ra[NR];
  {
    const int g = pos ? nmax : 0;
    const Box b0{d.isc - 1 - g, d.iec + 1 + g, d.jsc - 1 - g, d.jec + 1 + g};
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int lj = ty + r * DN_NY;
      const int j = jb + lj;
      const bool ok = lj < H && col_ok && j >= 0 && j < nj;
      const int64_t o2 = off2 + (int64_t)j * d.sj2 + i;
      qv[r] = (ok && b0.in(i, j)) ? q[off3 + (int64_t)j * d.sj + i] : T(0);
      dv[r] = ok ? del6_v[o2] : T(0);
      du[r] = ok ? del6_u[o2] : T(0);
      ra[r] = (ok && pos) ? rarea[o2] : T(0);
    }
  }
  // d2 preamble (d2_damp_interval / copy_stencil_interval)
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int lj = ty + r * DN_NY;
    if (lj < H) {
      sd2[lj][li] = use_damp ? dk * qv[r] : qv[r];
      sfx[lj][li] = T(0);
      sfy[lj][li] = T(0);
    }
  }
  // does any corner-copy destination fall inside this block's window?
  bool mine = false;
  for (int e = tid; e < tab.n; e += DN_NT) {
    const int di = tab.dst_i[e] - ib, dj = tab.dst_j[e] - jb;
    mine = mine || (di >= 0 && di < W && dj >= 0 && dj < H);
  }
  const bool corners = pos && __syncthreads_or(mine);
  if (!pos) __syncthreads();

  auto corner_copy = [&](int dir) {
    // sources are never corner cells, so the in-place copy has no ordering hazard
    for (int e = tid; e < tab.n; e += DN_NT) {
      if (tab.dir[e] != dir) continue;
      const int di = tab.dst_i[e] - ib, dj = tab.dst_j[e] - jb;
      if (di < 0 || di >= W || dj < 0 || dj >= H) continue;
      const int si = tab.src_i[e] - ib, sj = tab.src_j[e] - jb;
      const T v = (si >= 0 && si < W && sj >= 0 && sj < H) ? sd2[sj][si] : T(0);
      sd2[dj][di] = v;
    }
  };
  auto flux_x = [&](const Box& bx, T sgnv) {
    if (li < 1 || !col_ok) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int lj = ty + r * DN_NY;
      const int j = jb + lj;
      if (lj < H && bx.in(i, j) && j < nj)
        sfx[lj][li] = sgnv * dv[r] * (sd2[lj][li - 1] - sd2[lj][li]);
    }
  };
  auto flux_y = [&](const Box& bx, T sgnv) {
    if (!col_ok) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int lj = ty + r * DN_NY;
      const int j = jb + lj;
      if (lj >= 1 && lj < H && bx.in(i, j) && j < nj)
        sfy[lj][li] = sgnv * du[r] * (sd2[lj - 1][li] - sd2[lj][li]);
    }
  };
  // one del-2 flux pass: with corner cells in the window the x and y fluxes
  // see the x- and y-corner-copied d2 (two more barriers); otherwise both
  // read the same d2 in one phase
  auto fluxes = [&](const Box& bx, const Box& by, T sgnv) {
    if (corners) {
      corner_copy(0);
      __syncthreads();
      flux_x(bx, sgnv);
      __syncthreads();
      corner_copy(1);
      __syncthreads();
      flux_y(by, sgnv);
    } else {
      flux_x(bx, sgnv);
      flux_y(by, sgnv);
    }
    __syncthreads();
  };

  if (pos) {
    fluxes(Box{d.isc - nmax, d.iec + 1 + nmax, d.jsc - nmax, d.jec + nmax},
           Box{d.isc - nmax, d.iec + nmax, d.jsc - nmax, d.jec + 1 + nmax}, T(1));
    for (int n = 0; n < nmax; ++n) {
      const int nt = nmax - 1 - n;
      // d2_highorder (delnflux.py:178-181)
      const Box bd{d.isc - nt - 1, d.iec + nt + 1, d.jsc - nt - 1, d.jec + nt + 1};
      if (li + 1 < W && col_ok) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int lj = ty + r * DN_NY;
          const int j = jb + lj;
          if (lj + 1 < H && bd.in(i, j) && j < nj)
            sd2[lj][li] = (sfx[lj][li] - sfx[lj][li + 1] + sfy[lj][li] - sfy[lj + 1][li]) * ra[r];
        }
      }
      __syncthreads();
      fluxes(Box{d.isc - nt, d.iec + 1 + nt, d.jsc - nt, d.jec + nt},
             Box{d.isc - nt, d.iec + nt, d.jsc - nt, d.jec + 1 + nt}, T(-1));
    }
  } else {
    fluxes(Box{d.isc, d.iec + 1, d.jsc, d.jec}, Box{d.isc, d.iec, d.jsc, d.jec + 1}, T(1));
  }

  // apply to the caller's fluxes on the block's interfaces
  if (li < 3 || li >= 3 + TI) return;
  for (int lj = 3 + ty; lj < 3 + TJ; lj += DN_NY) {
    const int j = jb + lj;
    const int64_t o = off3 + (int64_t)j * d.sj + i;
    if (i <= d.iec + 1 && j <= d.jec) {
      const T f2 = sfx[lj][li];
      if (mode == FV3_DELN_ADD) {
        fx[o] = fx[o] + f2;
      } else if (mode == FV3_DELN_MASS) {
        fx[o] = fx[o] + T(0.5) * dk * (mass[o - 1] + mass[o]) * f2;
      } else {
        fx[o] = f2;
      }
    }
    if (i <= d.iec && j <= d.jec + 1) {
      const T f2 = sfy[lj][li];
      if (mode == FV3_DELN_ADD) {
        fy[o] = fy[o] + f2;
      } else if (mode == FV3_DELN_MASS) {
        fy[o] = fy[o] + T(0.5) * dk * (mass[o - d.sj] + mass[o]) * f2;
      } else {
        fy[o] = f2;
      }
    }
  }
}

template <typename T>
static int delnflux_launch(const fv3_domain* dom, int mode, const int32_t* nord4, const T* q,
                           const T* mass, const T* damp, const T* del6_u, const T* del6_v,
                           const T* rarea, T* fx, T* fy, void* stream) {
  const char* who = "fv3_delnflux";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(nord4 && q && damp && del6_u && del6_v && rarea && fx && fy, FV3_E_ARG,
              "%s: required pointer is NULL", who);
  FV3_REQUIRE(mode >= 0 && mode <= 3, FV3_E_ARG, "%s: bad mode %d", who, mode);
  FV3_REQUIRE(mode != FV3_DELN_MASS || mass, FV3_E_ARG, "%s: mass is required", who);
  int nmax = 0;
  for (int b = 0; b < 4; ++b) {
    // The reference states nord in {0, 2, 3} but its check accepts any value
    // up to 3 (delnflux.py:1024-1029), so 1 is accepted here too.
    FV3_REQUIRE(nord4[b] >= 0 && nord4[b] <= 3, FV3_E_UNSUPPORTED,
                "%s: nord must be in [0, 3] (got %d)", who, nord4[b]);
    nmax = nord4[b] > nmax ? nord4[b] : nmax;
  }
  // nmax = 3 puts the reference's d2 origin at isc - 4, outside the 3-cell
  // halo storage (delnflux.py:1031-1036), so it cannot run there either.
  FV3_REQUIRE(nmax <= 2, FV3_E_UNSUPPORTED, "%s: nord = 3 needs a 4-cell halo", who);
  if (dom->nk == 0 || dom->ntiles == 0) return FV3_OK;
  FV3_REQUIRE(dom->nk <= 65535 && dom->ntiles <= 65535, FV3_E_ARG, "%s: grid too large", who);
  Dom d = to_dom(dom);
  CornerTable tab;
  build_corner_table(dom, 3, &tab);
  const int nfi = d.iec + 2 - d.isc, nfj = d.jec + 2 - d.jsc;
  const int nbi = (nfi + DN_TI - 1) / DN_TI;
  // the tall shape for FP32 when it computes at least a quarter fewer LDS
  // rows per column of blocks (192 C48 tiles: 2 x 31 rows instead of 7 x 14,
  // 602 -> 494 us per launch; in FP64 six C96 tiles keep the default, 86 us
  // against 105 us tall)
  const int nbj_d = (nfj + DN_TJ - 1) / DN_TJ, nbj_t = (nfj + DN_TJ_TALL - 1) / DN_TJ_TALL;
  const bool tall = sizeof(T) == 4 &&
                    4 * nbj_t * (DN_TJ_TALL + 6) <= 3 * nbj_d * (DN_TJ + 6);
  const int nbj = tall ? nbj_t : nbj_d;
  dim3 grid(nbi * nbj, d.nk, d.ntiles);
  int4 n4 = make_int4(nord4[0], nord4[1], nord4[2], nord4[3]);
  // q read, fx / fy read and written (per cell-level of the compute domain)
  TimedLaunch timed("delnflux_kernel", as_stream(stream),
                    5.0 * sizeof(T) * (double)(d.iec - d.isc + 1) * (d.jec - d.jsc + 1) * d.nk *
                        d.ntiles);
  if (tall)
    hipLaunchKernelGGL((delnflux_kernel<T, DN_TJ_TALL, DN_NY_TALL>), grid,
                       dim3(DN_W, DN_NY_TALL), 0, as_stream(stream), d, mode, n4, nmax, tab, q,
                       mass, damp, del6_u, del6_v, rarea, fx, fy, nbi);
  else
    hipLaunchKernelGGL((delnflux_kernel<T, DN_TJ, DN_NY>), grid, dim3(DN_W, DN_NY), 0,
                       as_stream(stream), d, mode, n4, nmax, tab, q, mass, damp, del6_u, del6_v,
                       rarea, fx, fy, nbi);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // namespace fv3

extern "C" {

int32_t fv3_delnflux_f64(const fv3_domain* d, int32_t mode, const int32_t nord4[4],
                         const double* q, const double* mass, const double* damp,
                         const double* del6_u, const double* del6_v, const double* rarea,
                         double* fx, double* fy, void* stream) {
  return fv3::delnflux_launch<double>(d, mode, nord4, q, mass, damp, del6_u, del6_v, rarea,
                                      fx, fy, stream);
}

int32_t fv3_delnflux_f32(const fv3_domain* d, int32_t mode, const int32_t nord4[4],
                         const float* q, const float* mass, const float* damp,
                         const float* del6_u, const float* del6_v, const float* rarea,
                         float* fx, float* fy, void* stream) {
  return fv3::delnflux_launch<float>(d, mode, nord4, q, mass, damp, del6_u, del6_v, rarea, fx,
                                     fy, stream);
}
}
