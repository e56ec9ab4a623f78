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

namespace fv3 {


constexpr int DN_TI = 32;
constexpr int DN_TJ = 16;
constexpr int DN_NT = 256;

struct Box {
  int i0, i1, j0, j1;  // inclusive absolute bounds
  __device__ __forceinline__ bool in(int i, int j) const {
    return i >= i0 && i <= i1 && j >= j0 && j <= j1;
  }
};

template <typename T>
__global__ void __launch_bounds__(DN_NT)
    delnflux_kernel(Dom d, int mode, int4 nord4, int nmax, CornerTable tab,
                    const T* __restrict__ q, const T* __restrict__ mass,
                    const T* __restrict__ damp, const T* __restrict__ del6_u,
                    const T* __restrict__ del6_v, const T* __restrict__ rarea,
                    T* __restrict__ fx, T* __restrict__ fy, int nbi) {
  constexpr int TI = DN_TI, TJ = DN_TJ;
  constexpr int W = TI + 6, H = TJ + 6;
  __shared__ T sd2[H][W];
  __shared__ T sfx[H][W];
  __shared__ T sfy[H][W];

  const int tid = threadIdx.x;
  const int bi = blockIdx.x % nbi, bj = blockIdx.x / nbi;
  const int k = blockIdx.y, t = blockIdx.z;
  const int i0 = d.isc + bi * TI, j0 = d.jsc + bj * TJ;
  const int ib = i0 - 3, jb = j0 - 3;  // absolute index of LDS [0][0]
  const int64_t off3 = (int64_t)t * d.st + (int64_t)k * d.sk;
  const int64_t off2 = (int64_t)t * d.st2;
  const int band = k < 3 ? k : 3;
  const int nord_k = band == 0 ? nord4.x : band == 1 ? nord4.y : band == 2 ? nord4.z : nord4.w;
  const bool pos = nord_k > 0;
  const bool use_damp = (mode == FV3_DELN_ADD || mode == FV3_DELN_STORE);
  const T dk = damp[k];
  const int ni = d.ni, nj = d.nj;

  // d2 preamble (d2_damp_interval / copy_stencil_interval)
  {
    const int g = pos ? nmax : 0;
    const Box b0{d.isc - 1 - g, d.iec + 1 + g, d.jsc - 1 - g, d.jec + 1 + g};
    for (int n = tid; n < W * H; n += DN_NT) {
      const int li = n % W, lj = n / W;
      const int i = ib + li, j = jb + lj;
      T v = T(0);
      if (b0.in(i, j) && i >= 0 && i < ni && j >= 0 && j < nj) {
        const T qv = q[off3 + (int64_t)j * d.sj + i];
        v = use_damp ? dk * qv : qv;
      }
      sd2[lj][li] = v;
      sfx[lj][li] = T(0);
      sfy[lj][li] = T(0);
    }
  }
  __syncthreads();

  auto corner_copy = [&](int dir) {
    // sources are never corner cells, so the in-place copy has no ordering hazard
    for (int e = tid; e < tab.n; e += DN_NT) {
      if (tab.dir[e] != dir) continue;
      const int li = tab.dst_i[e] - ib, lj = tab.dst_j[e] - jb;
      if (li < 0 || li >= W || lj < 0 || lj >= H) continue;
      const int si = tab.src_i[e] - ib, sj = tab.src_j[e] - jb;
      const T v = (si >= 0 && si < W && sj >= 0 && sj < H) ? sd2[sj][si] : T(0);
      sd2[lj][li] = v;
    }
  };
  auto flux_x = [&](const Box& bx, T sgnv) {
    for (int n = tid; n < W * H; n += DN_NT) {
      const int li = n % W, lj = n / W;
      const int i = ib + li, j = jb + lj;
      if (li >= 1 && bx.in(i, j) && i < ni && j < nj) {
        const T c = del6_v[off2 + (int64_t)j * d.sj2 + i];
        sfx[lj][li] = sgnv * c * (sd2[lj][li - 1] - sd2[lj][li]);
      }
    }
  };
  auto flux_y = [&](const Box& bx, T sgnv) {
    for (int n = tid; n < W * H; n += DN_NT) {
      const int li = n % W, lj = n / W;
      const int i = ib + li, j = jb + lj;
      if (lj >= 1 && bx.in(i, j) && i < ni && j < nj) {
        const T c = del6_u[off2 + (int64_t)j * d.sj2 + i];
        sfy[lj][li] = sgnv * c * (sd2[lj - 1][li] - sd2[lj][li]);
      }
    }
  };

  if (pos) corner_copy(0);
  __syncthreads();
  {
    const Box bx = pos ? Box{d.isc - nmax, d.iec + 1 + nmax, d.jsc - nmax, d.jec + nmax}
                       : Box{d.isc, d.iec + 1, d.jsc, d.jec};
    flux_x(bx, T(1));
  }
  __syncthreads();
  if (pos) corner_copy(1);
  __syncthreads();
  {
    const Box by = pos ? Box{d.isc - nmax, d.iec + nmax, d.jsc - nmax, d.jec + 1 + nmax}
                       : Box{d.isc, d.iec, d.jsc, d.jec + 1};
    flux_y(by, T(1));
  }
  __syncthreads();

  if (pos) {
    for (int n = 0; n < nmax; ++n) {
      const int nt = nmax - 1 - n;
      // d2_highorder (delnflux.py:178-181)
      const Box bd{d.isc - nt - 1, d.iec + nt + 1, d.jsc - nt - 1, d.jec + nt + 1};
      for (int m = tid; m < W * H; m += DN_NT) {
        const int li = m % W, lj = m / W;
        const int i = ib + li, j = jb + lj;
        if (li + 1 < W && lj + 1 < H && bd.in(i, j) && i < ni && j < nj) {
          const T ra = rarea[off2 + (int64_t)j * d.sj2 + i];
          sd2[lj][li] = (sfx[lj][li] - sfx[lj][li + 1] + sfy[lj][li] - sfy[lj + 1][li]) * ra;
        }
      }
      __syncthreads();
      corner_copy(0);
      __syncthreads();
      flux_x(Box{d.isc - nt, d.iec + 1 + nt, d.jsc - nt, d.jec + nt}, T(-1));
      __syncthreads();
      corner_copy(1);
      __syncthreads();
      flux_y(Box{d.isc - nt, d.iec + nt, d.jsc - nt, d.jec + 1 + nt}, T(-1));
      __syncthreads();
    }
  }

  // apply to the caller's fluxes on the block's interfaces
  for (int n = tid; n < TI * TJ; n += DN_NT) {
    const int li = n % TI, lj = n / TI;
    const int i = i0 + li, j = j0 + lj;
    const int64_t o = off3 + (int64_t)j * d.sj + i;
    if (i <= d.iec + 1 && j <= d.jec) {
      const T f2 = sfx[lj + 3][li + 3];
      if (mode == FV3_DELN_ADD) {
        fx[o] = fx[o] + f2;
      } else if (mode == FV3_DELN_MASS) {
        fx[o] = fx[o] + T(0.5) * dk * (mass[o - 1] + mass[o]) * f2;
      } else {
        fx[o] = f2;
      }
    }
    if (i <= d.iec && j <= d.jec + 1) {
      const T f2 = sfy[lj + 3][li + 3];
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
  const int nbi = (nfi + DN_TI - 1) / DN_TI, nbj = (nfj + DN_TJ - 1) / DN_TJ;
  dim3 grid(nbi * nbj, d.nk, d.ntiles);
  int4 n4 = make_int4(nord4[0], nord4[1], nord4[2], nord4[3]);
  hipLaunchKernelGGL((delnflux_kernel<T>), grid, dim3(DN_NT), 0, as_stream(stream), d, mode, n4,
                     nmax, tab, q, mass, damp, del6_u, del6_v, rarea, fx, fy, nbi);
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
