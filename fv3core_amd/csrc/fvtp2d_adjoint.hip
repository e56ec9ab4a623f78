// Tangent-linear model and adjoint of fv_tp_2d (FiniteVolumeTransport,
// fv3core/stencils/fvtp2d.py:151-362) with respect to the transported field,
// at the branch decisions (limiters, smoothness masks, upwind side) of the
// forward state.  The reference has no derivative code; these kernels
// linearise the same statements the forward kernel evaluates.
//
// Structure.  The forward is a chain of linear maps once its branches are
// fixed:
//   fy2 = PPM_y(q_y)                            y-inner faces
//   q_i = (q_y A + ay(j) fy2(j) - ay(j+1) fy2(j+1)) / Dy      fvtp2d.py:33-47
//   fxo = PPM_x(q_i)                            x-outer faces
//   fx2 = PPM_x(q_x)                            x-inner faces
//   q_j = (q_x A + ax(i) fx2(i) - ax(i+1) fx2(i+1)) / Dx      fvtp2d.py:50-60
//   fyo = PPM_y(q_j)                            y-outer faces
//   fx = (fxo + fx2) ux / 2,  fy = (fyo + fy2) uy / 2         fvtp2d.py:63-92
// Each 1-D PPM face depends on the six cells around it.  Its six partial
// derivatives are obtained by running the forward's own face function
// (ppm_face, every hord and tile-edge branch) on Dual<6> numbers, whose value
// part repeats the double arithmetic exactly.  The linearisation pass stores
// the partials and the intermediate fields in workspace slabs; the TLM then
// applies the chain forward, the adjoint applies its transpose as gathers
// (each output point sums its own contributions, no atomics), so both are
// deterministic.
#include "common.hpp"
#include "dual.hpp"
#include "ppm.hpp"
#include "stencil.hpp"
#include "timing.hpp"

namespace fv3 {
namespace {

using D6 = Dual<6>;

// workspace slabs (each a field of the transport domain)
enum : int {
  kFy2 = 0,   // y-inner face means
  kFx2 = 1,   // x-inner face means
  kQi = 2,    // q advected in y
  kQj = 3,    // q advected in x
  kAy = 4,    // 6 partials of the y-inner faces w.r.t. q_y(i, f-3+m)
  kAx = 10,   // 6 partials of the x-inner faces w.r.t. q_x(f-3+m, j)
  kAxo = 16,  // 6 partials of the x-outer faces w.r.t. q_i(f-3+m, j)
  kAyo = 22,  // 6 partials of the y-outer faces w.r.t. q_j(i, f-3+m)
  kLi = 28,   // q_i perturbation (TLM) or cotangent (adjoint)
  kLj = 29,   // q_j perturbation or cotangent
  kLy2 = 30,  // y-inner face perturbation or cotangent
  kLx2 = 31,  // x-inner face perturbation or cotangent
  kNumSlabs = 32
};

struct Lin {
  Dom d;
  int hord;
  const double *qx, *qy, *crx, *cry, *ax, *ay, *ux, *uy, *area, *dxa, *dya;
  double* w;
  int64_t slab;
  __device__ __forceinline__ double* s(int n) const { return w + (int64_t)n * slab; }
};

template <int ORD>
__device__ __forceinline__ D6 face_dual(const double* q, int64_t stride, double c, int f, int is,
                                        int ie, const double* dx, int64_t dstride) {
  D6 q6[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) q6[m] = D6::seed(q[(m - 3) * stride], m);
  DxRow<D6, double> dr{dx - 3 * dstride, dstride};
  return ppm_face<D6, ORD>(q6, D6(c), f, is, ie, dr);
}

// Inner faces on the input fields: values and partials.
template <int ORD>
void launch_inner(hipStream_t s, const Lin& L) {
  const Dom& d = L.d;
  const int I0 = d.isc, I1 = d.iec, J0 = d.jsc, J1 = d.jec;
  launch_box(s, d, Box{I0 - 3, I1 + 3, J0 - 3, J1 + 3}, 0, d.nk,
             [=] __device__(int i, int j, int k, int t) {
               const int64_t o = at3(d, i, j, k, t);
               const int64_t o2 = at2(d, i, j, t);
               if (j >= J0 && j <= J1 + 1) {  // y-inner face (i, j)
                 const D6 v = face_dual<ORD>(L.qy + o, d.sj, L.cry[o], j, d.js, d.je,
                                             L.dya + o2, d.sj2);
                 L.s(kFy2)[o] = v.v;
#pragma unroll
                 for (int m = 0; m < 6; ++m) L.s(kAy + m)[o] = v.d[m];
               }
               if (i >= I0 && i <= I1 + 1) {  // x-inner face (i, j)
                 const D6 v = face_dual<ORD>(L.qx + o, 1, L.crx[o], i, d.is, d.ie, L.dxa + o2, 1);
                 L.s(kFx2)[o] = v.v;
#pragma unroll
                 for (int m = 0; m < 6; ++m) L.s(kAx + m)[o] = v.d[m];
               }
             });
}

// Cross updates q_i, q_j with the forward's expressions.
void launch_cross(hipStream_t s, const Lin& L) {
  const Dom& d = L.d;
  const int I0 = d.isc, I1 = d.iec, J0 = d.jsc, J1 = d.jec;
  launch_box(s, d, Box{I0 - 3, I1 + 3, J0 - 3, J1 + 3}, 0, d.nk,
             [=] __device__(int i, int j, int k, int t) {
               const int64_t o = at3(d, i, j, k, t);
               const double a = L.area[at2(d, i, j, t)];
               if (j >= J0 && j <= J1) {
                 const double g0 = L.ay[o], g1 = L.ay[o + d.sj];
                 L.s(kQi)[o] = (L.qy[o] * a + g0 * L.s(kFy2)[o] - g1 * L.s(kFy2)[o + d.sj]) /
                               (a + g0 - g1);
               }
               if (i >= I0 && i <= I1) {
                 const double f0 = L.ax[o], f1 = L.ax[o + 1];
                 L.s(kQj)[o] = (L.qx[o] * a + f0 * L.s(kFx2)[o] - f1 * L.s(kFx2)[o + 1]) /
                               (a + f0 - f1);
               }
             });
}

// Outer faces: partials only.
template <int ORD>
void launch_outer(hipStream_t s, const Lin& L) {
  const Dom& d = L.d;
  const int I0 = d.isc, I1 = d.iec, J0 = d.jsc, J1 = d.jec;
  launch_box(s, d, Box{I0, I1 + 1, J0, J1 + 1}, 0, d.nk,
             [=] __device__(int i, int j, int k, int t) {
               const int64_t o = at3(d, i, j, k, t);
               const int64_t o2 = at2(d, i, j, t);
               if (j <= J1) {  // x-outer face (i, j) on q_i
                 const D6 v = face_dual<ORD>(L.s(kQi) + o, 1, L.crx[o], i, d.is, d.ie,
                                             L.dxa + o2, 1);
#pragma unroll
                 for (int m = 0; m < 6; ++m) L.s(kAxo + m)[o] = v.d[m];
               }
               if (i <= I1) {  // y-outer face (i, j) on q_j
                 const D6 v = face_dual<ORD>(L.s(kQj) + o, d.sj, L.cry[o], j, d.js, d.je,
                                             L.dya + o2, d.sj2);
#pragma unroll
                 for (int m = 0; m < 6; ++m) L.s(kAyo + m)[o] = v.d[m];
               }
             });
}

template <int ORD>
void linearize_ord(hipStream_t s, const Lin& L) {
  launch_inner<ORD>(s, L);
  launch_cross(s, L);
  launch_outer<ORD>(s, L);
}

void linearize(hipStream_t s, const Lin& L) {
  switch (L.hord) {
    case 5: linearize_ord<5>(s, L); break;
    case 6: linearize_ord<6>(s, L); break;
    case 7: linearize_ord<7>(s, L); break;
    default: linearize_ord<8>(s, L); break;
  }
}

// TLM: dfx, dfy from dq_x, dq_y.
void launch_tlm(hipStream_t s, const Lin& L, const double* dqx, const double* dqy, double* dfx,
                double* dfy) {
  const Dom d = L.d;
  const int I0 = d.isc, I1 = d.iec, J0 = d.jsc, J1 = d.jec;
  const Box wide{I0 - 3, I1 + 3, J0 - 3, J1 + 3};
  // inner face perturbations
  launch_box(s, d, wide, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    if (j >= J0 && j <= J1 + 1) {
      double v = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) v += L.s(kAy + m)[o] * dqy[o + (m - 3) * d.sj];
      L.s(kLy2)[o] = v;
    }
    if (i >= I0 && i <= I1 + 1) {
      double v = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) v += L.s(kAx + m)[o] * dqx[o + (m - 3)];
      L.s(kLx2)[o] = v;
    }
  });
  // cross-update perturbations
  launch_box(s, d, wide, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const double a = L.area[at2(d, i, j, t)];
    if (j >= J0 && j <= J1) {
      const double g0 = L.ay[o], g1 = L.ay[o + d.sj];
      L.s(kLi)[o] = (dqy[o] * a + g0 * L.s(kLy2)[o] - g1 * L.s(kLy2)[o + d.sj]) / (a + g0 - g1);
    }
    if (i >= I0 && i <= I1) {
      const double f0 = L.ax[o], f1 = L.ax[o + 1];
      L.s(kLj)[o] = (dqx[o] * a + f0 * L.s(kLx2)[o] - f1 * L.s(kLx2)[o + 1]) / (a + f0 - f1);
    }
  });
  // outer faces and final fluxes
  launch_box(s, d, Box{I0, I1 + 1, J0, J1 + 1}, 0, d.nk,
             [=] __device__(int i, int j, int k, int t) {
               const int64_t o = at3(d, i, j, k, t);
               if (j <= J1) {
                 double v = 0.0;
#pragma unroll
                 for (int m = 0; m < 6; ++m) v += L.s(kAxo + m)[o] * L.s(kLi)[o + (m - 3)];
                 dfx[o] = 0.5 * (v + L.s(kLx2)[o]) * L.ux[o];
               }
               if (i <= I1) {
                 double v = 0.0;
#pragma unroll
                 for (int m = 0; m < 6; ++m) v += L.s(kAyo + m)[o] * L.s(kLj)[o + (m - 3) * d.sj];
                 dfy[o] = 0.5 * (v + L.s(kLy2)[o]) * L.uy[o];
               }
             });
}

// Adjoint: g_qx, g_qy from g_fx, g_fy (overwrites the whole storage of both).
void launch_adjoint(hipStream_t s, const Lin& L, const double* gfx, const double* gfy,
                    double* gqx, double* gqy) {
  const Dom d = L.d;
  const int I0 = d.isc, I1 = d.iec, J0 = d.jsc, J1 = d.jec;
  const Box wide{I0 - 3, I1 + 3, J0 - 3, J1 + 3};
  // cotangents of the outer-sweep inputs q_i and q_j
  launch_box(s, d, wide, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    if (j >= J0 && j <= J1) {  // q_i(i, j) feeds x-outer faces f = i + 3 - m
      double g = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const int f = i + 3 - m;
        if (f >= I0 && f <= I1 + 1) {
          const int64_t of = o + (3 - m);
          g += L.s(kAxo + m)[of] * (0.5 * L.ux[of] * gfx[of]);
        }
      }
      L.s(kLi)[o] = g;
    }
    if (i >= I0 && i <= I1) {  // q_j(i, j) feeds y-outer faces f = j + 3 - m
      double g = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const int f = j + 3 - m;
        if (f >= J0 && f <= J1 + 1) {
          const int64_t of = o + (int64_t)(3 - m) * d.sj;
          g += L.s(kAyo + m)[of] * (0.5 * L.uy[of] * gfy[of]);
        }
      }
      L.s(kLj)[o] = g;
    }
  });
  // cotangents of the inner faces
  launch_box(s, d, wide, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    if (j >= J0 && j <= J1 + 1) {  // y-inner face (i, j)
      double g = (i >= I0 && i <= I1) ? 0.5 * L.uy[o] * gfy[o] : 0.0;
      const double ayf = L.ay[o];
      if (j <= J1) {
        const double a = L.area[at2(d, i, j, t)];
        g += ayf * L.s(kLi)[o] / (a + ayf - L.ay[o + d.sj]);
      }
      if (j - 1 >= J0) {
        const double a = L.area[at2(d, i, j - 1, t)];
        g -= ayf * L.s(kLi)[o - d.sj] / (a + L.ay[o - d.sj] - ayf);
      }
      L.s(kLy2)[o] = g;
    }
    if (i >= I0 && i <= I1 + 1) {  // x-inner face (i, j)
      double g = (j >= J0 && j <= J1) ? 0.5 * L.ux[o] * gfx[o] : 0.0;
      const double axf = L.ax[o];
      if (i <= I1) {
        const double a = L.area[at2(d, i, j, t)];
        g += axf * L.s(kLj)[o] / (a + axf - L.ax[o + 1]);
      }
      if (i - 1 >= I0) {
        const double a = L.area[at2(d, i - 1, j, t)];
        g -= axf * L.s(kLj)[o - 1] / (a + L.ax[o - 1] - axf);
      }
      L.s(kLx2)[o] = g;
    }
  });
  // cotangents of q_y and q_x over the whole storage
  launch_box(s, d, Box{0, d.ni - 1, 0, d.nj - 1}, 0, d.nk,
             [=] __device__(int i, int j, int k, int t) {
               const int64_t o = at3(d, i, j, k, t);
               const bool iw = i >= I0 - 3 && i <= I1 + 3, jw = j >= J0 - 3 && j <= J1 + 3;
               double gy = 0.0, gx = 0.0;
               if (iw && jw) {
                 const double a = L.area[at2(d, i, j, t)];
                 if (j >= J0 && j <= J1) gy = L.s(kLi)[o] * a / (a + L.ay[o] - L.ay[o + d.sj]);
#pragma unroll
                 for (int m = 0; m < 6; ++m) {  // y-inner faces f = j + 3 - m
                   const int f = j + 3 - m;
                   if (f >= J0 && f <= J1 + 1) {
                     const int64_t of = o + (int64_t)(3 - m) * d.sj;
                     gy += L.s(kAy + m)[of] * L.s(kLy2)[of];
                   }
                 }
                 if (i >= I0 && i <= I1) gx = L.s(kLj)[o] * a / (a + L.ax[o] - L.ax[o + 1]);
#pragma unroll
                 for (int m = 0; m < 6; ++m) {  // x-inner faces f = i + 3 - m
                   const int f = i + 3 - m;
                   if (f >= I0 && f <= I1 + 1) {
                     const int64_t of = o + (3 - m);
                     gx += L.s(kAx + m)[of] * L.s(kLx2)[of];
                   }
                 }
               }
               gqy[o] = gy;
               gqx[o] = gx;
             });
}

int setup(const fv3_domain* dom, const char* who, int hord, const double* q_x, const double* q_y,
          const double* crx, const double* cry, const double* xfx, const double* yfx,
          const double* xmf, const double* ymf, const double* area, const double* dxa,
          const double* dya, double* work, int64_t work_slab, Lin* L) {
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(q_x && q_y && crx && cry && xfx && yfx && area && dxa && dya && work, FV3_E_ARG,
              "%s: required pointer is NULL", who);
  FV3_REQUIRE(hord == 5 || hord == 6 || hord == 7 || hord == 8, FV3_E_UNSUPPORTED,
              "%s: hord=%d not supported (5, 6, 7, 8)", who, hord);
  FV3_REQUIRE(dom->isc >= 3 && dom->jsc >= 3 && dom->iec + 3 <= dom->ni - 1 &&
                  dom->jec + 3 <= dom->nj - 1,
              FV3_E_ARG, "%s: the compute domain needs a 3-cell halo inside the storage", who);
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + (int64_t)dom->nk * dom->sk,
              FV3_E_ARG, "%s: work_slab too small", who);
  L->d = to_dom(dom);
  L->hord = hord;
  L->qx = q_x;
  L->qy = q_y;
  L->crx = crx;
  L->cry = cry;
  L->ax = xfx;
  L->ay = yfx;
  L->ux = xmf ? xmf : xfx;
  L->uy = ymf ? ymf : yfx;
  L->area = area;
  L->dxa = dxa;
  L->dya = dya;
  L->w = work;
  L->slab = work_slab;
  return FV3_OK;
}

}  // namespace
}  // namespace fv3

extern "C" {

int32_t fv3_fvtp2d_workspace_slabs(void) { return fv3::kNumSlabs; }

int32_t fv3_fvtp2d_tlm_f64(const fv3_domain* d, int32_t hord, const double* q_x,
                           const double* q_y, const double* crx, const double* cry,
                           const double* x_area_flux, const double* y_area_flux,
                           const double* x_mass_flux, const double* y_mass_flux,
                           const double* area, const double* dxa, const double* dya,
                           const double* dq_x, const double* dq_y, double* dq_x_flux,
                           double* dq_y_flux, double* work, int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_fvtp2d_tlm";
  Lin L;
  int rc = setup(d, who, hord, q_x, q_y, crx, cry, x_area_flux, y_area_flux, x_mass_flux,
                 y_mass_flux, area, dxa, dya, work, work_slab, &L);
  if (rc) return rc;
  FV3_REQUIRE(dq_x && dq_y && dq_x_flux && dq_y_flux, FV3_E_ARG, "%s: NULL perturbation", who);
  if (d->nk == 0 || d->ntiles == 0) return FV3_OK;
  hipStream_t s = as_stream(stream);
  linearize(s, L);
  launch_tlm(s, L, dq_x, dq_y, dq_x_flux, dq_y_flux);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_fvtp2d_adjoint_f64(const fv3_domain* d, int32_t hord, const double* q_x,
                               const double* q_y, const double* crx, const double* cry,
                               const double* x_area_flux, const double* y_area_flux,
                               const double* x_mass_flux, const double* y_mass_flux,
                               const double* area, const double* dxa, const double* dya,
                               const double* g_x_flux, const double* g_y_flux, double* g_q_x,
                               double* g_q_y, double* work, int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_fvtp2d_adjoint";
  Lin L;
  int rc = setup(d, who, hord, q_x, q_y, crx, cry, x_area_flux, y_area_flux, x_mass_flux,
                 y_mass_flux, area, dxa, dya, work, work_slab, &L);
  if (rc) return rc;
  FV3_REQUIRE(g_x_flux && g_y_flux && g_q_x && g_q_y, FV3_E_ARG, "%s: NULL cotangent", who);
  if (d->nk == 0 || d->ntiles == 0) return FV3_OK;
  hipStream_t s = as_stream(stream);
  linearize(s, L);
  launch_adjoint(s, L, g_x_flux, g_y_flux, g_q_x, g_q_y);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}
