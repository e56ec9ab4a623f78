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
// part repeats the double arithmetic exactly.  The TLM stores the partials and
// the intermediate fields in workspace slabs and applies the chain forward
// (deterministic).  The adjoint keeps no partials: each face evaluates them
// and adds its contributions into the cotangent fields (see
// launch_adjoint_fused).
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

// Adjoint without stored partials.  (1) the forward values q_i and q_j are
// recomputed (plain double faces) and the accumulators zeroed; (2) every
// outer face evaluates its six partials on Dual<6> and adds its cotangent
// times each partial into the cotangents of q_i / q_j; (3) every inner face
// forms its cotangent (its own flux share plus the transposed cross update,
// gathered from the completed q_i / q_j cotangents), evaluates its partials
// and adds into g_qy / g_qx, and each cell adds the direct cross-update
// term.  The adds are hardware FP64 atomics (at most 13 per cell), so the
// summation order, and the last bit of the result, may differ between runs.
template <int ORD>
void launch_adjoint_fused(hipStream_t s, const Lin& L, const double* gfx, const double* gfy,
                          double* gqx, double* gqy) {
  const Dom d = L.d;
  const int I0 = d.isc, I1 = d.iec, J0 = d.jsc, J1 = d.jec;
  const Box wide{I0 - 3, I1 + 3, J0 - 3, J1 + 3};
  launch_box(s, d, Box{0, d.ni - 1, 0, d.nj - 1}, 0, d.nk,
             [=] __device__(int i, int j, int k, int t) {
               const int64_t o = at3(d, i, j, k, t);
               gqx[o] = 0.0;
               gqy[o] = 0.0;
               L.s(kLi)[o] = 0.0;
               L.s(kLj)[o] = 0.0;
               if (i < I0 - 3 || i > I1 + 3 || j < J0 - 3 || j > J1 + 3) return;
               const int64_t o2 = at2(d, i, j, t);
               const double a = L.area[o2];
               if (j >= J0 && j <= J1) {  // q_i from the y-inner faces j, j+1
                 double f[2];
#pragma unroll
                 for (int u = 0; u < 2; ++u) {
                   const int64_t of = o + u * d.sj;
                   double q6[6];
#pragma unroll
                   for (int m = 0; m < 6; ++m) q6[m] = L.qy[of + (m - 3) * d.sj];
                   DxRow<double> dr{L.dya + o2 + (int64_t)(u - 3) * d.sj2, d.sj2};
                   f[u] = ppm_face<double, ORD>(q6, L.cry[of], j + u, d.js, d.je, dr);
                 }
                 const double g0 = L.ay[o], g1 = L.ay[o + d.sj];
                 L.s(kQi)[o] = (L.qy[o] * a + g0 * f[0] - g1 * f[1]) / (a + g0 - g1);
               }
               if (i >= I0 && i <= I1) {  // q_j from the x-inner faces i, i+1
                 double f[2];
#pragma unroll
                 for (int u = 0; u < 2; ++u) {
                   const int64_t of = o + u;
                   double q6[6];
#pragma unroll
                   for (int m = 0; m < 6; ++m) q6[m] = L.qx[of + (m - 3)];
                   DxRow<double> dr{L.dxa + o2 + (u - 3), 1};
                   f[u] = ppm_face<double, ORD>(q6, L.crx[of], i + u, d.is, d.ie, dr);
                 }
                 const double f0 = L.ax[o], f1 = L.ax[o + 1];
                 L.s(kQj)[o] = (L.qx[o] * a + f0 * f[0] - f1 * f[1]) / (a + f0 - f1);
               }
             });
  // outer faces -> cotangents of q_i, q_j
  launch_box(s, d, Box{I0, I1 + 1, J0, J1 + 1}, 0, d.nk,
             [=] __device__(int i, int j, int k, int t) {
               const int64_t o = at3(d, i, j, k, t);
               const int64_t o2 = at2(d, i, j, t);
               if (j <= J1) {
                 const double g = 0.5 * L.ux[o] * gfx[o];
                 const D6 v = face_dual<ORD>(L.s(kQi) + o, 1, L.crx[o], i, d.is, d.ie,
                                             L.dxa + o2, 1);
#pragma unroll
                 for (int m = 0; m < 6; ++m) unsafeAtomicAdd(L.s(kLi) + o + (m - 3), g * v.d[m]);
               }
               if (i <= I1) {
                 const double g = 0.5 * L.uy[o] * gfy[o];
                 const D6 v = face_dual<ORD>(L.s(kQj) + o, d.sj, L.cry[o], j, d.js, d.je,
                                             L.dya + o2, d.sj2);
#pragma unroll
                 for (int m = 0; m < 6; ++m)
                   unsafeAtomicAdd(L.s(kLj) + o + (int64_t)(m - 3) * d.sj, g * v.d[m]);
               }
             });
  // inner faces and the cross updates -> g_qy, g_qx
  launch_box(s, d, wide, 0, d.nk, [=] __device__(int i, int j, int k, int t) {
    const int64_t o = at3(d, i, j, k, t);
    const int64_t o2 = at2(d, i, j, t);
    const double a = L.area[o2];
    if (j >= J0 && j <= J1) {  // direct term of q_i(i, j) on q_y(i, j)
      unsafeAtomicAdd(gqy + o, L.s(kLi)[o] * a / (a + L.ay[o] - L.ay[o + d.sj]));
    }
    if (i >= I0 && i <= I1) {  // direct term of q_j(i, j) on q_x(i, j)
      unsafeAtomicAdd(gqx + o, L.s(kLj)[o] * a / (a + L.ax[o] - L.ax[o + 1]));
    }
    if (j >= J0 && j <= J1 + 1) {  // y-inner face (i, j)
      double g = (i >= I0 && i <= I1) ? 0.5 * L.uy[o] * gfy[o] : 0.0;
      const double ayf = L.ay[o];
      if (j <= J1) g += ayf * L.s(kLi)[o] / (a + ayf - L.ay[o + d.sj]);
      if (j - 1 >= J0) {
        const double am = L.area[o2 - d.sj2];
        g -= ayf * L.s(kLi)[o - d.sj] / (am + L.ay[o - d.sj] - ayf);
      }
      const D6 v = face_dual<ORD>(L.qy + o, d.sj, L.cry[o], j, d.js, d.je, L.dya + o2, d.sj2);
#pragma unroll
      for (int m = 0; m < 6; ++m) unsafeAtomicAdd(gqy + o + (int64_t)(m - 3) * d.sj, g * v.d[m]);
    }
    if (i >= I0 && i <= I1 + 1) {  // x-inner face (i, j)
      double g = (j >= J0 && j <= J1) ? 0.5 * L.ux[o] * gfx[o] : 0.0;
      const double axf = L.ax[o];
      if (i <= I1) g += axf * L.s(kLj)[o] / (a + axf - L.ax[o + 1]);
      if (i - 1 >= I0) {
        const double am = L.area[o2 - 1];
        g -= axf * L.s(kLj)[o - 1] / (am + L.ax[o - 1] - axf);
      }
      const D6 v = face_dual<ORD>(L.qx + o, 1, L.crx[o], i, d.is, d.ie, L.dxa + o2, 1);
#pragma unroll
      for (int m = 0; m < 6; ++m) unsafeAtomicAdd(gqx + o + (m - 3), g * v.d[m]);
    }
  });
}

void adjoint_fused(hipStream_t s, const Lin& L, const double* gfx, const double* gfy,
                   double* gqx, double* gqy) {
  switch (L.hord) {
    case 5: launch_adjoint_fused<5>(s, L, gfx, gfy, gqx, gqy); break;
    case 6: launch_adjoint_fused<6>(s, L, gfx, gfy, gqx, gqy); break;
    case 7: launch_adjoint_fused<7>(s, L, gfx, gfy, gqx, gqy); break;
    default: launch_adjoint_fused<8>(s, L, gfx, gfy, gqx, gqy); break;
  }
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
  adjoint_fused(s, L, g_x_flux, g_y_flux, g_q_x, g_q_y);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}
