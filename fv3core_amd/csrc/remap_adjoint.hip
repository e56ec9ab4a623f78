// Tangent-linear model and adjoint of MapSingle (map_scalar) with kord 9
// with respect to the remapped field, at the branch decisions of the forward
// (fv3core/stencils/map_single.py:69-173 with remap_profile.py:498-638; the
// reference has no derivative code).  The source and target interfaces
// pe1 / pe2, qs and qmin are held fixed.
//
// TLM: the forward column kernel itself (remap_profile.hpp) instantiated on
// Dual<1>, whose value part repeats the double arithmetic; the fields are
// packed into (value, perturbation) pairs and unpacked afterwards.
//
// Adjoint, one thread per column, after recomputing the forward edge values:
//   (4) lagrangian_contributions is linear in the layer means a4_1 and the
//       parabola coefficients a4_2..a4_4 with weights from pe1 / pe2 only:
//       its transpose re-walks the source layers and scatters the output
//       cotangents;
//   (3) set_interpolation_coefficients acts level by level on (a4_1 of the
//       level and its two neighbours each side, the two clamped edge values):
//       its local Jacobian comes from running the same statements on Dual<7>;
//   (2) apply_constraints clamps each edge value between the neighbouring
//       means: local Jacobian on Dual<3>;
//   (1) set_initial_vals is a tridiagonal solve for the edge values, linear
//       in a4_1 with coefficients from pe1: its transpose reverses the back
//       substitution and then the forward elimination.
#include "common.hpp"
#include "dual.hpp"
#include "remap_profile.hpp"
#include "timing.hpp"

namespace fv3 {
namespace {

// Edge value kk (1 <= kk <= km-1) after apply_constraints
// (remap_profile.py:210-280): qn clamped between the means of the layers
// on either side; gm1 / gp1 are the neighbouring mean differences that
// select the clamp (values only).
template <typename T>
__device__ __forceinline__ T edge_clamp(int kk, int km, int iv, T qn, T a0, T ak, double gm1,
                                        double gp1) {
  const T tmp = a0 > ak ? a0 : ak;
  const T tmp2 = a0 < ak ? a0 : ak;
  if (kk == 1 || kk == km - 1) {
    if (qn >= tmp) qn = tmp;
    if (qn <= tmp2) qn = tmp2;
  } else {
    if ((gm1 * gp1) > 0.0) {
      if (qn >= tmp) qn = tmp;
      if (qn <= tmp2) qn = tmp2;
    } else if (gm1 > 0.0) {
      if (qn <= tmp2) qn = tmp2;
    } else {
      if (qn >= tmp) qn = tmp;
      if (iv == 0) {
        if (qn < T(0)) qn = T(0);
      }
    }
  }
  return qn;
}

// set_interpolation_coefficients (remap_profile.py:283-495) for kord 9 at
// level k: a1 = a4_1(k), a2 / a3 its clamped edge values, am2..ap2 the means
// of levels k-2, k-1, k+1, k+2 (used at 2 <= k < km-2), em* the extremum
// flags of levels k-1, k, k+1.  Returns a4_2, a4_3, a4_4.
template <typename T>
__device__ __forceinline__ void level_coeffs9(int k, int km, int iv, double qmin, T a1, T a2, T a3,
                                              T am2, T am1, T ap1, T ap2, bool em_m1, bool em,
                                              bool em_p1, T& o2, T& o3, T& o4) {
  T a4 = T(0);
  if (k < 2) {
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
      remap_constraint(a1, a2, a3, a4, em);
    }
  } else if (k < km - 2) {
    const T gk = a1 - am1, gp1 = ap1 - a1, gp2 = ap2 - ap1, gm1 = am1 - am2;
    const T pmp_1 = a1 - T(2) * gp1;
    const T lac_1 = pmp_1 + T(1.5) * gp2;
    const T pmp_2 = a1 + T(2) * gk;
    const T lac_2 = pmp_2 - T(1.5) * gm1;
    if ((em && em_m1) || (em && em_p1) || (em && (T(qmin) > T(0) && a1 < T(qmin)))) {
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
    if (iv == 0) posdef_iv0(a1, a2, a3, a4);
  } else {
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
      remap_constraint(a1, a2, a3, a4, em);
    else
      posdef_iv1(a1, a2, a3, a4);
  }
  o2 = a2;
  o3 = a3;
  o4 = a4;
}

// workspace slabs of the adjoint: 0 edge values (km+1), 1 gam (km+1),
// 2 clamped edge values (km+1), 3 extremum flags (km), 4 edge cotangents
// (km+1), 5..7 cotangents of a4_2..a4_4 (km)
constexpr int kAdjSlabs = 8;

__global__ void __launch_bounds__(256)
    map_single_adjoint_kernel(Dom d, RemapArgs a, const double* __restrict__ q1,
                              const double* __restrict__ pe1, const double* __restrict__ pe2,
                              const double* __restrict__ qs, const double* __restrict__ gout,
                              double* __restrict__ gin, double* __restrict__ work) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)a.ni_box * a.nj_box) return;
  const int i = a.i1 + (int)(n % a.ni_box);
  const int j = a.j1 + (int)(n / a.ni_box);
  const int t = blockIdx.y;
  const int64_t c = (int64_t)t * d.st + (int64_t)j * d.sj + i;
  const int64_t sk = d.sk;
  const int km = a.km, iv = a.iv;
  Col<const double> A1{q1 + c, sk}, P1{pe1 + c, sk}, P2{pe2 + c, sk}, GO{gout + c, sk};
  Col<double> GA{gin + c, sk};
  Col<double> W0{work + c, sk}, W1{work + a.slab + c, sk}, E{work + 2 * a.slab + c, sk},
      FL{work + 3 * a.slab + c, sk}, GE{work + 4 * a.slab + c, sk},
      G2{work + 5 * a.slab + c, sk}, G3{work + 6 * a.slab + c, sk},
      G4{work + 7 * a.slab + c, sk};
  const double qsv = qs ? qs[(int64_t)t * d.st2 + (int64_t)j * d.sj2 + i] : 0.0;
  auto dp = [&](int k) -> double { return P1[k + 1] - P1[k]; };
  auto gc = [&](int k) -> double { return A1[k] - A1[k - 1]; };

  // ---- forward: edge values (set_initial_vals), as the forward kernel
  if (iv == -2) {
    W0[0] = 1.5 * A1[0];
    W1[1] = 0.5;
    {
      const double gr = dp(0) / dp(1);
      const double bet = 2.0 + gr + gr - W1[1];
      W0[1] = (3.0 * (A1[0] + A1[1]) - W0[0]) / bet;
    }
    for (int k = 2; k < km; ++k) {
      const double ogr = dp(k - 2) / dp(k - 1);
      const double obet = 2.0 + ogr + ogr - W1[k - 1];
      const double g = ogr / obet;
      W1[k] = g;
      const double gr = dp(k - 1) / dp(k);
      if (k < km - 1) {
        const double bet = 2.0 + gr + gr - g;
        W0[k] = (3.0 * (A1[k - 1] + A1[k]) - W0[k - 1]) / bet;
      } else {
        W0[k] = (3.0 * (A1[k - 1] + A1[k]) - gr * qsv - W0[k - 1]) / (2.0 + gr + gr - g);
      }
    }
    W0[km] = qsv;
    for (int k = km - 2; k >= 0; --k) W0[k] = W0[k] - W1[k + 1] * W0[k + 1];
  } else {
    {
      const double gr = dp(1) / dp(0);
      const double bet = gr * (gr + 0.5);
      W0[0] = ((gr + gr) * (gr + 1.0) * A1[0] + A1[1]) / bet;
      W1[0] = (1.0 + gr * (gr + 1.5)) / bet;
    }
    for (int k = 1; k < km; ++k) {
      const double d4 = dp(k - 1) / dp(k);
      const double bet = 2.0 + d4 + d4 - W1[k - 1];
      W0[k] = (3.0 * (A1[k - 1] + d4 * A1[k]) - W0[k - 1]) / bet;
      W1[k] = d4 / bet;
    }
    {
      const double d4 = dp(km - 2) / dp(km - 1);
      const double a_bot = 1.0 + d4 * (d4 + 1.5);
      W0[km] = (2.0 * d4 * (d4 + 1.0) * A1[km - 1] + A1[km - 2] - a_bot * W0[km - 1]) /
               (d4 * (d4 + 0.5) - a_bot * W1[km - 1]);
    }
    for (int k = km - 1; k >= 0; --k) W0[k] = W0[k] - W1[k] * W0[k + 1];
  }
  // ---- forward: clamped edges and extremum flags (apply_constraints)
  E[0] = W0[0];
  E[km] = W0[km];
  for (int kk = 1; kk <= km - 1; ++kk) {
    const double gm1 = (kk >= 2) ? gc(kk - 1) : 0.0;
    const double gp1 = (kk + 1 <= km - 1) ? gc(kk + 1) : 0.0;
    E[kk] = edge_clamp<double>(kk, km, iv, W0[kk], A1[kk - 1], A1[kk], gm1, gp1);
  }
  for (int k = 0; k < km; ++k) {
    bool extm;
    if (k == 0 || k == km - 1) {
      extm = (E[k] - A1[k]) * (E[k + 1] - A1[k]) > 0.0;
    } else {
      extm = gc(k) * gc(k + 1) < 0.0;
    }
    FL[k] = extm ? 1.0 : 0.0;
    G2[k] = 0.0;
    G3[k] = 0.0;
    G4[k] = 0.0;
    GA[k] = 0.0;
  }
  for (int k = 0; k <= km; ++k) GE[k] = 0.0;

  // ---- (4) transpose of lagrangian_contributions
  int L = 0;
  for (int k = 0; k < km; ++k) {
    const double g = GO[k];
    const double p2k = P2[k], p2n = P2[k + 1];
    const double pL = P1[L];
    const double dpl = P1[L + 1] - pL;
    const double pl = (p2k - pL) / dpl;
    if (p2n <= P1[L + 1]) {
      const double pr = (p2n - pL) / dpl;
      // out = b2 + 0.5 (b4 + b3 - b2)(pr + pl) - b4/3 (pr (pr + pl) + pl^2)
      G2[L] += g * (1.0 - 0.5 * (pr + pl));
      G3[L] += g * (0.5 * (pr + pl));
      G4[L] += g * (0.5 * (pr + pl) - (pr * (pr + pl) + pl * pl) / 3.0);
    } else {
      const double w = g / (p2n - p2k);
      const double wf = (P1[L + 1] - p2k) * w;
      G2[L] += wf * (1.0 - 0.5 * (1.0 + pl));
      G3[L] += wf * (0.5 * (1.0 + pl));
      G4[L] += wf * (0.5 * (1.0 + pl) - (1.0 + pl * (1.0 + pl)) / 3.0);
      L += 1;
      if (L > km - 1) L = km - 1;
      while (L < km - 1 && P1[L + 1] < p2n) {
        GA[L] += (P1[L + 1] - P1[L]) * w;
        L += 1;
      }
      const double dpx = p2n - P1[L];
      const double esl = dpx / (P1[L + 1] - P1[L]);
      const double u = dpx * w;
      // dpx (a2 + 0.5 esl (a3 - a2 + a4 (1 - 2/3 esl)))
      G2[L] += u * (1.0 - 0.5 * esl);
      G3[L] += u * (0.5 * esl);
      G4[L] += u * (0.5 * esl * (1.0 - (2.0 / 3.0) * esl));
    }
  }

  // ---- (3) transpose of set_interpolation_coefficients, level by level
  using D7 = Dual<7>;
  for (int k = 0; k < km; ++k) {
    const bool mid = k >= 2 && k < km - 2;
    const D7 a1 = D7::seed(A1[k], 0);
    const D7 a2 = D7::seed(E[k], 1);
    const D7 a3 = D7::seed(E[k + 1], 2);
    const D7 am2 = mid ? D7::seed(A1[k - 2], 3) : D7(0.0);
    const D7 am1 = mid ? D7::seed(A1[k - 1], 4) : D7(0.0);
    const D7 ap1 = mid ? D7::seed(A1[k + 1], 5) : D7(0.0);
    const D7 ap2 = mid ? D7::seed(A1[k + 2], 6) : D7(0.0);
    const bool em = FL[k] != 0.0;
    const bool em_m1 = k >= 1 && FL[k - 1] != 0.0;
    const bool em_p1 = k + 1 < km && FL[k + 1] != 0.0;
    D7 o2, o3, o4;
    level_coeffs9<D7>(k, km, iv, a.qmin, a1, a2, a3, am2, am1, ap1, ap2, em_m1, em, em_p1, o2,
                      o3, o4);
    const double g2 = G2[k], g3 = G3[k], g4 = G4[k];
    double gat[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) gat[m] = g2 * o2.d[m] + g3 * o3.d[m] + g4 * o4.d[m];
    GA[k] += gat[0];
    GE[k] += gat[1];
    GE[k + 1] += gat[2];
    if (mid) {
      GA[k - 2] += gat[3];
      GA[k - 1] += gat[4];
      GA[k + 1] += gat[5];
      GA[k + 2] += gat[6];
    }
  }

  // ---- (2) transpose of the edge clamps: GE becomes the edge-value cotangent
  using D3 = Dual<3>;
  for (int kk = 1; kk <= km - 1; ++kk) {
    const double gm1 = (kk >= 2) ? gc(kk - 1) : 0.0;
    const double gp1 = (kk + 1 <= km - 1) ? gc(kk + 1) : 0.0;
    const D3 e = edge_clamp<D3>(kk, km, iv, D3::seed(W0[kk], 0), D3::seed(A1[kk - 1], 1),
                                D3::seed(A1[kk], 2), gm1, gp1);
    const double g = GE[kk];
    GE[kk] = g * e.d[0];
    GA[kk - 1] += g * e.d[1];
    GA[kk] += g * e.d[2];
  }

  // ---- (1) transpose of the tridiagonal solve for the edge values
  if (iv == -2) {
    for (int k = 0; k <= km - 2; ++k) GE[k + 1] = GE[k + 1] - W1[k + 1] * GE[k];
    // GE[km] belongs to qs, held fixed
    for (int k = km - 1; k >= 2; --k) {
      const double gr = dp(k - 1) / dp(k);
      const double bet = 2.0 + gr + gr - W1[k];
      const double h = GE[k] / bet;
      GA[k - 1] += 3.0 * h;
      GA[k] += 3.0 * h;
      GE[k - 1] -= h;
    }
    {
      const double gr = dp(0) / dp(1);
      const double bet = 2.0 + gr + gr - W1[1];
      const double h = GE[1] / bet;
      GA[0] += 3.0 * h;
      GA[1] += 3.0 * h;
      GE[0] -= h;
    }
    GA[0] += 1.5 * GE[0];
  } else {
    for (int k = 0; k <= km - 1; ++k) GE[k + 1] = GE[k + 1] - W1[k] * GE[k];
    {
      const double d4 = dp(km - 2) / dp(km - 1);
      const double a_bot = 1.0 + d4 * (d4 + 1.5);
      const double den = d4 * (d4 + 0.5) - a_bot * W1[km - 1];
      const double h = GE[km] / den;
      GA[km - 1] += 2.0 * d4 * (d4 + 1.0) * h;
      GA[km - 2] += h;
      GE[km - 1] -= a_bot * h;
    }
    for (int k = km - 1; k >= 1; --k) {
      const double d4 = dp(k - 1) / dp(k);
      const double bet = 2.0 + d4 + d4 - W1[k - 1];
      const double h = GE[k] / bet;
      GA[k - 1] += 3.0 * h;
      GA[k] += 3.0 * d4 * h;
      GE[k - 1] -= h;
    }
    {
      const double gr = dp(1) / dp(0);
      const double bet = gr * (gr + 0.5);
      const double h = GE[0] / bet;
      GA[0] += (gr + gr) * (gr + 1.0) * h;
      GA[1] += h;
    }
  }
}

// (value, perturbation) packing for the Dual<1> TLM
__global__ void __launch_bounds__(256)
    pack_kernel(int64_t n, const double* __restrict__ v, const double* __restrict__ dv,
                Dual<1>* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  Dual<1> x(v[e]);
  x.d[0] = dv ? dv[e] : 0.0;
  out[e] = x;
}

__global__ void __launch_bounds__(256)
    unpack_kernel(int64_t n, const Dual<1>* __restrict__ in, double* __restrict__ v,
                  double* __restrict__ dv) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  v[e] = in[e].v;
  dv[e] = in[e].d[0];
}

int check_remap_args(const fv3_domain* dom, const char* who, int kord, int i1, int i2, int j1,
                     int j2, int64_t work_slab) {
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(kord == 9 || kord == -9, FV3_E_UNSUPPORTED,
              "%s: only kord 9 has a derivative here (got %d)", who, kord);
  FV3_REQUIRE(dom->nk >= 6, FV3_E_ARG, "%s: need at least 6 layers (got %d)", who, dom->nk);
  FV3_REQUIRE(i1 >= 0 && j1 >= 0 && i2 < dom->ni && j2 < dom->nj && i2 >= i1 - 1 &&
                  j2 >= j1 - 1,
              FV3_E_ARG, "%s: box outside the storage", who);
  FV3_REQUIRE(work_slab >= (int64_t)(dom->ntiles - 1) * dom->st + dom->sk * (dom->nk + 1),
              FV3_E_ARG, "%s: workspace slab too small", who);
  return FV3_OK;
}

}  // namespace
}  // namespace fv3

extern "C" {

int32_t fv3_map_single_tlm_workspace_slabs(void) { return 22; }
int32_t fv3_map_single_adjoint_workspace_slabs(void) { return fv3::kAdjSlabs; }

int32_t fv3_map_single_tlm_f64(const fv3_domain* dom, int32_t kord, int32_t mode, int32_t i1,
                               int32_t i2, int32_t j1, int32_t j2, double* q1, double* dq1,
                               const double* pe1, const double* pe2, const double* qs,
                               double qmin, double* work, int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_map_single_tlm";
  int rc = check_remap_args(dom, who, kord, i1, i2, j1, j2, work_slab);
  if (rc) return rc;
  FV3_REQUIRE(q1 && dq1 && pe1 && pe2 && work, FV3_E_ARG, "%s: NULL pointer", who);
  const int nib = i2 - i1 + 1, njb = j2 - j1 + 1;
  if (nib <= 0 || njb <= 0 || dom->ntiles == 0) return FV3_OK;
  hipStream_t s = as_stream(stream);
  // Dual<1> arrays: each slab of Dual elements spans two double slabs
  Dual<1>* W = reinterpret_cast<Dual<1>*>(work);
  Dual<1>* kwork = W;                 // 7 kernel slabs (14 double slabs)
  Dual<1>* dq = W + 7 * work_slab;    // q1 (2 double slabs)
  Dual<1>* dp1 = W + 8 * work_slab;   // pe1
  Dual<1>* dp2 = W + 9 * work_slab;   // pe2
  Dual<1>* dqs = W + 10 * work_slab;  // qs (IJ, fits in one slab)
  // exact extents the kernel addresses: q1 on km layers, pe1 / pe2 on km + 1
  // interfaces, qs on one IJ plane per tile
  const int64_t tail = (int64_t)(dom->nj - 1) * dom->sj + dom->ni;
  const int64_t base = (int64_t)(dom->ntiles - 1) * dom->st;
  const int64_t nq = base + (int64_t)(dom->nk - 1) * dom->sk + tail;
  const int64_t npe = base + (int64_t)dom->nk * dom->sk + tail;
  const int64_t n2 = (int64_t)(dom->ntiles - 1) * dom->st2 + (int64_t)(dom->nj - 1) * dom->sj2 +
                     dom->ni;
  FV3_REQUIRE(npe <= work_slab && n2 <= work_slab, FV3_E_ARG,
              "%s: the fields do not fit a workspace slab", who);
  auto blocks = [](int64_t n) { return dim3((unsigned)((n + 255) / 256)); };
  hipLaunchKernelGGL(pack_kernel, blocks(nq), dim3(256), 0, s, nq, q1, dq1, dq);
  hipLaunchKernelGGL(pack_kernel, blocks(npe), dim3(256), 0, s, npe, pe1, nullptr, dp1);
  hipLaunchKernelGGL(pack_kernel, blocks(npe), dim3(256), 0, s, npe, pe2, nullptr, dp2);
  if (qs) {
    hipLaunchKernelGGL(pack_kernel, blocks(n2), dim3(256), 0, s, n2, qs, nullptr, dqs);
  }
  RemapArgs a{9, mode, dom->nk, i1, j1, nib, njb, qmin, work_slab};
  Dom d = to_dom(dom);
  dim3 grid((unsigned)(((int64_t)nib * njb + 255) / 256), d.ntiles);
  hipLaunchKernelGGL((map_single_kernel<Dual<1>>), grid, dim3(256), 0, s, d, a, dq, dp1, dp2,
                     qs ? dqs : nullptr, kwork);
  hipLaunchKernelGGL(unpack_kernel, blocks(nq), dim3(256), 0, s, nq, dq, q1, dq1);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_map_single_adjoint_f64(const fv3_domain* dom, int32_t kord, int32_t mode, int32_t i1,
                                   int32_t i2, int32_t j1, int32_t j2, const double* q1,
                                   const double* pe1, const double* pe2, const double* qs,
                                   double qmin, const double* g_out, double* g_q1,
                                   double* work, int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_map_single_adjoint";
  int rc = check_remap_args(dom, who, kord, i1, i2, j1, j2, work_slab);
  if (rc) return rc;
  FV3_REQUIRE(q1 && pe1 && pe2 && g_out && g_q1 && work, FV3_E_ARG, "%s: NULL pointer", who);
  const int nib = i2 - i1 + 1, njb = j2 - j1 + 1;
  if (nib <= 0 || njb <= 0 || dom->ntiles == 0) return FV3_OK;
  RemapArgs a{9, mode, dom->nk, i1, j1, nib, njb, qmin, work_slab};
  Dom d = to_dom(dom);
  dim3 grid((unsigned)(((int64_t)nib * njb + 255) / 256), d.ntiles);
  hipLaunchKernelGGL(map_single_adjoint_kernel, grid, dim3(256), 0, as_stream(stream), d, a, q1,
                     pe1, pe2, qs, g_out, g_q1, work);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}
