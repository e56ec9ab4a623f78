// Semi-implicit vertical sound-wave solvers: sim1_solver, RiemannSolverC,
// RiemannSolver3.
//
// Reference: fv3core/stencils/sim1_solver.py:11-101 (two tridiagonal
// solves, for the pressure perturbation and for w, plus the dz update),
// riem_solver_c.py:13-165 and riem_solver3.py:24-242 (precompute / sim1 /
// finalize stencils).  Here each solver is one kernel: one thread owns one
// column and runs every FORWARD/BACKWARD sweep of precompute, sim1 and
// finalize in sequence, always in FP64.  Lanes of a wave own neighbouring
// i, so each k-indexed access is coalesced.  Column intermediates that a
// later sweep reads back live in a caller-provided workspace of field-shaped
// slabs; cheap ones (gm, peg, peln) are recomputed instead of stored.
#include "common.hpp"
#include "timing.hpp"

namespace fv3 {

constexpr double kRdgas = 287.05;
constexpr double kGrav = 9.80665;
constexpr double kCpAir = 1004.6;

template <typename T>
struct KCol {
  T* p;
  int64_t sk;
  __device__ __forceinline__ T& operator[](int k) const { return p[(int64_t)k * sk]; }
};

// Arrays sim1 works on (sim1_solver.py:126-139): w and dz are in/out, pe is
// out; gm is 1 / (1 - cp3) of the same column (recomputed on the fly).
struct Sim1Cols {
  KCol<double> w, dm, dz, ptr, pm, pe, pem, cp3;
  KCol<double> w1, gam, pp, aa;  // workspace
};

__device__ void sim1_column(const Sim1Cols& c, int nk, double wsr, double dt, double p_fac) {
  const int nz = nk - 1;
  const double t1g = 2.0 * dt * dt;
  const double rdt = 1.0 / dt;
  auto gm = [&](int k) -> double { return 1.0 / (1.0 - c.cp3[k]); };
  // pe = exp(gm*log(-dm/dz*Rd*ptr)) - pm; w1 = w
  for (int k = 0; k < nz; ++k) {
    c.pe[k] = exp(gm(k) * log(-c.dm[k] / c.dz[k] * kRdgas * c.ptr[k])) - c.pm[k];
    c.w1[k] = c.w[k];
  }
  auto g_rat = [&](int k) -> double { return c.dm[k] / c.dm[k + 1]; };
  auto bb = [&](int k) -> double { return k < nz - 1 ? 2.0 * (1.0 + g_rat(k)) : 2.0; };
  auto dd = [&](int k) -> double {
    return k < nz - 1 ? 3.0 * (c.pe[k] + g_rat(k) * c.pe[k + 1]) : 3.0 * c.pe[k];
  };
  // forward elimination for pp (pp[0] = 0, pp[1] = dd[0]/bb[0])
  double bet_prev = bb(0);  // bet[0]
  c.pp[0] = 0.0;
  c.pp[1] = dd(0) / bet_prev;
  for (int k = 1; k < nz; ++k) {
    const double g = g_rat(k - 1) / bet_prev;
    c.gam[k] = g;
    const double bet_k = bb(k) - g;
    // pp[k+1] = (dd[k] - pp[k]) / bet[k]
    c.pp[k + 1] = (dd(k) - c.pp[k]) / bet_k;
    bet_prev = bet_k;
  }
  // back substitution and aa
  for (int k = nz - 1; k >= 1; --k) {
    const double ppk = c.pp[k] - c.gam[k] * c.pp[k + 1];
    c.pp[k] = ppk;
    c.aa[k] = t1g * 0.5 * (gm(k - 1) + gm(k)) / (c.dz[k - 1] + c.dz[k]) * (c.pem[k] + ppk);
  }
  // w tridiagonal (gam reused)
  double bet = c.dm[0] - c.aa[1];
  c.w[0] = (c.dm[0] * c.w1[0] + dt * c.pp[1]) / bet;
  for (int k = 1; k < nz - 1; ++k) {
    const double g = c.aa[k] / bet;
    c.gam[k] = g;
    const double b = c.dm[k] - (c.aa[k] + c.aa[k + 1] + c.aa[k] * g);
    c.w[k] = (c.dm[k] * c.w1[k] + dt * (c.pp[k + 1] - c.pp[k]) - c.aa[k] * c.w[k - 1]) / b;
    bet = b;
  }
  {
    const int k = nz - 1;
    const double p1 = t1g * gm(k) / c.dz[k] * (c.pem[k + 1] + c.pp[k + 1]);
    const double g = c.aa[k] / bet;
    c.gam[k] = g;
    const double b = c.dm[k] - (c.aa[k] + p1 + c.aa[k] * g);
    c.w[k] = (c.dm[k] * c.w1[k] + dt * (c.pp[k + 1] - c.pp[k]) - p1 * wsr - c.aa[k] * c.w[k - 1]) /
             b;
  }
  for (int k = nz - 2; k >= 0; --k) c.w[k] = c.w[k] - c.gam[k + 1] * c.w[k + 1];
  // pe from the change of w
  c.pe[0] = 0.0;
  for (int k = 1; k < nk; ++k)
    c.pe[k] = c.pe[k - 1] + c.dm[k - 1] * (c.w[k - 1] - c.w1[k - 1]) * rdt;
  // p1 (BACKWARD) and the dz update
  double p1_next = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    double p1;
    if (k == nz - 1) {
      p1 = (c.pe[k] + 2.0 * c.pe[k + 1]) * 1.0 / 3.0;
    } else {
      p1 = (c.pe[k] + bb(k) * c.pe[k + 1] + g_rat(k) * c.pe[k + 2]) * 1.0 / 3.0 -
           g_rat(k) * p1_next;
    }
    p1_next = p1;
    const double pm = c.pm[k];
    const double maxp = p_fac * c.dm[k] > p1 + pm ? p_fac * pm : p1 + pm;
    c.dz[k] = -c.dm[k] * kRdgas * c.ptr[k] * exp((c.cp3[k] - 1.0) * log(maxp));
  }
}

struct Box2 {
  int i1, j1, nib, njb;
};

__global__ void __launch_bounds__(256)
    riem_c_kernel(Dom d, Box2 bx, double dt2, double ptop, double p_fac,
                  const double* __restrict__ cappa, const double* __restrict__ hs,
                  const double* __restrict__ ws, const double* __restrict__ ptc,
                  const double* __restrict__ q_con, const double* __restrict__ delpc,
                  double* __restrict__ gz, double* __restrict__ pef,
                  const double* __restrict__ w3, double* __restrict__ work, int64_t slab) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)bx.nib * bx.njb) return;
  const int i = bx.i1 + (int)(n % bx.nib), j = bx.j1 + (int)(n / bx.nib);
  const int t = blockIdx.y;
  const int64_t c = (int64_t)t * d.st + (int64_t)j * d.sj + i;
  const int64_t c2 = (int64_t)t * d.st2 + (int64_t)j * d.sj2 + i;
  const int64_t sk = d.sk;
  const int nk = d.nk, nz = nk - 1;
  auto S = [&](int s) { return KCol<double>{work + s * slab + c, sk}; };
  KCol<double> dm = S(0), w = S(1), dz = S(2), pm = S(3), pem = S(4), pe = S(5);
  KCol<const double> Cap{cappa + c, sk}, Qc{q_con + c, sk}, Dp{delpc + c, sk}, W3{w3 + c, sk};
  KCol<double> Gz{gz + c, sk}, Pef{pef + c, sk};
  // precompute (riem_solver_c.py:13-43)
  double pem_k = ptop, peg_k = ptop;
  for (int k = 0; k < nk; ++k) {
    const double dmk = Dp[k];
    w[k] = W3[k];
    pem[k] = pem_k;
    dm[k] = dmk / kGrav;
    if (k < nz) {
      dz[k] = Gz[k + 1] - Gz[k];
      const double peg_n = peg_k + dmk * (1.0 - Qc[k]);
      pm[k] = (peg_n - peg_k) / log(peg_n / peg_k);
      peg_k = peg_n;
      pem_k = pem_k + dmk;
    }
  }
  Sim1Cols sc{w, dm, dz, KCol<double>{const_cast<double*>(ptc) + c, sk}, pm, pe, pem,
              KCol<double>{const_cast<double*>(cappa) + c, sk}, S(6), S(7), S(8), S(9)};
  (void)Cap;
  sim1_column(sc, nk, ws[c2], dt2, p_fac);
  // finalize (riem_solver_c.py:46-64)
  Pef[0] = ptop;
  for (int k = 1; k < nk; ++k) Pef[k] = pe[k] + pem[k];
  double g = hs[c2];
  Gz[nz] = g;
  for (int k = nz - 1; k >= 0; --k) {
    g = g - dz[k] * kGrav;
    Gz[k] = g;
  }
}

struct Riem3Opts {
  int last_call, use_logp;
  double beta, dt, ptop, p_fac, peln1, ptk, kappa, rgrav;
};

__global__ void __launch_bounds__(256)
    riem3_kernel(Dom d, Box2 bx, Riem3Opts o, const double* __restrict__ cappa,
                 const double* __restrict__ zs, const double* __restrict__ wsd,
                 double* __restrict__ delz, const double* __restrict__ q_con,
                 const double* __restrict__ delp, const double* __restrict__ pt,
                 double* __restrict__ zh, double* __restrict__ pe, double* __restrict__ ppe,
                 double* __restrict__ pk3, double* __restrict__ pk, double* __restrict__ peln,
                 double* __restrict__ w, double* __restrict__ work, int64_t slab) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)bx.nib * bx.njb) return;
  const int i = bx.i1 + (int)(n % bx.nib), j = bx.j1 + (int)(n / bx.nib);
  const int t = blockIdx.y;
  const int64_t c = (int64_t)t * d.st + (int64_t)j * d.sj + i;
  const int64_t c2 = (int64_t)t * d.st2 + (int64_t)j * d.sj2 + i;
  const int64_t sk = d.sk;
  const int nk = d.nk, nz = nk - 1;
  auto S = [&](int s) { return KCol<double>{work + s * slab + c, sk}; };
  KCol<double> dm = S(0), pm = S(1), pem = S(2), pe_init = S(3);
  KCol<const double> Qc{q_con + c, sk}, Dp{delp + c, sk};
  KCol<double> Dz{delz + c, sk}, Zh{zh + c, sk}, Pe{pe + c, sk}, Ppe{ppe + c, sk},
      Pk3{pk3 + c, sk}, Pk{pk + c, sk}, Peln{peln + c, sk}, W{w + c, sk};
  // precompute (riem_solver3.py:24-68)
  double pem_k = o.ptop, peg_k = o.ptop, pelng_k = o.peln1;
  for (int k = 0; k < nk; ++k) {
    const double dmk = Dp[k];
    pe_init[k] = Pe[k];
    dm[k] = dmk * o.rgrav;
    if (k == 0) {
      pem[0] = o.ptop;
      Pk3[0] = o.ptk;
    }
    if (k < nz) {
      const double pem_n = pem_k + dmk;
      const double peln_n = log(pem_n);
      const double peg_n = peg_k + dmk * (1.0 - Qc[k]);
      const double pelng_n = log(peg_n);
      pem[k + 1] = pem_n;
      Pk3[k + 1] = exp(o.kappa * peln_n);
      pm[k] = (peg_n - peg_k) / (pelng_n - pelng_k);
      Dz[k] = Zh[k + 1] - Zh[k];
      pem_k = pem_n;
      peg_k = peg_n;
      pelng_k = pelng_n;
    }
  }
  Sim1Cols sc{W, dm, Dz, KCol<double>{const_cast<double*>(pt) + c, sk}, pm, Pe, pem,
              KCol<double>{const_cast<double*>(cappa) + c, sk}, S(4), S(5), S(6), S(7)};
  sim1_column(sc, nk, wsd[c2], o.dt, o.p_fac);
  // finalize (riem_solver3.py:70-103)
  for (int k = 0; k < nk; ++k) {
    const double peln_run = k == 0 ? o.peln1 : log(pem[k]);
    if (o.use_logp) Pk3[k] = peln_run;
    const double pek = Pe[k];
    Ppe[k] = o.beta < -0.1 ? pek + pem[k] : pek;
    if (o.last_call) {
      Peln[k] = peln_run;
      Pk[k] = Pk3[k];
      Pe[k] = pem[k];
    } else {
      Pe[k] = pe_init[k];
    }
  }
  double z = zs[c2];
  Zh[nz] = z;
  for (int k = nz - 1; k >= 0; --k) {
    z = z - Dz[k];
    Zh[k] = z;
  }
}

static int check_riem(const fv3_domain* dom, const char* who, int i1, int i2, int j1, int j2,
                      int64_t slab, int nslabs) {
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(dom->nk >= 4, FV3_E_ARG, "%s: need at least 4 levels", who);
  FV3_REQUIRE(i1 >= 0 && j1 >= 0 && i2 < dom->ni && j2 < dom->nj, FV3_E_ARG,
              "%s: box outside the storage", who);
  FV3_REQUIRE(slab >= (int64_t)(dom->ntiles - 1) * dom->st + dom->sk * dom->nk, FV3_E_ARG,
              "%s: workspace slab too small", who);
  (void)nslabs;
  return FV3_OK;
}

}  // namespace fv3

extern "C" {

int32_t fv3_riem_solver_c_f64(const fv3_domain* dom, double dt2, double ptop, double p_fac,
                              const double* cappa, const double* hs, const double* ws,
                              const double* ptc, const double* q_con, const double* delpc,
                              double* gz, double* pef, const double* w3, double* work,
                              int64_t work_slab, void* stream) {
  using namespace fv3;
  const char* who = "fv3_riem_solver_c";
  if (!dom) {
    set_error("%s: domain is NULL", who);
    return FV3_E_ARG;
  }
  const int i1 = dom->isc - 1, i2 = dom->iec + 1, j1 = dom->jsc - 1, j2 = dom->jec + 1;
  int rc = check_riem(dom, who, i1, i2, j1, j2, work_slab, 10);
  if (rc) return rc;
  FV3_REQUIRE(cappa && hs && ws && ptc && q_con && delpc && gz && pef && w3 && work, FV3_E_ARG,
              "%s: NULL pointer", who);
  if (dom->ntiles == 0) return FV3_OK;
  Box2 bx{i1, j1, i2 - i1 + 1, j2 - j1 + 1};
  Dom d = to_dom(dom);
  dim3 grid((unsigned)(((int64_t)bx.nib * bx.njb + 255) / 256), d.ntiles);
  // cappa, ptc, q_con, delpc, gz, w3 read; gz, pef written (per cell-level)
  TimedLaunch timed("riem_c_kernel", as_stream(stream),
                    8.0 * sizeof(double) * bx.nib * bx.njb * (double)dom->nk * d.ntiles);
  hipLaunchKernelGGL(riem_c_kernel, grid, dim3(256), 0, as_stream(stream), d, bx, dt2, ptop,
                     p_fac, cappa, hs, ws, ptc, q_con, delpc, gz, pef, w3, work, work_slab);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

int32_t fv3_riem_solver3_f64(const fv3_domain* dom, int32_t last_call, double dt, double ptop,
                             double p_fac, int32_t use_logp, double beta, const double* cappa,
                             const double* zs, const double* wsd, double* delz,
                             const double* q_con, const double* delp, const double* pt,
                             double* zh, double* pe, double* ppe, double* pk3, double* pk,
                             double* peln, double* w, double* work, int64_t work_slab,
                             void* stream) {
  using namespace fv3;
  const char* who = "fv3_riem_solver3";
  if (!dom) {
    set_error("%s: domain is NULL", who);
    return FV3_E_ARG;
  }
  const int i1 = dom->isc, i2 = dom->iec, j1 = dom->jsc, j2 = dom->jec;
  int rc = check_riem(dom, who, i1, i2, j1, j2, work_slab, 8);
  if (rc) return rc;
  FV3_REQUIRE(cappa && zs && wsd && delz && q_con && delp && pt && zh && pe && ppe && pk3 &&
                  pk && peln && w && work,
              FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(ptop > 0.0, FV3_E_ARG, "%s: ptop must be positive", who);
  if (dom->ntiles == 0) return FV3_OK;
  Box2 bx{i1, j1, i2 - i1 + 1, j2 - j1 + 1};
  Riem3Opts o;
  o.last_call = last_call ? 1 : 0;
  o.use_logp = use_logp ? 1 : 0;
  o.beta = beta;
  o.dt = dt;
  o.ptop = ptop;
  o.p_fac = p_fac;
  o.peln1 = log(ptop);
  o.kappa = kRdgas / kCpAir;
  o.ptk = exp(o.kappa * o.peln1);
  o.rgrav = 1.0 / kGrav;
  Dom d = to_dom(dom);
  dim3 grid((unsigned)(((int64_t)bx.nib * bx.njb + 255) / 256), d.ntiles);
  // cappa, delz, q_con, delp, pt, zh, pe, w read; zh, pe, ppe, pk3, delz, w
  // written, plus pk and peln on the last call (per cell-level)
  TimedLaunch timed("riem3_kernel", as_stream(stream),
                    (14.0 + 2.0 * o.last_call) * sizeof(double) * bx.nib * bx.njb *
                        (double)dom->nk * d.ntiles);
  hipLaunchKernelGGL(riem3_kernel, grid, dim3(256), 0, as_stream(stream), d, bx, o, cappa, zs,
                     wsd, delz, q_con, delp, pt, zh, pe, ppe, pk3, pk, peln, w, work, work_slab);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}
