// Semi-implicit vertical sound-wave solvers: sim1_solver, RiemannSolverC,
// RiemannSolver3.
//
// Reference: fv3core/stencils/sim1_solver.py:11-101 (two tridiagonal
// solves, for the pressure perturbation and for w, plus the dz update),
// riem_solver_c.py:13-165 and riem_solver3.py:24-242 (precompute / sim1 /
// finalize stencils).  Here each solver is one kernel: one thread owns one
// column, always in FP64, and lanes of a wave own neighbouring i, so each
// k-indexed access is coalesced.
//
// The stencils' ten FORWARD/BACKWARD passes become six sweeps: precompute,
// the pe evaluation and the pp forward elimination run as one upward sweep;
// then the pp back substitution (down), the w forward elimination (up), the
// w back substitution (down), the pe update with the finalize of pe / pk /
// ppe (up) and the dz update with the height integration (down).  Of the
// two column arrays a sweep hands to the next one in the opposite
// direction, one (pp, then w) lives in LDS, 40 KB per wave at 80 levels so
// that four waves fit a CU and every wave of a C96 cube is resident at once,
// and the other (gam, aa, the w factors, the new pe) in one workspace slab;
// dm, dz, pem and pk3 are recomputed from the inputs wherever a sweep needs
// them (by the same operations, so with the same bits), and pem / pm, which
// only an upward running sum produces, go to two more slabs.
//
// A C96 cube has ~55k columns, under one wave per SIMD, so every sweep
// walks k in chunks of KC levels: the chunk's loads (which never depend on
// the values the sweep carries from level to level) are issued together
// into registers, then the recurrence runs on them and the results are
// stored.  Each chunk waits for memory once instead of once per level.
#include "common.hpp"
#include "timing.hpp"

namespace fv3 {

constexpr double kRdgas = 287.05;
constexpr double kGrav = 9.80665;
constexpr double kCpAir = 1004.6;
constexpr int KC = 8;       // levels per chunk of a column sweep
constexpr int RW = 64;      // columns per workgroup (one wave)

__device__ __forceinline__ double gm_of(double cp3) { return 1.0 / (1.0 - cp3); }

struct Box2 {
  int i1, j1, nib, njb;
};

struct Riem3Opts {
  int last_call, use_logp;
  double beta, dt, ptop, p_fac, peln1, ptk, kappa, rgrav;
};

// Fields of one solver call.  RiemannSolverC: dp = delpc, qc = q_con,
// z = gz (geopotential, written at the end), pt = ptc, w = w3 (read only);
// RiemannSolver3: dp = delp, z = zh (height, written at the end), w in/out,
// plus the pressure outputs.
// S is the storage type of cappa, q_con, delp, pt and w: FP64, or FP32 in
// the mixed-precision acoustic step (loaded into FP64, w stored back
// rounded); the solve itself is always FP64.
template <typename S>
struct RiemFields {
  const S *cappa, *qc, *dp, *pt;
  const double* surf;  // hs (C) or zs (3), IJ
  const double* wsd;   // ws (C) or wsd (3), IJ
  double* z;
  S* w;
  double* pef;                                  // C only
  double *delz, *pe, *ppe, *pk3, *pk, *peln;    // 3 only
  double* work;
  int64_t slab;
};

// C = true: RiemannSolverC (riem_solver_c.py:13-64); false: RiemannSolver3
// (riem_solver3.py:24-103).  The sim1 arithmetic (sim1_solver.py:11-101) is
// shared and evaluated in the stencils' order.
template <bool C, typename S>
__global__ void __launch_bounds__(RW)
    riem_kernel(Dom d, Box2 bx, Riem3Opts o, double dt, double p_fac, RiemFields<S> f) {
  extern __shared__ double lds_y[];
  const int64_t n = (int64_t)blockIdx.x * RW + threadIdx.x;
  if (n >= (int64_t)bx.nib * bx.njb) return;
  const int i = bx.i1 + (int)(n % bx.nib), j = bx.j1 + (int)(n / bx.nib);
  const int t = blockIdx.y;
  const int64_t c = (int64_t)t * d.st + (int64_t)j * d.sj + i;
  const int64_t c2 = (int64_t)t * d.st2 + (int64_t)j * d.sj2 + i;
  const int64_t sk = d.sk;
  const int nk = d.nk, nz = nk - 1;
  auto at = [sk](int k) { return (int64_t)k * sk; };
  double* const Y = lds_y + threadIdx.x;  // Y[k * RW]: pp, then w
  auto y = [Y](int k) -> double& { return Y[k * RW]; };
  double* const X = f.work + c;                 // gam / aa / gam2 / new pe
  double* const PEM = f.work + f.slab + c;      // pem (interfaces)
  double* const PM = f.work + 2 * f.slab + c;   // pm (layers)
  const S* const CP = f.cappa + c;
  const S* const QC = f.qc + c;
  const S* const DP = f.dp + c;
  const S* const PT = f.pt + c;
  double* const Z = f.z + c;
  S* const W = f.w + c;
  const double ptop = o.ptop;
  const double t1g = 2.0 * dt * dt;
  const double rdt = 1.0 / dt;
  const double rgrav = o.rgrav;
  auto dm_of = [rgrav](double dpk) -> double {
    if constexpr (C)
      return dpk / kGrav;
    else
      return dpk * rgrav;
  };

  // ---- S1 (up): precompute, pe = exp(gm log(-dm/dz Rd pt)) - pm, and the pp
  // forward elimination, which lags one level behind (it needs pe and dm of
  // the level below)
  double z_k = Z[0];
  double pem_k = ptop, peg_k = ptop, pelng_k = o.peln1;
  PEM[0] = ptop;
  double gr_prev = 0.0, bet_prev = 2.0, pp_k = 0.0, dm_k = 0.0, pe_k = 0.0;
  y(0) = 0.0;
  for (int k0 = 0; k0 < nz; k0 += KC) {
    double dp[KC], qc[KC], z1[KC], pt[KC], cp[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = min(k0 + u, nz - 1);
      dp[u] = DP[at(k)];
      qc[u] = QC[at(k)];
      z1[u] = Z[at(k + 1)];
      pt[u] = PT[at(k)];
      cp[u] = CP[at(k)];
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int m = k0 + u;
      if (m < nz) {
        const double dmk = dp[u];
        const double dm = dm_of(dmk);
        const double dz = z1[u] - z_k;
        const double peg_n = peg_k + dmk * (1.0 - qc[u]);
        double pm;
        if constexpr (C) {
          pm = (peg_n - peg_k) / log(peg_n / peg_k);
          pem_k = pem_k + dmk;
        } else {
          const double pem_n = pem_k + dmk;
          const double pelng_n = log(peg_n);
          pm = (peg_n - peg_k) / (pelng_n - pelng_k);
          pem_k = pem_n;
          pelng_k = pelng_n;
        }
        peg_k = peg_n;
        PEM[at(m + 1)] = pem_k;
        PM[at(m)] = pm;
        z_k = z1[u];
        const double pe = exp(gm_of(cp[u]) * log(-dm / dz * kRdgas * pt[u])) - pm;
        // pp forward elimination for level m - 1 (sim1_solver.py:42-56)
        if (m == 0) {
          dm_k = dm;
          pe_k = pe;
        } else if (m == 1) {
          gr_prev = dm_k / dm;
          bet_prev = 2.0 * (1.0 + gr_prev);
          pp_k = 3.0 * (pe_k + gr_prev * pe) / bet_prev;
          y(1) = pp_k;
          dm_k = dm;
          pe_k = pe;
        } else {
          const int k = m - 1;
          const double g = gr_prev / bet_prev;
          X[at(k)] = g;
          const double gr = dm_k / dm;
          const double bb = 2.0 * (1.0 + gr);
          const double dd = 3.0 * (pe_k + gr * pe);
          const double bet_k = bb - g;
          pp_k = (dd - pp_k) / bet_k;
          y(k + 1) = pp_k;
          bet_prev = bet_k;
          gr_prev = gr;
          dm_k = dm;
          pe_k = pe;
        }
      }
    }
  }
  {
    // last level of the elimination (bb = 2, dd = 3 pe)
    const int k = nz - 1;
    const double g = gr_prev / bet_prev;
    X[at(k)] = g;
    const double bet_k = 2.0 - g;
    pp_k = (3.0 * pe_k - pp_k) / bet_k;
    y(k + 1) = pp_k;
  }

  // ---- S2 (down): pp back substitution and aa (k = nz-1 .. 1)
  {
    double pp_next = y(nz);
    double cpk = CP[at(nz - 1)];
    double zk = Z[at(nz - 1)];
    double dzk = Z[at(nz)] - zk;
    for (int k0 = nz - 1; k0 >= 1; k0 -= KC) {
      double gam[KC], cpm[KC], zm[KC], pem[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = max(k0 - u, 1);
        gam[u] = X[at(k)];
        cpm[u] = CP[at(k - 1)];
        zm[u] = Z[at(k - 1)];
        pem[u] = PEM[at(k)];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 - u;
        if (k >= 1) {
          const double dzm = zk - zm[u];
          const double ppk = y(k) - gam[u] * pp_next;
          y(k) = ppk;
          X[at(k)] = t1g * 0.5 * (gm_of(cpm[u]) + gm_of(cpk)) / (dzm + dzk) * (pem[u] + ppk);
          pp_next = ppk;
          cpk = cpm[u];
          zk = zm[u];
          dzk = dzm;
        }
      }
    }
  }

  // ---- S3 (up): w forward elimination (w1 = w on entry, gam into X)
  const double wsr = f.wsd[c2];
  double w_prev;
  {
    double aa_k = X[at(1)];
    double pp_c = y(1);
    const double dm0 = dm_of(DP[at(0)]);
    double bet = dm0 - aa_k;
    w_prev = (dm0 * W[at(0)] + dt * pp_c) / bet;
    y(0) = w_prev;
    for (int k0 = 1; k0 < nz; k0 += KC) {
      double aan[KC], dp[KC], w1[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = min(k0 + u, nz - 1);
        aan[u] = X[at(min(k + 1, nz - 1))];
        dp[u] = DP[at(k)];
        w1[u] = W[at(k)];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < nz) {
          const double dm = dm_of(dp[u]);
          const double ppn = y(k + 1);
          const double g = aa_k / bet;
          X[at(k)] = g;
          double wk, b;
          if (k < nz - 1) {
            b = dm - (aa_k + aan[u] + aa_k * g);
            wk = (dm * w1[u] + dt * (ppn - pp_c) - aa_k * w_prev) / b;
          } else {
            const double dzl = Z[at(k + 1)] - Z[at(k)];
            const double p1 = t1g * gm_of(CP[at(k)]) / dzl * (PEM[at(k + 1)] + ppn);
            b = dm - (aa_k + p1 + aa_k * g);
            wk = (dm * w1[u] + dt * (ppn - pp_c) - p1 * wsr - aa_k * w_prev) / b;
          }
          y(k) = wk;
          w_prev = wk;
          bet = b;
          pp_c = ppn;
          if (k + 1 < nz) aa_k = aan[u];
        }
      }
    }
  }

  // ---- S4 (down): w back substitution (k = nz-2 .. 0)
  {
    double w_next = w_prev;  // w[nz-1]
    for (int k0 = nz - 2; k0 >= 0; k0 -= KC) {
      double gam[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) gam[u] = X[at(max(k0 - u, 0) + 1)];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 - u;
        if (k >= 0) {
          w_next = y(k) - gam[u] * w_next;
          y(k) = w_next;
        }
      }
    }
  }

  // ---- S5 (up): pe from the change of w (new pe into X), and the finalize
  // of the interface pressures; pem is re-summed as the precompute summed it
  {
    double pe_run = 0.0;
    X[at(0)] = 0.0;
    double pem = ptop;
    if constexpr (C) {
      f.pef[c] = ptop;
    } else {
      const double pk3k = o.use_logp ? o.peln1 : o.ptk;
      f.pk3[c] = pk3k;
      f.ppe[c] = 0.0;  // new pe is 0 at the top; beta < -0.1 adds pem
      if (o.beta < -0.1) f.ppe[c] = 0.0 + pem;
      if (o.last_call) {
        f.peln[c] = o.peln1;
        f.pk[c] = pk3k;
        f.pe[c] = pem;
      }
    }
    for (int k0 = 1; k0 < nk; k0 += KC) {
      double dp[KC], w1[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = min(k0 + u, nk - 1);
        dp[u] = DP[at(k - 1)];
        w1[u] = W[at(k - 1)];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < nk) {
          const double wk = y(k - 1);
          pe_run = pe_run + dm_of(dp[u]) * (wk - w1[u]) * rdt;
          X[at(k)] = pe_run;
          pem = pem + dp[u];
          if constexpr (C) {
            f.pef[c + at(k)] = pe_run + pem;
          } else {
            W[at(k - 1)] = (S)wk;
            const double peln_run = log(pem);
            const double pk3k = o.use_logp ? peln_run : exp(o.kappa * peln_run);
            f.pk3[c + at(k)] = pk3k;
            f.ppe[c + at(k)] = o.beta < -0.1 ? pe_run + pem : pe_run;
            if (o.last_call) {
              f.peln[c + at(k)] = peln_run;
              f.pk[c + at(k)] = pk3k;
              f.pe[c + at(k)] = pem;
            }
          }
        }
      }
    }
  }

  // ---- S6 (down): p1 and the dz update (sim1_solver.py:90-101), then the
  // height / geopotential integration from the surface
  {
    double p1_next = 0.0;
    double pe1 = X[at(nz)], pe2 = 0.0, dm1 = 0.0;
    double zsum = f.surf[c2];
    Z[at(nz)] = zsum;
    for (int k0 = nz - 1; k0 >= 0; k0 -= KC) {
      double pe0[KC], dp[KC], pm[KC], pt[KC], cp[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = max(k0 - u, 0);
        pe0[u] = X[at(k)];
        dp[u] = DP[at(k)];
        pm[u] = PM[at(k)];
        pt[u] = PT[at(k)];
        cp[u] = CP[at(k)];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 - u;
        if (k >= 0) {
          const double dm0 = dm_of(dp[u]);
          double p1;
          if (k == nz - 1) {
            p1 = (pe0[u] + 2.0 * pe1) * 1.0 / 3.0;
          } else {
            const double gr = dm0 / dm1;
            const double bb = 2.0 * (1.0 + gr);
            p1 = (pe0[u] + bb * pe1 + gr * pe2) * 1.0 / 3.0 - gr * p1_next;
          }
          p1_next = p1;
          const double maxp = p_fac * dm0 > p1 + pm[u] ? p_fac * pm[u] : p1 + pm[u];
          const double dz = -dm0 * kRdgas * pt[u] * exp((cp[u] - 1.0) * log(maxp));
          if constexpr (C) {
            zsum = zsum - dz * kGrav;
          } else {
            f.delz[c + at(k)] = dz;
            zsum = zsum - dz;
          }
          Z[at(k)] = zsum;
          pe2 = pe1;
          pe1 = pe0[u];
          dm1 = dm0;
        }
      }
    }
  }
}

// one column array of RW columns per wave in LDS: 64 KB at most
constexpr int kRiemMaxLevels = 128;
static size_t riem_lds_bytes(int nk) { return (size_t)nk * RW * sizeof(double); }

static int check_riem(const fv3_domain* dom, const char* who, int i1, int i2, int j1, int j2,
                      int64_t slab, int nslabs) {
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(dom->nk >= 4, FV3_E_ARG, "%s: need at least 4 levels", who);
  FV3_REQUIRE(dom->nk <= kRiemMaxLevels, FV3_E_UNSUPPORTED,
              "%s: at most %d interface levels (got %d)", who, kRiemMaxLevels, dom->nk);
  FV3_REQUIRE(i1 >= 0 && j1 >= 0 && i2 < dom->ni && j2 < dom->nj, FV3_E_ARG,
              "%s: box outside the storage", who);
  FV3_REQUIRE(slab >= (int64_t)(dom->ntiles - 1) * dom->st + dom->sk * dom->nk, FV3_E_ARG,
              "%s: workspace slab too small", who);
  (void)nslabs;
  return FV3_OK;
}

template <typename S>
int riem_solver_c(const fv3_domain* dom, double dt2, double ptop, double p_fac, const S* cappa,
                  const double* hs, const double* ws, const S* ptc, const S* q_con,
                  const S* delpc, double* gz, double* pef, const S* w3, double* work,
                  int64_t work_slab, void* stream) {
  const char* who = "fv3_riem_solver_c";
  if (!dom) {
    set_error("%s: domain is NULL", who);
    return FV3_E_ARG;
  }
  const int i1 = dom->isc - 1, i2 = dom->iec + 1, j1 = dom->jsc - 1, j2 = dom->jec + 1;
  int rc = check_riem(dom, who, i1, i2, j1, j2, work_slab, 3);
  if (rc) return rc;
  FV3_REQUIRE(cappa && hs && ws && ptc && q_con && delpc && gz && pef && w3 && work, FV3_E_ARG,
              "%s: NULL pointer", who);
  if (dom->ntiles == 0) return FV3_OK;
  Box2 bx{i1, j1, i2 - i1 + 1, j2 - j1 + 1};
  Dom d = to_dom(dom);
  Riem3Opts o{};
  o.ptop = ptop;
  RiemFields<S> f{};
  f.cappa = cappa;
  f.qc = q_con;
  f.dp = delpc;
  f.pt = ptc;
  f.surf = hs;
  f.wsd = ws;
  f.z = gz;
  f.w = const_cast<S*>(w3);  // read only in this solver
  f.pef = pef;
  f.work = work;
  f.slab = work_slab;
  dim3 grid((unsigned)(((int64_t)bx.nib * bx.njb + RW - 1) / RW), d.ntiles);
  // cappa, ptc, q_con, delpc, gz, w3 read; gz, pef written (per cell-level)
  TimedLaunch timed("riem_c_kernel", as_stream(stream),
                    (4.0 * sizeof(S) + 4.0 * sizeof(double)) * bx.nib * bx.njb *
                        (double)dom->nk * d.ntiles);
  hipLaunchKernelGGL((riem_kernel<true, S>), grid, dim3(RW), riem_lds_bytes(dom->nk),
                     as_stream(stream), d, bx, o, dt2, p_fac, f);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

template <typename S>
int riem_solver3(const fv3_domain* dom, int last_call, double dt, double ptop, double p_fac,
                 int use_logp, double beta, const S* cappa, const double* zs, const double* wsd,
                 double* delz, const S* q_con, const S* delp, const S* pt, double* zh, double* pe,
                 double* ppe, double* pk3, double* pk, double* peln, S* w, double* work,
                 int64_t work_slab, void* stream) {
  const char* who = "fv3_riem_solver3";
  if (!dom) {
    set_error("%s: domain is NULL", who);
    return FV3_E_ARG;
  }
  const int i1 = dom->isc, i2 = dom->iec, j1 = dom->jsc, j2 = dom->jec;
  int rc = check_riem(dom, who, i1, i2, j1, j2, work_slab, 3);
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
  RiemFields<S> f{};
  f.cappa = cappa;
  f.qc = q_con;
  f.dp = delp;
  f.pt = pt;
  f.surf = zs;
  f.wsd = wsd;
  f.z = zh;
  f.w = w;
  f.delz = delz;
  f.pe = pe;
  f.ppe = ppe;
  f.pk3 = pk3;
  f.pk = pk;
  f.peln = peln;
  f.work = work;
  f.slab = work_slab;
  dim3 grid((unsigned)(((int64_t)bx.nib * bx.njb + RW - 1) / RW), d.ntiles);
  // cappa, q_con, delp, pt, zh, w read; zh, ppe, pk3, delz, w written, plus
  // pe, pk and peln on the last call (per cell-level)
  TimedLaunch timed("riem3_kernel", as_stream(stream),
                    ((5.0 * sizeof(S) + (6.0 + 3.0 * o.last_call) * sizeof(double))) * bx.nib * bx.njb *
                        (double)dom->nk * d.ntiles);
  hipLaunchKernelGGL((riem_kernel<false, S>), grid, dim3(RW), riem_lds_bytes(dom->nk),
                     as_stream(stream), d, bx, o, dt, p_fac, f);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}
}  // namespace fv3

extern "C" {

int32_t fv3_riem_solver_c_f64(const fv3_domain* dom, double dt2, double ptop, double p_fac,
                              const double* cappa, const double* hs, const double* ws,
                              const double* ptc, const double* q_con, const double* delpc,
                              double* gz, double* pef, const double* w3, double* work,
                              int64_t work_slab, void* stream) {
  return fv3::riem_solver_c<double>(dom, dt2, ptop, p_fac, cappa, hs, ws, ptc, q_con, delpc, gz,
                                    pef, w3, work, work_slab, stream);
}

int32_t fv3_riem_solver_c_mixed(const fv3_domain* dom, double dt2, double ptop, double p_fac,
                                const float* cappa, const double* hs, const double* ws,
                                const float* ptc, const float* q_con, const float* delpc,
                                double* gz, double* pef, const float* w3, double* work,
                                int64_t work_slab, void* stream) {
  return fv3::riem_solver_c<float>(dom, dt2, ptop, p_fac, cappa, hs, ws, ptc, q_con, delpc, gz,
                                   pef, w3, work, work_slab, stream);
}

int32_t fv3_riem_solver3_f64(const fv3_domain* dom, int32_t last_call, double dt, double ptop,
                             double p_fac, int32_t use_logp, double beta, const double* cappa,
                             const double* zs, const double* wsd, double* delz,
                             const double* q_con, const double* delp, const double* pt,
                             double* zh, double* pe, double* ppe, double* pk3, double* pk,
                             double* peln, double* w, double* work, int64_t work_slab,
                             void* stream) {
  return fv3::riem_solver3<double>(dom, last_call, dt, ptop, p_fac, use_logp, beta, cappa, zs,
                                   wsd, delz, q_con, delp, pt, zh, pe, ppe, pk3, pk, peln, w,
                                   work, work_slab, stream);
}

int32_t fv3_riem_solver3_mixed(const fv3_domain* dom, int32_t last_call, double dt, double ptop,
                               double p_fac, int32_t use_logp, double beta, const float* cappa,
                               const double* zs, const double* wsd, double* delz,
                               const float* q_con, const float* delp, const float* pt,
                               double* zh, double* pe, double* ppe, double* pk3, double* pk,
                               double* peln, float* w, double* work, int64_t work_slab,
                               void* stream) {
  return fv3::riem_solver3<float>(dom, last_call, dt, ptop, p_fac, use_logp, beta, cappa, zs,
                                  wsd, delz, q_con, delp, pt, zh, pe, ppe, pk3, pk, peln, w,
                                  work, work_slab, stream);
}
}
