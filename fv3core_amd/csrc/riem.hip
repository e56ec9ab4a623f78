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
//
// A C96 cube has ~55k columns, under one wave per SIMD, so the sweeps are
// bound by memory latency rather than bandwidth.  Every sweep therefore
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
constexpr int KC = 8;  // levels per chunk of a column sweep

// Arrays sim1 works on (sim1_solver.py:126-139): w and dz are in/out, pe is
// out; gm is 1 / (1 - cp3) of the same column (recomputed on the fly).
// All pointers address level 0 of one column; levels are sk apart.
struct Sim1Cols {
  double *w, *dm, *dz, *ptr, *pm, *pe, *pem, *cp3;
  double *w1, *gam, *pp, *aa;  // workspace
};

__device__ __forceinline__ double gm_of(double cp3) { return 1.0 / (1.0 - cp3); }

__device__ void sim1_column(const Sim1Cols& c, int nk, int64_t sk, double wsr, double dt,
                            double p_fac) {
  const int nz = nk - 1;
  const double t1g = 2.0 * dt * dt;
  const double rdt = 1.0 / dt;
  double* __restrict__ W = c.w;
  double* __restrict__ DM = c.dm;
  double* __restrict__ DZ = c.dz;
  double* __restrict__ PTR = c.ptr;
  double* __restrict__ PM = c.pm;
  double* __restrict__ PE = c.pe;
  double* __restrict__ PEM = c.pem;
  double* __restrict__ CP3 = c.cp3;
  double* __restrict__ W1 = c.w1;
  double* __restrict__ GAM = c.gam;
  double* __restrict__ PP = c.pp;
  double* __restrict__ AA = c.aa;
  auto at = [sk](int k) { return (int64_t)k * sk; };

  // A. pe = exp(gm*log(-dm/dz*Rd*ptr)) - pm; w1 = w
  for (int k0 = 0; k0 < nz; k0 += KC) {
    double dm[KC], dz[KC], pt[KC], cp[KC], pm[KC], w[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nz) {
        dm[u] = DM[at(k)];
        dz[u] = DZ[at(k)];
        pt[u] = PTR[at(k)];
        cp[u] = CP3[at(k)];
        pm[u] = PM[at(k)];
        w[u] = W[at(k)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nz) {
        PE[at(k)] = exp(gm_of(cp[u]) * log(-dm[u] / dz[u] * kRdgas * pt[u])) - pm[u];
        W1[at(k)] = w[u];
      }
    }
  }

  // B. forward elimination for pp (pp[0] = 0, pp[1] = dd[0]/bb[0]) with
  // bb(k) = 2(1 + g_rat(k)) and dd(k) = 3(pe[k] + g_rat(k) pe[k+1]) below
  // the last layer, 2 and 3 pe[k] on it; g_rat(k) = dm[k] / dm[k+1]
  double dm_k = DM[at(0)], pe_k = PE[at(0)];
  double gr_prev, bet_prev, pp_k;
  if (nz > 1) {
    const double dm1 = DM[at(1)], pe1 = PE[at(1)];
    gr_prev = dm_k / dm1;
    bet_prev = 2.0 * (1.0 + gr_prev);
    pp_k = 3.0 * (pe_k + gr_prev * pe1) / bet_prev;
    dm_k = dm1;
    pe_k = pe1;
  } else {
    gr_prev = 0.0;
    bet_prev = 2.0;
    pp_k = 3.0 * pe_k / bet_prev;
  }
  PP[at(0)] = 0.0;
  PP[at(1)] = pp_k;
  for (int k0 = 1; k0 < nz; k0 += KC) {
    double dmn[KC], pen[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k + 1 < nz) {
        dmn[u] = DM[at(k + 1)];
        pen[u] = PE[at(k + 1)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nz) {
        const double g = gr_prev / bet_prev;
        GAM[at(k)] = g;
        double bb, dd, gr = 0.0;
        if (k < nz - 1) {
          gr = dm_k / dmn[u];
          bb = 2.0 * (1.0 + gr);
          dd = 3.0 * (pe_k + gr * pen[u]);
        } else {
          bb = 2.0;
          dd = 3.0 * pe_k;
        }
        const double bet_k = bb - g;
        pp_k = (dd - pp_k) / bet_k;
        PP[at(k + 1)] = pp_k;
        bet_prev = bet_k;
        gr_prev = gr;
        if (k < nz - 1) {
          dm_k = dmn[u];
          pe_k = pen[u];
        }
      }
    }
  }

  // C. back substitution and aa (k = nz-1 .. 1)
  double pp_next = PP[at(nz)];
  for (int k0 = nz - 1; k0 >= 1; k0 -= KC) {
    double pp[KC], gam[KC], cpk[KC], cpm[KC], dzk[KC], dzm[KC], pem[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 1) {
        pp[u] = PP[at(k)];
        gam[u] = GAM[at(k)];
        cpk[u] = CP3[at(k)];
        cpm[u] = CP3[at(k - 1)];
        dzk[u] = DZ[at(k)];
        dzm[u] = DZ[at(k - 1)];
        pem[u] = PEM[at(k)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 1) {
        const double ppk = pp[u] - gam[u] * pp_next;
        PP[at(k)] = ppk;
        AA[at(k)] = t1g * 0.5 * (gm_of(cpm[u]) + gm_of(cpk[u])) / (dzm[u] + dzk[u]) *
                    (pem[u] + ppk);
        pp_next = ppk;
      }
    }
  }

  // D. w tridiagonal, forward (gam reused)
  double aa_k = AA[at(1)];
  double pp_c = PP[at(1)];  // pp[k+1] of the level before becomes pp[k]
  double bet = DM[at(0)] - aa_k;
  double w_prev = (DM[at(0)] * W1[at(0)] + dt * pp_c) / bet;
  W[at(0)] = w_prev;
  for (int k0 = 1; k0 < nz; k0 += KC) {
    double aan[KC], dm[KC], w1[KC], ppn[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nz) {
        if (k + 1 < nz) aan[u] = AA[at(k + 1)];
        dm[u] = DM[at(k)];
        w1[u] = W1[at(k)];
        ppn[u] = PP[at(k + 1)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nz) {
        const double g = aa_k / bet;
        GAM[at(k)] = g;
        double wk, b;
        if (k < nz - 1) {
          b = dm[u] - (aa_k + aan[u] + aa_k * g);
          wk = (dm[u] * w1[u] + dt * (ppn[u] - pp_c) - aa_k * w_prev) / b;
        } else {
          const double p1 = t1g * gm_of(CP3[at(k)]) / DZ[at(k)] * (PEM[at(k + 1)] + ppn[u]);
          b = dm[u] - (aa_k + p1 + aa_k * g);
          wk = (dm[u] * w1[u] + dt * (ppn[u] - pp_c) - p1 * wsr - aa_k * w_prev) / b;
        }
        W[at(k)] = wk;
        w_prev = wk;
        bet = b;
        pp_c = ppn[u];
        if (k + 1 < nz) aa_k = aan[u];
      }
    }
  }

  // E. w back substitution (k = nz-2 .. 0)
  double w_next = w_prev;  // w[nz-1]
  for (int k0 = nz - 2; k0 >= 0; k0 -= KC) {
    double w[KC], gam[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        w[u] = W[at(k)];
        gam[u] = GAM[at(k + 1)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        w_next = w[u] - gam[u] * w_next;
        W[at(k)] = w_next;
      }
    }
  }

  // F. pe from the change of w (pe[0] = 0)
  double pe_run = 0.0;
  PE[at(0)] = 0.0;
  for (int k0 = 1; k0 < nk; k0 += KC) {
    double dm[KC], w[KC], w1[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        dm[u] = DM[at(k - 1)];
        w[u] = W[at(k - 1)];
        w1[u] = W1[at(k - 1)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        pe_run = pe_run + dm[u] * (w[u] - w1[u]) * rdt;
        PE[at(k)] = pe_run;
      }
    }
  }

  // G. p1 (BACKWARD) and the dz update
  double p1_next = 0.0;
  for (int k0 = nz - 1; k0 >= 0; k0 -= KC) {
    double pe0[KC], pe1[KC], pe2[KC], dm0[KC], dm1[KC], pm[KC], pt[KC], cp[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        pe0[u] = PE[at(k)];
        pe1[u] = PE[at(k + 1)];
        if (k < nz - 1) {
          pe2[u] = PE[at(k + 2)];
          dm1[u] = DM[at(k + 1)];
        }
        dm0[u] = DM[at(k)];
        pm[u] = PM[at(k)];
        pt[u] = PTR[at(k)];
        cp[u] = CP3[at(k)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        double p1;
        if (k == nz - 1) {
          p1 = (pe0[u] + 2.0 * pe1[u]) * 1.0 / 3.0;
        } else {
          const double gr = dm0[u] / dm1[u];
          const double bb = k < nz - 1 ? 2.0 * (1.0 + gr) : 2.0;
          p1 = (pe0[u] + bb * pe1[u] + gr * pe2[u]) * 1.0 / 3.0 - gr * p1_next;
        }
        p1_next = p1;
        const double maxp = p_fac * dm0[u] > p1 + pm[u] ? p_fac * pm[u] : p1 + pm[u];
        DZ[at(k)] = -dm0[u] * kRdgas * pt[u] * exp((cp[u] - 1.0) * log(maxp));
      }
    }
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
  auto at = [sk](int k) { return (int64_t)k * sk; };
  double* S[10];
#pragma unroll
  for (int s = 0; s < 10; ++s) S[s] = work + s * slab + c;
  double* __restrict__ DM = S[0];
  double* __restrict__ W = S[1];
  double* __restrict__ DZ = S[2];
  double* __restrict__ PM = S[3];
  double* __restrict__ PEM = S[4];
  double* __restrict__ PE = S[5];
  const double* __restrict__ QC = q_con + c;
  const double* __restrict__ DP = delpc + c;
  const double* __restrict__ W3 = w3 + c;
  double* __restrict__ GZ = gz + c;
  double* __restrict__ PEF = pef + c;
  // precompute (riem_solver_c.py:13-43)
  double pem_k = ptop, peg_k = ptop;
  for (int k0 = 0; k0 < nk; k0 += KC) {
    double dp[KC], w[KC], qc[KC], g0[KC], g1[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        dp[u] = DP[at(k)];
        w[u] = W3[at(k)];
        if (k < nz) {
          qc[u] = QC[at(k)];
          g0[u] = GZ[at(k)];
          g1[u] = GZ[at(k + 1)];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        const double dmk = dp[u];
        W[at(k)] = w[u];
        PEM[at(k)] = pem_k;
        DM[at(k)] = dmk / kGrav;
        if (k < nz) {
          DZ[at(k)] = g1[u] - g0[u];
          const double peg_n = peg_k + dmk * (1.0 - qc[u]);
          PM[at(k)] = (peg_n - peg_k) / log(peg_n / peg_k);
          peg_k = peg_n;
          pem_k = pem_k + dmk;
        }
      }
    }
  }
  Sim1Cols sc{W,  DM, DZ, const_cast<double*>(ptc) + c, PM, PE, PEM,
              const_cast<double*>(cappa) + c, S[6], S[7], S[8], S[9]};
  sim1_column(sc, nk, sk, ws[c2], dt2, p_fac);
  // finalize (riem_solver_c.py:46-64)
  PEF[0] = ptop;
  for (int k0 = 1; k0 < nk; k0 += KC) {
    double pe[KC], pem[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        pe[u] = PE[at(k)];
        pem[u] = PEM[at(k)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) PEF[at(k)] = pe[u] + pem[u];
    }
  }
  double g = hs[c2];
  GZ[at(nz)] = g;
  for (int k0 = nz - 1; k0 >= 0; k0 -= KC) {
    double dz[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) dz[u] = DZ[at(k)];
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        g = g - dz[u] * kGrav;
        GZ[at(k)] = g;
      }
    }
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
  auto at = [sk](int k) { return (int64_t)k * sk; };
  double* S[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) S[s] = work + s * slab + c;
  double* __restrict__ DM = S[0];
  double* __restrict__ PM = S[1];
  double* __restrict__ PEM = S[2];
  double* __restrict__ PEI = S[3];  // pe on entry
  const double* __restrict__ QC = q_con + c;
  const double* __restrict__ DP = delp + c;
  double* __restrict__ DZ = delz + c;
  double* __restrict__ ZH = zh + c;
  double* __restrict__ PE = pe + c;
  double* __restrict__ PPE = ppe + c;
  double* __restrict__ PK3 = pk3 + c;
  double* __restrict__ PK = pk + c;
  double* __restrict__ PELN = peln + c;
  // precompute (riem_solver3.py:24-68)
  double pem_k = o.ptop, peg_k = o.ptop, pelng_k = o.peln1;
  PEM[0] = o.ptop;
  PK3[0] = o.ptk;
  for (int k0 = 0; k0 < nk; k0 += KC) {
    double dp[KC], pe0[KC], qc[KC], z0[KC], z1[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        dp[u] = DP[at(k)];
        pe0[u] = PE[at(k)];
        if (k < nz) {
          qc[u] = QC[at(k)];
          z0[u] = ZH[at(k)];
          z1[u] = ZH[at(k + 1)];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        const double dmk = dp[u];
        PEI[at(k)] = pe0[u];
        DM[at(k)] = dmk * o.rgrav;
        if (k < nz) {
          const double pem_n = pem_k + dmk;
          const double peln_n = log(pem_n);
          const double peg_n = peg_k + dmk * (1.0 - qc[u]);
          const double pelng_n = log(peg_n);
          PEM[at(k + 1)] = pem_n;
          PK3[at(k + 1)] = exp(o.kappa * peln_n);
          PM[at(k)] = (peg_n - peg_k) / (pelng_n - pelng_k);
          DZ[at(k)] = z1[u] - z0[u];
          pem_k = pem_n;
          peg_k = peg_n;
          pelng_k = pelng_n;
        }
      }
    }
  }
  Sim1Cols sc{w + c, DM, DZ, const_cast<double*>(pt) + c, PM, PE, PEM,
              const_cast<double*>(cappa) + c, S[4], S[5], S[6], S[7]};
  sim1_column(sc, nk, sk, wsd[c2], o.dt, o.p_fac);
  // finalize (riem_solver3.py:70-103)
  for (int k0 = 0; k0 < nk; k0 += KC) {
    double pem[KC], pek[KC], pk3[KC], pei[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        pem[u] = PEM[at(k)];
        pek[u] = PE[at(k)];
        pk3[u] = PK3[at(k)];
        pei[u] = PEI[at(k)];
      }
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 + u;
      if (k < nk) {
        const double peln_run = k == 0 ? o.peln1 : log(pem[u]);
        double pk3k = pk3[u];
        if (o.use_logp) {
          pk3k = peln_run;
          PK3[at(k)] = pk3k;
        }
        PPE[at(k)] = o.beta < -0.1 ? pek[u] + pem[u] : pek[u];
        if (o.last_call) {
          PELN[at(k)] = peln_run;
          PK[at(k)] = pk3k;
          PE[at(k)] = pem[u];
        } else {
          PE[at(k)] = pei[u];
        }
      }
    }
  }
  double z = zs[c2];
  ZH[at(nz)] = z;
  for (int k0 = nz - 1; k0 >= 0; k0 -= KC) {
    double dz[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) dz[u] = DZ[at(k)];
    }
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        z = z - dz[u];
        ZH[at(k)] = z;
      }
    }
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
