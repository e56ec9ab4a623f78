/*
 * fv3core_amd: C ABI of the MI355X (gfx950) FV3 dynamical-core kernels.
 *
 * Every entry point
 *   - takes plain device pointers to caller-owned storages, never frees them;
 *   - only enqueues work on the given HIP stream (a hipStream_t passed as
 *     void*, NULL = default stream) and never synchronises, so it can be
 *     captured in a HIP graph;
 *   - returns 0 on success or a negative FV3_E* code; fv3_last_error()
 *     returns a thread-local message for the last failure on this thread.
 *
 * Storage layout (fv3_domain): a field of the reference's storage shape
 * (ni, nj, nk) is addressed as base[t*st + k*sk + j*sj + i] (i fastest, unit
 * stride), t being the tile/member index of a batch.  IJ metric fields use
 * base[t*st2 + j*sj2 + i].  Indices are the reference's storage indices:
 * the compute domain is [isc, iec] x [jsc, jec] (isc = jsc = n_halo).
 *
 * The reference exposes these operations as Python classes; the comment on
 * each entry point names the class method it replaces.
 */
#ifndef FV3CORE_AMD_H
#define FV3CORE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FV3_OK 0
#define FV3_E_ARG (-1)          /* invalid argument (null pointer, bad size) */
#define FV3_E_UNSUPPORTED (-2)  /* option the reference also rejects */
#define FV3_E_HIP (-3)          /* a HIP runtime call failed */

/* Anchors for a tile edge this domain does not own: far enough outside any
 * storage that no region test can match (the reference shifts the anchor by
 * the int16 maximum for the same purpose, fv3core/utils/stencil.py:457-495). */
#define FV3_NO_EDGE_LO (-(1 << 28))
#define FV3_NO_EDGE_HI (1 << 28)

typedef struct fv3_domain {
  int32_t isc, iec, jsc, jec;           /* compute domain (storage indices) */
  int32_t i_start, i_end, j_start, j_end; /* tile-edge anchors or sentinels */
  int32_t ni, nj;                       /* storage extents in i and j */
  int32_t nk;                           /* number of levels processed */
  int32_t ntiles;                       /* batch extent (tiles x members) */
  int64_t sj, sk, st;                   /* 3-D strides in elements */
  int64_t sj2, st2;                     /* IJ strides in elements */
} fv3_domain;

/* Library version string. */
const char* fv3_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* fv3_last_error(void);

/* Number of gfx950 devices visible (0 on a machine without a GPU). */
int32_t fv3_device_count(void);

/* Launch timing (no counterpart in the reference; for benchmarks):
 * fv3_timing_arm(name) makes every later launch of the kernel called
 * `name` record HIP events around itself on its stream (NULL or "" stops
 * it); fv3_timing_read waits for those events and returns the number of
 * timed launches, their summed duration and their summed algorithmic HBM
 * bytes (inputs read once, outputs written once), then starts a new count.
 * fv3_timing_hold(1) suspends the recording without clearing the count and
 * fv3_timing_hold(0) resumes it, so that a benchmark can time the launches
 * of one part of its timed region.
 * Timed kernels: fvtp2d_kernel, map_single_kernel, riem_c_kernel,
 * riem3_kernel. */
int32_t fv3_timing_arm(const char* kernel_name);
int32_t fv3_timing_hold(int32_t hold);
int32_t fv3_timing_read(int64_t* count, double* total_ms, double* total_bytes);

/*
 * CopyCornersXY.__call__ (fv3core/utils/corners.py:64-117, stencil
 * copy_corners_xy_stencil_defn corners.py:425-540):
 * q_x <- q with x-corner copies (in place when q_x == q), q_y <- q with
 * y-corner copies.  Processes the full storage extent in i and j and nk levels.
 */
int32_t fv3_copy_corners_xy_f64(const fv3_domain* d, const double* q, double* q_x,
                                double* q_y, void* stream);
int32_t fv3_copy_corners_xy_f32(const fv3_domain* d, const float* q, float* q_x,
                                float* q_y, void* stream);

/*
 * FiniteVolumeTransport.__call__ without damping (fv3core/stencils/fvtp2d.py:249-360),
 * i.e. y_ppm inner -> q_i -> x_ppm outer, x_ppm inner -> q_j -> y_ppm outer,
 * final_fluxes, fused in one kernel.  hord in {5, 6, 7, 8}.
 * x_mass_flux / y_mass_flux may be NULL (then the area fluxes are the unit
 * fluxes).  Writes q_x_flux on [isc, iec+1] x [jsc, jec] and q_y_flux on
 * [isc, iec] x [jsc, jec+1] for nk levels.
 */
int32_t fv3_fvtp2d_f64(const fv3_domain* d, int32_t hord, const double* q_x,
                       const double* q_y, const double* crx, const double* cry,
                       const double* x_area_flux, const double* y_area_flux,
                       const double* x_mass_flux, const double* y_mass_flux,
                       const double* area, const double* dxa, const double* dya,
                       double* q_x_flux, double* q_y_flux, void* stream);
int32_t fv3_fvtp2d_f32(const fv3_domain* d, int32_t hord, const float* q_x,
                       const float* q_y, const float* crx, const float* cry,
                       const float* x_area_flux, const float* y_area_flux,
                       const float* x_mass_flux, const float* y_mass_flux,
                       const float* area, const float* dxa, const float* dya,
                       float* q_x_flux, float* q_y_flux, void* stream);

/*
 * The corner copy with q_x = q followed by the transport of q_x = q and q_y,
 * in one call (corners.py:425-540 then fvtp2d.py:249-360), as d_sw and
 * updatedzd use them: q receives the x corner fill in place, and the y corner fill is
 * applied as the transport reads q instead of into a full copy.  Returns the
 * same fluxes, bit for bit, as the two calls.  qy_scratch (laid out like q)
 * is written only when the domain's x and y corner sets differ, which no
 * cube tile has.  No counterpart as one call in the reference.
 */
int32_t fv3_fvtp2d_corner_fill_f64(const fv3_domain* d, int32_t hord, double* q,
                                   double* qy_scratch, const double* crx, const double* cry,
                                   const double* x_area_flux, const double* y_area_flux,
                                   const double* x_mass_flux, const double* y_mass_flux,
                                   const double* area, const double* dxa, const double* dya,
                                   double* q_x_flux, double* q_y_flux, void* stream);

/* Mixed-precision form (no counterpart in the
 * reference, which computes in FP64 only): FP64 storages, as fv3_fvtp2d_f64,
 * with the PPM and cross-update arithmetic of the kernel in FP32 (inputs
 * rounded to FP32 on load, fluxes widened to FP64 on store). */
int32_t fv3_fvtp2d_mixed_f64(const fv3_domain* d, int32_t hord, const double* q_x,
                             const double* q_y, const double* crx, const double* cry,
                             const double* x_area_flux, const double* y_area_flux,
                             const double* x_mass_flux, const double* y_mass_flux,
                             const double* area, const double* dxa, const double* dya,
                             double* q_x_flux, double* q_y_flux, void* stream);

/*
 * Tangent-linear model and adjoint of fv3_fvtp2d with respect to the
 * transported field, at the branch decisions (limiters, smoothness masks,
 * upwind side) of the state q_x / q_y.  No counterpart in the reference,
 * which has no derivative code: they linearise FiniteVolumeTransport.__call__
 * (fv3core/stencils/fvtp2d.py:249-362) for the adjoint of PPM transport.
 * Arguments as fv3_fvtp2d; Courant numbers, area and mass fluxes and the
 * metric terms are held fixed.
 *   tlm:     dq_x, dq_y (perturbations of q_x, q_y) -> dq_x_flux, dq_y_flux
 *            (written on the flux domain of fv3_fvtp2d).
 *   adjoint: g_x_flux, g_y_flux (cotangents on the flux domain; values
 *            outside it are ignored) -> g_q_x, g_q_y (the whole storage is
 *            overwritten; for q_x == q_y the gradient is their sum).
 * work: fv3_fvtp2d_workspace_slabs() slabs of work_slab elements, each a
 * field with the strides of q_x ((ntiles - 1) * st + nk * sk elements).
 * The compute domain needs its 3-cell halo inside the storage.  FP64 only.
 */
int32_t fv3_fvtp2d_workspace_slabs(void);
int32_t fv3_fvtp2d_tlm_f64(const fv3_domain* d, int32_t hord, const double* q_x,
                           const double* q_y, const double* crx, const double* cry,
                           const double* x_area_flux, const double* y_area_flux,
                           const double* x_mass_flux, const double* y_mass_flux,
                           const double* area, const double* dxa, const double* dya,
                           const double* dq_x, const double* dq_y, double* dq_x_flux,
                           double* dq_y_flux, double* work, int64_t work_slab, void* stream);
int32_t fv3_fvtp2d_adjoint_f64(const fv3_domain* d, int32_t hord, const double* q_x,
                               const double* q_y, const double* crx, const double* cry,
                               const double* x_area_flux, const double* y_area_flux,
                               const double* x_mass_flux, const double* y_mass_flux,
                               const double* area, const double* dxa, const double* dya,
                               const double* g_x_flux, const double* g_y_flux, double* g_q_x,
                               double* g_q_y, double* work, int64_t work_slab, void* stream);

/*
 * XPiecewiseParabolic.__call__ / YPiecewiseParabolic.__call__
 * (fv3core/stencils/xppm.py:283-348, yppm.py): mean of q advected through
 * the x (direction 0) or y (direction 1) interfaces of the box
 * [oi, oi+ei) x [oj, oj+ej), for nk levels; ord = iord/jord in {5,6,7,8};
 * dxa is dxa (x) or dya (y).
 */
int32_t fv3_ppm_flux_f64(const fv3_domain* d, int32_t direction, int32_t ord, int32_t oi,
                         int32_t oj, int32_t ei, int32_t ej, const double* q, const double* c,
                         const double* dxa, double* out, void* stream);
int32_t fv3_ppm_flux_f32(const fv3_domain* d, int32_t direction, int32_t ord, int32_t oi,
                         int32_t oj, int32_t ei, int32_t ej, const float* q, const float* c,
                         const float* dxa, float* out, void* stream);

/* DelnFlux modes */
#define FV3_DELN_ADD 0       /* fx += fx2, fy += fy2 (DelnFlux, mass is None) */
#define FV3_DELN_MASS 1      /* fx += 0.5*damp*(mass[i-1]+mass)*fx2 (DelnFlux with mass) */
#define FV3_DELN_STORE 2     /* fx = fx2, fy = fy2 (DelnFluxNoSG, damp*q preamble) */
#define FV3_DELN_STORE_MASS 3 /* fx = fx2, fy = fy2 (DelnFluxNoSG with mass: d2 = q) */

/*
 * DelnFlux.__call__ (fv3core/stencils/delnflux.py:955-998) and
 * DelnFluxNoSG.__call__ (delnflux.py:1156-1207), fused: del-(2*nord+2)
 * diffusive fluxes with the in-between corner copies, for levels [0, nk).
 * nord4 holds nord for the level bands k = 0, 1, 2 and k >= 3 (each in
 * [0, 3], max <= 2: nord = 3 would need a 4-cell halo); damp is a device array of nk per-level coefficients
 * ((damp_c*da_min)^(nord+1) for DelnFlux).  mass may be NULL unless the mode
 * needs it.  fx/fy are updated on [isc, iec+1] x [jsc, jec] and
 * [isc, iec] x [jsc, jec+1].
 */
int32_t fv3_delnflux_f64(const fv3_domain* d, int32_t mode, const int32_t nord4[4],
                         const double* q, const double* mass, const double* damp,
                         const double* del6_u, const double* del6_v,
                         const double* rarea, double* fx, double* fy, void* stream);
int32_t fv3_delnflux_f32(const fv3_domain* d, int32_t mode, const int32_t nord4[4],
                         const float* q, const float* mass, const float* damp,
                         const float* del6_u, const float* del6_v,
                         const float* rarea, float* fx, float* fy, void* stream);

/*
 * MapSingle.__call__ (fv3core/stencils/map_single.py:130-173) with its
 * RemapProfile (remap_profile.py:576-638): conservative remap of the layer
 * means q1 from interfaces pe1 to interfaces pe2 on the box
 * [i1, i2] x [j1, j2], km = d->nk layers (pe1/pe2 hold km + 1 interfaces).
 * kord <= 10 (its magnitude selects the limiter), mode = iv in {-2,...,2};
 * qs (IJ, bottom value for iv = -2) may be NULL (then 0); qmin is the
 * kord-9 floor.  work: 7 slabs, each of work_slab elements laid out like the
 * fields (work_slab >= (ntiles-1)*st + (km+1)*sk).
 */
int32_t fv3_map_single_f64(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                           int32_t i2, int32_t j1, int32_t j2, double* q1, const double* pe1,
                           const double* pe2, const double* qs, double qmin, double* work,
                           int64_t work_slab, void* stream);
int32_t fv3_map_single_f32(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                           int32_t i2, int32_t j1, int32_t j2, float* q1, const float* pe1,
                           const float* pe2, const float* qs, double qmin, float* work,
                           int64_t work_slab, void* stream);

/*
 * fv3_map_single_f64 without the work argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The six column slabs live in a
 * file-scope device array of 96 MiB held by every GPU the library runs on
 * until the process ends; a call needing more, i.e. with
 * 6 * 8 * ((ntiles-1)*st + (km+1)*sk) bytes above 96 MiB (one C128 tile,
 * one C96 tile or six C48 tiles at 79 levels fit), fails with FV3_E_ARG
 * before any launch.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other host threads) race.  The
 * kernel writes every scratch entry it reads, so nothing needs clearing.
 */
int32_t fv3_map_single_noalloc_f64(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                                   int32_t i2, int32_t j1, int32_t j2, double* q1,
                                   const double* pe1, const double* pe2, const double* qs,
                                   double qmin, void* stream);

/*
 * Tangent-linear model and adjoint of fv3_map_single_f64 (map_scalar), kord 9,
 * with respect to the remapped field, at the branch decisions of the forward;
 * pe1, pe2, qs and qmin are held fixed.  No counterpart in the reference,
 * which has no derivative code: they linearise MapSingle.__call__
 * (fv3core/stencils/map_single.py:130-173) for the adjoint of the vertical
 * remap.  Arguments as fv3_map_single.
 *   tlm:     q1 is remapped in place as by fv3_map_single, and dq1 (its
 *            perturbation) becomes the perturbation of the result.
 *            work: fv3_map_single_tlm_workspace_slabs() slabs.
 *   adjoint: q1 (the forward input, unchanged) and g_out (cotangent of the
 *            remapped field) -> g_q1 on the box columns, levels [0, km)
 *            (other entries are not written).  work:
 *            fv3_map_single_adjoint_workspace_slabs() slabs.
 * Slabs have work_slab elements, work_slab >= (ntiles-1)*st + (km+1)*sk.
 */
int32_t fv3_map_single_tlm_workspace_slabs(void);
int32_t fv3_map_single_adjoint_workspace_slabs(void);
int32_t fv3_map_single_tlm_f64(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                               int32_t i2, int32_t j1, int32_t j2, double* q1, double* dq1,
                               const double* pe1, const double* pe2, const double* qs,
                               double qmin, double* work, int64_t work_slab, void* stream);
int32_t fv3_map_single_adjoint_f64(const fv3_domain* d, int32_t kord, int32_t mode, int32_t i1,
                                   int32_t i2, int32_t j1, int32_t j2, const double* q1,
                                   const double* pe1, const double* pe2, const double* qs,
                                   double qmin, const double* g_out, double* g_q1,
                                   double* work, int64_t work_slab, void* stream);

/*
 * RiemannSolverC.__call__ (fv3core/stencils/riem_solver_c.py:119-165):
 * precompute + sim1_solver + finalize on [isc-1, iec+1] x [jsc-1, jec+1],
 * d->nk = nz + 1 levels (4 .. 128), FP64.  Outputs gz and pef; w3 is read
 * only.  work: 3 slabs of work_slab elements (>= (ntiles-1)*st + nk*sk).
 */
int32_t fv3_riem_solver_c_f64(const fv3_domain* d, double dt2, double ptop, double p_fac,
                              const double* cappa, const double* hs, const double* ws,
                              const double* ptc, const double* q_con, const double* delpc,
                              double* gz, double* pef, const double* w3, double* work,
                              int64_t work_slab, void* stream);

/*
 * RiemannSolver3.__call__ (fv3core/stencils/riem_solver3.py:145-242) on the
 * compute domain, d->nk = nz + 1 levels (4 .. 128), FP64 (a_imp > 0.999
 * path).  In/out: delz, zh, pe, ppe, pk3, pk, peln, w; pe is written only
 * on the last call.  work: 3 slabs.
 */
int32_t fv3_riem_solver3_f64(const fv3_domain* d, int32_t last_call, double dt, double ptop,
                             double p_fac, int32_t use_logp, double beta, const double* cappa,
                             const double* zs, const double* wsd, double* delz,
                             const double* q_con, const double* delp, const double* pt,
                             double* zh, double* pe, double* ppe, double* pk3, double* pk,
                             double* peln, double* w, double* work, int64_t work_slab,
                             void* stream);

/*
 * Halo update of all six tiles of batched storages (the scalar and vector
 * halo updaters of fv3gfs-util used at fv3core/stencils/dyn_core.py:216-286;
 * no counterpart code in the reference itself).  map: device array of
 * nentries x 8 int32 (dst_tile, dst_i, dst_j, dst_comp, src_tile, src_i,
 * src_j, src_comp*2 + negate), built by fv3core_amd/halo.py.  fields: host
 * pointer to 16 device pointers; nfields of them are used, in groups of
 * ncomp (1 scalar, 2 = x/y vector components).  d->ntiles is a multiple
 * of 6: a batch of cubes (ensemble members, six consecutive tiles each),
 * each updated with the same six-tile map (tile indices 0..5).
 */
int32_t fv3_halo_update_f64(const fv3_domain* d, int32_t nentries, const int32_t* map,
                            int32_t nfields, int32_t ncomp, const void* fields, void* stream);
int32_t fv3_halo_update_f32(const fv3_domain* d, int32_t nentries, const int32_t* map,
                            int32_t nfields, int32_t ncomp, const void* fields, void* stream);

/*
 * TracerAdvection.__call__ helpers (fv3core/stencils/tracer_2d_1l.py:199-262):
 * flux_compute (x/y area fluxes from cxd/cyd on [isc, iec+1] x [jsc-3, jec+3]
 * and [isc-3, iec+3] x [jsc, jec+1]) fused with cmax_multiply_by_frac
 * (cxd, cyd, mfxd, mfyd and the area fluxes times 1/n_split when n_split > 1),
 * over the full storage plus one point.
 */
int32_t fv3_tracer_flux_prep_f64(const fv3_domain* d, int32_t n_split, double* cxd,
                                 double* cyd, double* mfxd, double* mfyd, double* xfx,
                                 double* yfx, const double* dxa, const double* dya,
                                 const double* dx, const double* dy, const double* sin_sg1,
                                 const double* sin_sg2, const double* sin_sg3,
                                 const double* sin_sg4, void* stream);
/*
 * dp_fluxadjustment (mode 0: out = dp2 = dp1 + div(fx, fy) * rarea) and
 * q_adjust (mode 1: out = q = (q*dp1 + div(fx, fy) * rarea) / dp2) on the
 * compute domain (tracer_2d_1l.py:92-117).
 */
int32_t fv3_tracer_adjust_f64(const fv3_domain* d, int32_t mode, double* out,
                              const double* dp1, const double* fx, const double* fy,
                              const double* rarea, const double* dp2, void* stream);

/*
 * One sub-cycle of TracerAdvection's loop over tracers
 * (fv3core/stencils/tracer_2d_1l.py:264-296) for ntracers (1 .. FV3_MAX_TRACERS)
 * tracers at once: the x corner copies of every tracer in place, fv_tp_2d
 * of every tracer with the shared Courant numbers and area fluxes and the
 * mass fluxes mfx, mfy (hord; the y corner copies applied as the transport
 * reads q, as fv3_fvtp2d_corner_fill_f64 does), then q_adjust
 * q = (q dp1 + div(fx, fy) rarea) / dp2 on the compute domain, in three
 * launches.  q, fx, fy: host arrays of ntracers device pointers (fx, fy are
 * flux scratch storages laid out like q).  FV3_E_UNSUPPORTED when the
 * domain's x and y corner sets differ (partial tile edges): then transport
 * the tracers one by one with fv3_fvtp2d.
 */
int32_t fv3_tracer_transport_f64(const fv3_domain* d, int32_t hord, int32_t ntracers,
                                 double* const* q, double* const* fx, double* const* fy,
                                 const double* crx, const double* cry, const double* xfx,
                                 const double* yfx, const double* mfx, const double* mfy,
                                 const double* area, const double* dxa, const double* dya,
                                 const double* rarea, const double* dp1, const double* dp2,
                                 void* stream);

/*
 * IJ metric terms of GridData (fv3core/utils/grid.py:474-798), as device
 * pointers of the entry point's floating type, each laid out with the
 * domain's sj2/st2 strides.  Members an entry point does not read may be
 * NULL; each entry point checks the ones it reads.  agrid1/agrid2 and
 * bgrid1/bgrid2 are the longitude/latitude of cell centres and corners and
 * edge_w/e/s/n the edge weights of a2b_ord4 (a2b_ord4.py:474-490), stored
 * as IJ arrays.
 */
typedef struct fv3_metrics {
  const void *dx, *dy, *dxa, *dya, *dxc, *dyc;
  const void *rdx, *rdy, *rdxa, *rdya, *rdxc, *rdyc;
  const void *area, *rarea, *area_c, *rarea_c;
  const void *sin_sg1, *sin_sg2, *sin_sg3, *sin_sg4;
  const void *cos_sg1, *cos_sg2, *cos_sg3, *cos_sg4;
  const void *cosa, *sina, *cosa_u, *cosa_v, *cosa_s, *sina_u, *sina_v;
  const void *rsin_u, *rsin_v, *rsina, *rsin2;
  const void *fC, *f0, *del6_u, *del6_v, *divg_u, *divg_v;
  const void *agrid1, *agrid2, *bgrid1, *bgrid2;
  const void *edge_w, *edge_e, *edge_s, *edge_n;
} fv3_metrics;

/*
 * DGrid2AGrid2CGridVectors.__call__ (fv3core/stencils/d2a2c_vect.py:573-651):
 * D-grid winds u, v -> A-grid ua, va and C-grid uc, vc, utc, vtc, with the
 * tile-edge interpolations and corner fills.  work: 2 slabs of work_slab
 * elements, each addressed with the 3-D strides (utmp, vtmp).
 */
int32_t fv3_d2a2c_vect_f64(const fv3_domain* d, const fv3_metrics* g, double* uc, double* vc,
                           const double* u, const double* v, double* ua, double* va,
                           double* utc, double* vtc, double* work, int64_t work_slab,
                           void* stream);

/*
 * CGridShallowWaterDynamics.__call__ (fv3core/stencils/c_sw.py:534-659):
 * advances the C-grid winds uc, vc by dt2 and returns the half-step delp and
 * pt in delpc, ptc (zero outside [isc-1, iec+1]^2 in the full domain).  delp,
 * pt and w get the reference's in-place corner fills; omga receives the
 * transported w; divgd is written on the compute corners when nord > 0.
 * work: 7 slabs (utmp, vtmp, fx, fx1, fx2, ke, vort).
 */
int32_t fv3_c_sw_f64(const fv3_domain* d, const fv3_metrics* g, int32_t nord, double dt2,
                     double* delp, double* pt, const double* u, const double* v, double* w,
                     double* uc, double* vc, double* ua, double* va, double* ut, double* vt,
                     double* divgd, double* omga, double* delpc, double* ptc, double* work,
                     int64_t work_slab, void* stream);
/* The same with FP32 storages and metric terms, in FP32 arithmetic (the bulk
 * stencils of the mixed-precision acoustic step). */
int32_t fv3_c_sw_f32(const fv3_domain* d, const fv3_metrics* g, int32_t nord, float dt2,
                     float* delp, float* pt, const float* u, const float* v, float* w, float* uc,
                     float* vc, float* ua, float* va, float* ut, float* vt, float* divgd,
                     float* omga, float* delpc, float* ptc, float* work, int64_t work_slab,
                     void* stream);

/*
 * AGrid2BGridFourthOrder.__call__ (fv3core/stencils/a2b_ord4.py:617-695):
 * qout <- qin interpolated from cell centres to the corners [isc, iec+1]^2
 * (and qin <- qout there when replace != 0).  work: 3 slabs (qx, qy, edge
 * values).  Needs dxa, dya, agrid1/2, bgrid1/2 and edge_w/e/s/n.
 */
int32_t fv3_a2b_ord4_f64(const fv3_domain* d, const fv3_metrics* g, int32_t replace, double* qin,
                         double* qout, double* work, int64_t work_slab, void* stream);

/*
 * Namelist values and per-level columns of DGridShallowWaterLagrangianDynamics
 * (fv3core/stencils/d_sw.py:606-842, columns from get_column_namelist
 * d_sw.py:512-563).  Column pointers are device arrays of d->nk entries in
 * the storage type of the call: double for fv3_d_sw_f64, float for
 * fv3_d_sw_f32.
 */
typedef struct fv3_dsw_params {
  int32_t hord_dp, hord_tm, hord_mt, hord_vt;
  int32_t hydrostatic;                         /* nonzero: no w transport */
  int32_t nord_v4[4], nord_w4[4], nord_t4[4];  /* nord_v/w/t for k = 0, 1, 2, >= 3 */
  int32_t div_k0;                              /* first level with nord > 0 */
  int32_t div_nord;                            /* divergence-damping order there */
  int32_t transport_fp32;  /* nonzero: the fv_tp_2d calls use fv3_fvtp2d_mixed_f64
                              (FP32 arithmetic on the FP64 storages) */
  double dddmp, da_min_c;
  double dd8;                                  /* (da_min_c * d4_bg) ** (div_nord + 1) */
  double d_con;                                /* namelist d_con */
  const void* damp_vt_deln;  /* DelnFlux damping of delp and pt (calc_damp, da_min) */
  const void* damp_t_deln;   /* DelnFlux damping of q_con */
  const void* damp_w_nosg;   /* DelnFluxNoSG damping of w (calc_damp, da_min_c) */
  const void* damp_v_nosg;   /* DelnFluxNoSG damping of the vorticity */
  const void* damp_w;        /* columns damp_w, damp_vt, ke_bg, d_con, d2_divg */
  const void* damp_vt;
  const void* ke_bg;
  const void* d_con_col;
  const void* d2_divg;
} fv3_dsw_params;

/* The storages d_sw reads and updates (same layout, nk + 1 levels each). */
typedef struct fv3_dsw_state {
  double *delpc, *delp, *ptc, *pt, *u, *v, *w, *uc, *vc, *ua, *va, *divgd;
  double *mfx, *mfy, *cx, *cy, *crx, *cry, *xfx, *yfx, *q_con, *zh, *heat_source, *diss_est;
} fv3_dsw_state;
typedef struct fv3_dsw_state_f32 {
  float *delpc, *delp, *ptc, *pt, *u, *v, *w, *uc, *vc, *ua, *va, *divgd;
  float *mfx, *mfy, *cx, *cy, *crx, *cry, *xfx, *yfx, *q_con, *zh, *heat_source, *diss_est;
} fv3_dsw_state_f32;

/*
 * DGridShallowWaterLagrangianDynamics.__call__ (fv3core/stencils/d_sw.py:843-1131):
 * one full D-grid step of dt on levels [0, d->nk) (transport on nk + 1).
 * Writes the reference's in-place updates of every state member (uc, vc,
 * delpc and ptc are used as scratch by the divergence damping, as there).
 * work: 22 slabs of work_slab elements laid out like the fields (nk + 1 levels).
 */
int32_t fv3_d_sw_f64(const fv3_domain* d, const fv3_metrics* g, const fv3_dsw_params* p,
                     const fv3_dsw_state* st, double dt, double* work, int64_t work_slab,
                     void* stream);
/* The same with FP32 storages, metric terms and columns, in FP32 arithmetic
 * (the bulk stencils of the mixed-precision acoustic step); zh may be NULL
 * (d_sw does not read it); transport_fp32 is ignored. */
int32_t fv3_d_sw_f32(const fv3_domain* d, const fv3_metrics* g, const fv3_dsw_params* p,
                     const fv3_dsw_state_f32* st, float dt, float* work, int64_t work_slab,
                     void* stream);

/*
 * UpdateGeopotentialHeightOnCGrid.__call__ (fv3core/stencils/updatedzc.py:122-163):
 * gz (d->nk interface levels) advected by the C-grid winds ut, vt on
 * [isc-1, iec+1]^2 with the DZ_MIN floor; ws (IJ) = (zs - gz_surface) / dt.
 * dp_ref: device column of layer thicknesses.  work: 1 slab.
 */
int32_t fv3_update_dz_c_f64(const fv3_domain* d, const double* area, const double* zs,
                            const double* dp_ref, const double* ut, const double* vt,
                            double* gz, double* ws, double dt, double* work, int64_t work_slab,
                            void* stream);
/* The same with FP32 ut, vt (the mixed-precision acoustic step's storages,
 * widened exactly on load; gz, ws and the arithmetic stay FP64). */
int32_t fv3_update_dz_c_mixed(const fv3_domain* d, const double* area, const double* zs,
                              const double* dp_ref, const float* ut, const float* vt,
                              double* gz, double* ws, double dt, double* work,
                              int64_t work_slab, void* stream);

/*
 * UpdateHeightOnDGrid.__call__ (fv3core/stencils/updatedzd.py:323-393) on
 * d->nk interface levels: spline interpolation of crx, xfx, cry, yfx to the
 * interfaces (device columns gk, beta, gamma from the reference's
 * cubic_spline_interpolation_constants), fv_tp_2d of height (hord_tm),
 * DelnFluxNoSG damping (nord4, per-level damp), the flux update with the
 * DZ_MIN floor and ws.  work: 9 slabs.
 */
int32_t fv3_update_dz_d_f64(const fv3_domain* d, const fv3_metrics* g, int32_t hord_tm,
                            const int32_t nord4[4], const double* damp, const double* gk,
                            const double* beta, const double* gamma, const double* zs,
                            double* height, const double* crx, const double* cry,
                            const double* xfx, const double* yfx, double* ws, double dt,
                            double* work, int64_t work_slab, void* stream);
/* The same with FP32 crx, cry, xfx, yfx (widened exactly on load; the
 * metric terms g, height and the arithmetic stay FP64). */
int32_t fv3_update_dz_d_mixed(const fv3_domain* d, const fv3_metrics* g, int32_t hord_tm,
                              const int32_t nord4[4], const double* damp, const double* gk,
                              const double* beta, const double* gamma, const double* zs,
                              double* height, const float* crx, const float* cry,
                              const float* xfx, const float* yfx, double* ws, double dt,
                              double* work, int64_t work_slab, void* stream);

/*
 * NonHydrostaticPressureGradient.__call__ (fv3core/stencils/nh_p_grad.py:147-206)
 * on d->nk layers: pp, pk3 (levels 1..nk) and gz (0..nk) replaced by their
 * B-grid values, then u, v updated (and multiplied by rdx, rdy).
 * ptk = ptop ** akap.  work: 6 slabs of nk + 1 levels.
 */
int32_t fv3_nh_p_grad_f64(const fv3_domain* d, const fv3_metrics* g, double* u, double* v,
                          double* pp, double* gz, double* pk3, const double* delp, double dt,
                          double ptk, double* work, int64_t work_slab, void* stream);
/* Mixed-precision form (no counterpart in the reference, which is FP64
 * only): u, v FP32 storages, everything else and the arithmetic FP64. */
int32_t fv3_nh_p_grad_mixed(const fv3_domain* d, const fv3_metrics* g, float* u, float* v,
                            double* pp, double* gz, double* pk3, const double* delp, double dt,
                            double ptk, double* work, int64_t work_slab, void* stream);

/* PK3Halo.__call__ (fv3core/stencils/pk3_halo.py:11-68): pk3 = pe ** akap on
 * the 2-wide ring around the compute domain, levels 1..d->nk-1. */
int32_t fv3_pk3_halo_f64(const fv3_domain* d, double* pk3, const double* delp, double ptop,
                         double akap, void* stream);

/* edge_pe (fv3core/stencils/pe_halo.py:6-33): pe on the 1-wide ring. */
int32_t fv3_pe_halo_f64(const fv3_domain* d, double* pe, const double* delp, double ptop,
                        void* stream);

/* p_grad_c_stencil (fv3core/stencils/dyn_core.py:101-145) on [isc, iec+1]^2. */
int32_t fv3_p_grad_c_f64(const fv3_domain* d, const double* rdxc, const double* rdyc,
                         int32_t hydrostatic, double* uc, double* vc, const double* delpc,
                         const double* pkc, const double* gz, double dt2, void* stream);
/* Mixed-precision form: uc, vc, delpc FP32 storages; pkc, gz, the metric
 * terms and the arithmetic FP64. */
int32_t fv3_p_grad_c_mixed(const fv3_domain* d, const double* rdxc, const double* rdyc,
                           int32_t hydrostatic, float* uc, float* vc, const float* delpc,
                           const double* pkc, const double* gz, double dt2, void* stream);

/*
 * RayleighDamping.__call__ (fv3core/stencils/ray_fast.py:130-192) with the
 * per-level factors prepared on the host: rf, dp, dm (column mass of the
 * nudged levels) and the masks pfull < rf_cutoff, pfull < rf_cutoff_nudge.
 */
int32_t fv3_ray_fast_f64(const fv3_domain* d, int32_t hydrostatic, const double* rf,
                         const double* dp, const double* dm, const int32_t* use_rf,
                         const int32_t* use_nudge, double* u, double* v, double* w,
                         void* stream);
int32_t fv3_ray_fast_f32(const fv3_domain* d, int32_t hydrostatic, const float* rf,
                         const float* dp, const float* dm, const int32_t* use_rf,
                         const int32_t* use_nudge, float* u, float* v, float* w, void* stream);

/* HyperdiffusionDamping.__call__ (fv3core/stencils/del2cubed.py:158-181):
 * min(3, nmax) del-2 passes on qdel with coefficient cd.  work: 2 slabs. */
int32_t fv3_del2cubed_f64(const fv3_domain* d, const double* del6_u, const double* del6_v,
                          const double* rarea, int32_t nmax, double cd, double* qdel,
                          double* work, int64_t work_slab, void* stream);
int32_t fv3_del2cubed_f32(const fv3_domain* d, const float* del6_u, const float* del6_v,
                          const float* rarea, int32_t nmax, float cd, float* qdel, float* work,
                          int64_t work_slab, void* stream);

/* compute_pkz_tempadjust (fv3core/stencils/temperature_adjust.py:8-44) on
 * the compute domain, levels [0, d->nk). */
int32_t fv3_pkz_tempadjust_f64(const fv3_domain* d, const double* delp, const double* delz,
                               const double* cappa, const double* heat_source, double* pt,
                               double* pkz, double delt_time_factor, void* stream);
/* Mixed-precision form: delp, cappa, heat_source, pt FP32 storages; delz,
 * pkz and the arithmetic FP64. */
int32_t fv3_pkz_tempadjust_mixed(const fv3_domain* d, const float* delp, const double* delz,
                                 const float* cappa, const float* heat_source, float* pt,
                                 double* pkz, double delt_time_factor, void* stream);

/*
 * Element-wise conversion of n elements of one dense storage (no counterpart
 * in the reference, which is FP64 only): the mixed-precision acoustic step
 * hands its FP32 fields to the FP64 column solvers and back through these.
 */
int32_t fv3_cast_f32_to_f64(int64_t n, const float* src, double* dst, void* stream);
int32_t fv3_cast_f64_to_f32(int64_t n, const double* src, float* dst, void* stream);

/*
 * Small stencils of dyn_core (fv3core/stencils/dyn_core.py:47-98) by op:
 * 0 set_gz (out = gz, a = zs IJ, b = delz), 1 set_pem (out = pem, a = delp,
 * scalar = ptop), 2 compute_geopotential (out = a * GRAV on compute + 2),
 * 3 copy a -> out on the full domain, 4 fill out with scalar on the full
 * domain; d->nk levels.
 */
int32_t fv3_dyncore_column_f64(const fv3_domain* d, int32_t op, double scalar, const double* a,
                               const double* b, double* out, void* stream);
int32_t fv3_dyncore_column_f32(const fv3_domain* d, int32_t op, double scalar, const float* a,
                               const float* b, float* out, void* stream);

/* ---- fast saturation adjustment ----------------------------------------- */

/* Namelist values and per-call scalars of SatAdjust3d
 * (fv3core/stencils/saturation_adjustment.py:899-1024). */
typedef struct fv3_satadj_params {
  int32_t hydrostatic, rad_snow, rad_rain, rad_graupel, tintqs, icloud_f;
  int32_t do_qa, consv_te, last_step;
  double sat_adj0, ql_gen, qs_mlt, ql0_max, t_sub, qi_gen, qi_lim, qi0_max;
  double dw_ocean, dw_land, cld_min;
  double sdt, zvir, fac_i2s, c_air, c_vap, mdt, fac_r2g, fac_smlt, fac_l2r, fac_imlt;
  double d0_vap, lv00, fac_v2l, fac_l2v;
} fv3_satadj_params;

/*
 * The saturation vapour-pressure tables of saturation_adjustment.py:43-168
 * (qs_tablew_fn, qs_table2_fn, desw_table, des2_table) for the table indices
 * -1 .. 2620: four consecutive arrays of fv3_sat_adjust_table_entries()
 * doubles, entry e holding index e - 1.  Each value is the closed form the
 * stencil evaluates at that index.
 */
int32_t fv3_sat_adjust_table_entries(void);
int32_t fv3_sat_adjust_tables_f64(double* tables, void* stream);

/*
 * SatAdjust3d.__call__ (saturation_adjustment.py:936-1024): the satadjust
 * stencil (:554-896) on [isc, iec] x [jsc, jec] x [kmp, d->nk), FP64.
 * In/out: qv, ql, qi, qr, qs, qg (qvapor .. qgraupel), qa (qcld, written on
 * the last step), pt, q_con, cappa, pkz (nonhydrostatic), te (written when
 * consv_te), delz (shifted down one level in the hydrostatic form, as the
 * stencil does).  area, hs: IJ; peln: d->nk + 1 levels (hydrostatic only).
 */
int32_t fv3_sat_adjust_f64(const fv3_domain* d, int32_t kmp, const fv3_satadj_params* p,
                           const double* tables, double* te, double* qv, double* ql, double* qi,
                           double* qr, double* qs, double* qg, double* qa, const double* area,
                           const double* hs, const double* peln, const double* delp,
                           double* delz, double* q_con, double* pt, double* pkz, double* cappa,
                           void* stream);

/* ---- vertical remapping (Lagrangian -> Eulerian) ---------------------- */

/* Most tracers a remapping call handles (fv3core/utils/gt4py_utils.py:26-36). */
#define FV3_MAX_TRACERS 9

/*
 * FillNegativeTracerValues.__call__ / fix_tracer (fv3core/stencils/
 * fillz.py:12-150) on the compute box, levels [0, d->nk): borrows mass
 * from the layers above and below a negative layer, then rescales the
 * column below the top so its total mass is kept.  q: ntracers fields
 * (<= FV3_MAX_TRACERS), dp: layer thicknesses.
 */
int32_t fv3_fillz_f64(const fv3_domain* d, int32_t ntracers, double* const* q, const double* dp,
                      void* stream);

/* Scalars of LagrangianToEulerian (fv3core/_config.py RemappingConfig and
 * the call arguments of remapping.py:416-452). */
typedef struct fv3_remap_params {
  int32_t kord_tm, kord_tr, kord_wz, kord_mt; /* kord_tm < 0 */
  int32_t nq;                                 /* tracers remapped (<= FV3_MAX_TRACERS) */
  int32_t fill;                               /* fillz after the tracer remap */
  int32_t last_step;                          /* moist_pt_last_step instead of pt / pkz */
  int32_t defer_pt;  /* nonzero: leave the final pt step (moist_pt_last_step or pt / pkz) to
                        fv3_remap_final_pt_f64, so that a saturation adjustment runs first */
  double ptop, akap, zvir;
} fv3_remap_params;

/* Fields of LagrangianToEulerian.__call__ (remapping.py:416-452), in the
 * reference's tracer order qvapor, qliquid, qrain, qice, qsnow, qgraupel,
 * qo3mr, qsgs_tke, qcld.  The first six tracers are always read (moist
 * thermodynamics); the others only when n < nq.  ps and wsd are IJ fields;
 * ak and bk hold d->nk + 1 values. */
typedef struct fv3_remap_state {
  double* tracers[FV3_MAX_TRACERS];
  double *pt, *delp, *delz, *peln, *u, *v, *w, *ua, *cappa, *q_con, *pkz, *pk, *pe;
  double* ps;
  const double *wsd, *ak, *bk;
} fv3_remap_state;

/*
 * LagrangianToEulerian.__call__ (fv3core/stencils/remapping.py:416-628)
 * for kord_tm < 0, non-hydrostatic, nwat = 6 (the saturation adjustment,
 * when on, runs between this call with p->defer_pt set and
 * fv3_remap_final_pt_f64):
 * remaps pt (log-pressure, T >= 184 K), the nq tracers (then fillz), w,
 * delz, u and v from the Lagrangian interfaces pe to the hybrid levels
 * ak + bk * ps, updates delp, peln, pk, pkz, cappa, q_con, ps, pe, ua.
 * d->nk = number of layers (>= 6); every 3-D field holds d->nk + 1 levels.
 * work: 15 slabs of work_slab elements, each laid out like the fields.
 */
int32_t fv3_lagrangian_to_eulerian_f64(const fv3_domain* d, const fv3_remap_params* p,
                                       const fv3_remap_state* s, double* work,
                                       int64_t work_slab, void* stream);

/* The final pt step of LagrangianToEulerian.__call__ (remapping.py:613-628):
 * moist_pt_last_step (moist_cv.py:139-162, dtmp = 0) on the last step, else
 * pt / pkz, on the compute domain; for calls made with p->defer_pt. */
int32_t fv3_remap_final_pt_f64(const fv3_domain* d, const fv3_remap_params* p,
                               const fv3_remap_state* s, void* stream);

/* ---- full dynamics step (fv_dynamics) --------------------------------- */

/*
 * compute_preamble (fv3core/stencils/fv_dynamics.py:54-106): fv_setup
 * (moist_cv.py:197-228, moist_phys, nwat = 6) then pt_adjust
 * (fv_dynamics.py:32-34) on the compute domain, levels [0, d->nk).
 * q6: qvapor, qliquid, qrain, qice, qsnow, qgraupel.  Writes cvm, q_con,
 * dp1, cappa, pkz and pt.
 */
int32_t fv3_fv_setup_f64(const fv3_domain* d, const double* const* q6, const double* delp,
                         const double* delz, double* pt, double* q_con, double* cvm, double* pkz,
                         double* cappa, double* dp1, void* stream);

/* set_omega (fv3core/stencils/fv_dynamics.py:37-39): omga = delp / delz * w
 * on the compute domain, levels [0, d->nk). */
int32_t fv3_set_omega_f64(const fv3_domain* d, const double* delp, const double* delz,
                          const double* w, double* omga, void* stream);

/* ---- the step's wrap-up ----------------------------------------------- */

/*
 * AdjustNegativeTracerMixingRatio.__call__ (fv3core/stencils/neg_adj3.py:
 * 335-372), nonhydrostatic, check_negative off: fix_neg_water, fillq on
 * graupel then rain, fix_water_vapor_down and fix_neg_cloud on the compute
 * box, layers [0, d->nk) (d->nk >= 3).  Updates the seven water species
 * and pt in place; delp is read only.
 */
int32_t fv3_neg_adj3_f64(const fv3_domain* d, double* qvapor, double* qliquid, double* qrain,
                         double* qsnow, double* qice, double* qgraupel, double* qcld, double* pt,
                         const double* delp, void* stream);

/*
 * CubedToLatLon's stencil (fv3core/stencils/c2l_ord.py:17-72, called at
 * :118-141): D-grid winds u, v (halos already filled) to A-grid
 * latitude-longitude winds ua, va with the IJ transform a11..a22
 * (fv3core/grid/geometry.py:596-601).  order 4: ord4_transform on the
 * compute box; order 2: c2l_ord2 on the compute box grown by one.
 * Layers [0, d->nk).
 */
int32_t fv3_cubed_to_latlon_f64(const fv3_domain* d, int32_t order, const double* u,
                                const double* v, const double* dx, const double* dy,
                                const double* a11, const double* a12, const double* a21,
                                const double* a22, double* ua, double* va, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FV3CORE_AMD_H */
