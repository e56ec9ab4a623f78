// Launchers shared between translation units of the library.
#pragma once

#include "stencil.hpp"

namespace fv3 {

// AGrid2BGridFourthOrder.__call__ (a2b_ord4.py:617-695) on levels
// [k0, k0 + nk): qout on the corners [isc, iec+1]^2; qx (whose head holds
// the 12 corner extrapolation weights per tile) and edges are scratch slabs
// laid out like the fields.  With replace set, qin_w (the writable qin)
// receives qout as well.
template <typename T>
int launch_a2b_ord4(hipStream_t s, const fv3_domain* dom, const Met<T>& m, int k0, int nk,
                    const T* qin, T* qout, T* qx, T* edges, int replace, T* qin_w,
                    const char* who);

int check_a2b_metrics(const fv3_metrics* g, const char* who);

// fv_tp_2d of q after CopyCornersXY(q, q_x = q, q_y) (corners.py:425-540),
// with the y corner fill applied as the transport reads q instead of into a
// copy (qy_scratch is used only when the x and y corner sets differ).  q
// receives the x corner fill in place, as CopyCornersXY does with q_x = q.
int fvtp2d_corner_fill_f64(const fv3_domain* dom, int hord, double* q, double* qy_scratch,
                           const double* crx, const double* cry, const double* xfx,
                           const double* yfx, const double* xmf, const double* ymf,
                           const double* area, const double* dxa, const double* dya, double* fx,
                           double* fy, bool fp32_arithmetic, void* stream);

// The same on FP32 storages with FP32 arithmetic.
int fvtp2d_corner_fill_f32(const fv3_domain* dom, int hord, float* q, float* qy_scratch,
                           const float* crx, const float* cry, const float* xfx,
                           const float* yfx, const float* xmf, const float* ymf,
                           const float* area, const float* dxa, const float* dya, float* fx,
                           float* fy, void* stream);

// fv_tp_2d of n scalars (<= FV3_MAX_TRACERS) sharing the Courant numbers,
// area and mass fluxes, in one launch after one launch of their x corner
// fills in place (the corner handling of fvtp2d_corner_fill_f64).  q, fx, fy:
// host tables of n device pointers.  FV3_E_UNSUPPORTED when the domain's x
// and y corner sets differ (then the scalars go one by one).
int fvtp2d_batch_corner_fill_f64(const fv3_domain* dom, int hord, int n, double* const* q,
                                 double* const* fx, double* const* fy, const double* crx,
                                 const double* cry, const double* xfx, const double* yfx,
                                 const double* xmf, const double* ymf, const double* area,
                                 const double* dxa, const double* dya, void* stream);

}  // namespace fv3
