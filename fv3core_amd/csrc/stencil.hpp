// Building blocks of the pointwise stencil kernels of the acoustic step:
// a box launcher that runs a device lambda once per (i, j, k, tile) point,
// typed views of the IJ metric terms, and the ordered corner-fill tables of
// fv3core/utils/corners.py:127-300.
#pragma once

#include <string.h>

#include "common.hpp"

namespace fv3 {

// Inclusive index box in the storage index space.
struct Box {
  int i0, i1, j0, j1;
};

template <typename F>
__global__ void __launch_bounds__(256)
    box_kernel(int i0, int ni, int j0, int nij, int k0, F f) {
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= nij) return;
  f(i0 + n % ni, j0 + n / ni, k0 + (int)blockIdx.y, (int)blockIdx.z);
}

// Runs f(i, j, k, t) for every point of box b, levels [k0, k0 + nk) and
// every tile of the batch.  The caller has checked b against the storage.
template <typename F>
inline void launch_box(hipStream_t s, const Dom& d, Box b, int k0, int nk, F f) {
  const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
  if (ni <= 0 || nj <= 0 || nk <= 0 || d.ntiles <= 0) return;
  dim3 grid((unsigned)((ni * nj + 255) / 256), (unsigned)nk, (unsigned)d.ntiles);
  hipLaunchKernelGGL((box_kernel<F>), grid, dim3(256), 0, s, b.i0, ni, b.j0, ni * nj, k0, f);
}

__device__ __forceinline__ int64_t at3(const Dom& d, int i, int j, int k, int t) {
  return (int64_t)t * d.st + (int64_t)k * d.sk + (int64_t)j * d.sj + i;
}

__device__ __forceinline__ int64_t at2(const Dom& d, int i, int j, int t) {
  return (int64_t)t * d.st2 + (int64_t)j * d.sj2 + i;
}

// Typed copy of fv3_metrics (same member order, every member a pointer).
template <typename T>
struct Met {
  const T *dx, *dy, *dxa, *dya, *dxc, *dyc;
  const T *rdx, *rdy, *rdxa, *rdya, *rdxc, *rdyc;
  const T *area, *rarea, *area_c, *rarea_c;
  const T *sin_sg1, *sin_sg2, *sin_sg3, *sin_sg4;
  const T *cos_sg1, *cos_sg2, *cos_sg3, *cos_sg4;
  const T *cosa, *sina, *cosa_u, *cosa_v, *cosa_s, *sina_u, *sina_v;
  const T *rsin_u, *rsin_v, *rsina, *rsin2;
  const T *fC, *f0, *del6_u, *del6_v, *divg_u, *divg_v;
  const T *agrid1, *agrid2, *bgrid1, *bgrid2;
  const T *edge_w, *edge_e, *edge_s, *edge_n;
};
static_assert(sizeof(Met<double>) == sizeof(fv3_metrics), "fv3_metrics layout");

template <typename T>
inline Met<T> to_met(const fv3_metrics* g) {
  Met<T> m;
  memcpy(&m, g, sizeof(m));
  return m;
}

// True when the domain owns the tile edge the anchor stands for.
inline bool has_edge(int anchor) {
  return anchor > FV3_NO_EDGE_LO / 2 && anchor < FV3_NO_EDGE_HI / 2;
}

// Ordered corner fills: entry e sets field[dst_f][dst] = mult * field[src_f][src]
// on every level, one entry after the other (a source written by an earlier
// entry is read after that write, as in the reference's statement order).
constexpr int kMaxFill = 96;
constexpr int kMaxFillFields = 8;
struct FillTable {
  int n;
  int dst_i[kMaxFill], dst_j[kMaxFill], src_i[kMaxFill], src_j[kMaxFill];
  signed char dst_f[kMaxFill], src_f[kMaxFill];
  signed char mult[kMaxFill];
};

enum FillKind { kFill2X = 0, kFill3X = 1, kFill2Y = 2, kFill3Y = 3 };

// Appends fill_corners_{2,3}cells_mult_{x,y}(q = field dst_f, q_corner =
// field src_f) with multipliers (sw, se, nw, ne), limited to the box lim.
// Returns false when the table is full.
bool append_fill(FillTable* tab, const fv3_domain* d, FillKind kind, int dst_f, int src_f,
                 const int mult[4], Box lim);

template <typename T>
struct FillFields {
  T* f[kMaxFillFields];
};

// Entries grouped into waves: an entry reads or writes a cell that an
// earlier entry writes or reads only in a later wave, so the entries of one
// wave are independent and the statement order of the table is kept.
struct FillWaves {
  int nlev;
  signed char level[kMaxFill];
};

inline FillWaves fill_waves(const FillTable& tab) {
  FillWaves w;
  w.nlev = 0;
  for (int e = 0; e < tab.n; ++e) {
    int lev = 0;
    for (int p = 0; p < e; ++p) {
      const bool raw = tab.src_f[e] == tab.dst_f[p] && tab.src_i[e] == tab.dst_i[p] &&
                       tab.src_j[e] == tab.dst_j[p];
      const bool war = tab.dst_f[e] == tab.src_f[p] && tab.dst_i[e] == tab.src_i[p] &&
                       tab.dst_j[e] == tab.src_j[p];
      const bool waw = tab.dst_f[e] == tab.dst_f[p] && tab.dst_i[e] == tab.dst_i[p] &&
                       tab.dst_j[e] == tab.dst_j[p];
      if ((raw || war || waw) && w.level[p] + 1 > lev) lev = w.level[p] + 1;
    }
    w.level[e] = (signed char)lev;
    if (lev + 1 > w.nlev) w.nlev = lev + 1;
  }
  return w;
}

// One workgroup per (level, tile); thread e applies entry e in its wave.
template <typename T>
__global__ void __launch_bounds__(128)
    fill_table_kernel(Dom d, FillTable tab, FillWaves waves, FillFields<T> ff) {
  const int k = (int)blockIdx.x;
  const int t = (int)blockIdx.y;
  const int e = (int)threadIdx.x;
  const int64_t base = (int64_t)t * d.st + (int64_t)k * d.sk;
  for (int w = 0; w < waves.nlev; ++w) {
    if (e < tab.n && waves.level[e] == w) {
      const T v = ff.f[tab.src_f[e]][base + (int64_t)tab.src_j[e] * d.sj + tab.src_i[e]];
      ff.f[tab.dst_f[e]][base + (int64_t)tab.dst_j[e] * d.sj + tab.dst_i[e]] =
          (T)tab.mult[e] * v;
    }
    __syncthreads();
  }
}

template <typename T>
inline void launch_fill(hipStream_t s, const Dom& d, const FillTable& tab,
                        const FillFields<T>& ff) {
  if (tab.n == 0 || d.nk <= 0 || d.ntiles <= 0) return;
  static_assert(kMaxFill <= 128, "one thread per fill entry");
  const FillWaves waves = fill_waves(tab);
  dim3 grid((unsigned)d.nk, (unsigned)d.ntiles);
  hipLaunchKernelGGL((fill_table_kernel<T>), grid, dim3(128), 0, s, d, tab, waves, ff);
}

}  // namespace fv3
