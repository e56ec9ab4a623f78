// Host side of the corner-fill tables (fv3core/utils/corners.py:127-300).
#include "stencil.hpp"

namespace fv3 {

namespace {

// One fill statement: corner (0 sw, 1 se, 2 nw, 3 ne), destination offset
// from the (i, j) anchors of that corner, source offset from the destination.
struct FillRow {
  int corner, di, dj, sdi, sdj;
};

// fill_corners_2cells_mult_x / 3cells_mult_x, statement order (corners.py:127-204)
const FillRow k2X[] = {{0, -1, -1, 0, 1},  {0, -2, -1, 1, 2},  {1, 1, -1, 0, 1},
                       {1, 2, -1, -1, 2},  {2, -1, 1, 0, -1},  {2, -2, 1, 1, -2},
                       {3, 1, 1, 0, -1},   {3, 2, 1, -1, -2}};
const FillRow k3X[] = {{0, -3, -1, 2, 3}, {1, 3, -1, -2, 3}, {2, -3, 1, 2, -3},
                       {3, 3, 1, -2, -3}};
// fill_corners_2cells_mult_y / 3cells_mult_y (corners.py:207-300)
const FillRow k2Y[] = {{0, -1, -1, 1, 0},  {0, -1, -2, 2, 1},  {1, 1, -1, -1, 0},
                       {1, 1, -2, -2, 1},  {2, -1, 1, 1, 0},   {2, -1, 2, 2, -1},
                       {3, 1, 1, -1, 0},   {3, 1, 2, -2, -1}};
const FillRow k3Y[] = {{0, -1, -3, 3, 2}, {1, 1, -3, -3, 2}, {2, -1, 3, 3, -2},
                       {3, 1, 3, -3, -2}};

bool add_rows(FillTable* tab, const fv3_domain* d, const FillRow* rows, int nrows, int dst_f,
              int src_f, const int mult[4], Box lim) {
  for (int r = 0; r < nrows; ++r) {
    const FillRow& w = rows[r];
    const int ai = (w.corner == 0 || w.corner == 2) ? d->i_start : d->i_end;
    const int aj = (w.corner == 0 || w.corner == 1) ? d->j_start : d->j_end;
    if (!has_edge(ai) || !has_edge(aj)) continue;
    const int i = ai + w.di, j = aj + w.dj;
    if (i < lim.i0 || i > lim.i1 || j < lim.j0 || j > lim.j1) continue;
    const int si = i + w.sdi, sj = j + w.sdj;
    if (i < 0 || j < 0 || si < 0 || sj < 0 || i >= d->ni || si >= d->ni || j >= d->nj ||
        sj >= d->nj)
      continue;
    if (tab->n >= kMaxFill) return false;
    const int e = tab->n++;
    tab->dst_i[e] = i;
    tab->dst_j[e] = j;
    tab->src_i[e] = si;
    tab->src_j[e] = sj;
    tab->dst_f[e] = (signed char)dst_f;
    tab->src_f[e] = (signed char)src_f;
    tab->mult[e] = (signed char)mult[w.corner];
  }
  return true;
}

}  // namespace

bool append_fill(FillTable* tab, const fv3_domain* d, FillKind kind, int dst_f, int src_f,
                 const int mult[4], Box lim) {
  const bool x = kind == kFill2X || kind == kFill3X;
  const bool three = kind == kFill3X || kind == kFill3Y;
  // the 3-cell variants run the 2-cell statements first (corners.py:166-204)
  if (!add_rows(tab, d, x ? k2X : k2Y, 8, dst_f, src_f, mult, lim)) return false;
  if (three && !add_rows(tab, d, x ? k3X : k3Y, 4, dst_f, src_f, mult, lim)) return false;
  return true;
}

}  // namespace fv3
