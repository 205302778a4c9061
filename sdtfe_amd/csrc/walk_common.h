// walk_common.h -- device code shared by the walk kernels: entry lookup,
// sample placement and per-lane counters.
#pragma once

#include <climits>

#include "sdtfe_internal.h"

namespace sdtfe {

SDTFE_HD inline int grid_cell(double v, double v0, double scale, int n) {
  int c = (int)floor((v - v0) * scale);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// Point-in-triangle with the reference's inclusive convention
// (src/dtfe.c:528-545): outside only if the query and the opposite vertex
// lie strictly on opposite sides of an edge.  Same expressions, same order.
SDTFE_HD inline bool tri_contains(const double *xy, double qx, double qy) {
  const double ax = xy[0], ay = xy[1], bx = xy[2], by = xy[3], cx = xy[4], cy = xy[5];
  double d_ins = (cx - ax) * (by - ay) - (cy - ay) * (bx - ax);
  double d_qry = (qx - ax) * (by - ay) - (qy - ay) * (bx - ax);
  if (d_ins * d_qry < 0.0) return false;
  d_ins = (ax - bx) * (cy - by) - (ay - by) * (cx - bx);
  d_qry = (qx - bx) * (cy - by) - (qy - by) * (cx - bx);
  if (d_ins * d_qry < 0.0) return false;
  d_ins = (bx - cx) * (ay - cy) - (by - cy) * (ax - cx);
  d_qry = (qx - cx) * (ay - cy) - (qy - cy) * (ax - cx);
  return !(d_ins * d_qry < 0.0);
}

// Entry tet of the vertical ray through (x, y): the forward-hull triangle
// containing the point (lowest triangle number if it lies on a shared edge),
// -1 outside the hull's shadow.  Replaces the visibility walk of pt_loc_2d
// (src/dtfe.c:518-552); both return the same tet except for points on a
// shared edge, where either neighbour is a valid entry.
SDTFE_HD inline int locate_tri(const EntryGrid &g, double x, double y) {
  if (g.n_tri == 0) return -1;
  if (!(x >= g.x0 && x <= g.x1 && y >= g.y0 && y <= g.y1)) return -1;
  const int c = grid_cell(y, g.y0, g.sy, g.ny) * g.nx + grid_cell(x, g.x0, g.sx, g.nx);
  const int b = g.cell_off[c], e = g.cell_off[c + 1];
  int best = INT_MAX;
  for (int i = b; i < e; ++i) {
    const int k = g.cell_tri[i];
    if (k < best && tri_contains(g.tri_xy + 6 * (int64_t)k, x, y)) best = k;
  }
  return best == INT_MAX ? -1 : best;
}

SDTFE_HD inline int locate_entry(const EntryGrid &g, double x, double y) {
  const int k = locate_tri(g, x, y);
  return k < 0 ? -1 : g.tri_tet[k];
}

// The same entry as the face record the ray enters by (4 x tet + face).
SDTFE_HD inline int locate_entry_rid(const EntryGrid &g, double x, double y) {
  const int k = locate_tri(g, x, y);
  return k < 0 ? -1 : g.tri_rid[k];
}

// Lower-left corner of pixel (px, py)'s sampling square and one sample in
// it (src/dtfe.c:613-622: x drawn before y, r / RAND_MAX in [0, 1]).
SDTFE_HD inline void sample_point(const MapGeom &g, int px, int py, uint64_t key, double &xt,
                                    double &yt) {
  const double x = px * g.dxy + g.off_x - g.soff;
  const double y = py * g.dxy + g.off_y - g.soff;
  xt = x + ((double)draw31(key, 0) / 2147483647.0) * g.scale;
  yt = y + ((double)draw31(key, 1) / 2147483647.0) * g.scale;
}

struct LaneCounters {
  // per lane between flushes (a pixel, or a batch tile): 32 bits hold the
  // sample counts; tets crossed take 64
  unsigned samples = 0, missed = 0, failed = 0, no_isect = 0, pert_exh = 0, no_isect_pert = 0;
  unsigned pert_steps = 0, step_cap = 0, fp64_redo = 0;
  unsigned long long tet_steps = 0;

  __device__ void flush(unsigned long long *st) {
    add_stat(st, ST_SAMPLES, samples);
    add_stat(st, ST_MISSED, missed);
    add_stat(st, ST_FAILED, failed);
    add_stat(st, ST_NO_ISECT, no_isect);
    add_stat(st, ST_PERT_EXHAUSTED, pert_exh);
    add_stat(st, ST_NO_ISECT_PERT, no_isect_pert);
    add_stat(st, ST_PERT_STEPS, pert_steps);
    add_stat(st, ST_TET_STEPS, tet_steps);
    add_stat(st, ST_STEP_CAP, step_cap);
    add_stat(st, ST_FP64_REDO, fp64_redo);
  }
};

// v as hi + lo in FP32: hi = v rounded, lo = (v - hi) rounded (v - hi is
// exact in FP64), so hi + lo is within 2^-48 |v| of v.
SDTFE_HD inline void split32(double v, float &hi, float &lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}

// Wave tile of L lanes per pixel (1, 2, 4, 8, 16): 64 / L pixels as
// wave_bw x wave_bh, the L lanes of a pixel adjacent.
SDTFE_HD constexpr int wave_bw(int L) { return L <= 2 ? 8 : L <= 8 ? 4 : 2; }
SDTFE_HD constexpr int wave_bh(int L) { return L == 1 ? 8 : L <= 4 ? 4 : 2; }

// Lanes per pixel of a single map: the rays of a wave then start together
// on a 2x2 (L = 16) or 4x4 (L = 4) pixel patch and share more records.
// The mixed walk's hand-over window bounds the samples a split pixel may
// have: 64 with 16 lanes (one mask word), kDeferWindow with 4.
SDTFE_HD constexpr int map_lanes(int samples) {
  return samples > kDeferWindow ? 1 : samples > 64 ? 4 : samples >= 16 ? 16 : samples >= 4 ? 4 : 1;
}

// Mask words of the hand-over window a map kernel with L lanes per pixel
// keeps (map_lanes picks L accordingly).
SDTFE_HD constexpr int map_window_words(int L) { return L == 4 ? 4 : 1; }

// Pixel and sample slice j (0..L-1) of a lane with L lanes per pixel:
// blocks of 256 lanes cover 2 x 2 wave tiles; block ids are remapped so
// that each XCD takes whole rows of tiles in turn.
template <int L>
__device__ inline void split_pixel(int grid, int &px, int &py, int &j) {
  constexpr int bw = wave_bw(L), bh = wave_bh(L);
  const int nbx = (grid + 2 * bw - 1) / (2 * bw);
  const int tile = xcd_tile(blockIdx.x, gridDim.x, nbx);
  const int bx = tile % nbx, by = tile / nbx;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane / L;
  j = lane % L;
  px = bx * 2 * bw + (w & 1) * bw + q % bw;
  py = by * 2 * bh + (w >> 1) * bh + q / bw;
}

}  // namespace sdtfe
