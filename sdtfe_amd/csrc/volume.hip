// volume.hip -- the volumetric grid (interpolate_to_3d_grid and
// compute_3d_density, src/dtfe.c:376-516, 657-735) on gfx950.
//
// Column (i, j) of the grid (x index i, y index j) owns the grid^3 array's
// slice [(i*G + j)*G, +G): its ss x ss rays, in (is, js) order, walk the
// tessellation bottom to top and add, in each tet crossed, ss_weight times
// the tet's linear field at every fine z point the tet holds.  One lane
// owns one column and walks its rays in that order, so every cell receives
// its terms in the reference's order and the grid is the reference's bit
// for bit (this file is built with -ffp-contract=off).  A wave takes an
// 8 x 8 patch of columns, whose rays cross mostly the same tets.
#include "sdtfe_host.h"
#include "walk_fp64_core.h"

namespace sdtfe {

// Fine points of one tet: k in [ceil((z_in - off_z - dzs/2) / dzs),
// ceil((z_out - off_z - dzs/2) / dzs)) clamped to [0, nzs) (src/dtfe.c:
// 482-490).  The clamps act on the double before its conversion, which
// gives the reference's indices wherever they fit an int.
SDTFE_HD inline void fine_range(const VolGeom &v, double z_in, double z_out, int &e, int &l) {
  const double ce = ceil((z_in - v.off_z - v.dzs / 2.) / v.dzs);
  const double cl = ceil((z_out - v.off_z - v.dzs / 2.) / v.dzs);
  e = ce < 0.0 ? 0 : (ce > (double)v.nzs ? v.nzs : (int)ce);
  l = cl > (double)v.nzs ? v.nzs : (cl < 0.0 ? 0 : (int)cl);
}

// evaluate_3d_ray_on_grid (src/dtfe.c:376-516): the walk of shoot_3d_ray
// with its perturbation; col is the ray's column.  The column's current
// cell is held in a register (c_iz, c_val) and written back when the walk
// moves to another cell, so each cell is read and written once per visit.
// A failed walk keeps what it added, as the reference does.  Returns the
// tets crossed (0 on failure).
SDTFE_HD inline int column_walk64(const Rec64 *__restrict__ rec, int entry, double x, double y,
                                  const VolGeom &v, uint64_t key, double *col,
                                  LaneCounters &cnt) {
  double ox = x, oy = y;
  uint32_t draw = 2;
  int cur = entry, prev = -1, f_in = -1, f_out = 0, nxt = entry, steps = 0, crossed = 0;
  int c_iz = -1;
  double c_val = 0.0;
  bool first = true;
  Tet64 r;
  for (;;) {
    if (!first) nxt = nb_of(r, f_out);
    first = false;
    if (nxt == -1) break;
    if (nxt == prev) {  // turn around (src/dtfe.c:433-439)
      const int tmp = f_in;
      f_in = f_out;
      f_out = tmp;
      nxt = nb_of(r, f_out);
      if (nxt == -1) break;
    }
    if (++steps > v.max_steps) {
      cnt.step_cap++;
      crossed = 0;
      break;
    }
    prev = cur;
    cur = nxt;
    load_tet(rec, cur, r);
    Hit h = plucker64(r, v.lz, ox, oy);
    if (!h.ok) {
      cnt.no_isect++;
      crossed = 0;
      break;
    }
    if (h.special && !resolve_special(r, v.lz, v.pert_step, key, draw, ox, oy, h, cnt)) {
      crossed = 0;
      break;
    }
    f_in = h.enter;
    f_out = h.leave;
    ++crossed;
    int e, l;
    fine_range(v, h.z_in, h.z_out, e, l);
    if (l > e) {
      const double fixed = r.rho_a + r.gx * (ox - r.ax) + r.gy * (oy - r.ay) - r.gz * r.az;
      for (int k = e; k < l; ++k) {
        const int iz = k / v.ss;
        const double z = v.off_z + v.dzs * (double)k + v.dzs / 2.;
        if (iz != c_iz) {
          if (c_iz >= 0) col[c_iz] = c_val;
          c_iz = iz;
          c_val = col[iz];
        }
        c_val = c_val + (fixed + r.gz * z) * v.ss_weight;
      }
    }
  }
  if (c_iz >= 0) col[c_iz] = c_val;
  return crossed;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void
volume_fp64_kernel(const Rec64 *__restrict__ rec, EntryGrid eg, VolGeom v, double *grid,
                   unsigned long long *stats) {
  // 16 x 16 columns per block, an 8 x 8 patch per wave
  const int tiles = (v.grid + 15) / 16;
  const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  const int i = (b / tiles) * 16 + (w >> 1) * 8 + (lane >> 3);
  const int j = (b % tiles) * 16 + (w & 1) * 8 + (lane & 7);
  LaneCounters cnt;
  if (i < v.grid && j < v.grid) {
    const uint64_t c = (uint64_t)i * (uint64_t)v.grid + (uint64_t)j;
    double *col = grid + c * (uint64_t)v.grid;
    for (int is = 0; is < v.ss; ++is)
      for (int js = 0; js < v.ss; ++js) {
        // src/dtfe.c:709-710
        const double x = (double)i * v.dxy + v.off_x + (double)is * v.ds + v.ds / 2.;
        const double y = (double)j * v.dxy + v.off_y + (double)js * v.ds + v.ds / 2.;
        cnt.samples++;
        const int entry = locate_entry(eg, x, y);
        if (entry < 0) {
          cnt.missed++;
          continue;
        }
        const uint64_t key = sample_key(v.seed, c, (uint32_t)(is * v.ss + js));
        const int crossed = column_walk64(rec, entry, x, y, v, key, col, cnt);
        cnt.tet_steps += (unsigned)crossed;
        if (crossed == 0) cnt.failed++;
      }
  }
  cnt.flush(stats);
}

void launch_volume_fp64(sdtfe_tess *t, const VolGeom &v, double *grid) {
  const int tiles = (v.grid + 15) / 16;
  volume_fp64_kernel<<<tiles * tiles, 256, 0, t->stream>>>(t->rec64, t->grid, v, grid,
                                                          t->stats);
}

}  // namespace sdtfe
