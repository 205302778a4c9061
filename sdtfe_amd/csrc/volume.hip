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

// Sixteen consecutive cells of a column (one 128-byte line of the grid,
// base a multiple of 16) staged in this lane's slot of LDS: a visit reads
// and writes whole lines, and the cells a ray passes are then added in LDS.
// (With thousands of columns in flight the lines do not stay in L2 between
// a lane's consecutive cells, so cell-by-cell access would fetch each line
// again for every cell.)  Cells at or beyond n are neither read nor
// written.  pairs: every column starts 16-byte aligned (an even grid size
// and an aligned grid), so lines move as 16-byte pieces.
constexpr int kWinCells = 16;
constexpr int kWinStride = 18;  // doubles per lane slot (16-byte aligned, fewer bank conflicts)

struct CellWindow {
  double *w;
  int base = -1;

  SDTFE_HD void load(const double *col, int b, int n, bool pairs) {
    base = b;
    if (b + kWinCells <= n) {
      if (pairs) {
        const double2 *src = reinterpret_cast<const double2 *>(col + b);
        double2 *dst = reinterpret_cast<double2 *>(w);
#pragma unroll
        for (int q = 0; q < kWinCells / 2; ++q) dst[q] = src[q];
      } else {
#pragma unroll
        for (int q = 0; q < kWinCells; ++q) w[q] = col[b + q];
      }
    } else {
      for (int q = 0; q < kWinCells; ++q) w[q] = b + q < n ? col[b + q] : 0.0;
    }
  }
  SDTFE_HD void store(double *col, int n, bool pairs) const {
    if (base + kWinCells <= n) {
      if (pairs) {
        double2 *dst = reinterpret_cast<double2 *>(col + base);
        const double2 *src = reinterpret_cast<const double2 *>(w);
#pragma unroll
        for (int q = 0; q < kWinCells / 2; ++q) dst[q] = src[q];
      } else {
#pragma unroll
        for (int q = 0; q < kWinCells; ++q) col[base + q] = w[q];
      }
    } else {
      for (int q = 0; q < kWinCells && base + q < n; ++q) col[base + q] = w[q];
    }
  }
};

// evaluate_3d_ray_on_grid (src/dtfe.c:376-516): the walk of shoot_3d_ray
// with its perturbation; col is the ray's column, whose cells are read and
// written a line at a time through win (CellWindow).  Every term is added to
// its cell in the reference's order.  A failed walk keeps what it added, as the
// reference does.  Returns the tets crossed (0 on failure).
SDTFE_HD inline int column_walk64(const Rec64 *__restrict__ rec, int entry, double x, double y,
                                  const VolGeom &v, uint64_t key, double *col,
                                  CellWindow &win, bool pairs, LaneCounters &cnt) {
  double ox = x, oy = y;
  uint32_t draw = 2;
  int cur = entry, prev = -1, f_in = -1, f_out = 0, nxt = entry, steps = 0, crossed = 0;
  win.base = -1;
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
        if ((iz & ~(kWinCells - 1)) != win.base) {
          if (win.base >= 0) win.store(col, v.grid, pairs);
          win.load(col, iz & ~(kWinCells - 1), v.grid, pairs);
        }
        double &cell = win.w[iz & (kWinCells - 1)];
        cell = cell + (fixed + r.gz * z) * v.ss_weight;
      }
    }
  }
  if (win.base >= 0) win.store(col, v.grid, pairs);
  return crossed;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void
volume_fp64_kernel(const Rec64 *__restrict__ rec, EntryGrid eg, VolGeom v, double *grid,
                   unsigned long long *stats) {
  __shared__ __attribute__((aligned(16))) double lds[256 * kWinStride];
  CellWindow win;
  win.w = lds + threadIdx.x * kWinStride;
  const bool pairs = (v.grid & 1) == 0 && ((uintptr_t)grid & 15) == 0;
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
        const int crossed = column_walk64(rec, entry, x, y, v, key, col, win, pairs, cnt);
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
