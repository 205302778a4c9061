// volume.hip -- the volumetric grid (interpolate_to_3d_grid and
// compute_3d_density, src/dtfe.c:376-516, 657-735) on gfx950.
//
// Column (i, j) of the grid (x index i, y index j) owns the grid^3 array's
// slice [(i*G + j)*G, +G): its ss x ss rays, in (is, js) order, walk the
// tessellation bottom to top and add, in each tet crossed, ss_weight times
// the tet's linear field at every fine z point the tet holds.  A cell must
// receive its terms in that order (ray by ray, each ray's in walk order) for
// the grid to be the reference's bit for bit (this file is built with
// -ffp-contract=off).  Two ways keep the order:
//
//  * column kernel: one lane owns a column and walks its rays in order,
//    adding into the grid (used for ss = 1, where a column has one ray);
//  * ray kernels (ss > 1): one lane per ray writes its terms, one per fine
//    point, into a scratch row; a second kernel, one lane per cell, adds the
//    cell's terms into it ray by ray.  A ray whose fine points do not come
//    strictly upward (a point counted in two tets, or a turn-around) marks
//    its column, which the column kernel then computes instead.
#include <cstdlib>
#include <cstring>

#include "sdtfe_host.h"
#include "walk_fp64_core.h"

namespace sdtfe {

// src/dtfe.c:386-409, 686-695
VolGeom make_vol_geom(const sdtfe_volume_params &p) {
  VolGeom v{};
  const unsigned ss = p.supersampling;
  v.dxy = p.box_len / p.grid_dim;
  v.ds = v.dxy / ss;
  v.off_x = -(p.box_len / 2.0) + p.center[0];
  v.off_y = -(p.box_len / 2.0) + p.center[1];
  v.off_z = -(p.box_len / 2.0) + p.center[2];
  v.lz = p.box_len;
  const double dz = v.lz / p.grid_dim;
  v.dzs = dz / ss;
  v.ss_weight = 1. / (ss * ss * ss);
  v.pert_step = (v.dxy / 2.0) / kPerturbMax;
  v.seed = p.seed;
  v.grid = p.grid_dim;
  v.ss = (int)ss;
  v.nzs = p.grid_dim * (int)ss;
  v.max_steps = p.max_steps > 0 ? p.max_steps : SDTFE_MAX_STEPS_DEFAULT;
  return v;
}

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
// base a multiple of 16) staged in this lane's slot of LDS, so that a lane
// reads and writes whole lines.  (With thousands of columns in flight the
// lines do not stay in L2 between a lane's consecutive cells, so cell-by-
// cell reads would fetch each line again for every cell.)  Cells at or
// beyond n are neither read nor written.  pairs: every column starts
// 16-byte aligned, so lines move as 16-byte pieces.
constexpr int kWinCells = 16;
constexpr int kWinStride = 18;  // doubles per lane slot (16-byte aligned, fewer bank conflicts)

struct LineWindow {
  double *w;
  int base = -1;

  SDTFE_HD void load(const double *row, int b, int n, bool pairs) {
    base = b;
    if (b + kWinCells <= n) {
      if (pairs) {
        const double2 *src = reinterpret_cast<const double2 *>(row + b);
        double2 *dst = reinterpret_cast<double2 *>(w);
#pragma unroll
        for (int q = 0; q < kWinCells / 2; ++q) dst[q] = src[q];
      } else {
#pragma unroll
        for (int q = 0; q < kWinCells; ++q) w[q] = row[b + q];
      }
    } else {
      for (int q = 0; q < kWinCells; ++q) w[q] = b + q < n ? row[b + q] : 0.0;
    }
  }
  SDTFE_HD void store(double *row, int n, bool pairs) const {
    if (base + kWinCells <= n) {
      if (pairs) {
        double2 *dst = reinterpret_cast<double2 *>(row + base);
        const double2 *src = reinterpret_cast<const double2 *>(w);
#pragma unroll
        for (int q = 0; q < kWinCells / 2; ++q) dst[q] = src[q];
      } else {
#pragma unroll
        for (int q = 0; q < kWinCells; ++q) row[base + q] = w[q];
      }
    } else {
      for (int q = 0; q < kWinCells && base + q < n; ++q) row[base + q] = w[q];
    }
  }
};

// Column kernel's sink: term for fine point k is added to its cell k / ss of
// the column, through a window of 16 cells.
struct CellSink {
  LineWindow &win;
  double *col;
  int n, ss;
  bool pairs;

  SDTFE_HD void add(int k, double t) {
    const int iz = k / ss;
    if ((iz & ~(kWinCells - 1)) != win.base) {
      if (win.base >= 0) win.store(col, n, pairs);
      win.load(col, iz & ~(kWinCells - 1), n, pairs);
    }
    double &cell = win.w[iz & (kWinCells - 1)];
    cell = cell + t;
  }
  // fine points e..l-1 of one tet, term(k) each, upward
  template <class F>
  SDTFE_HD void run(int e, int l, const F &term) {
    for (int k = e; k < l; ++k) add(k, term(k));
  }
  SDTFE_HD void finish() {
    if (win.base >= 0) win.store(col, n, pairs);
  }
};

// An empty term slot of a ray's row: a signalling-NaN pattern, which no
// arithmetic result takes (a NaN a term computes is quiet).
constexpr uint64_t kNoTerm = 0x7FF4A5A5A5A5A5A5ull;

// Ray kernel's sink: term for fine point k goes to slot k of the ray's row;
// slots between two fine points of the ray that no tet holds get kNoTerm,
// and [first, last] records the ray's slots.  With pairs (16-byte aligned
// rows, nzs even) slots go out as aligned pairs (2m, 2m + 1): an even slot
// is held in p0 until its odd partner is written, half the stores of
// slot-by-slot writes.  A tet whose fine points do not start above the
// ray's last slot sets bad (the row is then not used).  Within a tet the
// slots are consecutive, so the checks run once per tet and the loop over
// its fine points only computes and places terms.
struct TermSink {
  double *row;
  bool pairs;
  int first = -1, last = -1;
  bool bad = false;
  double p0 = __builtin_bit_cast(double, kNoTerm);  // pairs: even slot last, if held

  // slot k, one past the previous slot put (or the ray's first)
  SDTFE_HD void put(int k, double t) {
    if (!pairs) {
      row[k] = t;
    } else if (k & 1) {
      *reinterpret_cast<double2 *>(row + (k - 1)) = make_double2(p0, t);
      p0 = __builtin_bit_cast(double, kNoTerm);
    } else {
      p0 = t;
    }
  }
  // fine points e..l-1 (l > e) of one tet, term(k) each
  template <class F>
  SDTFE_HD void run(int e, int l, const F &term) {
    if (bad) return;
    if (e <= last) {
      bad = true;
      return;
    }
    if (first < 0)
      first = e;
    else
      for (int g = last + 1; g < e; ++g) put(g, __builtin_bit_cast(double, kNoTerm));
    for (int k = e; k < l; ++k) put(k, term(k));
    last = l - 1;
  }
  SDTFE_HD void finish() {
    // an even last slot is still held: its pair goes out with an empty partner
    if (pairs && !bad && first >= 0 && (last & 1) == 0)
      *reinterpret_cast<double2 *>(row + last) =
          make_double2(p0, __builtin_bit_cast(double, kNoTerm));
  }
};

// evaluate_3d_ray_on_grid (src/dtfe.c:376-516): the walk of shoot_3d_ray
// with its perturbation; in each tet crossed, the fine points' terms
// ss_weight * (linear field at the point) go to sink.run in walk order.
// A failed walk keeps what it added, as the reference does.  Returns the
// tets crossed (0 on failure).
template <class Sink>
SDTFE_HD inline int ray_walk64(const Rec64 *__restrict__ rec, int entry, double x, double y,
                               const VolGeom &v, uint64_t key, Sink &sink, LaneCounters &cnt) {
  double ox = x, oy = y;
  uint32_t draw = 2;
  int cur = entry, prev = -1, f_in = -1, f_out = 0, nxt = entry, steps = 0, crossed = 0;
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
      const double gz = r.gz;
      sink.run(e, l, [&](int k) {
        const double z = v.off_z + v.dzs * (double)k + v.dzs / 2.;
        return (fixed + gz * z) * v.ss_weight;
      });
    }
  }
  sink.finish();
  return crossed;
}

// Ray (is, js) of column (i, j): position (src/dtfe.c:709-710) and key.
SDTFE_HD inline void column_ray(const VolGeom &v, int i, int j, int is, int js, double &x,
                                double &y, uint64_t &key) {
  x = (double)i * v.dxy + v.off_x + (double)is * v.ds + v.ds / 2.;
  y = (double)j * v.dxy + v.off_y + (double)js * v.ds + v.ds / 2.;
  key = sample_key(v.seed, (uint64_t)i * (uint64_t)v.grid + (uint64_t)j,
                   (uint32_t)(is * v.ss + js));
}

// All rays of column (i, j) in order, adding into the grid; locate(x, y)
// gives a ray's entry tet (-1 outside the hull's shadow).
template <class Locate>
SDTFE_HD inline void walk_column(const Rec64 *__restrict__ rec, const Locate &locate,
                                 const VolGeom &v, double *grid, int i, int j, LineWindow &win,
                                 bool pairs, LaneCounters &cnt) {
  double *col = grid + ((uint64_t)i * (uint64_t)v.grid + (uint64_t)j) * (uint64_t)v.grid;
  CellSink sink{win, col, v.grid, v.ss, pairs};
  for (int is = 0; is < v.ss; ++is)
    for (int js = 0; js < v.ss; ++js) {
      double x, y;
      uint64_t key;
      column_ray(v, i, j, is, js, x, y, key);
      cnt.samples++;
      const int entry = locate(x, y);
      if (entry < 0) {
        cnt.missed++;
        continue;
      }
      win.base = -1;
      const int crossed = ray_walk64(rec, entry, x, y, v, key, sink, cnt);
      cnt.tet_steps += (unsigned)crossed;
      if (crossed == 0) cnt.failed++;
    }
}

// 16 x 16 positions per block, an 8 x 8 patch per wave.
__device__ inline void block_patch(int tiles_y, int &px, int &py) {
  const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x, tiles_y);
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  px = (b / tiles_y) * 16 + (w >> 1) * 8 + (lane >> 3);
  py = (b % tiles_y) * 16 + (w & 1) * 8 + (lane & 7);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void
volume_fp64_kernel(const Rec64 *__restrict__ rec, EntryGrid eg, VolGeom v, double *grid,
                   unsigned long long *stats) {
  __shared__ __attribute__((aligned(16))) double lds[256 * kWinStride];
  LineWindow win;
  win.w = lds + threadIdx.x * kWinStride;
  const bool pairs = (v.grid & 1) == 0 && ((uintptr_t)grid & 15) == 0;
  int i, j;
  block_patch((v.grid + 15) / 16, i, j);
  LaneCounters cnt;
  auto locate = [&](double x, double y) { return locate_entry(eg, x, y); };
  if (i < v.grid && j < v.grid) walk_column(rec, locate, v, grid, i, j, win, pairs, cnt);
  cnt.flush(stats);
}


// Scratch of the ray kernels for grid rows i in [i0, i0 + rows): the rays
// form a (rows*ss) x (G*ss) lattice, ray (X, Y) = (i*ss + is, j*ss + js)
// at lattice index (X - i0*ss) * G*ss + Y; each has a row of nzs term
// slots and the range of windows it wrote; each column has a mark.
struct RayScratch {
  double *terms;  // rays x nzs
  int2 *range;    // rays: first and last slot written (first > last: none)
  int *mark;      // rows x G columns: nonzero = computed by the column kernel
  int *redo;      // marked columns, as i * G + j
  int *n_redo;
  int i0, rows;
  int mark_all;   // every column to the column kernel (a check of that path)
};

// Ray (X, Y) of the band (lattice position within it): its terms into its
// scratch row, its range, and its column's mark when the row cannot be used;
// locate(x, y) gives its entry tet.
template <class Locate>
SDTFE_HD inline void ray_terms(const Rec64 *__restrict__ rec, const Locate &locate,
                               const VolGeom &v, const RayScratch &s, int X, int Y,
                               LaneCounters &cnt) {
  const int gs = v.grid * v.ss;
  const int i = s.i0 + X / v.ss, is = X % v.ss, j = Y / v.ss, js = Y % v.ss;
  const uint64_t ray = (uint64_t)X * (uint64_t)gs + (uint64_t)Y;
  double x, y;
  uint64_t key;
  column_ray(v, i, j, is, js, x, y, key);
  cnt.samples++;
  int2 range = make_int2(1, 0);
  const int entry = locate(x, y);
  if (entry < 0) {
    cnt.missed++;
  } else {
    const bool pairs = (v.nzs & 1) == 0 && ((uintptr_t)s.terms & 15) == 0;
    TermSink sink{s.terms + ray * (uint64_t)v.nzs, pairs};
    const int crossed = ray_walk64(rec, entry, x, y, v, key, sink, cnt);
    cnt.tet_steps += (unsigned)crossed;
    if (crossed == 0) cnt.failed++;
    if (sink.bad || s.mark_all) {
      const int c = (i - s.i0) * v.grid + j;
#if defined(__HIP_DEVICE_COMPILE__)
      if (atomicOr(s.mark + c, 1) == 0) s.redo[atomicAdd(s.n_redo, 1)] = i * v.grid + j;
#else
      if (s.mark[c] == 0) {
        s.mark[c] = 1;
        s.redo[(*s.n_redo)++] = i * v.grid + j;
      }
#endif
    } else if (sink.first >= 0) {
      range = make_int2(sink.first, sink.last);
    }
  }
  s.range[ray] = range;
}

// Cell (i0 + ib, j, iz) of the band: its terms, ray by ray in (is, js) order
// and each ray's fine points upward, added onto val.
SDTFE_HD inline double fold_cell(const VolGeom &v, const RayScratch &s, int ib, int j, int iz,
                                 double val) {
  const uint64_t gs = (uint64_t)v.grid * (uint64_t)v.ss;
  for (int is = 0; is < v.ss; ++is)
    for (int js = 0; js < v.ss; ++js) {
      const uint64_t ray = (uint64_t)(ib * v.ss + is) * gs + (uint64_t)(j * v.ss + js);
      const int2 range = s.range[ray];
      const double *row = s.terms + ray * (uint64_t)v.nzs;
      for (int q = 0; q < v.ss; ++q) {
        const int k = iz * v.ss + q;
        if (k < range.x || k > range.y) continue;
        const double t = row[k];
        if (__builtin_bit_cast(uint64_t, t) != kNoTerm) val = val + t;
      }
    }
  return val;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void
volume_rays_kernel(const Rec64 *__restrict__ rec, EntryGrid eg, VolGeom v, RayScratch s,
                   unsigned long long *stats) {
  const int gs = v.grid * v.ss;
  int X, Y;  // within the band
  block_patch((gs + 15) / 16, X, Y);
  LaneCounters cnt;
  auto locate = [&](double x, double y) { return locate_entry(eg, x, y); };
  if (X < s.rows * v.ss && Y < gs) ray_terms(rec, locate, v, s, X, Y, cnt);
  cnt.flush(stats);
}

// One lane per cell (i, j, iz) of the band, unless its column is marked.
__global__ __launch_bounds__(256) void volume_fold_kernel(VolGeom v, RayScratch s,
                                                          double *grid) {
  const uint64_t cell = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t g = (uint64_t)v.grid;
  if (cell >= (uint64_t)s.rows * g * g) return;
  const int iz = (int)(cell % g);
  const uint64_t c = cell / g;  // column within the band
  const int j = (int)(c % g), ib = (int)(c / g);
  if (s.mark[c]) return;
  double *out = grid + ((uint64_t)(s.i0 + ib) * g + (uint64_t)j) * g + (uint64_t)iz;
  *out = fold_cell(v, s, ib, j, iz, *out);
}

// Marked columns of the band, by the column kernel's order of work.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void
volume_redo_kernel(const Rec64 *__restrict__ rec, EntryGrid eg, VolGeom v, RayScratch s,
                   double *grid) {
  __shared__ __attribute__((aligned(16))) double lds[256 * kWinStride];
  LineWindow win;
  win.w = lds + threadIdx.x * kWinStride;
  const bool pairs = (v.grid & 1) == 0 && ((uintptr_t)grid & 15) == 0;
  LaneCounters cnt;  // the ray kernel counted these rays
  auto locate = [&](double x, double y) { return locate_entry(eg, x, y); };
  const int n = *s.n_redo;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    walk_column(rec, locate, v, grid, s.redo[q] / v.grid, s.redo[q] % v.grid, win, pairs, cnt);
}

__global__ void volume_band_reset_kernel(int *mark, int n, int *n_redo) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) mark[q] = 0;
  if (q == 0) *n_redo = 0;
}

// Bytes of ray scratch one band may use (a 512^3 grid with supersampling 2
// takes 8.6 GB in one band).
constexpr uint64_t kRayScratchBytes = 16ull << 30;

// SDTFE_VOLUME_FORM selects the work for supersampling > 1: "columns" (the
// column kernel), "redo" (the ray kernels with every column marked, so the
// column kernel redoes all), else the ray kernels; SDTFE_VOLUME_BAND_ROWS
// caps the grid rows per band.  All give the same grid.
int volume_form() {
  const char *e = getenv("SDTFE_VOLUME_FORM");
  if (!e) return 0;
  if (strcmp(e, "columns") == 0) return 1;
  if (strcmp(e, "redo") == 0) return 2;
  return 0;
}
int volume_band_rows() {
  const char *e = getenv("SDTFE_VOLUME_BAND_ROWS");
  return e && *e ? atoi(e) : 0;
}

void launch_volume_fp64(sdtfe_tess *t, const VolGeom &v, double *grid) {
  hipStream_t st = t->stream;
  const int form = volume_form();
  const int col_tiles = (v.grid + 15) / 16;
  void *mem = nullptr;
  const uint64_t g = (uint64_t)v.grid, gs = g * (uint64_t)v.ss;
  // bands of whole grid rows i, each within the scratch budget
  const uint64_t row_bytes = (uint64_t)v.ss * gs * ((uint64_t)v.nzs * 8 + sizeof(int2));
  const uint64_t bands = std::max<uint64_t>(1, (row_bytes * g + kRayScratchBytes - 1) /
                                                   kRayScratchBytes);
  int rows = (int)((g + bands - 1) / bands);
  if (volume_band_rows() > 0) rows = std::min(rows, volume_band_rows());
  const uint64_t rays = (uint64_t)rows * (uint64_t)v.ss * gs;
  const size_t terms_b = sizeof(double) * rays * (uint64_t)v.nzs;
  const size_t range_b = sizeof(int2) * rays, cols = (size_t)rows * g;
  if (v.ss == 1 || form == 1 ||
      dev_malloc(&mem, terms_b + range_b + sizeof(int) * (2 * cols + 1), st) != hipSuccess) {
    // one ray per column, or no room for the scratch: the column kernel
    (void)hipGetLastError();
    volume_fp64_kernel<<<col_tiles * col_tiles, 256, 0, st>>>(t->rec64, t->grid, v, grid,
                                                             t->stats);
    return;
  }
  RayScratch s{};
  char *p = (char *)mem;
  s.terms = (double *)p;
  s.range = (int2 *)(p + terms_b);
  s.mark = (int *)(p + terms_b + range_b);
  s.redo = s.mark + cols;
  s.n_redo = s.redo + cols;
  s.mark_all = form == 2;
  for (int i0 = 0; i0 < v.grid; i0 += rows) {
    s.i0 = i0;
    s.rows = std::min(rows, v.grid - i0);
    const int nc = s.rows * v.grid;
    volume_band_reset_kernel<<<(nc + 255) / 256, 256, 0, st>>>(s.mark, nc, s.n_redo);
    const int tx = (s.rows * v.ss + 15) / 16, ty = (int)((gs + 15) / 16);
    volume_rays_kernel<<<tx * ty, 256, 0, st>>>(t->rec64, t->grid, v, s, t->stats);
    const uint64_t cells = (uint64_t)s.rows * g * g;
    volume_fold_kernel<<<(unsigned)((cells + 255) / 256), 256, 0, st>>>(v, s, grid);
    volume_redo_kernel<<<64, 256, 0, st>>>(t->rec64, t->grid, v, s, grid);
  }
  dev_free(mem, st);
}

}  // namespace sdtfe
