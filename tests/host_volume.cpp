// Host build of the volumetric device code (sdtfe_amd/csrc/volume.hip): the
// same functions the kernels run -- the column form (walk_column) and the
// ray + fold form (ray_terms, fold_cell, then walk_column for marked
// columns) -- compiled for the CPU and driven by tests/test_volume_host.py.
// Test infrastructure; the library never uses it.  Entry tets come from the
// input (one per ray in (column, is, js) order) instead of the grid lookup.
//
// Input (binary): int64 nt, grid, ss, mark_all, band_rows; float64 box_len,
// cx, cy, cz; uint64 seed; Rec64[nt]; int32 entries[grid^2 ss^2];
// float64 init[grid^3].  Output: float64 column-form grid, float64 ray-form
// grid, then int64 counters of each form (samples, missed, failed, tet
// steps, perturbation steps) and the number of marked columns.
#include "volume.hip"

#include <cstdio>
#include <vector>

// volume.hip's launch function refers to the library's stream-ordered
// allocator (capi.cpp); the harness never launches, so it links these.
namespace sdtfe {
hipError_t dev_malloc(void **, size_t, hipStream_t) { return hipErrorNotSupported; }
void dev_free(void *, hipStream_t) {}
}  // namespace sdtfe

namespace {

bool rd(FILE *f, void *p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

void put_counters(FILE *o, const sdtfe::LaneCounters &c) {
  const long long v[5] = {(long long)c.samples, (long long)c.missed, (long long)c.failed,
                          (long long)c.tet_steps, (long long)c.pert_steps};
  fwrite(v, 8, 5, o);
}

}  // namespace

int main(int argc, char **argv) {
  using namespace sdtfe;
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 3;
  long long hdr[5];
  double geo[4];
  unsigned long long seed = 0;
  if (!rd(f, hdr, sizeof hdr) || !rd(f, geo, sizeof geo) || !rd(f, &seed, 8)) return 4;
  const long long nt = hdr[0];
  const int G = (int)hdr[1], ss = (int)hdr[2], mark_all = (int)hdr[3], band_rows = (int)hdr[4];
  std::vector<Rec64> rec((size_t)nt);
  const size_t rays = (size_t)G * G * ss * ss, cells = (size_t)G * G * G;
  std::vector<int> entries(rays);
  std::vector<double> init(cells);
  if (!rd(f, rec.data(), sizeof(Rec64) * rec.size()) ||
      !rd(f, entries.data(), sizeof(int) * rays) || !rd(f, init.data(), 8 * cells))
    return 5;
  fclose(f);
  sdtfe_volume_params p{};
  p.grid_dim = G;
  p.supersampling = (unsigned)ss;
  p.box_len = geo[0];
  p.center[0] = geo[1];
  p.center[1] = geo[2];
  p.center[2] = geo[3];
  p.seed = seed;
  const VolGeom v = make_vol_geom(p);
  double lds[kWinStride];
  LineWindow win;
  win.w = lds;
  size_t next = 0;  // the ray whose entry the next locate() returns
  auto locate = [&](double, double) { return entries[next++]; };

  // column form
  std::vector<double> col_grid = init;
  LaneCounters cc;
  for (int i = 0; i < G; ++i)
    for (int j = 0; j < G; ++j) {
      next = ((size_t)i * G + j) * ss * ss;
      walk_column(rec.data(), locate, v, col_grid.data(), i, j, win, false, cc);
    }

  // ray + fold form, in bands of rows as the launch makes them
  std::vector<double> ray_grid = init;
  LaneCounters rc;
  const int rows = band_rows > 0 ? band_rows : G;
  const size_t gs = (size_t)G * ss;
  std::vector<double> terms((size_t)rows * ss * gs * v.nzs);
  std::vector<int2> range((size_t)rows * ss * gs);
  std::vector<int> mark((size_t)rows * G), redo((size_t)rows * G);
  int n_redo = 0;
  long long marked = 0;
  for (int i0 = 0; i0 < G; i0 += rows) {
    RayScratch s{};
    s.terms = terms.data();
    s.range = range.data();
    s.mark = mark.data();
    s.redo = redo.data();
    s.n_redo = &n_redo;
    s.i0 = i0;
    s.rows = std::min(rows, G - i0);
    s.mark_all = mark_all;
    std::fill(mark.begin(), mark.end(), 0);
    n_redo = 0;
    for (int X = 0; X < s.rows * ss; ++X)
      for (int Y = 0; Y < (int)gs; ++Y) {
        const int i = i0 + X / ss, is = X % ss, j = Y / ss, js = Y % ss;
        next = ((size_t)i * G + j) * ss * ss + (size_t)(is * ss + js);
        ray_terms(rec.data(), locate, v, s, X, Y, rc);
      }
    for (int ib = 0; ib < s.rows; ++ib)
      for (int j = 0; j < G; ++j) {
        if (s.mark[ib * G + j]) continue;
        double *col = ray_grid.data() + ((size_t)(i0 + ib) * G + j) * G;
        for (int iz = 0; iz < G; ++iz) col[iz] = fold_cell(v, s, ib, j, iz, col[iz]);
      }
    LaneCounters unused;
    for (int q = 0; q < n_redo; ++q) {
      const int i = s.redo[q] / G, j = s.redo[q] % G;
      next = ((size_t)i * G + j) * ss * ss;
      walk_column(rec.data(), locate, v, ray_grid.data(), i, j, win, false, unused);
    }
    marked += n_redo;
  }

  FILE *o = fopen(argv[2], "wb");
  if (!o) return 7;
  fwrite(col_grid.data(), 8, cells, o);
  fwrite(ray_grid.data(), 8, cells, o);
  put_counters(o, cc);
  put_counters(o, rc);
  fwrite(&marked, 8, 1, o);
  fclose(o);
  return 0;
}
