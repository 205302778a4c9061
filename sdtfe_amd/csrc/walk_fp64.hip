// walk_fp64.hip -- the FP64 ray walk (gfx950), one sample-ray per lane.
//
// Same results as the reference CPU path fed the same sample positions:
// the Plücker side products, face choice, barycentric crossing heights,
// perturbation and midpoint integrand follow src/dtfe.c:19-374 operation by
// operation, and this file is built with -ffp-contract=off.
//
// Layout: each step of a walk reads one 144 B record (sdtfe::Rec64) that
// holds the tet's neighbours, its four vertices and its field, so a step is a
// single dependent load.  A map launch gives each lane one pixel (8x8 pixels
// per wave, 16x16 per block) and walks its samples in order, which keeps the
// 64 rays of a wave in the same column of tets.
#include "sdtfe_host.h"
#include "walk_fp64_core.h"

namespace sdtfe {

// One pixel per lane: S samples in order, pixel value = sum / n_avg
// (src/dtfe.c:613-645).  accumulate: out += value, else out = value.
__global__ __launch_bounds__(256) void map_fp64_kernel(const Rec64 *__restrict__ rec,
                                                      EntryGrid eg, MapGeom g, double *out,
                                                      int *navg, int accumulate,
                                                      unsigned long long *stats) {
  int px, py;
  lane_pixel(g.grid, px, py);
  LaneCounters cnt;
  if (px < g.grid && py < g.grid) {
    const uint64_t pix = (uint64_t)py * (uint64_t)g.grid + (uint64_t)px;
    double sum = 0.0;
    int n_avg = g.samples;
    for (int s = 0; s < g.samples; ++s) {
      const uint64_t key = sample_key(g.seed, pix, (uint32_t)s);
      double xt, yt;
      sample_point(g, px, py, key, xt, yt);
      cnt.samples++;
      const int entry = locate_entry(eg, xt, yt);
      if (entry < 0) {
        cnt.missed++;
        --n_avg;
        continue;
      }
      int crossed;
      const double v = walk64(rec, entry, xt, yt, g, key, crossed, cnt);
      sum += v;
      cnt.tet_steps += (unsigned)crossed;
      if (crossed == 0) {
        cnt.failed++;
        --n_avg;
      }
    }
    const double val = n_avg > 0 ? sum / n_avg : 0.0;
    if (accumulate)
      out[pix] += val;
    else
      out[pix] = val;
    if (navg) navg[pix] = n_avg;
  }
  cnt.flush(stats);
}

__global__ void trace_fp64_kernel(const Rec64 *__restrict__ rec, EntryGrid eg, MapGeom g, int n,
                                  const double *x, const double *y, const int *entry_in,
                                  const uint64_t *pixel, const uint32_t *sample, double *value,
                                  int *crossed_out, unsigned long long *stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  LaneCounters cnt;
  if (i < n) {
    const uint64_t key = sample_key(g.seed, pixel[i], sample[i]);
    const int entry = entry_in ? entry_in[i] : locate_entry(eg, x[i], y[i]);
    int crossed;
    value[i] = walk64(rec, entry, x[i], y[i], g, key, crossed, cnt);
    crossed_out[i] = crossed;
    cnt.tet_steps += (unsigned)crossed;
  }
  cnt.flush(stats);
}

__global__ void locate_kernel(EntryGrid eg, int n, const double *x, const double *y, int *entry) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) entry[i] = locate_entry(eg, x[i], y[i]);
}

// ------------------------------------------------------------ launchers

void launch_map_fp64(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate) {
  const int nb = (g.grid + 15) / 16;
  (void)hipEventRecord(t->evk0, t->stream);
  map_fp64_kernel<<<nb * nb, 256, 0, t->stream>>>(t->rec64, t->grid, g, out, navg, accumulate,
                                                  t->stats);
  (void)hipEventRecord(t->evk1, t->stream);
}

void launch_trace_fp64(sdtfe_tess *t, const MapGeom &g, int n, const double *x, const double *y,
                       const int *entry, const uint64_t *pixel, const uint32_t *sample,
                       double *value, int *crossed) {
  trace_fp64_kernel<<<(n + 255) / 256, 256, 0, t->stream>>>(t->rec64, t->grid, g, n, x, y, entry,
                                                            pixel, sample, value, crossed,
                                                            t->stats);
}

void launch_locate(sdtfe_tess *t, int n, const double *x, const double *y, int *entry) {
  locate_kernel<<<(n + 255) / 256, 256, 0, t->stream>>>(t->grid, n, x, y, entry);
}

}  // namespace sdtfe
