// walk_mixed.hip -- map and trace kernels of the mixed-precision walk
// (walk_mixed_core.h).
#include "walk_mixed_core.h"

namespace sdtfe {

__global__ __launch_bounds__(256) void map_mixed_kernel(const Rec32 *__restrict__ rec32,
                                                       const Rec64 *__restrict__ rec64,
                                                       EntryGrid eg, MapGeom g, MixedConsts mc,
                                                       double *out, int *navg, int accumulate,
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
      const double v = walk_mixed(rec32, rec64, entry, xt, yt, g, mc.flt, key, crossed, cnt);
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

__global__ void trace_mixed_kernel(const Rec32 *__restrict__ rec32,
                                   const Rec64 *__restrict__ rec64, EntryGrid eg, MapGeom g,
                                   MixedConsts mc, int n, const double *x, const double *y,
                                   const int *entry_in, const uint64_t *pixel,
                                   const uint32_t *sample, double *value, int *crossed_out,
                                   unsigned long long *stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  LaneCounters cnt;
  if (i < n) {
    const uint64_t key = sample_key(g.seed, pixel[i], sample[i]);
    const int entry = entry_in ? entry_in[i] : locate_entry(eg, x[i], y[i]);
    int crossed;
    value[i] = walk_mixed(rec32, rec64, entry, x[i], y[i], g, mc.flt, key, crossed, cnt);
    crossed_out[i] = crossed;
    cnt.tet_steps += (unsigned)crossed;
  }
  cnt.flush(stats);
}

extern const bool kHaveMixed = true;

void launch_map_mixed(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate) {
  const int nb = (g.grid + 15) / 16;
  map_mixed_kernel<<<nb * nb, 256, 0, t->stream>>>(t->rec32, t->rec64, t->grid, g,
                                                   mixed_consts(t, g), out, navg, accumulate,
                                                   t->stats);
}

void launch_trace_mixed(sdtfe_tess *t, const MapGeom &g, int n, const double *x,
                        const double *y, const int *entry, const uint64_t *pixel,
                        const uint32_t *sample, double *value, int *crossed) {
  trace_mixed_kernel<<<(n + 255) / 256, 256, 0, t->stream>>>(
      t->rec32, t->rec64, t->grid, g, mixed_consts(t, g), n, x, y, entry, pixel,
      sample, value, crossed, t->stats);
}

}  // namespace sdtfe
