// walk_fp64.hip -- the FP64 ray walk (gfx950), one sample-ray per lane.
//
// Same results as the reference CPU path fed the same sample positions:
// the Plücker side products, face choice, barycentric crossing heights,
// perturbation and midpoint integrand follow src/dtfe.c:19-374 operation by
// operation, and this file is built with -ffp-contract=off.
//
// Layout: each step of a map's walk reads one 144 B face record
// (sdtfe::FRec64: the entered face's vertices, the fourth vertex, the tet's
// field and the three ways out), so a step is a single dependent load; the
// walk over per-tet records (sdtfe::Rec64, also 144 B) finishes what the
// face walk hands over.  A map launch spreads each pixel's samples over 16
// lanes (16-64 samples) or 4 (4-15 or 65-256), so the 64 rays of a wave start
// together on a 2x2- or 4x4-pixel patch of the map and cross mostly the same
// tets.
#include <cstdlib>
#include <string>

#include "sdtfe_host.h"
#include "walk_fp64_face.h"

namespace sdtfe {

// L lanes per pixel (lane j walks samples j, j + L, ...; L = 1: one pixel
// per lane, S samples in order).  The pixel's first lane takes the values in
// sample order after each round, so pixel value = sum / n_avg is the same
// for every L (src/dtfe.c:613-645).  accumulate: out += value, else out =
// value.
// kFace: the walk over face records (walk64f), else over Rec64 (walk64);
// the same results.
template <int L, bool kFace>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void map_fp64_kernel(
    const Rec64 *__restrict__ rec, const FRec64 *__restrict__ frec, EntryGrid eg, MapGeom g,
    double *out, int *navg, int accumulate, unsigned long long *stats) {
  int px, py, j;
  split_pixel<L>(g.grid, px, py, j);
  const bool active = px < g.grid && py < g.grid;
  const uint64_t pix = (uint64_t)py * (uint64_t)g.grid + (uint64_t)px;
  LaneCounters cnt;
  double sum = 0.0;
  int n_avg = g.samples;
  // one sample: 0 a miss, 1 a walk that crossed tets, 2 one that crossed
  // none (both add v, src/dtfe.c:630-636); -1 not walked
  auto sample = [&](int s, double &v) -> int {
    const uint64_t key = sample_key(g.seed, pix, (uint32_t)s);
    double xt, yt;
    sample_point(g, px, py, key, xt, yt);
    cnt.samples++;
    v = 0.0;
    const int entry = kFace ? locate_entry_rid(eg, xt, yt) : locate_entry(eg, xt, yt);
    if (entry < 0) {
      cnt.missed++;
      return 0;
    }
    int crossed;
    v = kFace ? walk64f(rec, frec, entry, xt, yt, g, key, crossed, cnt)
              : walk64(rec, entry, xt, yt, g, key, crossed, cnt);
    cnt.tet_steps += (unsigned)crossed;
    if (crossed == 0) cnt.failed++;
    return crossed > 0 ? 1 : 2;
  };
  auto take = [&](int k, double v) {
    if (k >= 1) sum += v;
    if (k == 0 || k == 2) --n_avg;
  };
  if (L == 1) {
    if (active)
      for (int s = 0; s < g.samples; ++s) {
        double v;
        const int k = sample(s, v);
        take(k, v);
      }
  } else {
    const int lead = (int)(threadIdx.x & 63) - j;
    for (int s0 = 0; s0 < g.samples; s0 += L) {
      double v = 0.0;
      int k = -1;
      if (active && s0 + j < g.samples) k = sample(s0 + j, v);
      for (int i = 0; i < L; ++i) {
        const double vi = __shfl(v, lead + i, 64);
        const int ki = __shfl(k, lead + i, 64);
        if (j == 0) take(ki, vi);
      }
    }
  }
  if (active && j == 0) {
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

// SDTFE_FP64_WALK=rec takes the walk over Rec64 for every FP64 map (the two
// walks give the same maps; tests run both).
bool fp64_face_walk() {
  const char *e = getenv("SDTFE_FP64_WALK");
  return !(e && std::string(e) == "rec");
}

template <int L>
void launch_map_fp64_l(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate,
                       unsigned long long *stats) {
  constexpr int tw = 2 * wave_bw(L), th = 2 * wave_bh(L);
  const int nb = ((g.grid + tw - 1) / tw) * ((g.grid + th - 1) / th);
  if (t->frec64 && fp64_face_walk())
    map_fp64_kernel<L, true><<<nb, 256, 0, t->stream>>>(t->rec64, t->frec64, t->grid, g, out,
                                                        navg, accumulate, stats);
  else
    map_fp64_kernel<L, false><<<nb, 256, 0, t->stream>>>(t->rec64, nullptr, t->grid, g, out,
                                                         navg, accumulate, stats);
}

void launch_map_fp64(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate,
                     unsigned long long *stats) {
  (void)hipEventRecord(t->evk0, t->stream);
  const int lanes = map_lanes(g.samples);
  if (lanes == 16)
    launch_map_fp64_l<16>(t, g, out, navg, accumulate, stats);
  else if (lanes >= 4)
    launch_map_fp64_l<4>(t, g, out, navg, accumulate, stats);
  else
    launch_map_fp64_l<1>(t, g, out, navg, accumulate, stats);
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
