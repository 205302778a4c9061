/*
 * ref_harness.c -- drives the reference C (compiled from /root/reference by
 * oracle/Makefile into oracle/_ref/libdtfe_ref.so) from the shared sample
 * stream.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference draws its Monte-Carlo sample positions and its perturbation
 * corners from libc rand() (src/dtfe.c:621-622 and :223).  This file defines
 * rand() inside the reference library (linked -Bsymbolic, so the reference's
 * calls bind here) and answers each call from the counter-based stream of
 * oracle/dtfe_oracle.c, keyed by (seed, pixel, sample, draw):
 *
 *   - a call whose return address lies inside compute_density is a sample
 *     draw; in a single-threaded run the reference visits pixels in 4x4-tile
 *     order (src/dtfe.c:607-614) and draws x then y per sample, so the
 *     harness replays that order to name the pixel and sample;
 *   - any other call comes from perturb_ray and takes the next perturbation
 *     draw (index 2, 3, ...) of the current sample.
 *
 * For direct calls of shoot_3d_ray, refh_set_ray() names the sample.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <link.h>
#include <stdint.h>
#include <stdlib.h>

#include "dtfe_oracle.h"

/* The reference's own functions, in the same shared object. */
extern void compute_density(double *, int, int *, int, int, double, double, double *, float,
                            double, double, double, const int, const double);
extern double shoot_3d_ray(int, int *, double *, double *, double, double, double, double, double,
                           int, int, int *);
extern int pt_loc_2d(int *, double *, int *, double, double);
extern void evaluate_3d_ray_on_grid(int, int *, double *, double *, double, double, double, double,
                                    int, double, int *, unsigned, double *);

static uintptr_t g_cd_lo, g_cd_hi;
static uint64_t g_seed;
static int g_samples = 1;
static int64_t g_sample_draws;
static int *g_visit; /* pixel index of the q-th pixel the reference visits */
static int64_t g_npix;
static uint64_t g_pixel;
static uint32_t g_sample, g_draw = 2;
static int64_t g_rand_calls;

static void refh_init_range(void) {
  if (g_cd_hi) return;
  Dl_info info;
  const ElfW(Sym) *sym = NULL;
  if (dladdr1((void *)&compute_density, &info, (void **)&sym, RTLD_DL_SYMENT) && sym) {
    g_cd_lo = (uintptr_t)info.dli_saddr;
    g_cd_hi = g_cd_lo + sym->st_size;
  }
}

/* Map mode: the next compute_density call draws for a grid_dim^2 map with
 * n_samp samples per pixel from stream `seed`. */
void refh_set_map(uint64_t seed, int grid_dim, int n_samp) {
  refh_init_range();
  g_seed = seed;
  g_samples = n_samp;
  g_sample_draws = 0;
  free(g_visit);
  g_npix = (int64_t)grid_dim * grid_dim;
  g_visit = (int *)malloc(sizeof(int) * (size_t)(g_npix > 0 ? g_npix : 1));
  int64_t q = 0;
  for (int band = 0; band < grid_dim / 4; ++band)
    for (int tx = 0; tx < grid_dim / 4; ++tx)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) g_visit[q++] = (band * 4 + j) * grid_dim + tx * 4 + k;
}

/* Ray mode: perturbation draws of the next shoot_3d_ray call belong to
 * (seed, pixel, sample). */
void refh_set_ray(uint64_t seed, uint64_t pixel, uint32_t sample) {
  refh_init_range();
  g_seed = seed;
  g_pixel = pixel;
  g_sample = sample;
  g_draw = 2;
}

int64_t refh_sample_draws(void) { return g_sample_draws; }
/* rand() calls since the last reset (a volumetric run that makes none took
 * no perturbation, so its result does not depend on the draws). */
int64_t refh_rand_calls(void) { return g_rand_calls; }
void refh_reset_rand_calls(void) { g_rand_calls = 0; }
uintptr_t refh_cd_size(void) {
  refh_init_range();
  return g_cd_hi - g_cd_lo;
}

int rand(void) {
  ++g_rand_calls;
  uintptr_t ra = (uintptr_t)__builtin_return_address(0);
  if (ra >= g_cd_lo && ra < g_cd_hi && g_visit) {
    int64_t q = g_sample_draws++;
    int64_t seq = q / 2;
    uint32_t coord = (uint32_t)(q % 2);
    int64_t visit = seq / g_samples;
    g_sample = (uint32_t)(seq % g_samples);
    g_pixel = visit < g_npix ? (uint64_t)g_visit[visit] : 0;
    g_draw = 2;
    return (int)sdo_draw31(g_seed, g_pixel, g_sample, coord);
  }
  return (int)sdo_draw31(g_seed, g_pixel, g_sample, g_draw++);
}

/* The pixels of the 4-row bands band_begin, band_begin + band_step, ... of
 * a grid_dim^2 map, each as compute_density forms it (src/dtfe.c:613-645)
 * with the reference's own pt_loc_2d and shoot_3d_ray, sample positions and
 * perturbation corners from the shared stream; single-threaded.  For timing
 * the reference's walk on a bounded part of a large map.  Returns the rays
 * traced. */
int64_t refh_map_bands(double *pd, int *tetra, double *nabla, int *tri, int grid_dim,
                       double box_len, double box_len_z, double *rho, double cx, double cy,
                       double cz, int n_samp, double delta_sample, uint64_t seed,
                       int band_begin, int band_step) {
  refh_init_range();
  const double dxy = box_len / grid_dim;
  const double off_x = -(box_len / 2.0) + cx, off_y = -(box_len / 2.0) + cy;
  const double off_z = -(box_len_z / 2.0) + cz;
  const double scale = delta_sample > dxy ? delta_sample : dxy;
  const double soff = delta_sample > dxy ? (delta_sample - dxy) / 2.0 : 0.0;
  int64_t rays = 0;
  if (band_step < 1) band_step = 1;
  for (int band = band_begin; band < grid_dim / 4; band += band_step)
    for (int tx = 0; tx < grid_dim / 4; ++tx)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) {
          const int px = tx * 4 + k, py = band * 4 + j;
          const uint64_t pix = (uint64_t)py * (uint64_t)grid_dim + (uint64_t)px;
          const double x = px * dxy + off_x - soff, y = py * dxy + off_y - soff;
          double sum = 0.0;
          int n_avg = n_samp;
          for (int r = 0; r < n_samp; ++r) {
            const double xt = x + ((double)sdo_draw31(seed, pix, (uint32_t)r, 0) / SDO_RAND_MAX) * scale;
            const double yt = y + ((double)sdo_draw31(seed, pix, (uint32_t)r, 1) / SDO_RAND_MAX) * scale;
            int start = 0;
            const int ft = pt_loc_2d(tri, pd, &start, xt, yt);
            ++rays;
            if (ft == -1) {
              --n_avg;
              continue;
            }
            int crossed = 0;
            refh_set_ray(seed, pix, (uint32_t)r);
            sum += shoot_3d_ray(ft, tetra, nabla, pd, xt, yt, off_z, box_len_z, dxy, grid_dim,
                                (int)pix, &crossed);
            if (crossed == 0) --n_avg;
          }
          rho[pix] += n_avg > 0 ? sum / n_avg : 0.0;
        }
  return rays;
}

/* interpolate_to_3d_grid's column loop (src/dtfe.c:686-731) over the
 * reference's own pt_loc_2d and evaluate_3d_ray_on_grid, single-threaded,
 * with each ray's perturbation corners keyed (seed, column, is * ss + js)
 * so that the grid can be compared with other builds bit for bit.  grid
 * (grid_dim^3) is accumulated into.  Returns the rays traced. */
int64_t refh_volume(double *pd, int *tetra, double *nabla, int *tri, int grid_dim,
                    double box_len, double *grid, double cx, double cy, double cz,
                    unsigned ss) {
  refh_init_range();
  const double dxy = box_len / grid_dim;
  const double off_x = -(box_len / 2.0) + cx, off_y = -(box_len / 2.0) + cy;
  const double off_z = -(box_len / 2.0) + cz;
  const double ds = dxy / ss;
  int64_t rays = 0;
  const uint64_t seed = g_seed;
  for (int i = 0; i < grid_dim; ++i)
    for (int j = 0; j < grid_dim; ++j) {
      const size_t index0 = ((size_t)i * grid_dim + j) * grid_dim;
      for (int is = 0; is < (int)ss; ++is)
        for (int js = 0; js < (int)ss; ++js) {
          const double x = i * dxy + off_x + is * ds + ds / 2.;
          const double y = j * dxy + off_y + js * ds + ds / 2.;
          int start = 0, crossed = 0;
          const int ft = pt_loc_2d(tri, pd, &start, x, y);
          refh_set_ray(seed, (uint64_t)i * grid_dim + j, (uint32_t)(is * (int)ss + js));
          evaluate_3d_ray_on_grid(ft, tetra, nabla, pd, x, y, off_z, box_len, grid_dim, dxy,
                                  &crossed, ss, grid + index0);
          ++rays;
        }
    }
  return rays;
}
