/*
 * sdtfe_amd.h -- C ABI of the MI355X (gfx950) DTFE surface-density library
 * (libsdtfe_amd.so).  Plain pointers and sizes only.
 *
 * Drop-in entry points (same names, prototypes and argument meaning as the
 * reference's headers, so existing C callers and the pydtfe binding link
 * against this library unchanged):
 *
 *   compute_density   replaces src/dtfe.h:13-15 (src/dtfe.c:554-654)
 *   pre_compute_vol   replaces src/triangulate.h:26 (src/triangulate.c:183-225)
 *   gradients         replaces src/triangulate.h:27 (src/triangulate.c:233-260)
 *   interpolate_to_3d_grid  replaces src/dtfe.h:17-18 (src/dtfe.c:667-735)
 *   compute_3d_density      replaces src/dtfe.h:20-21 (src/dtfe.c:657-665)
 *   find_forward_hull replaces src/triangulate.h:28 (src/triangulate.c:263-316)
 *   copy_fh_triangles replaces src/triangulate.h:29 (src/triangulate.c:318-395)
 *   convert_to_qhdata, rotate3d, shuffle_buff
 *                     replace src/utils.h:10-12 (src/utils.c:8-79)
 *   set_bit, clear_bit, get_bit, create_bitmap, delete_bitmap
 *                     replace src/bitmap.h:9-17 (src/bitmap.c:8-31)
 *   write_image, write_rho  replace src/io.h:35-36 (src/io.c:13-95)
 *
 * The reference's triangulate (src/triangulate.h:25, Qhull) is not
 * replaced: Qhull stays on the host.
 *
 * Behaviour kept from the reference: particle_data is N x 4 float64 AoS
 * (x, y, z, w) and column 3 is overwritten with the vertex density;
 * tetra_data is T x 8 int32 rows (A, B, C, D, N_BDC, N_ACD, N_ADB, N_ABC);
 * rho is grid_dim^2 float64, row-major rho[y * grid_dim + x], and is
 * ACCUMULATED into (+=); grid_dim must be a multiple of 4; samples that miss
 * the hull's shadow or whose walk fails are dropped from the pixel average.
 *
 * Differences, by design: Monte-Carlo sample positions come from a
 * counter-based stream keyed by (seed, pixel, sample) instead of libc rand(),
 * so maps are reproducible and independent of scheduling; failures are
 * counted (sdtfe_stats) instead of printed per sample; a walk is cut after
 * SDTFE_MAX_STEPS_DEFAULT tets (the reference has no bound) and counted.
 *
 * The extended API keeps a tessellation resident in HBM across maps
 * (sdtfe_tess_*), offers the faster mixed-precision walk, and reports
 * counters and kernel times.  Functions returning int return 0 on success
 * and a negative SDTFE_E* code on failure; sdtfe_last_error() describes the
 * last failure of the calling thread.
 */
#ifndef SDTFE_AMD_H
#define SDTFE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- drop-in */

/* src/dtfe.h:13-15.  Runs on the current HIP device in FP64 (results match
 * the reference path fed the same sample positions).  Sample stream seed:
 * sdtfe_set_seed(); each call uses the current seed and then advances it, so
 * repeated calls draw fresh samples as the reference's rand() would. */
void compute_density(double *particle_data, int n_particles, int *tetra_data, int n_tetra,
                     int grid_dim, double box_len, double box_len_z, double *rho,
                     float p_mass, double center_x, double center_y, double center_z,
                     const int n_mc_samp, const double delta_sample);

/* src/triangulate.h:26: writes rho_i = 4 m / W_i into (*particle_data)[:,3]. */
void pre_compute_vol(double **particle_data, int n_particles, int *tetra_data, int n_tetra,
                     float p_mass);

/* src/triangulate.h:27: returns a malloc'd n_tetra x 3 array (free() it),
 * the per-tet gradient of the vertex values in particle_data[:,3]. */
double *gradients(int *tetra_data, int n_tetra, double *particle_data);

/* src/dtfe.h:17-18: the vertex field in particle_data[:,3] sampled on a
 * grid_dim^3 grid, grid[(i * grid_dim + j) * grid_dim + k] for cell (x i,
 * y j, z k) of the cube of side box_len about center, ACCUMULATED into: each
 * cell adds 1/supersampling^3 times the field at supersampling^3 points
 * (supersampling x supersampling vertical rays per column, supersampling
 * points per cell along each).  FP64, the reference's arithmetic; the
 * perturbation's random corners come from the counter-based stream (seed as
 * for compute_density). */
void interpolate_to_3d_grid(double *particle_data, int n_particles, int *tetra_data,
                            int n_tetra, int grid_dim, double box_len, double *grid,
                            double center_x, double center_y, double center_z,
                            unsigned supersampling);

/* src/dtfe.h:20-21: pre_compute_vol (column 3 <- density), then
 * interpolate_to_3d_grid of the density into rho. */
void compute_3d_density(double *particle_data, int n_particles, int *tetra_data, int n_tetra,
                        int grid_dim, double box_len, double *rho, float p_mass,
                        double center_x, double center_y, double center_z,
                        unsigned supersampling);

/* src/triangulate.h:28: marks the hull faces whose outward normal points
 * to -z (where vertical rays enter the hull) in bitmap, bit 4 t + f for
 * face f of tet t in the reference's bitmap layout (bit f of word t), and
 * returns their count.  Computed on the GPU. */
int find_forward_hull(int *tetra_data, int n_tetra, double *particle_data, int n_particles,
                      uint32_t *bitmap);

/* src/triangulate.h:29: a malloc'd n_tri x 8 int array (free() it) with one
 * row (a, b, c, n_CB, n_AC, n_BA, tet, 0) per face bitmap marks, in tet
 * then face order; n_* is the row across each edge or -1. */
int *copy_fh_triangles(int *tetra_data, int n_tetra, double *particle_data, int n_particles,
                       int n_tri, uint32_t *bitmap);

/* src/utils.h:10: planar float32 positions (n_data x values, then y, then
 * z) to a malloc'd n_data x 4 float64 array (x, y, z, 0); free() it. */
double *convert_to_qhdata(const float *rm_buff, size_t n_data);
/* src/utils.h:11: rotates rows 0 .. n-1 about center by
 * Qx(theta[0]) Qy(theta[1]) Qz(theta[2]). */
void rotate3d(double *data, size_t n_particles, double theta[3], double center[3]);
/* src/utils.h:12: partial Fisher-Yates shuffle of the first n_shuffle rows
 * (N x 4) drawing from the C library's rand(), as the reference does. */
void shuffle_buff(double *data, int n_particles_in_buff, int n_shuffle);

/* src/bitmap.h:9-17: bitmaps of uint32_t words holding 4 bits each (the
 * reference's layout): bit n is bit n % 4 of word n / 4.  create_bitmap(n)
 * returns n / 4 + 1 zeroed words (free with delete_bitmap). */
void set_bit(uint32_t *words, int n);
void clear_bit(uint32_t *words, int n);
int get_bit(uint32_t *words, int n);
uint32_t *create_bitmap(int n);
void delete_bitmap(uint32_t *words);

/* src/io.h:35: the map as an 8-bit palette TIFF (log-scaled indices 0-128,
 * Jet palette), one strip, uncompressed; no file if a pixel is negative or
 * NaN. */
void write_image(char *file_name, double *rho, int grid_dim);
/* src/io.h:36: the raw map, rho_size float64 values; returns the count
 * written. */
int write_rho(char *fname, double **rho, int rho_size);

/* ---------------------------------------------------------- extended */

enum {
  SDTFE_OK = 0,
  SDTFE_EINVAL = -1,   /* bad argument */
  SDTFE_EHIP = -2,     /* HIP runtime error */
  SDTFE_ENOMEM = -3,   /* device allocation failed */
  SDTFE_ETOPO = -4,    /* tessellation indices out of range */
  SDTFE_ENODEV = -5    /* no HIP device */
};

enum { SDTFE_PREC_FP64 = 0, SDTFE_PREC_MIXED = 1 };

#define SDTFE_MAX_STEPS_DEFAULT (1 << 24)

typedef struct sdtfe_tess sdtfe_tess; /* device-resident tessellation */

typedef struct {
  int grid_dim;        /* pixels per side, multiple of 4 */
  int n_mc_samp;       /* Monte-Carlo samples per pixel */
  double box_len;      /* map width and height */
  double box_len_z;    /* slab depth */
  double center[3];
  double delta_sample; /* MC sampling width (pixel width if smaller) */
  uint64_t seed;       /* sample stream seed */
  int precision;       /* SDTFE_PREC_* */
  int max_steps;       /* per-ray tet bound, <= 0 for the default */
} sdtfe_map_params;

typedef struct {
  int64_t samples;           /* sample-rays requested */
  int64_t missed;            /* outside the hull's shadow */
  int64_t failed;            /* rays that ended with 0 tets crossed */
  int64_t no_intersection;   /* src/dtfe.c:321-325 */
  int64_t perturb_exhausted; /* src/dtfe.c:338-342 */
  int64_t no_isect_perturb;  /* src/dtfe.c:344-348 */
  int64_t perturb_steps;     /* perturbations taken */
  int64_t tet_steps;         /* tets crossed by all rays */
  int64_t step_cap;          /* rays cut by max_steps */
  int64_t fp64_redo;         /* mixed walk: sample-rays handed to the FP64 walk */
  double t_walk_ms;          /* walk time (HIP events; both passes of a mixed map) */
  double t_total_ms;         /* whole call, host wall clock */
  double t_kernel_ms;        /* the walk kernel alone (a mixed map's first pass) */
} sdtfe_stats;

typedef struct {
  int n_particles, n_tetra, n_tri; /* n_tri: forward-hull faces */
  int grid_nx, grid_ny;            /* entry-lookup grid */
  int64_t device_bytes;            /* HBM held by the tessellation */
  double t_setup_ms;               /* upload + field + entry structure */
} sdtfe_tess_info;

/* The reference's tetra_data (T x 8: A, B, C, D, then the tet across the
 * face opposite each, -1 on the hull) from a Delaunay tetrahedralisation's
 * simplices (T x 4 vertex ids, in the order Qhull lists a facet's vertices)
 * and particle_data (N x 4, columns 0-2 read), assembled on the GPU as
 * triangulate assembles Qhull's facets (src/triangulate.c:56-175): B and C
 * swapped where (AB x AC) . AD > 0, the reference's expression and
 * operation order.  SDTFE_ETOPO when a vertex id is out of range or a face
 * is shared by more than two tets. */
int sdtfe_tetra_from_simplices(const double *particle_data, int n_particles,
                               const int *simplices, int n_tetra, int *tetra_data);

const char *sdtfe_last_error(void);
/* Status of the calling thread's last call of a void or pointer-returning
 * drop-in above (compute_density, pre_compute_vol, gradients,
 * interpolate_to_3d_grid, compute_3d_density, find_forward_hull,
 * copy_fh_triangles): 0 on success, else its SDTFE_E* code, with
 * sdtfe_last_error() describing it.  The reference's prototypes carry no
 * status; they print the failure to stderr as well. */
int sdtfe_dropin_status(void);
void sdtfe_set_seed(uint64_t seed);
uint64_t sdtfe_get_seed(void);

/* Upload particles (N x 4, column 3 ignored) and tets once, then compute on
 * the GPU: tet volumes, vertex densities (4 m / W, W summed in ascending tet
 * order), gradients, packed walk records and the entry-lookup grid. */
int sdtfe_tess_create(const double *particle_data, int n_particles, const int *tetra_data,
                      int n_tetra, float p_mass, sdtfe_tess **out);
/* Same, but the vertex field is taken from particle_data[:,3] as given
 * (no density pass): the route by which the reference interpolates an
 * arbitrary particle field (python/pydtfe.cpp:156-158). */
int sdtfe_tess_create_field(const double *particle_data, int n_particles,
                            const int *tetra_data, int n_tetra, sdtfe_tess **out);
void sdtfe_tess_destroy(sdtfe_tess *t);
int sdtfe_tess_get_info(const sdtfe_tess *t, sdtfe_tess_info *info);
/* Copies out: vertex densities (N), gradients (T x 3), forward-hull
 * triangles as copy_fh_triangles (src/triangulate.c:318-395) writes them:
 * (a, b, c, n_CB, n_AC, n_BA, tet, 0) rows in the reference's order, the
 * neighbour across each edge or -1 (slot 7, unwritten there, is 0). */
int sdtfe_tess_vertex_density(const sdtfe_tess *t, double *rho);
int sdtfe_tess_gradients(const sdtfe_tess *t, double *nabla);
int sdtfe_tess_triangles(const sdtfe_tess *t, int *tri_rows, int capacity);
/* The FP64 walk records, 144 bytes per tet: neighbours (4 x int32), A, B,
 * C, D coordinates (12 x float64), rho_A and the gradient (4 x float64). */
int sdtfe_tess_walk_records(const sdtfe_tess *t, void *records, size_t bytes);

/* Entry tet of each query point (-1 outside the hull's shadow). */
int sdtfe_tess_locate(sdtfe_tess *t, int n, const double *x, const double *y, int *entry);

/* Walk n explicit rays (host arrays).  entry may be NULL (then located);
 * (pixel, sample) key the perturbation draws, which start at draw 2. */
int sdtfe_tess_trace(sdtfe_tess *t, int precision, int n, const double *x, const double *y,
                     const int *entry, const uint64_t *pixel, const uint32_t *sample,
                     double off_z, double lz, double delta_xy, uint64_t seed,
                     double *value, int *crossed, sdtfe_stats *stats);

/* One map.  rho (host, grid_dim^2) is accumulated into; n_avg (host,
 * nullable) receives each pixel's count of contributing samples. */
int sdtfe_map(sdtfe_tess *t, const sdtfe_map_params *p, double *rho, int *n_avg,
              sdtfe_stats *stats);

/* One map into device memory (accumulated into), asynchronous on the
 * tessellation's stream; with stats != NULL it waits and fills them. */
int sdtfe_map_device(sdtfe_tess *t, const sdtfe_map_params *p, double *rho_device,
                     sdtfe_stats *stats);
int sdtfe_tess_sync(sdtfe_tess *t);

/* No counterpart in the reference.  sdtfe_map_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same maps, bit for bit, with
 * no hipMalloc, hipFree or allocating library call.  Its counters and the
 * mixed form's second-pass queue live in scratch memory the library holds
 * per GPU (90,182,400 bytes, about 86 MiB, kept until the process ends).  Limits: a mixed map
 * holds at most 2^20 pixels (1024 x 1024), else SDTFE_EINVAL; FP64 maps
 * any size.  Calls on one GPU share that memory and must not overlap (they
 * race); the tessellation's device must be current. */
int sdtfe_map_device_noalloc(sdtfe_tess *t, const sdtfe_map_params *p, double *rho_device,
                             sdtfe_stats *stats);

/* A volumetric grid of the tessellation's vertex field (the densities, or
 * the field given to sdtfe_tess_create_field), as interpolate_to_3d_grid
 * computes it.  Ray (is, js) of column c = i * grid_dim + j draws its
 * perturbation corners from (seed, c, is * supersampling + js). */
typedef struct {
  int grid_dim;            /* cells per side (any size >= 1) */
  unsigned supersampling;  /* >= 1 */
  double box_len;          /* the cube's side; rays run its full height */
  double center[3];
  uint64_t seed;
  int max_steps;           /* per-ray tet bound, <= 0 for the default */
} sdtfe_volume_params;

/* grid (host, grid_dim^3 doubles) is accumulated into. */
int sdtfe_volume(sdtfe_tess *t, const sdtfe_volume_params *p, double *grid,
                 sdtfe_stats *stats);
/* Into device memory (accumulated into), asynchronous on the
 * tessellation's stream; with stats != NULL it waits and fills them. */
int sdtfe_volume_device(sdtfe_tess *t, const sdtfe_volume_params *p, double *grid_device,
                        sdtfe_stats *stats);

/* ------------------------------------------------------------ batches */

/* Many maps (e.g. halo cutouts of uneven size, each its own resident
 * tessellation) in one persistent launch, replayed from a captured hipGraph.
 * Item i maps tess[i] with params[i] into device buffer rho_device[i]
 * (grid_dim^2 doubles, accumulated into).  All items share one precision.
 * lanes_per_pixel (nullable; entries 1, 2, 4, 8 or 16) splits each pixel's
 * samples over that many lanes; the FP64 map does not depend on it.  Waits
 * for completion; item_stats (nullable, n entries) and total (nullable)
 * receive counters; total->t_walk_ms is the launch time. */
typedef struct sdtfe_batch sdtfe_batch;
int sdtfe_batch_create(sdtfe_batch **out);
void sdtfe_batch_destroy(sdtfe_batch *b);
int sdtfe_batch_run(sdtfe_batch *b, int n, sdtfe_tess *const *tess,
                    const sdtfe_map_params *params, double *const *rho_device,
                    const int *lanes_per_pixel, sdtfe_stats *item_stats, sdtfe_stats *total);

/* ------------------------------------------------------------ serving */

/* A map server with continuous batching.  Requests (particles, tets, map
 * parameters) are copied at submit and admitted in arrival order while their
 * estimated HBM footprint fits hbm_budget_bytes and fewer than max_active are
 * resident; setup threads build their tessellations while a walk thread
 * runs persistent launches over the next tiles (at most tile_budget per
 * launch, shared among resident requests), so requests arriving while others
 * are on the GPU join the next launch.  Each map is identical alone and in a
 * mix.  rho (host, grid_dim^2) is accumulated into when the request
 * completes; sdtfe_server_wait blocks until then.  sdtfe_server_destroy
 * drains all requests first. */
typedef struct sdtfe_server sdtfe_server;
typedef struct {
  int64_t hbm_budget_bytes;
  int max_active;
  int tile_budget;     /* 256-lane tiles per launch */
  int lanes_per_pixel; /* 1, 2, 4, 8, 16; <= 0: by samples per pixel S (16 for
                          16 <= S <= 256, 4 for 4 <= S < 16, else 1) */
  int precision;       /* SDTFE_PREC_*, for every request */
  int setup_threads;   /* requests set up concurrently; <= 0: 4 */
} sdtfe_server_config;
typedef struct {
  int status;
  double queue_ms;     /* submit -> admitted */
  double setup_ms;     /* admitted -> first launch */
  double latency_ms;   /* submit -> map complete */
  double submit_ms;    /* since server creation */
  double done_ms;
  int64_t footprint_bytes;
  sdtfe_stats stats;
} sdtfe_request_info;
typedef struct {
  int queued, active;
  int64_t used_bytes;
  int64_t launches;
  double walk_ms;      /* summed launch time */
} sdtfe_server_info;
int sdtfe_server_create(const sdtfe_server_config *cfg, sdtfe_server **out);
int sdtfe_server_submit(sdtfe_server *s, const double *particle_data, int n_particles,
                        const int *tetra_data, int n_tetra, float p_mass,
                        const sdtfe_map_params *p, double *rho, int64_t *request_id);
int sdtfe_server_wait(sdtfe_server *s, int64_t request_id, sdtfe_request_info *info);
int sdtfe_server_get_info(sdtfe_server *s, sdtfe_server_info *info);
void sdtfe_server_destroy(sdtfe_server *s);

/* Device buffers for callers without another HIP allocator.  The copies
 * and the fill return when they are complete. */
void *sdtfe_device_alloc(size_t bytes);
int sdtfe_device_free(void *p);
int sdtfe_memcpy_d2h(void *dst, const void *src, size_t bytes);
int sdtfe_memcpy_h2d(void *dst, const void *src, size_t bytes);
int sdtfe_memset_d(void *p, int value, size_t bytes);
int sdtfe_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
