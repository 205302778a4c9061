/*
 * dtfe_oracle.h -- CPU restatement of the SDTFE surface-density path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * and the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product (sdtfe_amd) never
 * links or calls it.
 *
 * Every function restates one piece of the reference, cited by file:line of
 * the reference tree (src/...).  The arithmetic is kept in the reference's
 * evaluation order so that a gcc build without FMA contraction reproduces the
 * reference bit for bit.  The only deliberate difference is the random
 * stream: the reference draws libc rand() inside an OpenMP region
 * (src/dtfe.c:621-622, 223); the oracle draws from a counter-based stream
 * keyed by (seed, pixel, sample, draw) so that results do not depend on the
 * thread count.  The reference build under oracle/_ref interposes rand() with
 * the same stream, which is how the two are compared.
 */
#ifndef SDTFE_ORACLE_H
#define SDTFE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest value of a stream draw; equals glibc RAND_MAX (2^31 - 1). */
#define SDO_RAND_MAX 2147483647

/* Counters the oracle reports instead of the reference's stderr warnings. */
typedef struct {
  int64_t samples;          /* sample-rays requested (G*G*S)              */
  int64_t missed;           /* entry lookup outside the hull shadow       */
  int64_t failed;           /* walk returned 0 with no tets (any reason)  */
  int64_t no_intersection;  /* src/dtfe.c:321-325                          */
  int64_t perturb_exhausted;/* src/dtfe.c:338-342                          */
  int64_t no_isect_perturb; /* src/dtfe.c:344-348                          */
  int64_t perturb_steps;    /* number of perturb_ray calls                 */
  int64_t tet_steps;        /* tets crossed, summed over all rays          */
} sdo_stats;

/* Counter-based sample stream: 31-bit draw in [0, SDO_RAND_MAX]. */
uint32_t sdo_draw31(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t draw);

/* A2  src/triangulate.c:183-225 -- volumes scattered to vertices, rho=4m/W. */
void sdo_pre_compute_vol(double *particle_data, int n_particles,
                         const int *tetra_data, int n_tetra, float p_mass);

/* Per-tet |det|/6 only (the first half of A2), for tests. */
void sdo_tet_volumes(const double *particle_data, const int *tetra_data,
                     int n_tetra, double *vol);

/* A3  src/triangulate.c:228-260 -- nabla is n_tetra x 3. */
void sdo_gradients(const int *tetra_data, int n_tetra,
                   const double *particle_data, double *nabla);

/* A4  src/triangulate.c:263-316 -- face_mask[t] bit f set when face f of
 * tet t lies on the hull and faces -z.  Returns the number of such faces. */
int sdo_forward_faces(const int *tetra_data, int n_tetra,
                      const double *particle_data, unsigned char *face_mask);

/* A5  src/triangulate.c:318-395 -- tri is n_tri x 8 int32 rows
 * (a, b, c, n_CB, n_AC, n_BA, tet, 0). */
void sdo_copy_fh_triangles(const int *tetra_data, int n_tetra,
                           const unsigned char *face_mask, int n_tri, int *tri);

/* A6  src/dtfe.c:518-552 -- walk from triangle 0; returns entry tet or -1.
 * max_steps bounds the walk (the reference has no bound); -2 when hit. */
int sdo_pt_loc_2d(const int *tri, const double *particle_data,
                  double qx, double qy, int max_steps);

/* A9  src/dtfe.c:28-215 for the z-directed ray through (rx, ry) of length
 * lz.  Returns 1 when an enter and a leave face were found. */
int sdo_plucker(double lz, double rx, double ry, const double *vert[4],
                int *enter_face, int *leave_face, double *enter_z,
                double *leave_z, int *special);

/* A8 + A10  src/dtfe.c:217-374 -- one ray.  (pixel, sample) key the
 * perturbation draws, which start at draw index 2.  *crossed gets the tets
 * crossed (0 on failure).  stats may be NULL. */
double sdo_shoot_ray(int entry_tet, const int *tetra_data, const double *nabla,
                     const double *particle_data, double x, double y,
                     double offset_z, double box_len_z, double delta_xy,
                     uint64_t seed, uint64_t pixel, uint32_t sample,
                     int *crossed, sdo_stats *stats);

/* A1  src/dtfe.c:554-654 with the counter-based stream.  Same argument
 * meaning as the reference compute_density plus seed, thread count
 * (<=0: OpenMP default), optional stats and optional per-pixel n_avg out. */
void sdo_compute_density(double *particle_data, int n_particles,
                         const int *tetra_data, int n_tetra, int grid_dim,
                         double box_len, double box_len_z, double *rho,
                         float p_mass, double center_x, double center_y,
                         double center_z, int n_mc_samp, double delta_sample,
                         uint64_t seed, int n_threads, sdo_stats *stats,
                         int *n_avg_out);

/* The walk part of A1 only, given precomputed rho (in particle_data[:,3]),
 * nabla and triangle list; processes the 4-row tile bands
 * band_begin, band_begin+band_step, ... (for bounded CPU baselines).
 * Returns the number of sample-rays traced. */
int64_t sdo_map_bands(const double *particle_data, const int *tetra_data,
                      const double *nabla, const int *tri, int grid_dim,
                      double box_len, double box_len_z, double *rho,
                      double center_x, double center_y, double center_z,
                      int n_mc_samp, double delta_sample, uint64_t seed,
                      int band_begin, int band_step, int n_threads,
                      sdo_stats *stats, int *n_avg_out);

/* F2  src/dtfe.c:376-516 -- one ray of the volumetric path: in each tet
 * crossed, adds ss_weight = 1/ss^3 times the linear field at the fine z
 * points offset_z + dzs*(k + 1/2) (dzs = lz/nz/ss, enter_z <= z < leave_z)
 * into rho_vector[k / ss].  (pixel, sample) key the perturbation draws,
 * which start at draw 2.  Returns tets crossed (0 on failure, when what was
 * added stays, as in the reference). */
int sdo_ray_3d(int entry, const int *tetra_data, const double *nabla,
               const double *particle_data, double x, double y, double offset_z,
               double lz, int nz, double delta_xy, unsigned supersampling,
               uint64_t seed, uint64_t pixel, uint32_t sample, double *rho_vector,
               sdo_stats *stats);

/* The column loop of interpolate_to_3d_grid (src/dtfe.c:686-731) given the
 * field (particle_data[:,3]), nabla and triangle list.  grid is G^3 in
 * [x][y][z] order and is accumulated into.  Ray (is, js) of column
 * c = i*G + j is keyed (seed, c, is*ss + js).  entries (nullable): entry tet
 * per ray in (c, is, js) order instead of pt_loc_2d. */
void sdo_volume_columns(const double *particle_data, const int *tetra_data,
                        const double *nabla, const int *tri, int grid_dim,
                        double box_len, double *grid, double center_x,
                        double center_y, double center_z, unsigned supersampling,
                        uint64_t seed, const int *entries, int n_threads,
                        sdo_stats *stats);

/* interpolate_to_3d_grid (src/dtfe.c:667-735) with the keyed stream. */
void sdo_interpolate_3d(const double *particle_data, int n_particles,
                        const int *tetra_data, int n_tetra, int grid_dim,
                        double box_len, double *grid, double center_x,
                        double center_y, double center_z, unsigned supersampling,
                        uint64_t seed, int n_threads, sdo_stats *stats);

/* compute_3d_density (src/dtfe.c:657-665): pre_compute_vol, then the above. */
void sdo_compute_3d_density(double *particle_data, int n_particles,
                            const int *tetra_data, int n_tetra, int grid_dim,
                            double box_len, double *rho, float p_mass,
                            double center_x, double center_y, double center_z,
                            unsigned supersampling, uint64_t seed, int n_threads,
                            sdo_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
