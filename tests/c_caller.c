/* A C caller written against the reference's prototypes (src/dtfe.h:13-15,
 * src/triangulate.h:26-29, src/utils.h:10-12, src/bitmap.h:9-17,
 * src/io.h:35-36), linked against libsdtfe_amd.so alone.  It runs what the
 * reference's CLI and binding call, in their order, and writes the results
 * as raw files for the test to compare:
 *
 *   c_caller pos.bin n tets.bin nt G box_len box_len_z p_mass cx cy cz S
 *            delta_sample seed outdir
 *
 * pos.bin: n x 4 float64 rows; tets.bin: nt x 8 int32 rows.  Writes
 * outdir/rho.bin (the map, rho zeroed first), vrho.bin (column 3 after
 * compute_density), nabla.bin (gradients of that column), tri.bin
 * (find_forward_hull + copy_fh_triangles rows) and bitmap.bin.  Exits
 * non-zero when a call reports a failure. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t word_t;

/* the reference's prototypes */
void compute_density(double *particle_data, int n_particles, int *tetra_data, int n_tetra,
                     int grid_dim, double box_len, double box_len_z, double *rho, float p_mass,
                     double center_x, double center_y, double center_z, const int n_mc_samp,
                     const double delta_sample);
void pre_compute_vol(double **, int, int *, int, float);
double *gradients(int *, int, double *);
int find_forward_hull(int *, int, double *, int, word_t *);
int *copy_fh_triangles(int *, int, double *, int, int, word_t *);
word_t *create_bitmap(int);
void delete_bitmap(word_t *);
int write_rho(char *, double **, int);

/* extensions of libsdtfe_amd.so: the sample stream's seed and the status
 * of the last drop-in call */
void sdtfe_set_seed(uint64_t seed);
int sdtfe_dropin_status(void);
const char *sdtfe_last_error(void);

static void *read_all(const char *path, size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  FILE *f = fopen(path, "rb");
  if (!p || !f || fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot read %s\n", path);
    exit(2);
  }
  fclose(f);
  return p;
}

static void check(const char *what) {
  if (sdtfe_dropin_status() != 0) {
    fprintf(stderr, "%s failed: %s\n", what, sdtfe_last_error());
    exit(3);
  }
}

static void write_raw(const char *dir, const char *name, const void *p, size_t bytes) {
  char path[1024];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot write %s\n", path);
    exit(2);
  }
  fclose(f);
}

int main(int argc, char **argv) {
  if (argc != 16) {
    fprintf(stderr, "usage: %s pos.bin n tets.bin nt G L Lz m cx cy cz S ds seed outdir\n",
            argv[0]);
    return 1;
  }
  const int n = atoi(argv[2]), nt = atoi(argv[4]), G = atoi(argv[5]);
  const double L = atof(argv[6]), Lz = atof(argv[7]);
  const float m = (float)atof(argv[8]);
  const double cx = atof(argv[9]), cy = atof(argv[10]), cz = atof(argv[11]);
  const int S = atoi(argv[12]);
  const double ds = atof(argv[13]);
  const uint64_t seed = strtoull(argv[14], NULL, 10);
  const char *out = argv[15];
  double *pd = read_all(argv[1], sizeof(double) * 4 * (size_t)n);
  int *tets = read_all(argv[3], sizeof(int) * 8 * (size_t)nt);

  double *rho = calloc((size_t)G * G, sizeof(double));
  sdtfe_set_seed(seed);
  compute_density(pd, n, tets, nt, G, L, Lz, rho, m, cx, cy, cz, S, ds);
  check("compute_density");
  char path[1024];
  snprintf(path, sizeof path, "%s/rho.bin", out);
  if (write_rho(path, &rho, G * G) != G * G) return 4;

  double *vrho = malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) vrho[i] = pd[4 * (size_t)i + 3];
  write_raw(out, "vrho.bin", vrho, sizeof(double) * (size_t)n);

  /* pre_compute_vol on a copy gives the same column */
  double *pd2 = malloc(sizeof(double) * 4 * (size_t)n);
  memcpy(pd2, pd, sizeof(double) * 4 * (size_t)n);
  for (int i = 0; i < n; ++i) pd2[4 * (size_t)i + 3] = -1.0;
  pre_compute_vol(&pd2, n, tets, nt, m);
  check("pre_compute_vol");
  for (int i = 0; i < n; ++i)
    if (pd2[4 * (size_t)i + 3] != vrho[i]) {
      fprintf(stderr, "pre_compute_vol differs from compute_density at %d\n", i);
      return 5;
    }

  double *nabla = gradients(tets, nt, pd);
  check("gradients");
  write_raw(out, "nabla.bin", nabla, sizeof(double) * 3 * (size_t)nt);
  free(nabla);

  word_t *bm = create_bitmap(4 * nt);
  const int nf = find_forward_hull(tets, nt, pd, n, bm);
  check("find_forward_hull");
  int *tri = copy_fh_triangles(tets, nt, pd, n, nf, bm);
  check("copy_fh_triangles");
  write_raw(out, "tri.bin", tri, sizeof(int) * 8 * (size_t)nf);
  write_raw(out, "bitmap.bin", bm, sizeof(word_t) * ((size_t)nt + 1));
  free(tri);
  delete_bitmap(bm);
  printf("c_caller ok: %d faces\n", nf);
  free(rho);
  free(vrho);
  free(pd2);
  free(pd);
  free(tets);
  return 0;
}
