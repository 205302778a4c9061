/*
 * dtfe_oracle.c -- CPU restatement of the SDTFE surface-density path.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline); never linked into the
 * product.  Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off, no
 * -march flags, so every double operation is one correctly rounded SSE2
 * operation, as in the reference's own x86-64 build).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - golden vectors written by the reference C compiled from its own sources
 *     (oracle/Makefile target `ref`, tests/golden/make_golden.py): per-tet
 *     density and gradient, forward-face triangle rows, entry lookups, single
 *     rays including crafted special cases that exercise the perturbation,
 *     and whole maps from the reference compute_density;
 *   - closed forms: single-tet volume and density, exact dyadic volumes,
 *     linear-field gradient recovery, chord integrals through one tet,
 *     mass conservation of the DTFE field, brute-force point location.
 */
#include "dtfe_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#if defined(_OPENMP)
#include <omp.h>
#endif

/* src/constants.h:6-8 */
#define SDO_ZERO (1e4 * 2.220446049250313080847e-16)
#define SDO_PERTURB_MAX 1000

/* Vertex slots of a tetra_data row, src/constants.h:12. */
enum { VA = 0, VB = 1, VC = 2, VD = 3 };
/* Faces, src/constants.h:10: face f is opposite vertex f and its neighbour
 * sits in row slot 4+f (src/triangulate.c:155-171). */
enum { F_BDC = 0, F_ACD = 1, F_ADB = 2, F_ABC = 3 };
/* Edges, src/constants.h:11. */
enum { E_CB = 0, E_AC = 1, E_BA = 2, E_DC = 3, E_BD = 4, E_DA = 5 };

/* ---------------------------------------------------------------- stream */

static inline uint64_t sdo_mix64(uint64_t z) {
  /* SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014). */
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

#define SDO_GAMMA 0x9E3779B97F4A7C15ULL

uint32_t sdo_draw31(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t draw) {
  uint64_t k = sdo_mix64(seed + SDO_GAMMA);
  k = sdo_mix64(k ^ pixel);
  k = sdo_mix64(k ^ (uint64_t)sample);
  return (uint32_t)(sdo_mix64(k + (uint64_t)(draw + 1u) * SDO_GAMMA) >> 33);
}

/* ------------------------------------------------------------ field (A2,3) */

static inline double tet_signed_vol6(const double *a, const double *b,
                                     const double *c, const double *d) {
  /* src/triangulate.c:191-213: expansion along the rows of (A-D), each
   * minor taken from (B-D) and (C-D), summed left to right. */
  double m0 = (b[1] - d[1]) * (c[2] - d[2]) - (b[2] - d[2]) * (c[1] - d[1]);
  double m1 = (b[2] - d[2]) * (c[0] - d[0]) - (b[0] - d[0]) * (c[2] - d[2]);
  double m2 = (b[0] - d[0]) * (c[1] - d[1]) - (b[1] - d[1]) * (c[0] - d[0]);
  return (a[0] - d[0]) * m0 + (a[1] - d[1]) * m1 + (a[2] - d[2]) * m2;
}

void sdo_tet_volumes(const double *pd, const int *tet, int nt, double *vol) {
  for (int t = 0; t < nt; ++t) {
    const int *r = tet + 8 * (size_t)t;
    double v = tet_signed_vol6(pd + 4 * (size_t)r[VA], pd + 4 * (size_t)r[VB],
                               pd + 4 * (size_t)r[VC], pd + 4 * (size_t)r[VD]) / 6;
    vol[t] = fabs(v);
  }
}

void sdo_pre_compute_vol(double *pd, int n, const int *tet, int nt, float p_mass) {
  for (int i = 0; i < n; ++i) pd[4 * (size_t)i + 3] = 0;
  /* Ascending tet order per vertex, as src/triangulate.c:190-221. */
  for (int t = 0; t < nt; ++t) {
    const int *r = tet + 8 * (size_t)t;
    double v = fabs(tet_signed_vol6(pd + 4 * (size_t)r[VA], pd + 4 * (size_t)r[VB],
                                    pd + 4 * (size_t)r[VC], pd + 4 * (size_t)r[VD]) / 6);
    for (int k = 0; k < 4; ++k) pd[4 * (size_t)r[k] + 3] += v;
  }
  /* src/triangulate.c:223-224: the mass factor is a float product. */
  const float m4 = p_mass * 4;
  for (int i = 0; i < n; ++i) pd[4 * (size_t)i + 3] = m4 / pd[4 * (size_t)i + 3];
}

static inline double det3_ref(double a, double b, double c, double d, double e,
                              double f, double g, double h, double i) {
  /* src/triangulate.c:228-230, left-to-right sum of the six terms. */
  return a * e * i + b * f * g + c * d * h - c * e * g - b * d * i - a * f * h;
}

void sdo_gradients(const int *tet, int nt, const double *pd, double *nabla) {
  for (int t = 0; t < nt; ++t) {
    const int *r = tet + 8 * (size_t)t;
    const double *a = pd + 4 * (size_t)r[VA];
    double dx[3], dy[3], dz[3], dr[3];
    for (int k = 0; k < 3; ++k) {
      const double *p = pd + 4 * (size_t)r[1 + k];
      dx[k] = p[0] - a[0];
      dy[k] = p[1] - a[1];
      dz[k] = p[2] - a[2];
      dr[k] = p[3] - a[3];
    }
    double den = det3_ref(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], dx[2], dy[2], dz[2]);
    nabla[3 * (size_t)t + 0] =
        det3_ref(dr[0], dy[0], dz[0], dr[1], dy[1], dz[1], dr[2], dy[2], dz[2]) / den;
    nabla[3 * (size_t)t + 1] =
        det3_ref(dx[0], dr[0], dz[0], dx[1], dr[1], dz[1], dx[2], dr[2], dz[2]) / den;
    nabla[3 * (size_t)t + 2] =
        det3_ref(dx[0], dy[0], dr[0], dx[1], dy[1], dr[1], dx[2], dy[2], dr[2]) / den;
  }
}

/* ------------------------------------------------------ entry (A4, A5, A6) */

/* For face f: the vertex triple emitted as a triangle (src/triangulate.c:
 * 328-343) and the two edge vectors whose z-cross decides "forward"
 * (src/triangulate.c:279-308): t1 = tri_edge[f][0] - base, t2 = ... */
static const int face_tri[4][3] = {
    {VB, VD, VC}, /* BDC */
    {VA, VC, VD}, /* ACD */
    {VA, VD, VB}, /* ADB */
    {VA, VB, VC}, /* ABC */
};

int sdo_forward_faces(const int *tet, int nt, const double *pd, unsigned char *mask) {
  int count = 0;
  for (int t = 0; t < nt; ++t) {
    const int *r = tet + 8 * (size_t)t;
    unsigned char m = 0;
    for (int f = 0; f < 4; ++f) {
      if (r[4 + f] != -1) continue;
      /* base, first and second point of the face's edge pair; for every face
       * the pair is (tri[1]-tri[0], tri[2]-tri[0]). */
      const double *p0 = pd + 4 * (size_t)r[face_tri[f][0]];
      const double *p1 = pd + 4 * (size_t)r[face_tri[f][1]];
      const double *p2 = pd + 4 * (size_t)r[face_tri[f][2]];
      double t1x = p1[0] - p0[0], t1y = p1[1] - p0[1];
      double t2x = p2[0] - p0[0], t2y = p2[1] - p0[1];
      if (t1x * t2y - t1y * t2x < 0.0) {
        m |= (unsigned char)(1u << f);
        ++count;
      }
    }
    if (mask) mask[t] = m;
  }
  return count;
}

typedef struct {
  int lo, hi, tri, slot;
} sdo_edge_ref;

static int cmp_edge_ref(const void *pa, const void *pb) {
  const sdo_edge_ref *a = (const sdo_edge_ref *)pa, *b = (const sdo_edge_ref *)pb;
  if (a->lo != b->lo) return a->lo < b->lo ? -1 : 1;
  if (a->hi != b->hi) return a->hi < b->hi ? -1 : 1;
  return a->tri < b->tri ? -1 : (a->tri > b->tri);
}

void sdo_copy_fh_triangles(const int *tet, int nt, const unsigned char *mask,
                           int n_tri, int *tri) {
  int k = 0;
  for (int t = 0; t < nt && k < n_tri; ++t) {
    for (int f = 0; f < 4; ++f) {
      if (!(mask[t] >> f & 1u)) continue;
      const int *r = tet + 8 * (size_t)t;
      int *o = tri + 8 * (size_t)k;
      o[0] = r[face_tri[f][0]];
      o[1] = r[face_tri[f][1]];
      o[2] = r[face_tri[f][2]];
      o[3] = o[4] = o[5] = -1;
      o[6] = t;
      o[7] = 0;
      ++k;
    }
  }
  /* Neighbours across shared edges (src/triangulate.c:358-392): slot 5 for
   * edge (a,b), 4 for (a,c), 3 for (b,c).  The reference scans all pairs and
   * the highest-numbered partner wins; a sort by edge reproduces that. */
  sdo_edge_ref *e = (sdo_edge_ref *)malloc(sizeof(sdo_edge_ref) * 3 * (size_t)(n_tri > 0 ? n_tri : 1));
  static const int pair[3][3] = {{0, 1, 5}, {0, 2, 4}, {1, 2, 3}};
  for (int i = 0; i < n_tri; ++i)
    for (int q = 0; q < 3; ++q) {
      int u = tri[8 * (size_t)i + pair[q][0]], v = tri[8 * (size_t)i + pair[q][1]];
      sdo_edge_ref *x = &e[3 * (size_t)i + q];
      x->lo = u < v ? u : v;
      x->hi = u < v ? v : u;
      x->tri = i;
      x->slot = pair[q][2];
    }
  size_t ne = 3 * (size_t)n_tri;
  qsort(e, ne, sizeof(sdo_edge_ref), cmp_edge_ref);
  for (size_t s = 0; s < ne;) {
    size_t g = s;
    while (g < ne && e[g].lo == e[s].lo && e[g].hi == e[s].hi) ++g;
    for (size_t a = s; a < g; ++a) {
      int best = -1;
      for (size_t b = s; b < g; ++b)
        if (e[b].tri != e[a].tri && e[b].tri > best) best = e[b].tri;
      if (best >= 0) tri[8 * (size_t)e[a].tri + e[a].slot] = best;
    }
    s = g;
  }
  free(e);
}

int sdo_pt_loc_2d(const int *tri, const double *pd, double qx, double qy, int max_steps) {
  /* src/dtfe.c:518-552: visibility walk from triangle 0.  For each edge the
   * query is compared with the triangle's third vertex; strictly opposite
   * sides means step across (points on an edge count as inside). */
  static const int edge_seq[3][3] = {
      /* edge start, edge end, opposite vertex, neighbour slot */
      {0, 1, 2}, /* AB -> slot 5 */
      {1, 2, 0}, /* BC -> slot 3 */
      {2, 0, 1}, /* CA -> slot 4 */
  };
  static const int edge_slot[3] = {5, 3, 4};
  int cur = 0;
  for (int step = 0; step < max_steps; ++step) {
    const int *r = tri + 8 * (size_t)cur;
    int moved = 0;
    for (int q = 0; q < 3; ++q) {
      const double *p = pd + 4 * (size_t)r[edge_seq[q][0]];
      const double *s = pd + 4 * (size_t)r[edge_seq[q][1]];
      const double *o = pd + 4 * (size_t)r[edge_seq[q][2]];
      double ex = s[0] - p[0], ey = s[1] - p[1];
      double d_ins = (o[0] - p[0]) * ey - (o[1] - p[1]) * ex;
      double d_qry = (qx - p[0]) * ey - (qy - p[1]) * ex;
      if (d_ins * d_qry < 0.0) {
        cur = r[edge_slot[q]];
        moved = 1;
        break;
      }
    }
    if (!moved) return r[6];
    if (cur == -1) return -1;
  }
  return -2;
}

/* -------------------------------------------------------------- walk (A9) */

/* Edge e is stored as the ordered pair (X, Y) whose Plücker coordinates the
 * reference forms as (X - Y, X.x*Y.y - X.y*Y.x) (src/dtfe.c:44-147). */
static const int edge_xy[6][2] = {
    {VC, VB}, /* CB */
    {VC, VA}, /* AC */
    {VB, VA}, /* BA */
    {VC, VD}, /* DC */
    {VB, VD}, /* BD */
    {VD, VA}, /* DA */
};

/* Per face, in test order (src/dtfe.c:64-213): the three edges with the sign
 * each carries in the face's orientation sum (the same signs make the enter
 * test s*sigma >= 0 and the leave test s*sigma <= 0), the two weighted edges
 * of the barycentric point and the three points (p0 gets 1 - w1 - w2). */
typedef struct {
  int e[3];
  int s[3];
  int w1_e, w1_s, w2_e, w2_s;
  int p[3];
} sdo_face;

static const sdo_face faces[4] = {
    {{E_CB, E_DC, E_BD}, {-1, +1, -1}, E_BD, -1, E_DC, +1, {VD, VC, VB}}, /* BDC */
    {{E_AC, E_DC, E_DA}, {+1, -1, -1}, E_AC, +1, E_DC, -1, {VC, VD, VA}}, /* ACD */
    {{E_DA, E_BD, E_BA}, {+1, +1, -1}, E_BD, +1, E_BA, -1, {VB, VA, VD}}, /* ADB */
    {{E_BA, E_CB, E_AC}, {+1, +1, -1}, E_AC, -1, E_BA, +1, {VA, VB, VC}}, /* ABC */
};

static inline int sdo_sign(double x) { return x > SDO_ZERO ? 1 : (x < -SDO_ZERO ? -1 : 0); }
static inline double sgn_apply(int s, double v) { return s < 0 ? -v : v; }

static inline double sdo_edge_u(double p2, double p3, double p4, const double *x,
                                 const double *y) {
  double t0 = x[0] - y[0], t1 = x[1] - y[1];
  double t5 = x[0] * y[1] - x[1] * y[0];
  return p2 * t5 + t0 * p3 + t1 * p4; /* src/dtfe.c:24-26 */
}

/* Test face f once its three edges are known; returns 1 when this face
 * completes the enter/leave pair. */
static inline int sdo_face_test(int f, const double u[6], const int sg[6],
                                const double *vert[4], int *enter_face, int *leave_face,
                                double *enter_z, double *leave_z, int *special, int *found) {
  const sdo_face *F = &faces[f];
  int a = F->s[0] * sg[F->e[0]], b = F->s[1] * sg[F->e[1]], c = F->s[2] * sg[F->e[2]];
  if (a == 0 && b == 0 && c == 0) return 0;
  int is_enter = (*enter_face == -1) && a >= 0 && b >= 0 && c >= 0;
  int is_leave = !is_enter && (*leave_face == -1) && a <= 0 && b <= 0 && c <= 0;
  if (!is_enter && !is_leave) return 0;
  double den = sgn_apply(F->s[0], u[F->e[0]]) + sgn_apply(F->s[1], u[F->e[1]]);
  den = den + sgn_apply(F->s[2], u[F->e[2]]);
  double inv = 1.0 / den;
  double w1 = sgn_apply(F->w1_s, u[F->w1_e]) * inv;
  double w2 = sgn_apply(F->w2_s, u[F->w2_e]) * inv;
  double z = (1 - w1 - w2) * vert[F->p[0]][2] + w1 * vert[F->p[1]][2] + w2 * vert[F->p[2]][2];
  if (is_enter) {
    *enter_face = f;
    *enter_z = z;
  } else {
    *leave_face = f;
    *leave_z = z;
  }
  if (a == 0 || b == 0 || c == 0) *special = 1;
  return ++*found == 2;
}

int sdo_plucker(double lz, double rx, double ry, const double *vert[4], int *enter_face,
                int *leave_face, double *enter_z, double *leave_z, int *special) {
  /* Ray Plücker coordinates: direction (0,0,lz), moment (-lz*y, lz*x, 0)
   * (src/dtfe.c:266-268, 296-299).  Faces are tested in the order BDC, ACD,
   * ADB, ABC and each edge product is formed just before the first face that
   * needs it; the search stops at the second face found. */
  const double p2 = lz, p3 = -lz * ry, p4 = lz * rx;
  double u[6];
  int sg[6];
  int found = 0;
  *enter_face = -1;
  *leave_face = -1;
  *special = 0;
#define SDO_EDGE(e)                                                               \
  do {                                                                            \
    u[e] = sdo_edge_u(p2, p3, p4, vert[edge_xy[e][0]], vert[edge_xy[e][1]]);     \
    sg[e] = sdo_sign(u[e]);                                                       \
  } while (0)
#define SDO_FACE(f)                                                                \
  if (sdo_face_test(f, u, sg, vert, enter_face, leave_face, enter_z, leave_z,      \
                    special, &found))                                              \
  return 1
  SDO_EDGE(E_BD);
  SDO_EDGE(E_DC);
  SDO_EDGE(E_CB);
  SDO_FACE(F_BDC);
  SDO_EDGE(E_AC);
  SDO_EDGE(E_DA);
  SDO_FACE(F_ACD);
  SDO_EDGE(E_BA);
  SDO_FACE(F_ADB);
  SDO_FACE(F_ABC);
#undef SDO_EDGE
#undef SDO_FACE
  return 0;
}

/* ------------------------------------------------------- walk (A8, A10) */

double sdo_shoot_ray(int entry, const int *tet, const double *nabla, const double *pd,
                     double x, double y, double off_z, double lz, double dxy,
                     uint64_t seed, uint64_t pixel, uint32_t sample, int *crossed,
                     sdo_stats *st) {
  *crossed = 0;
  if (entry == -1) return 0.0;
  const double max_dist = dxy / 2.0;              /* src/dtfe.c:273 */
  const double step = max_dist / SDO_PERTURB_MAX; /* src/dtfe.c:274 */
  double ox = x, oy = y; /* the ray's current (possibly perturbed) x, y */
  double acc = 0.0;
  uint32_t draw = 2; /* draws 0 and 1 placed the sample */
  const int *row = NULL; /* NULL: the seed record whose slot 4 is `entry` */
  int cur = entry, prev = -1;
  int f_in = -1, f_out = 0;
  for (;;) {
    int nxt = row ? row[4 + f_out] : entry;
    if (nxt == -1) break;
    if (nxt == prev) { /* turn around, src/dtfe.c:303-309 */
      int t = f_in;
      f_in = f_out;
      f_out = t;
      nxt = row[4 + f_out];
      if (nxt == -1) break;
    }
    prev = cur;
    cur = nxt;
    row = tet + 8 * (size_t)cur;
    const double *v[4] = {pd + 4 * (size_t)row[0], pd + 4 * (size_t)row[1],
                          pd + 4 * (size_t)row[2], pd + 4 * (size_t)row[3]};
    double z_in = 0, z_out = 0;
    int special = 0;
    if (!sdo_plucker(lz, ox, oy, v, &f_in, &f_out, &z_in, &z_out, &special)) {
      if (st) st->no_intersection++;
      *crossed = 0;
      return 0.0;
    }
    int tries = 0;
    while (special) {
      /* perturb_ray, src/dtfe.c:219-245: move toward a random corner. */
      int r = (int)(sdo_draw31(seed, pixel, sample, draw++) % 4u);
      double d0 = v[r][0] - ox, d1 = v[r][1] - oy;
      double mag = sqrt(d0 * d0 + d1 * d1);
      if (mag > step) {
        d0 = d0 / mag;
        d1 = d1 / mag;
        d0 = d0 * step;
        d1 = d1 * step;
      }
      ox = ox + d0;
      oy = oy + d1;
      if (st) st->perturb_steps++;
      if (++tries > SDO_PERTURB_MAX) {
        if (st) st->perturb_exhausted++;
        *crossed = 0;
        return 0.0;
      }
      if (!sdo_plucker(lz, ox, oy, v, &f_in, &f_out, &z_in, &z_out, &special)) {
        if (st) st->no_isect_perturb++;
        *crossed = 0;
        return 0.0;
      }
    }
    ++*crossed;
    /* Midpoint rule on the linear field, src/dtfe.c:356-369. */
    double dz = z_out - z_in;
    double zm = dz / 2 + z_in;
    if (zm > off_z && zm < off_z + lz) {
      const double *a = v[0];
      const double *g = nabla + 3 * (size_t)cur;
      double fixed = a[3] + g[0] * (ox - a[0]) + g[1] * (oy - a[1]) - g[2] * a[2];
      acc += (fixed + g[2] * zm) * dz;
    }
  }
  return acc;
}

/* --------------------------------------------------------------- map (A1) */

typedef struct {
  double dxy, off_x, off_y, off_z, scale, soff, lz;
} sdo_grid;

static sdo_grid sdo_make_grid(int G, double box_len, double box_len_z, double cx,
                              double cy, double cz, double delta_sample) {
  /* src/dtfe.c:579-589 */
  sdo_grid g;
  g.dxy = box_len / G;
  g.off_x = -(box_len / 2.0) + cx;
  g.off_y = -(box_len / 2.0) + cy;
  g.off_z = -(box_len_z / 2.0) + cz;
  g.scale = delta_sample > g.dxy ? delta_sample : g.dxy;
  g.soff = delta_sample > g.dxy ? (delta_sample - g.dxy) / 2.0 : 0.0;
  g.lz = box_len_z;
  return g;
}

static void sdo_stats_add(sdo_stats *dst, const sdo_stats *s) {
  dst->samples += s->samples;
  dst->missed += s->missed;
  dst->failed += s->failed;
  dst->no_intersection += s->no_intersection;
  dst->perturb_exhausted += s->perturb_exhausted;
  dst->no_isect_perturb += s->no_isect_perturb;
  dst->perturb_steps += s->perturb_steps;
  dst->tet_steps += s->tet_steps;
}

static void sdo_pixel(const double *pd, const int *tet, const double *nabla, const int *tri,
                      int G, const sdo_grid *g, int S, uint64_t seed, int px, int py,
                      double *rho, int *n_avg_out, sdo_stats *st) {
  /* src/dtfe.c:613-645 */
  const uint64_t pix = (uint64_t)py * (uint64_t)G + (uint64_t)px;
  double x = px * g->dxy + g->off_x - g->soff;
  double y = py * g->dxy + g->off_y - g->soff;
  double sum = 0.0;
  int n_avg = S;
  for (int r = 0; r < S; ++r) {
    double xt = x + ((double)sdo_draw31(seed, pix, (uint32_t)r, 0) / SDO_RAND_MAX) * g->scale;
    double yt = y + ((double)sdo_draw31(seed, pix, (uint32_t)r, 1) / SDO_RAND_MAX) * g->scale;
    st->samples++;
    int ft = sdo_pt_loc_2d(tri, pd, xt, yt, 1 << 24);
    if (ft < 0) {
      st->missed++;
      --n_avg;
      continue;
    }
    int crossed = 0;
    double v = sdo_shoot_ray(ft, tet, nabla, pd, xt, yt, g->off_z, g->lz, g->dxy, seed, pix,
                             (uint32_t)r, &crossed, st);
    sum += v;
    st->tet_steps += crossed;
    if (crossed == 0) {
      st->failed++;
      --n_avg;
    }
  }
  if (n_avg > 0) rho[pix] += sum / n_avg;
  else rho[pix] += 0.0;
  if (n_avg_out) n_avg_out[pix] = n_avg;
}

int64_t sdo_map_bands(const double *pd, const int *tet, const double *nabla, const int *tri,
                      int G, double box_len, double box_len_z, double *rho, double cx,
                      double cy, double cz, int S, double delta_sample, uint64_t seed,
                      int band_begin, int band_step, int n_threads, sdo_stats *stats,
                      int *n_avg_out) {
  const sdo_grid g = sdo_make_grid(G, box_len, box_len_z, cx, cy, cz, delta_sample);
  const int tiles_x = G / 4, bands = G / 4;
  if (band_step < 1) band_step = 1;
  int n_band_sel = band_begin < bands ? (bands - 1 - band_begin) / band_step + 1 : 0;
  const int64_t n_tiles = (int64_t)n_band_sel * tiles_x;
  sdo_stats total;
  memset(&total, 0, sizeof total);
#if defined(_OPENMP)
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
  {
    sdo_stats loc;
    memset(&loc, 0, sizeof loc);
    /* 4x4 pixel tiles handed out one at a time, as src/dtfe.c:601-608. */
#if defined(_OPENMP)
#pragma omp for schedule(dynamic, 1)
#endif
    for (int64_t it = 0; it < n_tiles; ++it) {
      int band = band_begin + (int)(it / tiles_x) * band_step;
      int tx = (int)(it % tiles_x);
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k)
          sdo_pixel(pd, tet, nabla, tri, G, &g, S, seed, tx * 4 + k, band * 4 + j, rho,
                    n_avg_out, &loc);
    }
#if defined(_OPENMP)
#pragma omp critical
#endif
    sdo_stats_add(&total, &loc);
  }
  if (stats) sdo_stats_add(stats, &total);
  return n_tiles * 16 * (int64_t)S;
}

void sdo_compute_density(double *pd, int n, const int *tet, int nt, int G, double box_len,
                         double box_len_z, double *rho, float p_mass, double cx, double cy,
                         double cz, int S, double delta_sample, uint64_t seed, int n_threads,
                         sdo_stats *stats, int *n_avg_out) {
  sdo_pre_compute_vol(pd, n, tet, nt, p_mass);
  double *nabla = (double *)malloc(sizeof(double) * 3 * (size_t)(nt > 0 ? nt : 1));
  sdo_gradients(tet, nt, pd, nabla);
  unsigned char *mask = (unsigned char *)malloc((size_t)(nt > 0 ? nt : 1));
  int nf = sdo_forward_faces(tet, nt, pd, mask);
  int *tri = (int *)malloc(sizeof(int) * 8 * (size_t)(nf > 0 ? nf : 1));
  sdo_copy_fh_triangles(tet, nt, mask, nf, tri);
  free(mask);
  if (nf > 0)
    sdo_map_bands(pd, tet, nabla, tri, G, box_len, box_len_z, rho, cx, cy, cz, S,
                  delta_sample, seed, 0, 1, n_threads, stats, n_avg_out);
  free(tri);
  free(nabla);
}

/* ------------------------------------------------------ volumetric (F2) */

int sdo_ray_3d(int entry, const int *tet, const double *nabla, const double *pd, double x,
               double y, double off_z, double lz, int nz, double dxy, unsigned ss,
               uint64_t seed, uint64_t pixel, uint32_t sample, double *rho_vector,
               sdo_stats *st) {
  /* evaluate_3d_ray_on_grid, src/dtfe.c:376-516: the walk of shoot_3d_ray;
   * in each tet crossed the fine z points offset_z + dzs*(k + 1/2) with
   * enter_z <= z < leave_z (k clamped to [0, nz*ss)) add ss_weight times
   * the tet's linear field to cell k / ss.  A failed walk returns 0 and
   * keeps what it added (:444-449, 462-472). */
  int crossed = 0;
  if (entry == -1) return 0;
  const double dz = lz / nz;           /* :401 */
  const double dzs = dz / ss;          /* :403 */
  const int nzs = nz * ss;             /* :404 */
  const double ss_weight = 1. / (ss * ss * ss); /* :406 */
  const double max_dist = dxy / 2.0;   /* :408 */
  const double step = max_dist / SDO_PERTURB_MAX;
  double ox = x, oy = y;
  uint32_t draw = 2;
  const int *row = NULL;
  int cur = entry, prev = -1;
  int f_in = -1, f_out = 0;
  for (;;) {
    int nxt = row ? row[4 + f_out] : entry;
    if (nxt == -1) break;
    if (nxt == prev) { /* turn around, :433-439 */
      int t = f_in;
      f_in = f_out;
      f_out = t;
      nxt = row[4 + f_out];
      if (nxt == -1) break;
    }
    prev = cur;
    cur = nxt;
    row = tet + 8 * (size_t)cur;
    const double *v[4] = {pd + 4 * (size_t)row[0], pd + 4 * (size_t)row[1],
                          pd + 4 * (size_t)row[2], pd + 4 * (size_t)row[3]};
    double z_in = 0, z_out = 0;
    int special = 0;
    if (!sdo_plucker(lz, ox, oy, v, &f_in, &f_out, &z_in, &z_out, &special)) {
      if (st) st->no_intersection++;
      return 0;
    }
    int tries = 0;
    while (special) { /* perturb_ray, src/dtfe.c:219-245 */
      int r = (int)(sdo_draw31(seed, pixel, sample, draw++) % 4u);
      double d0 = v[r][0] - ox, d1 = v[r][1] - oy;
      double mag = sqrt(d0 * d0 + d1 * d1);
      if (mag > step) {
        d0 = d0 / mag;
        d1 = d1 / mag;
        d0 = d0 * step;
        d1 = d1 * step;
      }
      ox = ox + d0;
      oy = oy + d1;
      if (st) st->perturb_steps++;
      if (++tries > SDO_PERTURB_MAX) {
        if (st) st->perturb_exhausted++;
        return 0;
      }
      if (!sdo_plucker(lz, ox, oy, v, &f_in, &f_out, &z_in, &z_out, &special)) {
        if (st) st->no_isect_perturb++;
        return 0;
      }
    }
    ++crossed;
    /* fine points of this tet, :482-513 */
    int e = (int)ceil((z_in - off_z - dzs / 2.) / dzs);
    e = e < 0 ? 0 : e;
    int l = (int)ceil((z_out - off_z - dzs / 2.) / dzs);
    l = l > nzs ? nzs : l;
    if (l > e) {
      const double *a = v[0];
      const double *g = nabla + 3 * (size_t)cur;
      double fixed = a[3] + g[0] * (ox - a[0]) + g[1] * (oy - a[1]) - g[2] * a[2];
      for (int k = e; k < l; ++k) {
        double z = off_z + dzs * k + dzs / 2.;
        rho_vector[k / (int)ss] += (fixed + g[2] * z) * ss_weight;
      }
    }
  }
  return crossed;
}

void sdo_volume_columns(const double *pd, const int *tet, const double *nabla, const int *tri,
                        int G, double box_len, double *grid, double cx, double cy, double cz,
                        unsigned ss, uint64_t seed, const int *entries, int n_threads,
                        sdo_stats *stats) {
  /* interpolate_to_3d_grid's loop, src/dtfe.c:686-731: column (i, j) at x
   * index i, y index j owns grid[(i*G + j)*G ...]; its ss x ss rays, in
   * (is, js) order, add into it.  Columns are independent, so the result
   * does not depend on the thread count. */
  const double dxy = box_len / G;         /* :688 */
  const double off_x = -(box_len / 2.0) + cx;
  const double off_y = -(box_len / 2.0) + cy;
  const double off_z = -(box_len / 2.0) + cz;
  const double ds = dxy / ss;             /* :695 */
  const int64_t n_col = (int64_t)G * G;
  sdo_stats total;
  memset(&total, 0, sizeof total);
#if defined(_OPENMP)
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
  {
    sdo_stats loc;
    memset(&loc, 0, sizeof loc);
#if defined(_OPENMP)
#pragma omp for schedule(dynamic, 16)
#endif
    for (int64_t c = 0; c < n_col; ++c) {
      const int i = (int)(c / G), j = (int)(c % G);
      double *col = grid + (size_t)c * (size_t)G;
      for (int is = 0; is < (int)ss; ++is)
        for (int js = 0; js < (int)ss; ++js) {
          double x = i * dxy + off_x + is * ds + ds / 2.;
          double y = j * dxy + off_y + js * ds + ds / 2.;
          const uint32_t r = (uint32_t)(is * (int)ss + js);
          loc.samples++;
          int ft = entries ? entries[(size_t)c * ss * ss + r] : sdo_pt_loc_2d(tri, pd, x, y, 1 << 24);
          if (ft < 0) {
            loc.missed++;
            continue;
          }
          int crossed = sdo_ray_3d(ft, tet, nabla, pd, x, y, off_z, box_len, G, dxy, ss, seed,
                                   (uint64_t)c, r, col, &loc);
          loc.tet_steps += crossed;
          if (crossed == 0) loc.failed++;
        }
    }
#if defined(_OPENMP)
#pragma omp critical
#endif
    sdo_stats_add(&total, &loc);
  }
  if (stats) sdo_stats_add(stats, &total);
}

void sdo_interpolate_3d(const double *pd, int n, const int *tet, int nt, int G, double box_len,
                        double *grid, double cx, double cy, double cz, unsigned ss,
                        uint64_t seed, int n_threads, sdo_stats *stats) {
  /* interpolate_to_3d_grid, src/dtfe.c:667-735 */
  (void)n;
  double *nabla = (double *)malloc(sizeof(double) * 3 * (size_t)(nt > 0 ? nt : 1));
  sdo_gradients(tet, nt, pd, nabla);
  unsigned char *mask = (unsigned char *)malloc((size_t)(nt > 0 ? nt : 1));
  int nf = sdo_forward_faces(tet, nt, pd, mask);
  int *tri = (int *)malloc(sizeof(int) * 8 * (size_t)(nf > 0 ? nf : 1));
  sdo_copy_fh_triangles(tet, nt, mask, nf, tri);
  free(mask);
  if (nf > 0)
    sdo_volume_columns(pd, tet, nabla, tri, G, box_len, grid, cx, cy, cz, ss, seed, NULL,
                       n_threads, stats);
  free(tri);
  free(nabla);
}

void sdo_compute_3d_density(double *pd, int n, const int *tet, int nt, int G, double box_len,
                            double *rho, float p_mass, double cx, double cy, double cz,
                            unsigned ss, uint64_t seed, int n_threads, sdo_stats *stats) {
  /* compute_3d_density, src/dtfe.c:657-665 */
  sdo_pre_compute_vol(pd, n, tet, nt, p_mass);
  sdo_interpolate_3d(pd, n, tet, nt, G, box_len, rho, cx, cy, cz, ss, seed, n_threads, stats);
}
