// walk_mixed_core.h -- the mixed-precision ray walk (gfx950).
//
// Same tets, same faces, same perturbations as the FP64 walk; faster
// arithmetic and a smaller record:
//
//  * Side tests in FP32 with a filter.  The Plücker product of the ray with
//    edge (X, Y) is lz * orient2d(p; X, Y) (src/dtfe.c:24-26 expanded), so it
//    is formed in FP32 as (X - p) x (Y - p) from coordinates stored relative
//    to the tessellation's origin.  An a-priori bound covers the rounding of
//    the stored coordinates, of the differences and of the products; the
//    FP32 sign is taken only when |c| exceeds that bound plus the reference's
//    zero band (ZERO = 1e4 DBL_EPSILON, src/constants.h:6, scaled by 1/lz)
//    plus the FP64 expression's own error.  Otherwise the product is redone
//    in FP64 with the reference expression and Sign().  Every side sign is
//    therefore the FP64 walk's sign, so the walk visits the same tets.
//  * A tet whose face choice is special (a zero side) or fails is handed to
//    the FP64 core, which perturbs the ray exactly as the FP64 walk does.
//  * The crossing height of each face is computed once (leave face, FP32) and
//    carried as the next tet's entry height; errors in a crossing height then
//    cancel to first order because the DTFE field is continuous across the
//    face.  Midpoints near the slab boundary are decided in FP64.
//  * The density at each crossing is the barycentric mix of the face's
//    vertex densities (same weights as the height) and is carried with it;
//    the field being linear along the chord, the trapezoid rule on the two
//    crossings is its exact integral.  This never forms rho_A + grad.(p - A),
//    which is ill-conditioned in FP32 for sliver tets.  Per-ray and per-pixel
//    sums stay FP64; tet volumes and vertex densities are the FP64 ones.
//
// Per step: one 64 B geometry record + one 16 B record of vertex densities.
#pragma once

#include <algorithm>

#include "sdtfe_host.h"
#include "walk_fp64_core.h"

namespace sdtfe {

// Largest tolerated relative uncertainty of FP32 barycentric weights.
constexpr float kCondMax = 1e-4f;

struct Tet32 {
  int nb0, nb1, nb2, nb3;
  float ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
};

SDTFE_HD inline void load_tet32(const Rec32 *__restrict__ rec, int t, Tet32 &r) {
  const float4 *p = reinterpret_cast<const float4 *>(rec + t);
  const int4 nb = *reinterpret_cast<const int4 *>(p);
  const float4 q0 = p[1], q1 = p[2], q2 = p[3];
  r.nb0 = nb.x; r.nb1 = nb.y; r.nb2 = nb.z; r.nb3 = nb.w;
  r.ax = q0.x; r.ay = q0.y; r.az = q0.z;
  r.bx = q0.w; r.by = q1.x; r.bz = q1.y;
  r.cx = q1.z; r.cy = q1.w; r.cz = q2.x;
  r.dx = q2.y; r.dy = q2.z; r.dz = q2.w;
}

SDTFE_HD inline int nb32(const Tet32 &r, int f) { return pick4i(f, r.nb0, r.nb1, r.nb2, r.nb3); }

// Per-map constants of the filter (host-computed, see filter_consts()).
struct Filter {
  float e;       // bound on |(X - p) stored - (X - p) exact| per component
  float k_prod;  // relative rounding of the products and difference
  float thr;     // (ZERO + FP64 expression error) / lz
};

// Sign of one side product, and its FP32 value for the barycentric weights.
// (x1, y1) = X - p, (x2, y2) = Y - p in FP32; l1 = |x|+|y| of each.
struct Side {
  float c;
  float b;  // bound on |c - exact|
  int s;
};

SDTFE_HD inline Side side32(float x1, float y1, float l1a, float x2, float y2, float l1b,
                              const Filter &f, bool &need64) {
  const float p1 = x1 * y2, p2 = y1 * x2;
  const float c = p1 - p2;
  const float bound = f.k_prod * (fabsf(p1) + fabsf(p2)) + f.e * (l1a + l1b) + f.thr;
  Side s;
  s.c = c;
  s.b = bound;
  s.s = c > bound ? 1 : (c < -bound ? -1 : 0);
  need64 = s.s == 0;
  return s;
}

// Crossing points of the four faces from FP32 side values: height and
// density there, both the barycentric mix of the face's three vertices with
// the weights the FP64 walk uses for the height, written relative to the
// face's first vertex.  rho holds the vertex densities (A, B, C, D).
// Straight-line for all faces; callers pick one (see pick4).
struct Cross {
  float z, rho;
};

struct Crosses {
  float z[4], rho[4];
};

SDTFE_HD inline void face_point32(float t0, float t1, float t2, float w1n, float w2n, float z0,
                                  float z1, float z2, float q0, float q1, float q2, float &z,
                                  float &q) {
  const float inv = 1.0f / (t0 + t1 + t2);
  const float w1 = w1n * inv, w2 = w2n * inv;
  z = fmaf(w1, z1 - z0, fmaf(w2, z2 - z0, z0));
  q = fmaf(w1, q1 - q0, fmaf(w2, q2 - q0, q0));
}

SDTFE_HD inline Crosses face_crosses32(float cCB, float cAC, float cBA, float cDC, float cBD,
                                       float cDA, const Tet32 &r, const float4 &rho) {
  Crosses c;
  // BDC: D + w1 (C - D) + w2 (B - D); ACD: C, D, A; ADB: B, A, D; ABC: A, B, C
  face_point32(-cCB, cDC, -cBD, -cBD, cDC, r.dz, r.cz, r.bz, rho.w, rho.z, rho.y, c.z[0], c.rho[0]);
  face_point32(cAC, -cDC, -cDA, cAC, -cDC, r.cz, r.dz, r.az, rho.z, rho.w, rho.x, c.z[1], c.rho[1]);
  face_point32(cDA, cBD, -cBA, cBD, -cBA, r.bz, r.az, r.dz, rho.y, rho.x, rho.w, c.z[2], c.rho[2]);
  face_point32(cBA, cCB, -cAC, -cAC, cBA, r.az, r.bz, r.cz, rho.x, rho.y, rho.z, c.z[3], c.rho[3]);
  return c;
}

SDTFE_HD inline Cross pick_cross(int f, const Crosses &c) {
  return Cross{pick4f(f, c.z[0], c.z[1], c.z[2], c.z[3]),
               pick4f(f, c.rho[0], c.rho[1], c.rho[2], c.rho[3])};
}

// The same crossing from the FP64 record (side products with the reference
// expression, weights and height in FP64; the density mixes the FP32 vertex
// densities).  Height returned relative to org_z.
SDTFE_HD inline Cross face_cross64(int f, const Tet64 &q, double lz, double ox, double oy,
                                   const float4 &rho, double org_z) {
  const double p2 = lz, p3 = -lz * oy, p4 = lz * ox;
  const double uCB = edge_u(p2, p3, p4, q.cx, q.cy, q.bx, q.by);
  const double uAC = edge_u(p2, p3, p4, q.cx, q.cy, q.ax, q.ay);
  const double uBA = edge_u(p2, p3, p4, q.bx, q.by, q.ax, q.ay);
  const double uDC = edge_u(p2, p3, p4, q.cx, q.cy, q.dx, q.dy);
  const double uBD = edge_u(p2, p3, p4, q.bx, q.by, q.dx, q.dy);
  const double uDA = edge_u(p2, p3, p4, q.dx, q.dy, q.ax, q.ay);
  double zz[4], rr[4];
  const double ra = rho.x, rb = rho.y, rc = rho.z, rd = rho.w;
  const double t[4][5] = {{-uCB, uDC, -uBD, -uBD, uDC},
                          {uAC, -uDC, -uDA, uAC, -uDC},
                          {uDA, uBD, -uBA, uBD, -uBA},
                          {uBA, uCB, -uAC, -uAC, uBA}};
  const double zv[4][3] = {{q.dz, q.cz, q.bz}, {q.cz, q.dz, q.az}, {q.bz, q.az, q.dz},
                           {q.az, q.bz, q.cz}};
  const double rv[4][3] = {{rd, rc, rb}, {rc, rd, ra}, {rb, ra, rd}, {ra, rb, rc}};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double inv = 1.0 / (t[k][0] + t[k][1] + t[k][2]);
    const double w1 = t[k][3] * inv, w2 = t[k][4] * inv;
    zz[k] = zv[k][0] + w1 * (zv[k][1] - zv[k][0]) + w2 * (zv[k][2] - zv[k][0]);
    rr[k] = rv[k][0] + w1 * (rv[k][1] - rv[k][0]) + w2 * (rv[k][2] - rv[k][0]);
  }
  Cross c;
  c.z = (float)(pick4(f, zz[0], zz[1], zz[2], zz[3]) - org_z);
  c.rho = (float)pick4(f, rr[0], rr[1], rr[2], rr[3]);
  return c;
}

// Relative uncertainty of face f's FP32 barycentric weights.
SDTFE_HD inline float face_cond(int f, const Side &CB, const Side &AC, const Side &BA,
                                const Side &DC, const Side &BD, const Side &DA) {
  const float c0 = (CB.b + DC.b + BD.b) / fabsf(-CB.c + DC.c - BD.c);
  const float c1 = (AC.b + DC.b + DA.b) / fabsf(AC.c - DC.c - DA.c);
  const float c2 = (DA.b + BD.b + BA.b) / fabsf(DA.c + BD.c - BA.c);
  const float c3 = (BA.b + CB.b + AC.b) / fabsf(BA.c + CB.c - AC.c);
  return pick4f(f, c0, c1, c2, c3);
}

// Face choice from the six side signs (src/dtfe.c:64-213, as plucker64).
SDTFE_HD inline void choose_faces(int sCB, int sAC, int sBA, int sDC, int sBD, int sDA,
                                    int &enter, int &leave, bool &ok, bool &special) {
  const int fa[4] = {-sCB, sAC, sDA, sBA};
  const int fb[4] = {sDC, -sDC, sBD, sCB};
  const int fc[4] = {-sBD, -sDA, -sBA, -sAC};
  enter = -1;
  leave = -1;
  special = false;
  int found = 0;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int a = fa[f], b = fb[f], c = fc[f];
    const bool active = found < 2 && (a | b | c) != 0;
    const bool en = active && enter == -1 && a >= 0 && b >= 0 && c >= 0;
    const bool le = active && !en && leave == -1 && a <= 0 && b <= 0 && c <= 0;
    if (en) enter = f;
    if (le) leave = f;
    if ((en || le) && (a == 0 || b == 0 || c == 0)) special = true;
    found += (en || le) ? 1 : 0;
  }
  ok = found == 2;
}

// One sample-ray.  (x, y) FP64 sample position; the walk keeps the FP64
// position (changed only by perturbations) and its FP32 offset from the
// origin.
SDTFE_HD inline double walk_mixed(const Rec32 *__restrict__ rec32, const Field32 *__restrict__ fld,
                             const Rec64 *__restrict__ rec64, int entry, double x, double y,
                             const MapGeom &g, const Filter &flt, float zlo, float zhi,
                             float zband, uint64_t key, int &crossed, LaneCounters &cnt) {
  crossed = 0;
  if (entry < 0) return 0.0;
  const double step = (g.dxy / 2.0) / kPerturbMax;
  double ox = x, oy = y, acc = 0.0;
  float rx = (float)(x - g.org_x), ry = (float)(y - g.org_y);
  uint32_t draw = 2;
  int cur = entry, prev = -1, f_in = -1, f_out = 0, nxt = entry, steps = 0;
  Cross carry{0.f, 0.f};
  bool have_carry = false;
  Tet32 r;
  for (;;) {
    if (steps > 0) nxt = nb32(r, f_out);
    if (nxt == -1) break;
    if (nxt == prev) {  // turn around (src/dtfe.c:303-309)
      const int tmp = f_in;
      f_in = f_out;
      f_out = tmp;
      nxt = nb32(r, f_out);
      if (nxt == -1) break;
      have_carry = false;
    }
    if (++steps > g.max_steps) {
      cnt.step_cap++;
      crossed = 0;
      return 0.0;
    }
    prev = cur;
    cur = nxt;
    load_tet32(rec32, cur, r);
    const float4 rho4 = *reinterpret_cast<const float4 *>(fld + cur);

    // vertex offsets from the ray, FP32
    const float ax = r.ax - rx, ay = r.ay - ry, bx = r.bx - rx, by = r.by - ry;
    const float cx = r.cx - rx, cy = r.cy - ry, dx = r.dx - rx, dy = r.dy - ry;
    const float la = fabsf(ax) + fabsf(ay), lb = fabsf(bx) + fabsf(by);
    const float lc = fabsf(cx) + fabsf(cy), ld = fabsf(dx) + fabsf(dy);
    bool n0, n1, n2, n3, n4, n5;
    Side CB = side32(cx, cy, lc, bx, by, lb, flt, n0);
    Side AC = side32(cx, cy, lc, ax, ay, la, flt, n1);
    Side BA = side32(bx, by, lb, ax, ay, la, flt, n2);
    Side DC = side32(cx, cy, lc, dx, dy, ld, flt, n3);
    Side BD = side32(bx, by, lb, dx, dy, ld, flt, n4);
    Side DA = side32(dx, dy, ld, ax, ay, la, flt, n5);
    if (n0 | n1 | n2 | n3 | n4 | n5) {
      // undecided sides: the reference expression in FP64
      Tet64 q;
      load_tet(rec64, cur, q);
      const double p2 = g.lz, p3 = -g.lz * oy, p4 = g.lz * ox;
      const float il = (float)(1.0 / g.lz);
      unsigned redo = 0;
#define SDTFE_REDO(S, N, XX, XY, YX, YY)                                 \
  if (N) {                                                               \
    const double u = edge_u(p2, p3, p4, q.XX, q.XY, q.YX, q.YY);         \
    S.s = sign_zero(u);                                                  \
    S.c = (float)u * il;                                                 \
    ++redo;                                                              \
  }
      SDTFE_REDO(CB, n0, cx, cy, bx, by)
      SDTFE_REDO(AC, n1, cx, cy, ax, ay)
      SDTFE_REDO(BA, n2, bx, by, ax, ay)
      SDTFE_REDO(DC, n3, cx, cy, dx, dy)
      SDTFE_REDO(BD, n4, bx, by, dx, dy)
      SDTFE_REDO(DA, n5, dx, dy, ax, ay)
#undef SDTFE_REDO
      cnt.fp64_redo += redo;
    }
    int enter, leave;
    bool ok, special;
    choose_faces(CB.s, AC.s, BA.s, DC.s, BD.s, DA.s, enter, leave, ok, special);
    if (!ok || special) {
      // the FP64 walk's own step, perturbation included
      Tet64 q;
      load_tet(rec64, cur, q);
      Hit h = plucker64(q, g.lz, ox, oy);
      if (!h.ok) {
        cnt.no_isect++;
        crossed = 0;
        return 0.0;
      }
      if (h.special && !resolve_special(q, g.lz, step, key, draw, ox, oy, h, cnt)) {
        crossed = 0;
        return 0.0;
      }
      rx = (float)(ox - g.org_x);
      ry = (float)(oy - g.org_y);
      f_in = h.enter;
      f_out = h.leave;
      ++crossed;
      acc += tet_integral64(q, ox, oy, h.z_in, h.z_out, g.off_z, g.lz);
      have_carry = false;  // the next tet starts from its own entry face
      continue;
    }
    f_in = enter;
    f_out = leave;
    ++crossed;
    // A crossing is taken in FP64 when its error would not cancel against a
    // neighbour's (the ray's first and last crossing) or when the face's FP32
    // weights are ill-conditioned (nearly vertical faces).
    const bool in64 = !have_carry;
    const bool out64 = nb32(r, leave) == -1 ||
                       face_cond(leave, CB, AC, BA, DC, BD, DA) > kCondMax;
    const Crosses cr = face_crosses32(CB.c, AC.c, BA.c, DC.c, BD.c, DA.c, r, rho4);
    Cross in = carry, out = pick_cross(leave, cr);
    if (in64 | out64) {
      Tet64 q;
      load_tet(rec64, cur, q);
      if (in64) in = face_cross64(enter, q, g.lz, ox, oy, rho4, g.org_z);
      if (out64) out = face_cross64(leave, q, g.lz, ox, oy, rho4, g.org_z);
      cnt.fp64_redo++;
    }
    carry = out;
    have_carry = true;
    const float dz = out.z - in.z;
    const float zm = fmaf(dz, 0.5f, in.z);
    bool inside = zm > zlo && zm < zhi;
    if (fabsf(zm - zlo) < zband || fabsf(zm - zhi) < zband) {
      // near the slab boundary: the FP64 walk's own midpoint decides
      Tet64 q;
      load_tet(rec64, cur, q);
      const Hit h = plucker64(q, g.lz, ox, oy);
      const double dz64 = h.z_out - h.z_in;
      const double zm64 = dz64 / 2 + h.z_in;
      inside = zm64 > g.off_z && zm64 < g.off_z + g.lz;
      cnt.fp64_redo++;
    }
    // the field is linear along the chord: trapezoid = exact integral
    if (inside) acc += (double)(0.5f * (in.rho + out.rho) * dz);
  }
  return acc;
}

struct MixedConsts {
  Filter flt;
  float zlo, zhi, zband;
};

// Filter constants for a launch: stored vertices and rays lie within `span`
// of the origin (per coordinate).
inline MixedConsts mixed_consts(const sdtfe_tess *t, const MapGeom &g) {
  const double u32 = 1.0 / (1 << 24);
  const double span = std::max(t->span, g.ray_span);
  // largest absolute coordinate the FP64 expression sees
  const double amax = span + std::max(std::fabs(g.org_x), std::fabs(g.org_y));
  MixedConsts m;
  // stored vertex and ray each carry u32*span; their FP32 difference adds
  // u32 * 2 span: E = 4 u32 span, padded for the rounding of the bound itself
  const double e = 4.25 * u32 * span;
  m.flt.e = (float)e;
  m.flt.k_prod = (float)(3.0 * u32);
  // reference zero band, the FP64 expression's own rounding (three products
  // and two sums of terms below 2 lz amax^2) and the 2 E^2 cross term
  const double err64 = 16.0 * std::ldexp(1.0, -53) * 2.0 * g.lz * amax * amax;
  m.flt.thr = (float)(1.01 * ((1e4 * 2.220446049250313080847e-16 + err64) / g.lz + 2 * e * e));
  m.zlo = (float)(g.off_z - g.org_z);
  m.zhi = (float)(g.off_z + g.lz - g.org_z);
  m.zband = (float)(1e-5 * std::max(g.lz, span));
  return m;
}

}  // namespace sdtfe
