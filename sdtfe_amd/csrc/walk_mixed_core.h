// walk_mixed_core.h -- the mixed-precision ray walk (gfx950).
//
// Same tets, same faces, same perturbations as the FP64 walk; less FP64
// arithmetic and a smaller record (sdtfe::Rec32, 96 B: neighbours, vertex A
// in FP64, B - A, C - A, D - A and the four vertex densities in FP32):
//
//  * Local coordinates.  The ray's offset from A is formed in FP64 and
//    rounded to FP32; the other vertices' offsets add the stored edges.  The
//    FP32 rounding error is therefore a fraction of the local tet size, not
//    of the box, so tiny tets in halo cores lose nothing.
//  * Side tests in FP32 with a filter.  The Plücker product of the ray with
//    edge (X, Y) is lz * orient2d(p; X, Y) (src/dtfe.c:24-26 expanded) and is
//    formed as (X - p) x (Y - p).  An a-priori bound covers the rounding of
//    the offsets and of the products; the FP32 sign is taken only when |c|
//    exceeds that bound plus the reference's zero band (ZERO = 1e4
//    DBL_EPSILON, src/constants.h:6, over lz) plus the FP64 expression's own
//    error.  Otherwise the product is redone in FP64 with the reference
//    expression and Sign().  Every side sign is therefore the FP64 walk's
//    sign, so the walk visits the same tets.
//  * A tet whose face choice is special (a zero side) or fails is handed to
//    the FP64 core, which perturbs the ray exactly as the FP64 walk does.
//  * Each face crossing is computed once (as the leave face, height and
//    density relative to A in FP32, height made absolute in FP64) and
//    carried as the next tet's entry; errors in a crossing then cancel to
//    first order because the DTFE field is continuous across the face.  The
//    field is linear along the chord, so the trapezoid rule on the two
//    crossings is its exact integral; this never forms
//    rho_A + grad.(p - A), which is ill-conditioned for sliver tets.
//  * The ray's first and last crossings (their errors have no neighbour to
//    cancel against), crossings with ill-conditioned FP32 weights (nearly
//    vertical faces) and midpoints at the slab boundary are done in FP64.  Per-ray and per-pixel
//    sums stay FP64; volumes and vertex densities are the FP64 ones.
#pragma once

#include <algorithm>

#include "sdtfe_host.h"
#include "walk_fp64_core.h"

namespace sdtfe {

// Largest tolerated relative uncertainty of FP32 barycentric weights.
constexpr float kCondMax = 1e-4f;

struct Tet32 {
  int nb0, nb1, nb2, nb3;
  double ax, ay, az;
  float ebx, eby, ebz, ecx, ecy, ecz, edx, edy, edz;  // B - A, C - A, D - A
  float ra, rb, rc, rd;                               // vertex densities
};

SDTFE_HD inline void load_tet32(const Rec32 *__restrict__ rec, int t, Tet32 &r) {
  const uint4 *p = reinterpret_cast<const uint4 *>(rec + t);
  const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3], q4 = p[4], q5 = p[5];
  r.nb0 = (int)q0.x; r.nb1 = (int)q0.y; r.nb2 = (int)q0.z; r.nb3 = (int)q0.w;
  r.ax = __builtin_bit_cast(double, ((uint64_t)q1.y << 32) | q1.x);
  r.ay = __builtin_bit_cast(double, ((uint64_t)q1.w << 32) | q1.z);
  r.az = __builtin_bit_cast(double, ((uint64_t)q2.y << 32) | q2.x);
  r.ebx = __builtin_bit_cast(float, q2.z); r.eby = __builtin_bit_cast(float, q2.w);
  r.ebz = __builtin_bit_cast(float, q3.x); r.ecx = __builtin_bit_cast(float, q3.y);
  r.ecy = __builtin_bit_cast(float, q3.z); r.ecz = __builtin_bit_cast(float, q3.w);
  r.edx = __builtin_bit_cast(float, q4.x); r.edy = __builtin_bit_cast(float, q4.y);
  r.edz = __builtin_bit_cast(float, q4.z); r.ra = __builtin_bit_cast(float, q4.w);
  r.rb = __builtin_bit_cast(float, q5.x); r.rc = __builtin_bit_cast(float, q5.y);
  r.rd = __builtin_bit_cast(float, q5.z);
}

SDTFE_HD inline int nb32(const Tet32 &r, int f) { return pick4i(f, r.nb0, r.nb1, r.nb2, r.nb3); }

// Per-launch constants of the filter (host-computed, mixed_consts()).
struct Filter {
  float k_prod;  // relative rounding of the two products and their difference
  float k_off;   // rounding of an offset, relative to the L1 norms it is built from
  float thr;     // (ZERO + FP64 expression error) / lz
};

// One side product: FP32 value, its error bound and its sign (0: undecided).
struct Side {
  float c;
  float b;
  int s;
};

// (x1, y1) = X - p, (x2, y2) = Y - p in FP32; l1 = |x| + |y| of each and
// e1, e2 their rounding bounds (per component).
SDTFE_HD inline Side side32(float x1, float y1, float l1, float e1, float x2, float y2, float l2,
                            float e2, const Filter &f) {
  const float p1 = x1 * y2, p2 = y1 * x2;
  const float c = p1 - p2;
  const float b = f.k_prod * (fabsf(p1) + fabsf(p2)) + e1 * l2 + e2 * l1 + e1 * e2 + f.thr;
  Side s;
  s.c = c;
  s.b = b;
  s.s = c > b ? 1 : (c < -b ? -1 : 0);
  return s;
}

struct Cross {
  double z;  // absolute height
  float rho;
};

// Crossing points of the four faces from FP32 side values: height and
// density relative to A, each the barycentric mix of the face's three
// vertices with the weights the FP64 walk uses for the height (face_z);
// straight-line for all faces, callers pick one (see pick4).
struct Crosses {
  float z[4], rho[4], inv[4];  // inv: 1 / (sum of the face's signed sides)
};

// Reciprocal to within 1 ulp (v_rcp_f32) on the device.
SDTFE_HD inline float rcp32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

SDTFE_HD inline void face_point32(float t0, float t1, float t2, float w1n, float w2n, float z0,
                                  float z1, float z2, float q0, float q1, float q2, float &z,
                                  float &q, float &inv) {
  inv = rcp32(t0 + t1 + t2);
  const float w1 = w1n * inv, w2 = w2n * inv;
  z = fmaf(w1, z1 - z0, fmaf(w2, z2 - z0, z0));
  q = fmaf(w1, q1 - q0, fmaf(w2, q2 - q0, q0));
}

SDTFE_HD inline Crosses face_crosses32(float cCB, float cAC, float cBA, float cDC, float cBD,
                                       float cDA, const Tet32 &r) {
  Crosses c;
  const float za = 0.f, zb = r.ebz, zc = r.ecz, zd = r.edz;
  // BDC: D + w1 (C - D) + w2 (B - D); ACD: C, D, A; ADB: B, A, D; ABC: A, B, C
  face_point32(-cCB, cDC, -cBD, -cBD, cDC, zd, zc, zb, r.rd, r.rc, r.rb, c.z[0], c.rho[0],
               c.inv[0]);
  face_point32(cAC, -cDC, -cDA, cAC, -cDC, zc, zd, za, r.rc, r.rd, r.ra, c.z[1], c.rho[1],
               c.inv[1]);
  face_point32(cDA, cBD, -cBA, cBD, -cBA, zb, za, zd, r.rb, r.ra, r.rd, c.z[2], c.rho[2],
               c.inv[2]);
  face_point32(cBA, cCB, -cAC, -cAC, cBA, za, zb, zc, r.ra, r.rb, r.rc, c.z[3], c.rho[3],
               c.inv[3]);
  return c;
}

// c ? a : b per lane, without control flow (see pick4).
SDTFE_HD inline Cross sel_cross(bool c, const Cross &a, const Cross &b) {
  const int f = c ? 0 : 1;
  return Cross{pick4(f, a.z, b.z, b.z, b.z), pick4f(f, a.rho, b.rho, b.rho, b.rho)};
}

SDTFE_HD inline Cross pick_cross(int f, const Crosses &c, double az) {
  return Cross{az + (double)pick4f(f, c.z[0], c.z[1], c.z[2], c.z[3]),
               pick4f(f, c.rho[0], c.rho[1], c.rho[2], c.rho[3])};
}

// The same crossing from the FP64 record: side products with the reference
// expression, weights and height in FP64; the density mixes the FP32 vertex
// densities.
SDTFE_HD inline Cross face_cross64(int f, const Tet64 &q, double lz, double ox, double oy,
                                   const Tet32 &r) {
  const double p2 = lz, p3 = -lz * oy, p4 = lz * ox;
  const double uCB = edge_u(p2, p3, p4, q.cx, q.cy, q.bx, q.by);
  const double uAC = edge_u(p2, p3, p4, q.cx, q.cy, q.ax, q.ay);
  const double uBA = edge_u(p2, p3, p4, q.bx, q.by, q.ax, q.ay);
  const double uDC = edge_u(p2, p3, p4, q.cx, q.cy, q.dx, q.dy);
  const double uBD = edge_u(p2, p3, p4, q.bx, q.by, q.dx, q.dy);
  const double uDA = edge_u(p2, p3, p4, q.dx, q.dy, q.ax, q.ay);
  const double ra = r.ra, rb = r.rb, rc = r.rc, rd = r.rd;
  const double t[4][5] = {{-uCB, uDC, -uBD, -uBD, uDC},
                          {uAC, -uDC, -uDA, uAC, -uDC},
                          {uDA, uBD, -uBA, uBD, -uBA},
                          {uBA, uCB, -uAC, -uAC, uBA}};
  const double zv[4][3] = {{q.dz, q.cz, q.bz}, {q.cz, q.dz, q.az}, {q.bz, q.az, q.dz},
                           {q.az, q.bz, q.cz}};
  const double rv[4][3] = {{rd, rc, rb}, {rc, rd, ra}, {rb, ra, rd}, {ra, rb, rc}};
  double zz[4], rr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double inv = 1.0 / (t[k][0] + t[k][1] + t[k][2]);
    const double w1 = t[k][3] * inv, w2 = t[k][4] * inv;
    zz[k] = zv[k][0] + w1 * (zv[k][1] - zv[k][0]) + w2 * (zv[k][2] - zv[k][0]);
    rr[k] = rv[k][0] + w1 * (rv[k][1] - rv[k][0]) + w2 * (rv[k][2] - rv[k][0]);
  }
  return Cross{pick4(f, zz[0], zz[1], zz[2], zz[3]), (float)pick4(f, rr[0], rr[1], rr[2], rr[3])};
}

// Relative uncertainty of face f's FP32 barycentric weights: the summed
// bounds of its sides over the magnitude of their signed sum.
SDTFE_HD inline float face_cond(int f, const Side &CB, const Side &AC, const Side &BA,
                                const Side &DC, const Side &BD, const Side &DA,
                                const Crosses &cr) {
  const float b0 = CB.b + DC.b + BD.b, b1 = AC.b + DC.b + DA.b;
  const float b2 = DA.b + BD.b + BA.b, b3 = BA.b + CB.b + AC.b;
  return pick4f(f, b0 * fabsf(cr.inv[0]), b1 * fabsf(cr.inv[1]), b2 * fabsf(cr.inv[2]),
                b3 * fabsf(cr.inv[3]));
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
// position, which only a perturbation changes.
SDTFE_HD inline double walk_mixed(const Rec32 *__restrict__ rec32,
                                  const Rec64 *__restrict__ rec64, int entry, double x,
                                  double y, const MapGeom &g, const Filter &flt, uint64_t key,
                                  int &crossed, LaneCounters &cnt) {
  crossed = 0;
  if (entry < 0) return 0.0;
  const double step = (g.dxy / 2.0) / kPerturbMax;
  const double z_lo = g.off_z, z_hi = g.off_z + g.lz;
  double ox = x, oy = y, acc = 0.0;
  uint32_t draw = 2;
  int cur = entry, prev = -1, f_in = -1, f_out = 0, nxt = entry, steps = 0;
  Cross carry{0.0, 0.f};
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

    // offsets of the vertices from the ray, FP32, relative to A
    const float ax = (float)(r.ax - ox), ay = (float)(r.ay - oy);
    const float bx = ax + r.ebx, by = ay + r.eby, cx = ax + r.ecx, cy = ay + r.ecy;
    const float dx = ax + r.edx, dy = ay + r.edy;
    const float la = fabsf(ax) + fabsf(ay), lb = fabsf(bx) + fabsf(by);
    const float lc = fabsf(cx) + fabsf(cy), ld = fabsf(dx) + fabsf(dy);
    // rounding bounds per component: A - p once, the others also carry the
    // stored edge and the sum
    const float ea = flt.k_off * la;
    const float eb = flt.k_off * (la + fabsf(r.ebx) + fabsf(r.eby) + lb);
    const float ec = flt.k_off * (la + fabsf(r.ecx) + fabsf(r.ecy) + lc);
    const float ed = flt.k_off * (la + fabsf(r.edx) + fabsf(r.edy) + ld);
    Side CB = side32(cx, cy, lc, ec, bx, by, lb, eb, flt);
    Side AC = side32(cx, cy, lc, ec, ax, ay, la, ea, flt);
    Side BA = side32(bx, by, lb, eb, ax, ay, la, ea, flt);
    Side DC = side32(cx, cy, lc, ec, dx, dy, ld, ed, flt);
    Side BD = side32(bx, by, lb, eb, dx, dy, ld, ed, flt);
    Side DA = side32(dx, dy, ld, ed, ax, ay, la, ea, flt);
    const bool undecided =
        (CB.s == 0) | (AC.s == 0) | (BA.s == 0) | (DC.s == 0) | (BD.s == 0) | (DA.s == 0);
    if (undecided) {
      // undecided sides: the reference expression in FP64
      Tet64 q;
      load_tet(rec64, cur, q);
      const double p2 = g.lz, p3 = -g.lz * oy, p4 = g.lz * ox;
      const float il = (float)(1.0 / g.lz);
      unsigned redo = 0;
#define SDTFE_REDO(S, XX, XY, YX, YY)                                    \
  if (S.s == 0) {                                                        \
    const double u = edge_u(p2, p3, p4, q.XX, q.XY, q.YX, q.YY);         \
    S.s = sign_zero(u);                                                  \
    S.c = (float)u * il;                                                 \
    S.b = 0.f;                                                           \
    ++redo;                                                              \
  }
      SDTFE_REDO(CB, cx, cy, bx, by)
      SDTFE_REDO(AC, cx, cy, ax, ay)
      SDTFE_REDO(BA, bx, by, ax, ay)
      SDTFE_REDO(DC, cx, cy, dx, dy)
      SDTFE_REDO(BD, bx, by, dx, dy)
      SDTFE_REDO(DA, dx, dy, ax, ay)
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
    const Crosses cr = face_crosses32(CB.c, AC.c, BA.c, DC.c, BD.c, DA.c, r);
    const Cross fresh = pick_cross(enter, cr, r.az);
    Cross in = sel_cross(have_carry, carry, fresh);
    Cross out = pick_cross(leave, cr, r.az);
    // FP64 for the crossings whose error no neighbour cancels (where the ray
    // enters and leaves the hull) and for ill-conditioned weights
    const bool in64 = !have_carry;
    const bool out64 =
        nb32(r, leave) == -1 || face_cond(leave, CB, AC, BA, DC, BD, DA, cr) > kCondMax;
    if (in64 | out64) {
      Tet64 q;
      load_tet(rec64, cur, q);
      const Cross in_64 = face_cross64(enter, q, g.lz, ox, oy, r);
      const Cross out_64 = face_cross64(leave, q, g.lz, ox, oy, r);
      in = sel_cross(in64, in_64, in);
      out = sel_cross(out64, out_64, out);
      cnt.fp64_redo++;
    }
    carry = out;
    have_carry = true;
    const double dz = out.z - in.z;
    const double zm = in.z + 0.5 * dz;
    bool inside = zm > z_lo && zm < z_hi;
    const double band = 1e-4 * fabs(dz) + 1e-12 * g.lz;
    if (fabs(zm - z_lo) < band || fabs(zm - z_hi) < band) {
      // at the slab boundary: the FP64 walk's own midpoint decides
      Tet64 q;
      load_tet(rec64, cur, q);
      const Hit h = plucker64(q, g.lz, ox, oy);
      const double dz64 = h.z_out - h.z_in;
      const double zm64 = dz64 / 2 + h.z_in;
      inside = zm64 > z_lo && zm64 < z_hi;
      cnt.fp64_redo++;
    }
    // the field is linear along the chord: trapezoid = exact integral
    if (inside) acc += (double)(0.5f * (in.rho + out.rho)) * dz;
  }
  return acc;
}

struct MixedConsts {
  Filter flt;
};

// Filter constants for a launch.
inline MixedConsts mixed_consts(const sdtfe_tess *t, const MapGeom &g) {
  const double u32 = 1.0 / (1 << 24);
  // largest absolute coordinate the FP64 expression sees
  const double amax = std::max(t->span + std::max(std::fabs(t->org[0]), std::fabs(t->org[1])),
                               g.ray_span + std::max(std::fabs(g.org_x), std::fabs(g.org_y)));
  MixedConsts m;
  m.flt.k_prod = (float)(3.0 * u32);
  // each offset is a rounded FP64 difference, a rounded stored edge and a
  // rounded sum: within u32 of the L1 norms involved, padded for the
  // rounding of the bound itself
  m.flt.k_off = (float)(1.01 * u32);
  // reference zero band and the FP64 expression's own rounding (three
  // products and two sums of terms below 2 lz amax^2), over lz
  const double err64 = 16.0 * std::ldexp(1.0, -53) * 2.0 * g.lz * amax * amax;
  m.flt.thr = (float)(1.01 * (1e4 * 2.220446049250313080847e-16 + err64) / g.lz);
  return m;
}

}  // namespace sdtfe
