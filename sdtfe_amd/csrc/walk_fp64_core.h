// walk_fp64_core.h -- FP64 pieces of the ray walk shared by the FP64 walk
// and the mixed walk's fallback.  Follows src/dtfe.c:19-245 operation by
// operation; every translation unit that includes it is built with
// -ffp-contract=off.
#pragma once

#include <cmath>

#include "walk_common.h"

namespace sdtfe {

constexpr double kZero = 1e4 * 2.220446049250313080847e-16;  // src/constants.h:6
constexpr int kPerturbMax = 1000;                              // src/constants.h:8

SDTFE_HD inline int sign_zero(double x) { return x > kZero ? 1 : (x < -kZero ? -1 : 0); }

struct Tet64 {
  int nb0, nb1, nb2, nb3;
  double ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
  double rho_a, gx, gy, gz;
};

SDTFE_HD inline void load_tet(const Rec64 *__restrict__ rec, int t, Tet64 &r) {
  const int4 *p = reinterpret_cast<const int4 *>(rec + t);
  const int4 nb = p[0];
  const double2 *q = reinterpret_cast<const double2 *>(p + 1);
  const double2 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6],
                q7 = q[7];
  r.nb0 = nb.x; r.nb1 = nb.y; r.nb2 = nb.z; r.nb3 = nb.w;
  r.ax = q0.x; r.ay = q0.y; r.az = q1.x;
  r.bx = q1.y; r.by = q2.x; r.bz = q2.y;
  r.cx = q3.x; r.cy = q3.y; r.cz = q4.x;
  r.dx = q4.y; r.dy = q5.x; r.dz = q5.y;
  r.rho_a = q6.x; r.gx = q6.y; r.gy = q7.x; r.gz = q7.y;
}

SDTFE_HD inline int pick4i(int f, int a, int b, int c, int d);
SDTFE_HD inline int nb_of(const Tet64 &r, int f) { return pick4i(f, r.nb0, r.nb1, r.nb2, r.nb3); }

// Plücker product of the z-directed ray with the edge (X, Y), formed as
// the reference forms it: (X - Y, X.x*Y.y - X.y*Y.x) against the ray's
// (0, 0, lz, -lz*y, lz*x, 0) (src/dtfe.c:24-26, 44-147).
SDTFE_HD inline double edge_u(double p2, double p3, double p4, double xx, double xy,
                                double yx, double yy) {
  const double t0 = xx - yx, t1 = xy - yy;
  const double t5 = xx * yy - xy * yx;
  return p2 * t5 + t0 * p3 + t1 * p4;
}

struct Hit {
  int enter, leave;
  double z_in, z_out;
  bool ok, special;
};

// Per-lane choice among four values without control flow.  hipcc (ROCm
// 7.2, gfx950) miscompiles an if/else (or ?:) chain that picks
// floating-point operands by a lane-varying face index: lanes on the last arm
// lose that arm's assignments.  Every such choice here computes all
// candidates in straight-line code and picks with bit masks: two bit-select
// levels (v_bfi_b32) on the index's two bits.  f must be in 0..3.
SDTFE_HD inline uint32_t bsel32(uint32_t m, uint32_t x, uint32_t y) { return (x & m) | (y & ~m); }

SDTFE_HD inline uint32_t pick4u(int f, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  const uint32_t m0 = 0u - (uint32_t)(f & 1), m1 = 0u - (uint32_t)((f >> 1) & 1);
  return bsel32(m1, bsel32(m0, d, c), bsel32(m0, b, a));
}

SDTFE_HD inline double pick4(int f, double a, double b, double c, double d) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const uint64_t uc = __builtin_bit_cast(uint64_t, c), ud = __builtin_bit_cast(uint64_t, d);
  const uint32_t lo = pick4u(f, (uint32_t)ua, (uint32_t)ub, (uint32_t)uc, (uint32_t)ud);
  const uint32_t hi = pick4u(f, (uint32_t)(ua >> 32), (uint32_t)(ub >> 32), (uint32_t)(uc >> 32),
                             (uint32_t)(ud >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

SDTFE_HD inline float pick4f(int f, float a, float b, float c, float d) {
  return __builtin_bit_cast(float, pick4u(f, __builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                          __builtin_bit_cast(uint32_t, c), __builtin_bit_cast(uint32_t, d)));
}

SDTFE_HD inline int pick4i(int f, int a, int b, int c, int d) {
  return (int)pick4u(f, (uint32_t)a, (uint32_t)b, (uint32_t)c, (uint32_t)d);
}

// Crossing heights of the four faces: barycentric weights from each face's
// three signed edge products, point = (1 - w1 - w2) p0 + w1 p1 + w2 p2
// (src/dtfe.c:67-74, 112-119, 151-158, 184-191), straight-line.
struct FaceHeights {
  double z[4];
};

SDTFE_HD inline double face_height(double t0, double t1, double t2, double w1n, double w2n,
                                   double z0, double z1, double z2) {
  double den = t0 + t1;
  den = den + t2;
  const double inv = 1.0 / den;
  const double w1 = w1n * inv, w2 = w2n * inv;
  return (1 - w1 - w2) * z0 + w1 * z1 + w2 * z2;
}

SDTFE_HD inline FaceHeights face_heights(double uCB, double uAC, double uBA, double uDC,
                                         double uBD, double uDA, const Tet64 &r) {
  FaceHeights h;
  h.z[0] = face_height(-uCB, uDC, -uBD, -uBD, uDC, r.dz, r.cz, r.bz);   // BDC
  h.z[1] = face_height(uAC, -uDC, -uDA, uAC, -uDC, r.cz, r.dz, r.az);   // ACD
  h.z[2] = face_height(uDA, uBD, -uBA, uBD, -uBA, r.bz, r.az, r.dz);    // ADB
  h.z[3] = face_height(uBA, uCB, -uAC, -uAC, uBA, r.az, r.bz, r.cz);    // ABC
  return h;
}

// Face choice of rayTetraPlucker_full (src/dtfe.c:64-213) from the six side
// signs: faces are tested in the order BDC, ACD, ADB, ABC; a face with any
// nonzero side sign is an enter face if all its signed sides are >= 0 and
// no enter face was found yet, else a leave face if all are <= 0; the search
// stops at the second face found; a zero side on a found face marks a
// special case.
SDTFE_HD inline void choose_faces(int sCB, int sAC, int sBA, int sDC, int sBD, int sDA, Hit &h) {
  const int fa[4] = {-sCB, sAC, sDA, sBA};
  const int fb[4] = {sDC, -sDC, sBD, sCB};
  const int fc[4] = {-sBD, -sDA, -sBA, -sAC};
  h.enter = -1;
  h.leave = -1;
  h.special = false;
  int found = 0;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int a = fa[f], b = fb[f], c = fc[f];
    const bool active = found < 2 && (a | b | c) != 0;
    const bool en = active && h.enter == -1 && a >= 0 && b >= 0 && c >= 0;
    const bool le = active && !en && h.leave == -1 && a <= 0 && b <= 0 && c <= 0;
    if (en) h.enter = f;
    if (le) h.leave = f;
    if ((en || le) && (a == 0 || b == 0 || c == 0)) h.special = true;
    found += (en || le) ? 1 : 0;
  }
  h.ok = found == 2;
}

// rayTetraPlucker_full (src/dtfe.c:28-215): side signs, face choice and the
// two crossing heights.
SDTFE_HD inline Hit plucker64(const Tet64 &r, double lz, double rx, double ry) {
  const double p2 = lz, p3 = -lz * ry, p4 = lz * rx;
  const double uCB = edge_u(p2, p3, p4, r.cx, r.cy, r.bx, r.by);
  const double uAC = edge_u(p2, p3, p4, r.cx, r.cy, r.ax, r.ay);
  const double uBA = edge_u(p2, p3, p4, r.bx, r.by, r.ax, r.ay);
  const double uDC = edge_u(p2, p3, p4, r.cx, r.cy, r.dx, r.dy);
  const double uBD = edge_u(p2, p3, p4, r.bx, r.by, r.dx, r.dy);
  const double uDA = edge_u(p2, p3, p4, r.dx, r.dy, r.ax, r.ay);
  Hit h;
  choose_faces(sign_zero(uCB), sign_zero(uAC), sign_zero(uBA), sign_zero(uDC), sign_zero(uBD),
               sign_zero(uDA), h);
  const FaceHeights zf = face_heights(uCB, uAC, uBA, uDC, uBD, uDA, r);
  h.z_in = pick4(h.enter, zf.z[0], zf.z[1], zf.z[2], zf.z[3]);
  h.z_out = pick4(h.leave, zf.z[0], zf.z[1], zf.z[2], zf.z[3]);
  return h;
}

// The perturbation loop of shoot_3d_ray (src/dtfe.c:329-349) for the tet r
// and first hit h: while the hit is special, move the ray toward a random
// corner by at most `step` (perturb_ray, src/dtfe.c:219-245) and test again.
// The moved ray persists.  Returns false when the walk must end with 0.
SDTFE_HD inline bool resolve_special(const Tet64 &r, double lz, double step, uint64_t key,
                                       uint32_t &draw, double &ox, double &oy, Hit &h,
                                       LaneCounters &cnt) {
  int tries = 0;
  while (h.special) {
    const int k = (int)(draw31(key, draw++) & 3u);
    const double vx = pick4(k, r.ax, r.bx, r.cx, r.dx);
    const double vy = pick4(k, r.ay, r.by, r.cy, r.dy);
    double d0 = vx - ox, d1 = vy - oy;
    const double mag = sqrt(d0 * d0 + d1 * d1);
    if (mag > step) {
      d0 = d0 / mag;
      d1 = d1 / mag;
      d0 = d0 * step;
      d1 = d1 * step;
    }
    ox = ox + d0;
    oy = oy + d1;
    cnt.pert_steps++;
    if (++tries > kPerturbMax) {
      cnt.pert_exh++;
      return false;
    }
    h = plucker64(r, lz, ox, oy);
    if (!h.ok) {
      cnt.no_isect_pert++;
      return false;
    }
  }
  return true;
}

// Midpoint-rule contribution of one tet (src/dtfe.c:356-369).
SDTFE_HD inline double tet_integral64(const Tet64 &r, double ox, double oy, double z_in,
                                        double z_out, double off_z, double lz) {
  const double dz = z_out - z_in;
  const double zm = dz / 2 + z_in;
  if (zm > off_z && zm < off_z + lz) {
    const double fixed = r.rho_a + r.gx * (ox - r.ax) + r.gy * (oy - r.ay) - r.gz * r.az;
    return (fixed + r.gz * zm) * dz;
  }
  return 0.0;
}

// shoot_3d_ray (src/dtfe.c:247-374) with perturb_ray (:217-245), from tet
// start entered from tet from with steps0 tets crossed before (a walk from
// the entry tet: start = from = entry, steps0 = 0, which matches the
// reference's dummy first record).  key names the sample; perturbation
// corners are draws 2, 3, ... of it.  crossed: the tets crossed by this
// call (0 when the walk fails).  acc0: the sum of the tets crossed before,
// which the walk adds to in order (0 when it starts at the entry tet).
SDTFE_HD inline double walk64_from(const Rec64 *__restrict__ rec, int start, int from,
                                   int steps0, double x, double y, const MapGeom &g,
                                   uint64_t key, int &crossed, LaneCounters &cnt,
                                   double acc0 = 0.0) {
  crossed = 0;
  if (start < 0) return 0.0;
  const double max_dist = g.dxy / 2.0;
  const double step = max_dist / kPerturbMax;
  double ox = x, oy = y, acc = acc0;
  uint32_t draw = 2;
  int cur = from, prev = -1, f_in = -1, f_out = 0, nxt = start, steps = steps0;
  bool first = true;
  Tet64 r;
  for (;;) {
    if (!first) nxt = nb_of(r, f_out);
    first = false;
    if (nxt == -1) break;
    if (nxt == prev) {  // turn around (src/dtfe.c:303-309)
      const int tmp = f_in;
      f_in = f_out;
      f_out = tmp;
      nxt = nb_of(r, f_out);
      if (nxt == -1) break;
    }
    if (++steps > g.max_steps) {
      cnt.step_cap++;
      crossed = 0;
      return 0.0;
    }
    prev = cur;
    cur = nxt;
    load_tet(rec, cur, r);
    Hit h = plucker64(r, g.lz, ox, oy);
    if (!h.ok) {
      cnt.no_isect++;
      crossed = 0;
      return 0.0;
    }
    if (h.special && !resolve_special(r, g.lz, step, key, draw, ox, oy, h, cnt)) {
      crossed = 0;
      return 0.0;
    }
    f_in = h.enter;
    f_out = h.leave;
    ++crossed;
    acc += tet_integral64(r, ox, oy, h.z_in, h.z_out, g.off_z, g.lz);
  }
  return acc;
}

// The whole walk from the entry tet.
SDTFE_HD inline double walk64(const Rec64 *__restrict__ rec, int entry, double x, double y,
                              const MapGeom &g, uint64_t key, int &crossed, LaneCounters &cnt) {
  return walk64_from(rec, entry, entry, 0, x, y, g, key, crossed, cnt);
}

}  // namespace sdtfe
