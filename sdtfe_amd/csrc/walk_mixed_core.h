// walk_mixed_core.h -- the mixed-precision ray walk (gfx950).
//
// Same tets, same faces, same perturbations as the FP64 walk; far fewer
// instructions per tet and a smaller record (sdtfe::Rec32, 96 B: neighbours,
// vertex A in FP64, B - A, C - A, D - A and the four vertex densities in
// FP32).  The walk is issue-bound, and on CDNA4 an unpacked FP32 operation
// costs what an FP64 one does, so the gain comes from doing less per tet:
//
//  * Local coordinates.  The ray's offset from A is formed in FP64 and
//    rounded to FP32; the other vertices' offsets add the stored edges.  The
//    FP32 rounding error is therefore a fraction of the local tet size.
//  * Side tests in FP32 with one filter per tet.  The Plücker product of the
//    ray with edge (X, Y) is lz * orient2d(p; X, Y) (src/dtfe.c:24-26
//    expanded), formed as (X - p) x (Y - p).  With L the largest L1 norm of
//    the four offsets, every side's error is below k_err L^2; a side whose
//    magnitude exceeds that plus the reference's zero band (ZERO =
//    1e4 DBL_EPSILON, src/constants.h:6, over lz) plus the FP64 expression's
//    own error has the FP64 walk's sign and is not zero there.
//  * Face choice by table.  With six nonzero signs the enter and leave faces
//    follow from the 6-bit sign mask (face_code); the table sits in LDS.
//  * Only the leave face's crossing is formed (height and density relative
//    to A, FP32 barycentric weights); it is carried as the next tet's entry,
//    so errors cancel to first order (the DTFE field is continuous across the
//    face).  The field is linear along the chord, so the trapezoid rule on
//    the two crossings is its exact integral.
//  * One FP64 block handles everything rare: a tet with an undecided side,
//    a failed face choice, an ill-conditioned leave face (nearly vertical),
//    a midpoint at the slab boundary, the first tet of a ray and a tet after
//    a turn-around.  It runs the FP64 walk's own step (perturbation
//    included) and its crossings.  The ray's last crossing (it has no
//    neighbour to cancel against) is redone in FP64 after the loop, where
//    the lanes of a wave are together again.
//  Per-ray and per-pixel sums stay FP64; volumes and vertex densities are the
//  FP64 ones.
#pragma once

#include <algorithm>

#include "sdtfe_host.h"
#include "walk_fp64_core.h"

namespace sdtfe {

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

// Face choice when no side is zero.  Bit e of w is set when side e is
// negative, sides in the order CB, AC, BA, DC, BD, DA.  Face f (BDC, ACD,
// ADB, ABC) uses the sides in kFaceSides[f]; it is an enter face when
// exactly the sides in kFaceEnterNeg[f] are negative among them and a leave
// face when exactly the others are.  The first enter and the first leave
// face in face order are taken, as src/dtfe.c:64-213 takes them (with no
// zero side a face is of one kind or neither, and the search only stops
// once both are found).  Code: enter in bits 0-1, leave in bits 2-3, bit 4
// set when both exist.
constexpr int kFaceSides[4] = {0x19, 0x2A, 0x34, 0x07};
constexpr int kFaceEnterNeg[4] = {0x11, 0x28, 0x04, 0x02};
constexpr int kFaceCodes = 64;

SDTFE_HD constexpr int face_code(int w) {
  int en = -1, le = -1;
  for (int f = 0; f < 4; ++f) {
    const int s = w & kFaceSides[f];
    if (en < 0 && s == kFaceEnterNeg[f])
      en = f;
    else if (le < 0 && s == (kFaceSides[f] ^ kFaceEnterNeg[f]))
      le = f;
  }
  return en >= 0 && le >= 0 ? (en | le << 2 | 16) : 0;
}

// Per-launch constants of the filter (host-computed, mixed_consts()).
struct Filter {
  float k_err;    // side error bound over L^2
  float thr;      // (ZERO + FP64 expression error) / lz
  float k_ill;    // a face whose weight sum is below k_ill L^2 is ill-conditioned
  float k_band;   // crossing height uncertainty over the tet's height range
  int full_slab;  // the slab holds the hull's z range: every midpoint is inside
};

// Reciprocal to within 1 ulp (v_rcp_f32) on the device.
SDTFE_HD inline float rcp32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// Crossing of face f from the FP32 side values: height (relative to A) and
// density, each the barycentric mix of the face's three vertices with the
// weights the FP64 walk uses for the height (face_z, src/dtfe.c:67-74 and
// its three siblings).  Per face the weight sum and the weighted differences
// from the face's first vertex are formed, then picked.
struct Sides32 {
  float cCB, cAC, cBA, cDC, cBD, cDA;
};

struct Cross32 {
  float z, rho, den;
};

SDTFE_HD inline Cross32 face_cross32(int f, const Sides32 &c, const Tet32 &r) {
  const float zb = r.ebz, zc = r.ecz, zd = r.edz;
  const float den0 = (-c.cCB + c.cDC) - c.cBD, den1 = (c.cAC - c.cDC) - c.cDA;
  const float den2 = (c.cDA + c.cBD) - c.cBA, den3 = (c.cBA + c.cCB) - c.cAC;
  const float nz0 = fmaf(-c.cBD, zc - zd, c.cDC * (zb - zd));
  const float nz1 = fmaf(c.cAC, zd - zc, c.cDC * zc);
  const float nz2 = fmaf(c.cBD, -zb, -c.cBA * (zd - zb));
  const float nz3 = fmaf(-c.cAC, zb, c.cBA * zc);
  const float nr0 = fmaf(-c.cBD, r.rc - r.rd, c.cDC * (r.rb - r.rd));
  const float nr1 = fmaf(c.cAC, r.rd - r.rc, -c.cDC * (r.ra - r.rc));
  const float nr2 = fmaf(c.cBD, r.ra - r.rb, -c.cBA * (r.rd - r.rb));
  const float nr3 = fmaf(-c.cAC, r.rb - r.ra, c.cBA * (r.rc - r.ra));
  Cross32 x;
  x.den = pick4f(f, den0, den1, den2, den3);
  const float inv = rcp32(x.den);
  x.z = fmaf(pick4f(f, nz0, nz1, nz2, nz3), inv, pick4f(f, zd, zc, zb, 0.f));
  x.rho = fmaf(pick4f(f, nr0, nr1, nr2, nr3), inv, pick4f(f, r.rd, r.rc, r.rb, r.ra));
  return x;
}

// The same crossing in FP64 from the FP64 record, for the ray's first and
// last crossings (their errors have no neighbour to cancel against; hull
// tets are large).  Face f's vertices (P0, P1, P2) are (f^3, f^2, f^1) in
// the order the reference mixes them; with s(X, Y) = (X - p) x (Y - p) the
// weights of P1 and P2 are s(P0, P2) and s(P1, P0) over the sum of the three
// edges' s (src/dtfe.c:67-74 and siblings, lz cancelled).  Loads only the
// three vertices.
struct Cross64 {
  double z;
  float rho;
  double cond;  // sum of the weights' magnitudes over their sum's
};

SDTFE_HD inline Cross64 face_cross64(const Rec64 *__restrict__ rec, int t, int f, double x,
                                     double y, const Tet32 &r) {
  const double *v = reinterpret_cast<const double *>(reinterpret_cast<const char *>(rec + t) + 16);
  const int i0 = f ^ 3, i1 = f ^ 2, i2 = f ^ 1;
  const double x0 = v[3 * i0] - x, y0 = v[3 * i0 + 1] - y, z0 = v[3 * i0 + 2];
  const double x1 = v[3 * i1] - x, y1 = v[3 * i1 + 1] - y, z1 = v[3 * i1 + 2];
  const double x2 = v[3 * i2] - x, y2 = v[3 * i2 + 1] - y, z2 = v[3 * i2 + 2];
  const double s0 = x2 * y1 - y2 * x1, s1 = x0 * y2 - y0 * x2, s2 = x1 * y0 - y1 * x0;
  const double sum = s0 + s1 + s2;
  const double inv = 1.0 / sum;
  const double w1 = s1 * inv, w2 = s2 * inv;
  const double r0 = pick4f(i0, r.ra, r.rb, r.rc, r.rd), r1 = pick4f(i1, r.ra, r.rb, r.rc, r.rd);
  const double r2 = pick4f(i2, r.ra, r.rb, r.rc, r.rd);
  Cross64 c;
  c.z = z0 + w1 * (z1 - z0) + w2 * (z2 - z0);
  c.rho = (float)(r0 + w1 * (r1 - r0) + w2 * (r2 - r0));
  c.cond = (fabs(s0) + fabs(s1) + fabs(s2)) / fabs(sum);
  return c;
}

// The FP64 walk's signs for the sides whose FP32 value is within the
// filter's bound: the reference expression (src/dtfe.c:24-26) on the FP64
// record and its zero band.  Updates the sign mask w; false when a side is
// zero there (the FP64 walk then perturbs the ray).
SDTFE_HD inline bool sides64(const Rec64 *__restrict__ rec, int t, double x, double y, double lz,
                             const Sides32 &c, float bound, uint32_t &w) {
  const double *v = reinterpret_cast<const double *>(reinterpret_cast<const char *>(rec + t) + 16);
  const double p2 = lz, p3 = -lz * y, p4 = lz * x;
  const float cs[6] = {c.cCB, c.cAC, c.cBA, c.cDC, c.cBD, c.cDA};
  constexpr int kX[6] = {2, 2, 1, 2, 1, 3}, kY[6] = {1, 0, 0, 3, 3, 0};  // CB AC BA DC BD DA
  bool ok = true;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    if (!(fabsf(cs[e]) > bound)) {
      const double u = edge_u(p2, p3, p4, v[3 * kX[e]], v[3 * kX[e] + 1], v[3 * kY[e]],
                              v[3 * kY[e] + 1]);
      const int sg = sign_zero(u);
      ok = ok && sg != 0;
      w = (w & ~(1u << e)) | ((uint32_t)(sg < 0) << e);
    }
  }
  return ok;
}

// Whether the chord from height z_in to z_out through tet t (faces enter,
// leave) counts: its midpoint strictly inside the slab, as the FP64 walk
// decides it (src/dtfe.c:356-369).  The FP32 crossing heights are within
// k_band of the tet's height range of the FP64 walk's; a midpoint that
// close to a slab bound is decided again from both crossings in FP64, with a
// band for their own rounding (a few ulps times the weights' conditioning).
// False when even that leaves it open (the FP64 walk then takes the ray).
SDTFE_HD inline bool slab_inside(const Rec64 *__restrict__ rec64, int t, int enter, int leave,
                                 double x, double y, const Tet32 &r, double z_in, double z_out,
                                 const MapGeom &g, const Filter &flt, bool &inside) {
  const double z_lo = g.off_z, z_hi = g.off_z + g.lz;
  const double zm = z_in + 0.5 * (z_out - z_in);
  inside = zm > z_lo && zm < z_hi;
  const float zext = fmaxf(fmaxf(fabsf(r.ebz), fabsf(r.ecz)), fabsf(r.edz));
  const double band = (double)(flt.k_band * zext) + 1e-12 * g.lz;
  if (!(fabs(zm - z_lo) < band || fabs(zm - z_hi) < band)) return true;
  // rare: the midpoint from FP64 crossings
  const Cross64 ci = face_cross64(rec64, t, enter, x, y, r);
  const Cross64 co = face_cross64(rec64, t, leave, x, y, r);
  const double zm64 = ci.z + 0.5 * (co.z - ci.z);
  const double band64 = 1e-12 * g.lz + 1e-14 * (double)zext * (ci.cond + co.cond);
  if (fabs(zm64 - z_lo) < band64 || fabs(zm64 - z_hi) < band64) return false;
  inside = zm64 > z_lo && zm64 < z_hi;
  return true;
}

// One sample-ray.  (x, y) FP64 sample position.  Returns false when the ray
// must be walked by the FP64 walk instead (walk64 with the same key): an
// undecided side (a possible zero, hence a possible perturbation), a failed
// face choice, a turn-around, an ill-conditioned crossing or a midpoint at
// the slab boundary.  Otherwise value and crossed are the ray's integral and
// tet count, and the tets are the FP64 walk's.  faces: the kFaceCodes
// face_code values (LDS on the device).  kFullSlab: the slab holds the
// hull's z range (Filter::full_slab), no midpoint test.
template <bool kFullSlab>
SDTFE_HD inline bool walk_mixed(const Rec32 *__restrict__ rec32, const Rec64 *__restrict__ rec64,
                                const uint8_t *faces, int entry, double x, double y,
                                const MapGeom &g, const Filter &flt, double &value, int &crossed,
                                LaneCounters &cnt) {
  value = 0.0;
  crossed = 0;
  if (entry < 0) return true;
  const double z_lo = g.off_z, z_hi = g.off_z + g.lz;
  double acc = 0.0;
  int cur = entry, prev = -1, f_in = 0, f_out = 0, steps = 0;
  double zc = 0.0;  // carried crossing: absolute height
  float rc = 0.f;   // and density
  Tet32 r;
  // why the loop ended: 0 the hull was left (last crossing pending), 1 the
  // FP64 walk takes the ray, 2 the step cap
  int end = 0;
  for (;;) {
    const int nxt = steps > 0 ? nb32(r, f_out) : entry;
    // a turn-around (src/dtfe.c:303-309) follows a perturbation only
    if (nxt == prev) {
      end = 1;
      break;
    }
    if (++steps > g.max_steps) {
      end = 2;
      break;
    }
    prev = cur;
    cur = nxt;
    load_tet32(rec32, cur, r);

    // offsets of the vertices from the ray, FP32, relative to A
    const float ax = (float)(r.ax - x), ay = (float)(r.ay - y);
    const float bx = ax + r.ebx, by = ay + r.eby, cx = ax + r.ecx, cy = ay + r.ecy;
    const float dx = ax + r.edx, dy = ay + r.edy;
    const float L = fmaxf(fmaxf(fabsf(ax) + fabsf(ay), fabsf(bx) + fabsf(by)),
                          fmaxf(fabsf(cx) + fabsf(cy), fabsf(dx) + fabsf(dy)));
    const float L2 = L * L;
    Sides32 c;
    c.cCB = fmaf(cx, by, -(cy * bx));
    c.cAC = fmaf(cx, ay, -(cy * ax));
    c.cBA = fmaf(bx, ay, -(by * ax));
    c.cDC = fmaf(cx, dy, -(cy * dx));
    c.cBD = fmaf(bx, dy, -(by * dx));
    c.cDA = fmaf(dx, ay, -(dy * ax));
    const float cmin = fminf(fminf(fminf(fabsf(c.cCB), fabsf(c.cAC)), fabsf(c.cBA)),
                             fminf(fminf(fabsf(c.cDC), fabsf(c.cBD)), fabsf(c.cDA)));
    const float bound = fmaf(flt.k_err, L2, flt.thr);
    uint32_t w = (__builtin_bit_cast(uint32_t, c.cCB) >> 31) |
                 ((__builtin_bit_cast(uint32_t, c.cAC) >> 31) << 1) |
                 ((__builtin_bit_cast(uint32_t, c.cBA) >> 31) << 2) |
                 ((__builtin_bit_cast(uint32_t, c.cDC) >> 31) << 3) |
                 ((__builtin_bit_cast(uint32_t, c.cBD) >> 31) << 4) |
                 ((__builtin_bit_cast(uint32_t, c.cDA) >> 31) << 5);
    if (!(cmin > bound)) {
      // rare: the sides the filter leaves open get the FP64 walk's sign; a
      // zero there means a perturbation, which the FP64 walk does
      if (!sides64(rec64, cur, x, y, g.lz, c, bound, w)) {
        end = 1;
        break;
      }
    }
    const int code = faces[w];
    if (!(code & 16)) {
      end = 1;
      break;
    }
    const int enter = code & 3, leave = (code >> 2) & 3;
    f_in = enter;
    if (steps == 1) {  // first tet: the entry crossing in FP64
      const Cross64 in = face_cross64(rec64, cur, enter, x, y, r);
      zc = in.z;
      rc = in.rho;
    }
    f_out = leave;
    ++crossed;
    if (nb32(r, leave) == -1) break;  // leaves the hull: last crossing below
    Cross32 out = face_cross32(leave, c, r);
    double z_out = r.az + (double)out.z;
    if (fabsf(out.den) < flt.k_ill * L2) {
      // rare: FP32 weights of a nearly vertical face are too uncertain
      const Cross64 o = face_cross64(rec64, cur, leave, x, y, r);
      z_out = o.z;
      out.rho = o.rho;
    }
    const double dz = z_out - zc;
    bool inside = true;
    if (!kFullSlab && !slab_inside(rec64, cur, enter, leave, x, y, r, zc, z_out, g, flt, inside)) {
      end = 1;
      break;
    }
    if (inside) acc += (double)(0.5f * (rc + out.rho)) * dz;
    zc = z_out;
    rc = out.rho;
  }
  if (end == 1) return false;
  if (end == 2) {
    cnt.step_cap++;
    crossed = 0;
    return true;
  }
  // the last crossing in FP64, where the lanes of a wave are together again
  const Cross64 out = face_cross64(rec64, cur, f_out, x, y, r);
  const double dz = out.z - zc;
  bool inside = true;
  if (!kFullSlab && !slab_inside(rec64, cur, f_in, f_out, x, y, r, zc, out.z, g, flt, inside))
    return false;
  if (inside) acc += (double)(0.5f * (rc + out.rho)) * dz;
  value = acc;
  return true;
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
  // offsets: A - p rounds twice (u32 + 2 u64), a stored edge once more and
  // the sum once: every component is within 4 u32 L.  A side's product and
  // fma add 2 u32 L^2, the offset errors 2 (4 u32 L) L; 11 pads the bound's
  // own rounding.
  m.flt.k_err = (float)(11.0 * u32);
  // reference zero band and the FP64 expression's own rounding (three
  // products and two sums of terms below 2 lz amax^2), over lz
  const double err64 = 16.0 * std::ldexp(1.0, -53) * 2.0 * g.lz * amax * amax;
  m.flt.thr = (float)(1.01 * (1e4 * 2.220446049250313080847e-16 + err64) / g.lz);
  // weight sums below 1e-3 L^2 (projected face area under 0.05 % of L^2)
  // leave the weights too uncertain; such rays take the FP64 walk
  m.flt.k_ill = 1e-3f;
  // a weight is within 4 k_err / k_ill (one side's error over the sum, and
  // the sum's error times the weight) and a face's height range is at most
  // twice the tet's: crossing heights are within 8 k_err / k_ill of the
  // tet's height range, plus the FP32 rounding of the mix
  m.flt.k_band = (float)(8.1 * m.flt.k_err / m.flt.k_ill + 1e-6);
  // every chord midpoint lies within the hull's z range: a midpoint on the
  // boundary belongs to a chord of zero length, which adds nothing
  m.flt.full_slab = t->n > 0 && g.off_z <= t->z_min && g.off_z + g.lz >= t->z_max;
  return m;
}

}  // namespace sdtfe
