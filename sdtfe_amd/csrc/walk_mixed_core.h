// walk_mixed_core.h -- the mixed-precision ray walk (gfx950).
//
// Same tets, same faces, same perturbations as the FP64 walk, with far fewer
// loads and instructions per tet.  On MI355X the walk is limited by
// instruction issue (an FP64 or conversion operation takes twice the issue
// time of an FP32 one) and by the vector memory pipeline (every lane of a
// wave gathers its own record), so the design cuts both:
//
//  * Carried face.  The ray enters each tet through a face whose three
//    vertices it already knows: their offsets from the ray (x, y in FP32,
//    rounded relative to the offset itself, so the error is a fraction of
//    the local tet size), heights above the carried crossing (FP32, rebased
//    every step) and densities, and the face's three edge sides.  A step
//    loads one 32-byte face record (sdtfe::FaceRec: the next tet and its
//    fourth vertex) and forms only the three sides of the new vertex.
//  * Side signs in FP32 with a filter.  The Plücker product of the ray with
//    edge (X, Y) is lz * orient2d(p; X, Y) (src/dtfe.c:24-26 expanded),
//    formed as (X - p) x (Y - p).  With M the largest offset coordinate of
//    the tet, every side's error is below k_err M^2; a side whose magnitude
//    exceeds that plus the reference's zero band (ZERO = 1e4 DBL_EPSILON,
//    src/constants.h:6, over lz) plus the FP64 expression's own error has the
//    FP64 walk's sign and is not zero there.  A side inside the band gets the
//    FP64 walk's sign from the FP64 record; a zero sign (a perturbation)
//    hands the ray to the FP64 walk.
//  * Exit face from two signs.  With the entry face oriented so that its
//    sides are positive, the ray leaves through the face that replaces the
//    vertex between a negative and a positive new side; with no zero side the
//    exit face is the one the reference's face search finds
//    (src/dtfe.c:64-213).  The first tet has no carried face: its six sides
//    go through the same filter and a 64-entry sign-mask table (face_code).
//  * Crossings.  The exit crossing (height and density, barycentric in the
//    face's three vertices with the weights the FP64 walk uses,
//    src/dtfe.c:67-74) is formed in FP32 relative to the entry crossing and
//    carried as the next tet's entry, so errors cancel to first order (the
//    DTFE field is continuous across the face); the field is linear along the
//    chord, so the trapezoid rule on the two crossings is its exact integral.
//    The ray's first and last crossings (no neighbour to cancel against; hull
//    tets are large), crossings of nearly vertical faces and midpoints at the
//    slab boundary are formed in FP64.
//  * Error bound.  A crossing's FP32 height and density are off by at most
//    sum_i kSideErr side_ext_i |h_i - h| / den (and the same with densities)
//    to first order, the weights' sensitivity to their sides' errors, where
//    side_ext_i = |x1 y2| + |y1 x2| of side i's own offsets.  Since each
//    crossing is carried into the next tet, its height error cancels except
//    for the density difference across the two chords; the walk sums these
//    terms into a bound on the ray's integral and hands the ray to the FP64
//    walk when the bound exceeds Filter::tol of the value (rays through few,
//    large, nearly vertical tets near the hull, whose value is small).
//  * One exit per step, at its top, so the loop's control flow stays simple
//    (a hand-over decided inside a step is carried to the next one's test).
//  Per-ray and per-pixel sums stay FP64; volumes and vertex densities are the
//  FP64 ones.
#pragma once

#include <algorithm>
#include <cstdlib>

#include "sdtfe_host.h"
#include "walk_fp64_core.h"

// Waves per SIMD the mixed walk kernels are built for: the register
// allocator then keeps the walk loop's values in registers and spills only
// around the once-per-ray FP64 blocks (more waves hide more load latency).
#ifndef SDTFE_MIXED_WAVES
#define SDTFE_MIXED_WAVES 5
#endif

namespace sdtfe {

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
  float k_err;    // side error bound over M^2 (M: largest offset coordinate)
  float thr;      // (ZERO + FP64 expression error) / lz
  float k_ill;    // a face whose weight sum is below k_ill M^2 is ill-conditioned
  float k_band;   // crossing height uncertainty over the tet's height range
  float tol;      // a ray whose error bound exceeds tol of its value goes to FP64
  int full_slab;  // the slab holds the hull's z range: every midpoint is inside
};

// Rounding allowance per FP32 crossing (heights rebased, the new vertex's
// height, the mix and its weight sum), in units of the values rounded.
constexpr float kCrossRound = 4.0f / (1 << 24);

// The filter's bound on a side's FP32 error, plus the zero band.
SDTFE_HD inline float side_bound(const Filter &flt, float M) {
  return fmaf(flt.k_err, M * M, flt.thr);
}

// Reciprocal to within 1 ulp (v_rcp_f32) on the device.
SDTFE_HD inline float rcp32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// (X - p) x (Y - p) from the two offsets.
SDTFE_HD inline float side32(float x1, float y1, float x2, float y2) {
  return fmaf(x1, y2, -(y1 * x2));
}

// |x1 y2| + |y1 x2|: a side's FP32 error is within kSideErr times this
// (offsets within 2 u of themselves, the product and the fma rounded once
// each: 4 u |x1 y2| + 5 u |y1 x2| + u |side|, first order).
SDTFE_HD inline float side_ext(float x1, float y1, float x2, float y2) {
  return fmaf(fabsf(x1), fabsf(y2), fabsf(y1 * x2));
}
constexpr float kSideErr = 6.0f / (1 << 24);

// Largest of two values; NaN-propagating, so the compiler adds no operand
// canonicalisation and forms three-input maxima (v_maximum3_f32).
SDTFE_HD inline float amax(float a, float b) { return __builtin_elementwise_maximum(a, b); }

SDTFE_HD inline float absmax3(float a, float b, float c) {
  return amax(amax(fabsf(a), fabsf(b)), fabsf(c));
}

// Face record (t, f), f8 = 8 f, as two 16-byte loads issued together.  Below
// kFaceRecNarrowTets tets its byte offset fits 32 bits, and the load adds it
// to the base address itself.
template <bool kWide>
SDTFE_HD inline FaceRec load_facerec(const FaceRec *__restrict__ frec, int t, int f8) {
  const char *base = reinterpret_cast<const char *>(frec);
  const uint4 *p =
      kWide ? reinterpret_cast<const uint4 *>(base + (((uint64_t)(uint32_t)t << 7) |
                                                      (uint32_t)(f8 << 2)))
            : reinterpret_cast<const uint4 *>(base + (((uint32_t)t << 7) | (uint32_t)(f8 << 2)));
  const uint4 a = p[0], b = p[1];
  FaceRec r;
  r.qxh = __builtin_bit_cast(float, a.x);
  r.qxl = __builtin_bit_cast(float, a.y);
  r.qyh = __builtin_bit_cast(float, a.z);
  r.qyl = __builtin_bit_cast(float, a.w);
  r.nb = (int)b.x;
  r.code = b.y;
  r.dqz = __builtin_bit_cast(float, b.z);
  r.rho = __builtin_bit_cast(float, b.w);
  return r;
}

// The byte of a face record's code at offset lab8 = 8 x label (one
// v_bfe_u32 on the device): 8 x that vertex's label in the next tet, or
// for the exit face's own label 8 x the face the ray enters the next tet by.
SDTFE_HD inline int code_field(unsigned code, int lab8) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (int)__builtin_amdgcn_ubfe(code, (unsigned)lab8, 5u);
#else
  return (int)((code >> lab8) & 31u);
#endif
}

// Vertex densities of tet t (FP32, from the mixed record).
struct Rho4 {
  float v[4];
};

SDTFE_HD inline Rho4 load_rho4(const Rec32 *__restrict__ rec32, int t) {
  const Rec32 *r = rec32 + t;
  return Rho4{{r->rho[0], r->rho[1], r->rho[2], r->rho[3]}};
}

// Face f's crossing in FP64 from the FP64 record: face f's vertices (P0, P1,
// P2) are (f^3, f^2, f^1) in the order the reference mixes them; with
// s(X, Y) = (X - p) x (Y - p) the weights of P1 and P2 are s(P0, P2) and
// s(P1, P0) over the sum of the three edges' s (src/dtfe.c:67-74 and
// siblings, lz cancelled).  Loads only the three vertices.
struct Cross64 {
  double z;
  float rho;
  double cond;  // sum of the weights' magnitudes over their sum's
  double zsum;  // the three vertices' heights, (z0 + z1) + z2
};

SDTFE_HD inline Cross64 face_cross64(const Rec64 *__restrict__ rec, int t, int f, double x,
                                     double y, const Rho4 &rho) {
  const double *v = reinterpret_cast<const double *>(reinterpret_cast<const char *>(rec + t) + 16);
  const int i0 = f ^ 3, i1 = f ^ 2, i2 = f ^ 1;
  const double x0 = v[3 * i0] - x, y0 = v[3 * i0 + 1] - y, z0 = v[3 * i0 + 2];
  const double x1 = v[3 * i1] - x, y1 = v[3 * i1 + 1] - y, z1 = v[3 * i1 + 2];
  const double x2 = v[3 * i2] - x, y2 = v[3 * i2 + 1] - y, z2 = v[3 * i2 + 2];
  const double s0 = x2 * y1 - y2 * x1, s1 = x0 * y2 - y0 * x2, s2 = x1 * y0 - y1 * x0;
  const double sum = s0 + s1 + s2;
  const double inv = 1.0 / sum;
  const double w1 = s1 * inv, w2 = s2 * inv;
  const double r0 = pick4f(i0, rho.v[0], rho.v[1], rho.v[2], rho.v[3]);
  const double r1 = pick4f(i1, rho.v[0], rho.v[1], rho.v[2], rho.v[3]);
  const double r2 = pick4f(i2, rho.v[0], rho.v[1], rho.v[2], rho.v[3]);
  Cross64 c;
  c.z = z0 + w1 * (z1 - z0) + w2 * (z2 - z0);
  c.rho = (float)(r0 + w1 * (r1 - r0) + w2 * (r2 - r0));
  c.cond = (fabs(s0) + fabs(s1) + fabs(s2)) / fabs(sum);
  c.zsum = (z0 + z1) + z2;
  return c;
}

// The FP64 walk's sign of edge (a, b) of tet t (labels), from the FP64
// record: the reference expression (src/dtfe.c:24-26) and its zero band.
SDTFE_HD inline int sign64(const Rec64 *__restrict__ rec, int t, int a, int b, double x,
                           double y, double lz) {
  const double *v = reinterpret_cast<const double *>(reinterpret_cast<const char *>(rec + t) + 16);
  return sign_zero(edge_u(lz, -lz * y, lz * x, v[3 * a], v[3 * a + 1], v[3 * b], v[3 * b + 1]));
}

// Whether the chord from height z_in to z_out through tet t (faces enter,
// leave) counts: its midpoint strictly inside the slab, as the FP64 walk
// decides it (src/dtfe.c:356-369).  The FP32 crossing heights are within
// k_band * zext (zext: the tet's height range) of the FP64 walk's; a
// midpoint that close to a slab bound is decided again from both crossings
// in FP64, with a band for their own rounding (a few ulps times the weights'
// conditioning).  False when even that leaves it open (the FP64 walk then
// takes the ray).
SDTFE_HD inline bool slab_inside(const Rec64 *__restrict__ rec64, const Rec32 *__restrict__ rec32,
                                 int t, int enter, int leave, double x, double y, float zext,
                                 double z_in, double z_out, const MapGeom &g, const Filter &flt,
                                 bool &inside) {
  const double z_lo = g.off_z, z_hi = g.off_z + g.lz;
  const double zm = z_in + 0.5 * (z_out - z_in);
  inside = zm > z_lo && zm < z_hi;
  const double band = (double)(flt.k_band * zext) + 1e-12 * g.lz;
  if (!(fabs(zm - z_lo) < band || fabs(zm - z_hi) < band)) return true;
  // rare: the midpoint from FP64 crossings
  const Rho4 rho = load_rho4(rec32, t);
  const Cross64 ci = face_cross64(rec64, t, enter, x, y, rho);
  const Cross64 co = face_cross64(rec64, t, leave, x, y, rho);
  const double zm64 = ci.z + 0.5 * (co.z - ci.z);
  const double band64 = 1e-12 * g.lz + 1e-14 * (double)zext * (ci.cond + co.cond);
  if (fabs(zm64 - z_lo) < band64 || fabs(zm64 - z_hi) < band64) return false;
  inside = zm64 > z_lo && zm64 < z_hi;
  return true;
}

// One vertex of the carried face: offset from the ray, height above the
// carried crossing (both FP32, so their rounding is a fraction of the local
// tet size), density and 8 x its label (0..3 for A..D) in the current tet.
struct Slot {
  float ox, oy, z;
  float rho;
  int lab8;
};

// The carried crossing's height from the FP64 heights of the carried face's
// vertices and their FP32 heights above it (the full-slab walk does not
// keep it: nothing else needs it).
SDTFE_HD inline double carried_z(const Cross64 &c, const Slot &p0, const Slot &p1,
                                 const Slot &p2) {
  return (c.zsum - (double)((p0.z + p1.z) + p2.z)) * (1.0 / 3.0);
}

// m ? a : b per lane, without control flow (see pick4).
SDTFE_HD inline float self(bool m, float a, float b) {
  return __builtin_bit_cast(float, bsel32(0u - (uint32_t)m, __builtin_bit_cast(uint32_t, a),
                                          __builtin_bit_cast(uint32_t, b)));
}

SDTFE_HD inline int seli(bool m, int a, int b) {
  return (int)bsel32(0u - (uint32_t)m, (uint32_t)a, (uint32_t)b);
}

SDTFE_HD inline Slot sel_slot(bool m, const Slot &a, const Slot &b) {
  return Slot{self(m, a.ox, b.ox), self(m, a.oy, b.oy), self(m, a.z, b.z), self(m, a.rho, b.rho),
              seli(m, a.lab8, b.lab8)};
}

// Where the FP64 walk takes over a ray the mixed walk hands over: tet t,
// entered through its face enter after steps tets, with acc (twice the
// integral so far, as the walk keeps it); enter < 0: the whole ray.  The
// mixed walk crossed the FP64 walk's tets up to t and no perturbation
// happened before it, so walk64_from(t) continues the FP64 walk exactly.
struct Resume {
  int t, enter, steps;
  double acc;
};

// One sample-ray.  (x, y) FP64 sample position.  Returns false when the ray
// must be walked by the FP64 walk instead, from where rs says (walk64_from
// with the same key): a zero side (a perturbation), a failed face choice, an
// inconsistent face record or a midpoint on the slab boundary.  Otherwise value and crossed are the
// ray's integral and tet count, and the tets are the FP64 walk's.  faces:
// the kFaceCodes face_code values (LDS on the device).  kFullSlab: the slab
// holds the hull's z range (Filter::full_slab), no midpoint test.  kWide:
// more than kFaceRecNarrowTets tets.
template <bool kFullSlab, bool kWide>
SDTFE_HD inline bool walk_mixed(const Rec32 *__restrict__ rec32, const Rec64 *__restrict__ rec64,
                                const FaceRec *__restrict__ frec, const uint8_t *faces, int entry,
                                double x, double y, const MapGeom &g, const Filter &flt,
                                double &value, int &crossed, LaneCounters &cnt, Resume &rs) {
  value = 0.0;
  crossed = 0;
  if (entry < 0) return true;
  if (g.max_steps < 1) {
    cnt.step_cap++;
    return true;
  }

  // ---- first tet: all four vertices from the FP64 record
  int t = entry;
  Slot p0, p1, p2;
  float e0, e1, e2;  // sides of the carried face: (P0, P1), (P1, P2), (P2, P0)
  int enter, f8;     // t's entry face, 8 x its exit face
  double zc;         // carried crossing: absolute height (kept by the slab walk)
  float rc;          // and density
  float zext;        // t's height range
  {
    Tet64 q;
    load_tet(rec64, t, q);
    const float ax = (float)(q.ax - x), ay = (float)(q.ay - y);
    const float bx = (float)(q.bx - x), by = (float)(q.by - y);
    const float cx = (float)(q.cx - x), cy = (float)(q.cy - y);
    const float dx = (float)(q.dx - x), dy = (float)(q.dy - y);
    const float M = amax(amax(absmax3(ax, ay, bx), absmax3(by, cx, cy)), absmax3(dx, dy, 0.f));
    const float bound = side_bound(flt, M);
    const float cs[6] = {side32(cx, cy, bx, by), side32(cx, cy, ax, ay), side32(bx, by, ax, ay),
                         side32(cx, cy, dx, dy), side32(bx, by, dx, dy), side32(dx, dy, ax, ay)};
    constexpr int kX[6] = {2, 2, 1, 2, 1, 3}, kY[6] = {1, 0, 0, 3, 3, 0};  // CB AC BA DC BD DA
    uint32_t w = 0;
    bool zero = false;
#pragma unroll
    for (int e = 0; e < 6; ++e) {
      int sg = cs[e] < 0.f ? -1 : 1;
      if (!(fabsf(cs[e]) > bound)) sg = sign64(rec64, t, kX[e], kY[e], x, y, g.lz);
      zero = zero || sg == 0;
      w |= (uint32_t)(sg < 0) << e;
    }
    const int code = faces[w];
    if (zero || !(code & 16)) {
      rs = Resume{entry, -1, 0, 0.0};
      return false;
    }
    enter = code & 3;
    const int f = (code >> 2) & 3;
    f8 = 8 * f;
    zext = (float)(fmax(fmax(q.az, q.bz), fmax(q.cz, q.dz)) -
                   fmin(fmin(q.az, q.bz), fmin(q.cz, q.dz)));
    const Rho4 rho = load_rho4(rec32, t);
    const Cross64 in = face_cross64(rec64, t, enter, x, y, rho);
    zc = in.z;
    rc = in.rho;
    // carried face: the exit face's vertices in the order f^1, f^3, f^2, in
    // which its sides are positive (the reference orients every tet so that
    // a leave face's sides in the order f^1, f^2, f^3 are negative,
    // src/triangulate.c:72-99 with the face signs of src/dtfe.c:64-213); the
    // magnitudes keep that sign also where a side's FP32 value was too small
    // to be decided
    const float oxs[4] = {ax, bx, cx, dx}, oys[4] = {ay, by, cy, dy};
    const float zs[4] = {(float)(q.az - zc), (float)(q.bz - zc), (float)(q.cz - zc),
                         (float)(q.dz - zc)};
    const int l0 = f ^ 1, l1 = f ^ 3, l2 = f ^ 2;
    p0 = Slot{pick4f(l0, oxs[0], oxs[1], oxs[2], oxs[3]), pick4f(l0, oys[0], oys[1], oys[2], oys[3]),
              pick4f(l0, zs[0], zs[1], zs[2], zs[3]),
              pick4f(l0, rho.v[0], rho.v[1], rho.v[2], rho.v[3]), 8 * l0};
    p1 = Slot{pick4f(l1, oxs[0], oxs[1], oxs[2], oxs[3]), pick4f(l1, oys[0], oys[1], oys[2], oys[3]),
              pick4f(l1, zs[0], zs[1], zs[2], zs[3]),
              pick4f(l1, rho.v[0], rho.v[1], rho.v[2], rho.v[3]), 8 * l1};
    p2 = Slot{pick4f(l2, oxs[0], oxs[1], oxs[2], oxs[3]), pick4f(l2, oys[0], oys[1], oys[2], oys[3]),
              pick4f(l2, zs[0], zs[1], zs[2], zs[3]),
              pick4f(l2, rho.v[0], rho.v[1], rho.v[2], rho.v[3]), 8 * l2};
    e0 = fabsf(side32(p0.ox, p0.oy, p1.ox, p1.oy));
    e1 = fabsf(side32(p1.ox, p1.oy, p2.ox, p2.oy));
    e2 = fabsf(side32(p2.ox, p2.oy, p0.ox, p0.oy));
  }

  // ---- the walk: leave t through face f (the carried face).  The next
  // face record is requested as soon as f is known, so its load overlaps
  // the face update and the exit crossing.  Each step has one exit, at its
  // top: the hull, a record without back link, the step bound, or a
  // hand-over decided in the step before (steps = -1).
  float xh, xl, yh, yl;  // the ray as FP32 pairs (split32)
  split32(x, xh, xl);
  split32(y, yh, yl);
  double acc = 0.0;  // sum of (rho_in + rho_out) dz: twice the integral
  int steps = 1;     // tets entered; ~(tets before t) once handed over
  bool slab_fail = false;
  // Bound on acc's error against the FP64 walk's.  Crossing k's height
  // error eps_k and density error erho_k enter acc as
  //   eps_k |a_k - a_{k+1}| + erho_k (d_k + d_{k+1}),
  // a_k = I_k (rho_{k-1} + rho_k), d_k = I_k |dz_k| (I_k: chord k counts);
  // the parts with chord k + 1 are added once it is known.  The densities'
  // own FP32 rounding (kCrossRound rho_k per crossing) sums to exactly
  // kCrossRound acc and is added at the end.  Crossing 0 is the FP64 one.
  float E = 0.f, eps_p = 0.f, erho_p = 0.f, a_p = 0.f;
  FaceRec R = load_facerec<kWide>(frec, t, f8);
  // one step: false at the walk's exit; two per iteration, so that the
  // values one step hands the next alternate between registers
  auto step = [&]() -> bool {
    // exit crossing in FP32, height above the entry crossing (does not wait
    // for R)
    const float den = (e0 + e1) + e2;
    const float M = amax(absmax3(p0.ox, p0.oy, p1.ox), absmax3(p1.oy, p2.ox, p2.oy));
    const float inv = rcp32(den);
    float dz = fmaf(e1, p0.z, fmaf(e2, p1.z, e0 * p2.z)) * inv;
    float r_out = fmaf(e1, p0.rho, fmaf(e2, p1.rho, e0 * p2.rho)) * inv;
    const bool stop = (int)R.code < 0 || (unsigned)steps >= (unsigned)g.max_steps;
    // first-order error of this crossing from its weights: side e (error
    // within kSideErr side_ext) moves it by (value at the side's vertex -
    // crossing value) e / den; heights are rounded relative to hab (the
    // crossing lies within the face's heights, so |dz| <= hab)
    const float hab = amax(amax(fabsf(p0.z), fabsf(p1.z)), fabsf(p2.z));
    float q = kSideErr * inv;
    if (!stop && fabsf(den) < flt.k_ill * M * M) {
      // rare: FP32 weights of a nearly vertical face are too uncertain
      const Cross64 o = face_cross64(rec64, t, f8 >> 3, x, y, load_rho4(rec32, t));
      dz = (float)(o.z - (kFullSlab ? carried_z(o, p0, p1, p2) : zc));
      r_out = o.rho;
      q = 0.f;
    }
    if (stop) return false;
    bool ok = true;
    bool inside = true;
    if (kFullSlab) {
      acc += (double)((rc + r_out) * dz);
    } else {
      ok = slab_inside(rec64, rec32, t, enter, f8 >> 3, x, y, zext, zc, zc + (double)dz, g, flt,
                       inside);
      if (inside) acc += (double)((rc + r_out) * dz);
    }
    {
      const float a = self(inside, rc + r_out, 0.f), d = self(inside, fabsf(dz), 0.f);
      // the entry crossing's terms with this chord
      E = fmaf(eps_p, fabsf(a_p - a), fmaf(erho_p, d, E));
      // the exit crossing's, its chord-k part now
      const float x0 = side_ext(p0.ox, p0.oy, p1.ox, p1.oy);
      const float x1 = side_ext(p1.ox, p1.oy, p2.ox, p2.oy);
      const float x2 = side_ext(p2.ox, p2.oy, p0.ox, p0.oy);
      eps_p = fmaf(q, fmaf(x1, fabsf(p0.z - dz), fmaf(x2, fabsf(p1.z - dz), x0 * fabsf(p2.z - dz))),
                   2.f * kCrossRound * hab);
      erho_p = q * fmaf(x1, fabsf(p0.rho - r_out),
                        fmaf(x2, fabsf(p1.rho - r_out), x0 * fabsf(p2.rho - r_out)));
      E = fmaf(erho_p, d, E);
      a_p = a;
    }
    if (!kFullSlab) zc += (double)dz;
    rc = r_out;
    p0.z -= dz;
    p1.z -= dz;
    p2.z -= dz;

    // enter the next tet through its face `back`; its fourth vertex Q
    const int back8 = code_field(R.code, f8);
    // offsets from the FP32 pairs: hi - hi rounds once relative to the
    // offset itself, lo - lo is far below it
    Slot qv{(R.qxh - xh) + (R.qxl - xl), (R.qyh - yh) + (R.qyl - yl),
            fmaf((p0.z + p1.z) + p2.z, 1.0f / 3.0f, R.dqz), R.rho, back8};
    // labels of the carried vertices in the next tet
    p0.lab8 = code_field(R.code, p0.lab8);
    p1.lab8 = code_field(R.code, p1.lab8);
    p2.lab8 = code_field(R.code, p2.lab8);
    t = R.nb;
    enter = back8 >> 3;

    // sides of the new vertex and the exit face
    float s0 = side32(p0.ox, p0.oy, qv.ox, qv.oy);
    float s1 = side32(p1.ox, p1.oy, qv.ox, qv.oy);
    float s2 = side32(p2.ox, p2.oy, qv.ox, qv.oy);
    const float Mt = amax(amax(M, fabsf(qv.ox)), fabsf(qv.oy));
    const float bound = side_bound(flt, Mt);
    bool zero = false;
    if (!(fminf(fminf(fabsf(s0), fabsf(s1)), fabsf(s2)) > bound)) {
      // rare: the FP64 walk's signs for the sides the filter leaves open
      if (!(fabsf(s0) > bound)) {
        const int sg = sign64(rec64, t, p0.lab8 >> 3, enter, x, y, g.lz);
        zero = zero || sg == 0;
        s0 = copysignf(fabsf(s0), (float)sg);
      }
      if (!(fabsf(s1) > bound)) {
        const int sg = sign64(rec64, t, p1.lab8 >> 3, enter, x, y, g.lz);
        zero = zero || sg == 0;
        s1 = copysignf(fabsf(s1), (float)sg);
      }
      if (!(fabsf(s2) > bound)) {
        const int sg = sign64(rec64, t, p2.lab8 >> 3, enter, x, y, g.lz);
        zero = zero || sg == 0;
        s2 = copysignf(fabsf(s2), (float)sg);
      }
    }
    // the face replacing P_j: j = 2 if s0 < 0 < s1, 0 if s1 < 0 < s2, 1 if s2 < 0 < s0
    const bool n0 = __builtin_signbit(s0), n1 = __builtin_signbit(s1), n2 = __builtin_signbit(s2);
    const bool r2 = n0 && !n1, r0 = n1 && !n2, r1 = n2 && !n0;
    f8 = seli(r0, p0.lab8, seli(r1, p1.lab8, p2.lab8));
    R = load_facerec<kWide>(frec, t, f8);
    if (!kFullSlab)
      zext = amax(amax(p0.z, p1.z), amax(p2.z, qv.z)) - fminf(fminf(p0.z, p1.z), fminf(p2.z, qv.z));
    // new face: P_j := Q; sides (P_{j-1}, Q) = s_{j-1}, (Q, P_{j+1}) = -s_{j+1}
    const float ne0 = self(r1, s0, self(r0, -s1, e0));
    const float ne1 = self(r2, s1, self(r1, -s2, e1));
    const float ne2 = self(r0, s2, self(r2, -s0, e2));
    e0 = ne0;
    e1 = ne1;
    e2 = ne2;
    p0 = sel_slot(r0, qv, p0);
    p1 = sel_slot(r1, qv, p1);
    p2 = sel_slot(r2, qv, p2);
    // with exact signs exactly one of r0, r1, r2 holds
    if (!kFullSlab) slab_fail = slab_fail || !ok;
    const bool fail = zero || !(r0 || r1 || r2) || !ok;
    steps = fail ? ~steps : steps + 1;
    return true;
  };
  for (;;) {
    if (!step()) break;
    if (!step()) break;
  }
  // A hand-over in mid-ray keeps the mixed prefix acc when its error bound
  // is within tol of it: the crossings' terms so far (E), the last
  // crossing's height error against the last chord (it does not cancel: the
  // FP64 walk forms that crossing again for the next chord), and the
  // densities' rounding.  The ray's value is at least its prefix's (the
  // integrand is not negative), so the whole ray then keeps the bound;
  // otherwise the FP64 walk takes the whole ray.
  auto resume_mid = [&](int before) {
    const float bound = fmaf(eps_p, a_p, fmaf(kCrossRound, (float)acc, E));
    return bound <= flt.tol * (float)acc ? Resume{t, enter, before, acc}
                                         : Resume{entry, -1, 0, 0.0};
  };
  if (steps < 0) {
    // t is the tet just entered (a midpoint left open: the whole ray)
    rs = (!kFullSlab && slab_fail) ? Resume{entry, -1, 0, 0.0} : resume_mid(~steps);
    return false;
  }
  if ((int)R.code < 0) {
    if (R.nb >= 0) {  // the mesh lacks the back link: the FP64 walk from t
      rs = resume_mid(steps - 1);
      return false;
    }
    // the last crossing in FP64, where the lanes of a wave are together again
    const Cross64 out = face_cross64(rec64, t, f8 >> 3, x, y, load_rho4(rec32, t));
    const double dz = out.z - (kFullSlab ? carried_z(out, p0, p1, p2) : zc);
    bool inside = true;
    if (!kFullSlab) {
      const float ze = amax(amax(p0.z, p1.z), p2.z) - fminf(fminf(p0.z, p1.z), p2.z);
      if (!slab_inside(rec64, rec32, t, enter, f8 >> 3, x, y, ze, zc, out.z, g, flt, inside)) {
        rs = Resume{t, enter, steps - 1, acc};
        return false;
      }
    }
    if (inside) acc += (double)(rc + out.rho) * dz;
    {
      const float a = self(inside, rc + out.rho, 0.f), d = self(inside, fabsf((float)dz), 0.f);
      const float hab = amax(amax(fabsf(p0.z), fabsf(p1.z)), fabsf(p2.z));
      E = fmaf(eps_p, fabsf(a_p - a), fmaf(erho_p, d, E));
      // this crossing is FP64; dz's reference height comes from the FP32
      // heights above the carried crossing; the densities' rounding
      E = fmaf(kCrossRound * (hab + d), a, fmaf(kCrossRound, (float)acc, E));
    }
    if (E > flt.tol * (float)acc) {
      // the FP32 crossings leave the value less certain than tol: the FP64
      // walk takes the whole ray
      rs = Resume{entry, -1, 0, 0.0};
      return false;
    }
    value = 0.5 * acc;
    crossed = steps;
    return true;
  }
  cnt.step_cap++;  // the step bound
  return true;
}

// The mixed walk, and where it hands the ray over, the FP64 walk from
// there: the value and tets crossed (0 for a failed walk).  resumed: the
// FP64 walk took part.
template <bool kFullSlab, bool kWide>
SDTFE_HD inline double walk_mixed_resume(const Rec32 *__restrict__ rec32,
                                         const Rec64 *__restrict__ rec64,
                                         const FaceRec *__restrict__ frec, const uint8_t *faces,
                                         int entry, double x, double y, const MapGeom &g,
                                         const Filter &flt, uint64_t key, int &crossed,
                                         bool &resumed, LaneCounters &cnt) {
  double v;
  Resume rs;
  resumed = false;
  if (walk_mixed<kFullSlab, kWide>(rec32, rec64, frec, faces, entry, x, y, g, flt, v, crossed,
                                   cnt, rs))
    return v;
  resumed = true;
  if (rs.enter < 0 || rs.steps == 0) return walk64(rec64, entry, x, y, g, key, crossed, cnt);
  Tet64 q;
  load_tet(rec64, rs.t, q);
  int c;
  const double w = walk64_from(rec64, rs.t, nb_of(q, rs.enter), rs.steps, x, y, g, key, c, cnt);
  crossed = c > 0 ? rs.steps + c : 0;
  return c > 0 ? 0.5 * rs.acc + w : 0.0;
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
  // An offset X - p is within 2.0001 u32 M of exact: from FP64 it rounds
  // once (plus u64 terms); from the face records' FP32 pairs (split32) the
  // hi difference and the final sum round once each relative to the offset,
  // the lo terms and the pairs' own error are below 2^-46 of the
  // coordinates.  A side's fma and product add 3 u32 M^2, the offset errors
  // 4 (2.0001 u32 M) M; 12 pads the bound's own rounding.
  m.flt.k_err = (float)(12.0 * u32);
  // reference zero band and the FP64 expression's own rounding (three
  // products and two sums of terms below 2 lz amax^2), over lz
  const double err64 = 16.0 * std::ldexp(1.0, -53) * 2.0 * g.lz * amax * amax;
  m.flt.thr = (float)(1.01 * (1e4 * 2.220446049250313080847e-16 + err64) / g.lz);
  // weight sums below 4e-3 M^2 (projected face area under 0.2 % of M^2)
  // leave the FP32 weights too uncertain: that crossing is formed in FP64
  m.flt.k_ill = 4e-3f;
  // a weight is within 4 k_err / k_ill (one side's error over the sum, and
  // the sum's error times the weight; each side's error bound uses its two
  // vertices, within the face's M): crossing heights are within
  // 4 k_err / k_ill of the face's height range, itself within the tet's;
  // doubled, plus the FP32 rounding of the mix
  m.flt.k_band = (float)(8.1 * m.flt.k_err / m.flt.k_ill + 1e-6);
  // per-ray error bound allowed (relative): below the 1e-5 per pixel the
  // mixed form promises against the FP64 walk, which a pixel's mean of ray
  // values then keeps; SDTFE_MIXED_TOL overrides it (tests)
  const char *tol = std::getenv("SDTFE_MIXED_TOL");
  m.flt.tol = tol && *tol ? (float)std::atof(tol) : 9.5e-6f;
  // every chord midpoint lies within the hull's z range: a midpoint on the
  // boundary belongs to a chord of zero length, which adds nothing
  m.flt.full_slab = t->n > 0 && g.off_z <= t->z_min && g.off_z + g.lz >= t->z_max;
  return m;
}

}  // namespace sdtfe
