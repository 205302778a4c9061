// walk_fp64_face.h -- the FP64 walk over face records (FRec64): the results
// of walk64 (shoot_3d_ray, src/dtfe.c:247-374) bit for bit, with about half
// of its arithmetic per tet.
//
// A ray enters each tet through a face it has just left the previous tet
// by.  When the reference decides a tet without a special case, that face
// is its enter face, and:
//  * The record (t, f) lists face f's vertices P0, P1, P2 in the order the
//    reference forms that face's crossing height, and the vertex N opposite
//    it.  In that order the face's three sides (the reference's Plücker
//    products for its edges, signed as its face test signs them) are
//    E_i = u(P_{i+2}, P_{i+1}), u(X, Y) the product with the edge from X to
//    Y; they are all beyond the zero band and positive on the enter face.
//  * Every other face shares an edge with the enter face, on which its side
//    is negative, so none can be an enter face; the three faces meeting at N
//    have sides n_{k+1}, -n_{k+2}, -E_k (n_j = u(N, P_j), the face opposite
//    P_k), and the ray leaves through the one whose three sides are all
//    negative beyond the band: n_{k+1} < -ZERO < ZERO < n_{k+2}.  Such a face
//    shares an edge with each other face, on which their side is positive,
//    so it is the only leave face; the reference's sequential test over
//    BDC, ACD, ADB, ABC then picks the same two faces, without a special case.
//  * Only the enter and leave faces' heights are formed, each from its three
//    sides in the reference's order of addition (face_z).
//  * The enter face's sides are the previous tet's leave-face sides, negated
//    and permuted (the two tets see the face with opposite orientations, and
//    u(Y, X) = -u(X, Y) exactly), so only the three products n_j are new.
// Every value the reference forms is formed here from the same operands in
// the same order.  Whatever the fast step does not decide -- a side in the
// zero band, no leave face, a record marked for hand-over -- hands the ray,
// at that tet, to walk64_from, which continues with the reference's full
// face test and perturbation.  (A Plücker product formed with its edge
// reversed can differ from the reference's only in the sign of a zero, and
// the fast step never uses a zero side.)
#pragma once

#include "walk_fp64_core.h"

namespace sdtfe {

// Vertex i of face f in the reference's order for its crossing height (face
// f is opposite vertex f): BDC as D, C, B; ACD as C, D, A; ADB as B, A, D;
// ABC as A, B, C (src/dtfe.c:67-74, 112-119, 151-158, 184-191).
SDTFE_HD constexpr int face_vertex(int f, int i) {
  return f == 0 ? 3 - i
       : f == 1 ? (i == 0 ? 2 : i == 1 ? 3 : 0)
       : f == 2 ? (i == 0 ? 1 : i == 1 ? 0 : 3)
                : i;
}

// Which side face f's denominator adds last: BDC (E0 + E2) + E1, ACD
// (E1 + E2) + E0, ADB (E0 + E1) + E2, ABC (E2 + E0) + E1.
SDTFE_HD constexpr int face_last(int f) { return f == 1 ? 0 : f == 2 ? 2 : 1; }

// Rotation r with face_vertex(g, i) = Q[(i + r) % 3] for the face g opposite
// P_k of face f, Q = (P_{k+2}, P_{k+1}, N) (3: no rotation fits).
SDTFE_HD constexpr int leave_rotation(int f, int k) {
  const int g = face_vertex(f, k);
  const int q[3] = {face_vertex(f, (k + 2) % 3), face_vertex(f, (k + 1) % 3), f};
  for (int r = 0; r < 3; ++r) {
    bool ok = true;
    for (int i = 0; i < 3; ++i) ok = ok && face_vertex(g, i) == q[(i + r) % 3];
    if (ok) return r;
  }
  return 3;
}

SDTFE_HD constexpr bool leave_rotations_fit() {
  for (int f = 0; f < 4; ++f)
    for (int k = 0; k < 3; ++k)
      if (leave_rotation(f, k) > 2) return false;
  return true;
}
// the four faces are oriented alike: each leave face is a rotation of the
// cycle the enter face induces on it
static_assert(leave_rotations_fit(), "face orders");

// Per-lane choice among three or four values by an index k in 0..3, like
// pick4, with the two bit masks (bit 0 and bit 1 of k, each 0 or ~0) formed
// once per index: every 64-bit choice is then four bit-selects.
struct Sel {
  uint32_t m0, m1;
};

SDTFE_HD inline Sel sel_of(int k) {
  return Sel{0u - (uint32_t)(k & 1), 0u - (uint32_t)((k >> 1) & 1)};
}

// (m & x) | (~m & y) as one v_bfi_b32 on the GPU (the compiler otherwise
// splits many of these into two instructions).
SDTFE_HD inline uint32_t bfi32(uint32_t m, uint32_t x, uint32_t y) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(x), "v"(y));
  return r;
#else
  return (m & x) | (~m & y);
#endif
}

SDTFE_HD inline uint32_t lo32(double v) { return (uint32_t)__builtin_bit_cast(uint64_t, v); }
SDTFE_HD inline uint32_t hi32(double v) { return (uint32_t)(__builtin_bit_cast(uint64_t, v) >> 32); }
SDTFE_HD inline double join64(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// k = 0: a, 1: b, 2: c
SDTFE_HD inline double sel3(const Sel &s, double a, double b, double c) {
  return join64(bfi32(s.m1, lo32(c), bfi32(s.m0, lo32(b), lo32(a))),
                bfi32(s.m1, hi32(c), bfi32(s.m0, hi32(b), hi32(a))));
}

// k = 0: a, 1: b, 2: c, 3: d
SDTFE_HD inline double sel4(const Sel &s, double a, double b, double c, double d) {
  return join64(bfi32(s.m1, bfi32(s.m0, lo32(d), lo32(c)), bfi32(s.m0, lo32(b), lo32(a))),
                bfi32(s.m1, bfi32(s.m0, hi32(d), hi32(c)), bfi32(s.m0, hi32(b), hi32(a))));
}

SDTFE_HD inline int sel3i(const Sel &s, int a, int b, int c) {
  return (int)bfi32(s.m1, (uint32_t)c, bfi32(s.m0, (uint32_t)b, (uint32_t)a));
}

// Crossing height of a face from its sides in the reference's order, the
// side its denominator adds last, and its vertices' heights: the
// reference's barycentric point (src/dtfe.c:67-74 and its three siblings;
// every face's weights are E1 and E2 over the sum).
SDTFE_HD inline double face_z(double e0, double e1, double e2, const Sel &last, double z0,
                              double z1, double z2) {
  const double pa = sel3(last, e1, e0, e0), pb = sel3(last, e2, e2, e1);
  const double pl = sel3(last, e0, e1, e2);
  double den = pa + pb;
  den = den + pl;
  const double inv = 1.0 / den;
  const double w1 = e1 * inv, w2 = e2 * inv;
  return (1 - w1 - w2) * z0 + w1 * z1 + w2 * z2;
}

// The single face of tet n whose neighbour is t, -1 if none or several.
SDTFE_HD inline int back_face(const int *tets, int n, int t) {
  int b = -1, c = 0;
  for (int k = 0; k < 4; ++k)
    if (tets[8 * (int64_t)n + 4 + k] == t) {
      if (b < 0) b = k;
      ++c;
    }
  return c == 1 ? b : -1;
}

// Face record (t, f) from the particles (x, y, z, rho), the tetra_data rows
// (four vertex ids, then the tet across the face opposite each) and the
// tets' gradients (T x 3).
SDTFE_HD inline void make_frec64(const double4 *pos, const int *tets, const double *nabla, int t,
                                 int f, FRec64 &r) {
  const int *row = tets + 8 * (int64_t)t;
  const double4 nv = pos[row[f]];
  r.n[0] = nv.x;
  r.n[1] = nv.y;
  r.n[2] = nv.z;
  for (int i = 0; i < 3; ++i) {
    const double4 q = pos[row[face_vertex(f, i)]];
    r.p[3 * i] = q.x;
    r.p[3 * i + 1] = q.y;
    r.p[3 * i + 2] = q.z;
  }
  r.rho_a = pos[row[0]].w;
  for (int i = 0; i < 3; ++i) r.g[i] = nabla[3 * (int64_t)t + i];
  unsigned code = (unsigned)face_last(f);
  for (int k = 0; k < 3; ++k) {
    const int g = face_vertex(f, k);  // the face opposite P_k
    const int nb = row[4 + g];
    int next = -1;
    if (nb >= 0) {
      const int back = back_face(tets, nb, t);
      next = back < 0 ? -2 : 4 * nb + back;
    }
    r.next[k] = next;
    code |= (unsigned)leave_rotation(f, k) << (2 + 4 * k);
    code |= (unsigned)face_last(g) << (4 + 4 * k);
  }
  // the tet a ray entering here comes from, and the same face's order there
  const int pv = row[4 + f];
  const int b = pv >= 0 ? back_face(tets, pv, t) : -1;
  if (pv < 0) {
    code |= kFRecFresh;
  } else if (b < 0) {
    code |= kFRecHand;
  } else {
    int pi[3];
    bool ok = true;
    for (int i = 0; i < 3; ++i) {
      const int id = row[face_vertex(f, i)];
      pi[i] = -1;
      for (int j = 2; j >= 0; --j)
        if (tets[8 * (int64_t)pv + face_vertex(b, j)] == id) pi[i] = j;
      ok = ok && pi[i] >= 0;
    }
    ok = ok && pi[0] != pi[1] && pi[1] != pi[2] && pi[0] != pi[2];
    // opposite orientations: the permutation is a transposition (odd)
    const bool odd = ok && (pi[0] == 0 ? pi[1] == 2 : pi[0] == 1 ? pi[1] == 0 : pi[1] == 1);
    if (!odd)
      code |= kFRecHand;
    else
      for (int i = 0; i < 3; ++i) code |= (unsigned)pi[i] << (14 + 2 * i);
  }
  r.code = code;
}

SDTFE_HD inline void load_frec(const FRec64 *__restrict__ fr, int rid, FRec64 &r) {
  const double2 *q = reinterpret_cast<const double2 *>(fr + rid);
  const double2 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6],
                q7 = q[7];
  const int4 w = reinterpret_cast<const int4 *>(fr + rid)[8];
  r.n[0] = q0.x; r.n[1] = q0.y; r.n[2] = q1.x;
  r.p[0] = q1.y; r.p[1] = q2.x; r.p[2] = q2.y;
  r.p[3] = q3.x; r.p[4] = q3.y; r.p[5] = q4.x;
  r.p[6] = q4.y; r.p[7] = q5.x; r.p[8] = q5.y;
  r.rho_a = q6.x; r.g[0] = q6.y; r.g[1] = q7.x; r.g[2] = q7.y;
  r.next[0] = w.x; r.next[1] = w.y; r.next[2] = w.z;
  r.code = (unsigned)w.w;
}

// Where the face walk leaves a ray to the walk over Rec64: the tet it
// could not decide (-1: none, the walk ended on the hull), the tet the ray
// came from and the tets entered before it.
struct FaceHand {
  int t, prev, steps;
};

// The walk over face records from record rid (4 x tet + face: the entry
// tet seen from the forward-hull face the ray enters by) along the vertical
// ray through (x, y), lz long: act(r, z_in, z_out) for each tet crossed, in
// order, with r holding the tet's vertex A, rho_a and gradient.  Returns
// the tets crossed; hand says where the walk over Rec64 must take over (the
// entry tet itself when the ray does not enter by that face without a
// special case).
template <class Act>
SDTFE_HD inline int face_walk(const FRec64 *__restrict__ frec, int rid, double x, double y,
                              double lz, int max_steps, Act &act, FaceHand &hand) {
  const double p2 = lz, p3 = -lz * y, p4 = lz * x;
  int prev = rid >> 2, steps = 0, n_cross = 0, hand_t = -1;
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;  // the previous leave face's sides
  bool fresh = true;
  Tet64 r;
  // one step: false at the walk's end or a hand-over; two per iteration,
  // so that the values one step hands the next alternate between registers
  auto step = [&]() -> bool {
    const int t = rid >> 2;
    FRec64 f;
    load_frec(frec, rid, f);
    const unsigned code = f.code;
    bool hand_here = steps >= max_steps || (code & kFRecHand) != 0;
    double e0, e1, e2;
    if (fresh || (code & kFRecFresh)) {
      e0 = edge_u(p2, p3, p4, f.p[6], f.p[7], f.p[3], f.p[4]);  // u(P2, P1)
      e1 = edge_u(p2, p3, p4, f.p[0], f.p[1], f.p[6], f.p[7]);  // u(P0, P2)
      e2 = edge_u(p2, p3, p4, f.p[3], f.p[4], f.p[0], f.p[1]);  // u(P1, P0)
      hand_here = hand_here || !(e0 > kZero && e1 > kZero && e2 > kZero);
    } else {
      e0 = -sel3(sel_of((int)(code >> 14) & 3), c0, c1, c2);
      e1 = -sel3(sel_of((int)(code >> 16) & 3), c0, c1, c2);
      e2 = -sel3(sel_of((int)(code >> 18) & 3), c0, c1, c2);
    }
    const double n0 = edge_u(p2, p3, p4, f.n[0], f.n[1], f.p[0], f.p[1]);
    const double n1 = edge_u(p2, p3, p4, f.n[0], f.n[1], f.p[3], f.p[4]);
    const double n2 = edge_u(p2, p3, p4, f.n[0], f.n[1], f.p[6], f.p[7]);
    const bool l0 = n1 < -kZero && n2 > kZero;
    const bool l1 = n2 < -kZero && n0 > kZero;
    const bool l2 = n0 < -kZero && n1 > kZero;
    const int k = (l1 ? 1 : 0) | (l2 ? 2 : 0);
    const Sel sk = sel_of(k);
    const int nxt = sel3i(sk, f.next[0], f.next[1], f.next[2]);
    hand_here = hand_here || !(l0 || l1 || l2) || nxt == -2 || (nxt >= 0 && (nxt >> 2) == prev);
    if (hand_here) {
      hand_t = t;
      return false;
    }
    ++steps;
    const double z_in = face_z(e0, e1, e2, sel_of((int)(code & 3)), f.p[2], f.p[5], f.p[8]);
    // the leave face as (P_{k+2}, P_{k+1}, N): sides n_{k+1}, -n_{k+2}, -E_k
    const double q0 = sel3(sk, n1, n2, n0), q1 = -sel3(sk, n2, n0, n1), q2 = -sel3(sk, e0, e1, e2);
    const double y0 = sel3(sk, f.p[8], f.p[2], f.p[5]), y1 = sel3(sk, f.p[5], f.p[8], f.p[2]);
    const double y2 = f.n[2];
    const Sel rot = sel_of((int)(code >> (2 + 4 * k)) & 3);
    const Sel last = sel_of((int)(code >> (4 + 4 * k)) & 3);
    // ... rotated into the reference's order for that face
    const double g0 = sel3(rot, q0, q1, q2), g1 = sel3(rot, q1, q2, q0), g2 = sel3(rot, q2, q0, q1);
    const double z_out = face_z(g0, g1, g2, last, sel3(rot, y0, y1, y2), sel3(rot, y1, y2, y0),
                                sel3(rot, y2, y0, y1));
    // vertex A: N on face 0, else P2, P1, P0 on faces 1, 2, 3
    const Sel sa = sel_of(rid & 3);
    r.ax = sel4(sa, f.n[0], f.p[6], f.p[3], f.p[0]);
    r.ay = sel4(sa, f.n[1], f.p[7], f.p[4], f.p[1]);
    r.az = sel4(sa, f.n[2], f.p[8], f.p[5], f.p[2]);
    r.rho_a = f.rho_a;
    r.gx = f.g[0];
    r.gy = f.g[1];
    r.gz = f.g[2];
    act(r, z_in, z_out);
    ++n_cross;
    if (nxt < 0) return false;
    c0 = g0;
    c1 = g1;
    c2 = g2;
    fresh = false;
    prev = t;
    rid = nxt;
    return true;
  };
  for (;;) {
    if (!step()) break;
    if (!step()) break;
  }
  hand = FaceHand{hand_t, prev, steps};
  return n_cross;
}

// walk64 over face records from record rid (see face_walk): same value,
// tets crossed and counters.  The walk over Rec64 takes over where the face
// walk cannot decide, with the sum so far.
SDTFE_HD inline double walk64f(const Rec64 *__restrict__ rec, const FRec64 *__restrict__ frec,
                               int rid, double x, double y, const MapGeom &g, uint64_t key,
                               int &crossed, LaneCounters &cnt) {
  crossed = 0;
  if (rid < 0) return 0.0;
  double acc = 0.0;
  auto integrate = [&](const Tet64 &r, double z_in, double z_out) {
    acc += tet_integral64(r, x, y, z_in, z_out, g.off_z, g.lz);
  };
  FaceHand hand;
  const int n = face_walk(frec, rid, x, y, g.lz, g.max_steps, integrate, hand);
  if (hand.t >= 0) {
    int c;
    const double v = walk64_from(rec, hand.t, hand.prev, hand.steps, x, y, g, key, c, cnt, acc);
    crossed = c ? n + c : 0;
    return v;
  }
  crossed = n;
  return acc;
}

}  // namespace sdtfe
