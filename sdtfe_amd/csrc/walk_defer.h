// walk_defer.h -- pixels whose mixed walk handed sample-rays to the FP64
// walk.
//
// The mixed walk returns a ray it cannot take to the FP64 walk with the
// same tets and draws (walk_mixed's false return).  Walking that ray inside
// the mixed kernel would pull the FP64 walk's registers into every wave, so
// the pixel is queued instead: the first pass stores the sum over the samples
// it finished (in sample order) and which samples it handed over; a second
// pass walks those in FP64, adds them in sample order and writes the pixel.
// Every pixel thus has one writer and a fixed summation order, so results do
// not depend on scheduling.
#pragma once

#include "walk_mixed_core.h"

namespace sdtfe {

// Per-pixel bookkeeping of the first pass; the window of handed-over samples
// is 64 kWords wide (1 for a lane that walks its pixel's samples in order and
// can stop at a tail, 4 when the samples were spread over several lanes).
template <int kWords>
struct PixelSamples {
  double sum = 0.0;
  uint64_t mask[kWords] = {};
  int n_ok = 0;
  int base = -1;
  int tail;

  SDTFE_HD explicit PixelSamples(int samples) : tail(samples) {}

  // A finished sample (crossed == 0: a miss or a failed walk, which adds
  // nothing and does not count, src/dtfe.c:626-636).
  SDTFE_HD void add(double v, int crossed) {
    sum += v;
    n_ok += crossed > 0;
  }

  // A handed-over sample.  Returns false when it falls outside the window:
  // it and every later sample then go to the FP64 walk (the caller stops).
  SDTFE_HD bool hand_over(int s) {
    if (base < 0) base = s;
    const int k = s - base;
    if (k >= 64 * kWords) {
      tail = s;
      return false;
    }
    const uint64_t bit = 1ull << (k & 63);
    // bit-masked word select: the word index is lane-varying
    for (int j = 0; j < kWords; ++j) mask[j] |= bit & (0ull - (uint64_t)((k >> 6) == j));
    return true;
  }

  SDTFE_HD bool clean(int samples) const { return base < 0 && tail == samples; }
};

#if defined(__HIPCC__)
template <int kWords>
__device__ inline void push_deferred(const DeferQueue &dq, const PixelSamples<kWords> &ps,
                                     uint64_t pix, int item) {
  const int i = atomicAdd(dq.count, 1);
  if (i >= dq.cap) return;  // overflow: the host repeats the launch
  Deferred d;
  d.sum = ps.sum;
  d.pix = pix;
  for (int k = 0; k < 4; ++k) d.mask[k] = k < kWords ? ps.mask[k] : 0ull;
  d.base = ps.base;
  d.tail = ps.tail;
  d.n_ok = ps.n_ok;
  d.item = item;
  dq.q[i] = d;
}
#endif

// Samples of an entry: the handed-over ones in order, then tail .. samples-1.
SDTFE_HD inline int deferred_count(const Deferred *__restrict__ d, int samples) {
  int n = samples - d->tail;
  if (d->base >= 0)
    for (int w = 0; w < 4; ++w) n += __builtin_popcountll(d->mask[w]);
  return n;
}

SDTFE_HD inline int deferred_sample(const Deferred *__restrict__ d, int k) {
  if (d->base >= 0)
    for (int w = 0; w < 4; ++w) {
      uint64_t m = d->mask[w];
      const int c = __builtin_popcountll(m);
      if (k < c) {
        for (int i = 0; i < k; ++i) m &= m - 1;  // drop the k lowest set bits
        return d->base + 64 * w + __builtin_ctzll(m);
      }
      k -= c;
    }
  return d->tail + k;
}

// What the second pass walks with: the tessellation's records, the face
// table (LDS on the device) and the map.
struct DeferCtx {
  const Rec32 *rec32;
  const Rec64 *rec64;
  const FaceRec *frec;
  const uint8_t *faces;
  EntryGrid eg;
  MapGeom g;
  Filter flt;
};

// One sample of an entry: the mixed walk again (it hands the ray over where
// it did in the first pass) and the FP64 walk from there; its value (0 for
// a miss or a failed walk) and whether it crossed a tet.
template <bool kWide>
SDTFE_HD inline double deferred_walk(const Deferred *__restrict__ d, int s, const DeferCtx &c,
                                     bool &ok, LaneCounters &cnt) {
  const MapGeom &g = c.g;
  const uint64_t pix = d->pix;
  const int px = (int)(pix % (uint64_t)g.grid), py = (int)(pix / (uint64_t)g.grid);
  const uint64_t key = sample_key(g.seed, pix, (uint32_t)s);
  double xt, yt;
  sample_point(g, px, py, key, xt, yt);
  if (s > d->tail) cnt.samples++;  // the first pass counted the others
  ok = false;
  const int entry = locate_entry(c.eg, xt, yt);
  if (entry < 0) {  // only tail samples can miss
    cnt.missed++;
    return 0.0;
  }
  int crossed;
  bool resumed;
  const double v =
      c.flt.full_slab
          ? walk_mixed_resume<true, kWide>(c.rec32, c.rec64, c.frec, c.faces, entry, xt, yt, g,
                                           c.flt, key, crossed, resumed, cnt)
          : walk_mixed_resume<false, kWide>(c.rec32, c.rec64, c.frec, c.faces, entry, xt, yt, g,
                                            c.flt, key, crossed, resumed, cnt);
  if (resumed) cnt.fp64_redo++;
  cnt.tet_steps += (unsigned)crossed;
  if (crossed == 0) cnt.failed++;
  ok = crossed > 0;
  return v;
}

#if defined(__HIPCC__)
// Second pass for one entry by one wave: lane j walks the entry's samples
// j, j + 64, ...; the values are added in sample order.  Every lane returns
// the pixel's sample sum and n_ok (samples that crossed a tet).
template <bool kWide>
__device__ inline double finish_deferred_wave(const Deferred *__restrict__ d, const DeferCtx &c,
                                              int &n_ok, LaneCounters &cnt) {
  const int lane = threadIdx.x & 63;
  const int total = deferred_count(d, c.g.samples);
  double sum = d->sum;
  n_ok = d->n_ok;
  for (int r0 = 0; r0 < total; r0 += 64) {
    double v = 0.0;
    bool ok = false;
    if (r0 + lane < total) v = deferred_walk<kWide>(d, deferred_sample(d, r0 + lane), c, ok, cnt);
    const int m = min(64, total - r0);
    for (int i = 0; i < m; ++i) {
      sum += __shfl(v, i, 64);
      n_ok += __shfl((int)ok, i, 64);
    }
  }
  return sum;
}
#endif

}  // namespace sdtfe
