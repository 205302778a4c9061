// walk_mixed.hip -- map and trace kernels of the mixed-precision walk
// (walk_mixed_core.h).  A map is two passes: the mixed walk over every
// pixel, then the FP64 walk over the queued pixels' handed-over samples
// (walk_defer.h).
#include <climits>
#include <cstdlib>
#include <mutex>
#include <string>

#include "walk_defer.h"

namespace sdtfe {

// LDS copy of the face_code table; every thread of the block calls this.
__device__ inline void load_faces(uint8_t *faces) {
  if (threadIdx.x < kFaceCodes) faces[threadIdx.x] = (uint8_t)face_code((int)threadIdx.x);
  __syncthreads();
}

__global__ void defer_reset_kernel(int *count) { *count = 0; }

// First pass.  L lanes per pixel: lane j of a pixel walks its samples j,
// j + L, ...; the L lanes of a wave's pixel walk in step (a wave's rays
// start together at the top of one small patch of the map, so they share
// more face records), and after each round the pixel's first lane takes the
// L values in sample order, so the map does not depend on L.  With L > 1
// the samples must fit the window of handed-over samples (map_lanes).
template <bool kFullSlab, bool kWide, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SDTFE_MIXED_WAVES))) void
map_mixed_kernel(const Rec32 *__restrict__ rec32, const Rec64 *__restrict__ rec64,
                 const FaceRec *__restrict__ frec, EntryGrid eg, MapGeom g, MixedConsts mc,
                 double *out, int *navg, int accumulate, DeferQueue dq,
                 unsigned long long *stats) {
  __shared__ uint8_t faces[kFaceCodes];
  load_faces(faces);
  int px, py, j;
  split_pixel<L>(g.grid, px, py, j);
  const bool active = px < g.grid && py < g.grid;
  const uint64_t pix = (uint64_t)py * (uint64_t)g.grid + (uint64_t)px;
  LaneCounters cnt;
  PixelSamples<map_window_words(L)> ps(g.samples);  // the pixel's first lane's
  // one sample: 1 walked (value v, crossed > 0), 0 adds nothing (a miss, a
  // failed walk), 2 handed over; -1 not walked
  auto sample = [&](int s, double &v) -> int {
    const uint64_t key = sample_key(g.seed, pix, (uint32_t)s);
    double xt, yt;
    sample_point(g, px, py, key, xt, yt);
    cnt.samples++;
    v = 0.0;
    const int entry = locate_entry(eg, xt, yt);
    if (entry < 0) {
      cnt.missed++;
      return 0;
    }
    int crossed;
    Resume rs;  // the second pass walks the sample again and resumes there
    if (!walk_mixed<kFullSlab, kWide>(rec32, rec64, frec, faces, entry, xt, yt, g, mc.flt, v,
                                      crossed, cnt, rs))
      return 2;
    cnt.tet_steps += (unsigned)crossed;
    if (crossed == 0) cnt.failed++;
    return crossed > 0 ? 1 : 0;
  };
  if (L == 1) {
    if (active)
      for (int s = 0; s < g.samples; ++s) {
        double v;
        const int k = sample(s, v);
        if (k == 1)
          ps.add(v, 1);
        else if (k == 2 && !ps.hand_over(s))
          break;  // s and every later sample go to the FP64 walk
      }
  } else {
    const int lead = (int)(threadIdx.x & 63) - j;
    for (int s0 = 0; s0 < g.samples; s0 += L) {
      double v = 0.0;
      int k = -1;
      if (active && s0 + j < g.samples) k = sample(s0 + j, v);
      for (int i = 0; i < L; ++i) {
        const double vi = __shfl(v, lead + i, 64);
        const int ki = __shfl(k, lead + i, 64);
        if (j == 0) {
          if (ki == 1)
            ps.add(vi, 1);
          else if (ki == 2)
            ps.hand_over(s0 + i);  // within the window (map_lanes)
        }
      }
    }
  }
  if (active && j == 0) {
    if (ps.clean(g.samples)) {
      const double val = ps.n_ok > 0 ? ps.sum / ps.n_ok : 0.0;
      if (accumulate)
        out[pix] += val;
      else
        out[pix] = val;
      if (navg) navg[pix] = ps.n_ok;
    } else {
      push_deferred(dq, ps, pix, 0);
    }
  }
  cnt.flush(stats);
}

// Second pass: one wave per queued pixel, grid-stride over the queue; does
// nothing when the queue holds more than max_entries pixels (the host then
// runs the per-sample form) or overflowed (the host repeats the map), so it
// can be queued before the host has read the queue's length.
template <bool kWide>
__global__ __launch_bounds__(256) void map_deferred_kernel(const Rec32 *__restrict__ rec32,
                                                          const Rec64 *__restrict__ rec64,
                                                          const FaceRec *__restrict__ frec,
                                                          EntryGrid eg, MapGeom g, MixedConsts mc,
                                                          double *out, int *navg, int accumulate,
                                                          DeferQueue dq, int max_entries,
                                                          unsigned long long *stats) {
  const int n = *dq.count;
  if (n > max_entries || n > dq.cap) return;  // the same for every thread
  __shared__ uint8_t faces[kFaceCodes];
  load_faces(faces);
  const DeferCtx c{rec32, rec64, frec, faces, eg, g, mc.flt};
  LaneCounters cnt;
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
    int n_ok;
    const double sum = finish_deferred_wave<kWide>(dq.q + i, c, n_ok, cnt);
    if ((threadIdx.x & 63) == 0) {
      const uint64_t pix = dq.q[i].pix;
      const double val = n_ok > 0 ? sum / n_ok : 0.0;
      if (accumulate)
        out[pix] += val;
      else
        out[pix] = val;
      if (navg) navg[pix] = n_ok;
    }
  }
  cnt.flush(stats);
}

// Second pass with one lane per queued sample, for a queue longer than the
// chip holds waves (most queued pixels hand over one or two samples, so a
// wave per pixel would leave most lanes idle): per-entry counts, their
// scan, the walks, then each entry's values added in sample order (the same
// additions as map_deferred_kernel's, so results are identical).
__global__ void defer_count_kernel(const Deferred *__restrict__ q, int n, int samples,
                                   int *cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = deferred_count(q + i, samples);
}

template <bool kWide>
__global__ __launch_bounds__(256) void defer_walk_kernel(const Rec32 *__restrict__ rec32,
                                                        const Rec64 *__restrict__ rec64,
                                                        const FaceRec *__restrict__ frec,
                                                        EntryGrid eg, MapGeom g, MixedConsts mc,
                                                        const Deferred *__restrict__ q, int n,
                                                        const int *__restrict__ off, int total,
                                                        double *val, unsigned char *ok,
                                                        unsigned long long *stats) {
  __shared__ uint8_t faces[kFaceCodes];
  load_faces(faces);
  const DeferCtx c{rec32, rec64, frec, faces, eg, g, mc.flt};
  LaneCounters cnt;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) {
    int lo = 0, hi = n - 1;  // the entry: the last e with off[e] <= i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= i)
        lo = mid;
      else
        hi = mid - 1;
    }
    bool k_ok;
    val[i] = deferred_walk<kWide>(q + lo, deferred_sample(q + lo, i - off[lo]), c, k_ok, cnt);
    ok[i] = (unsigned char)k_ok;
  }
  cnt.flush(stats);
}

__global__ void defer_combine_kernel(const Deferred *__restrict__ q, int n,
                                     const int *__restrict__ off, const double *__restrict__ val,
                                     const unsigned char *__restrict__ ok, double *out, int *navg,
                                     int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sum = q[i].sum;
  int n_ok = q[i].n_ok;
  for (int k = off[i]; k < off[i + 1]; ++k) {
    sum += val[k];
    n_ok += ok[k];
  }
  const uint64_t pix = q[i].pix;
  const double v = n_ok > 0 ? sum / n_ok : 0.0;
  if (accumulate)
    out[pix] += v;
  else
    out[pix] = v;
  if (navg) navg[pix] = n_ok;
}

// Rays one by one; a handed-over ray goes on in the FP64 walk in place
// (this kernel serves inspection, not maps).
__global__ void trace_mixed_kernel(const Rec32 *__restrict__ rec32,
                                   const Rec64 *__restrict__ rec64,
                                   const FaceRec *__restrict__ frec, EntryGrid eg, MapGeom g,
                                   MixedConsts mc, int n, const double *x, const double *y,
                                   const int *entry_in, const uint64_t *pixel,
                                   const uint32_t *sample, double *value, int *crossed_out,
                                   unsigned long long *stats) {
  __shared__ uint8_t faces[kFaceCodes];
  load_faces(faces);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  LaneCounters cnt;
  if (i < n) {
    const uint64_t key = sample_key(g.seed, pixel[i], sample[i]);
    const int entry = entry_in ? entry_in[i] : locate_entry(eg, x[i], y[i]);
    int crossed;
    bool resumed;
    const double v =
        mc.flt.full_slab
            ? walk_mixed_resume<true, true>(rec32, rec64, frec, faces, entry, x[i], y[i], g, mc.flt,
                                            key, crossed, resumed, cnt)
            : walk_mixed_resume<false, true>(rec32, rec64, frec, faces, entry, x[i], y[i], g,
                                             mc.flt, key, crossed, resumed, cnt);
    if (resumed) cnt.fp64_redo++;
    value[i] = v;
    crossed_out[i] = crossed;
    cnt.tet_steps += (unsigned)crossed;
  }
  cnt.flush(stats);
}

extern const bool kHaveMixed = true;

namespace {

__global__ void set_int_kernel(int *p, int v) { *p = v; }

int ensure_defer(sdtfe_tess *t, size_t entries) {
  if (!t->defer_count &&
      dev_malloc((void **)&t->defer_count, sizeof(int), t->stream) != hipSuccess)
    return fail(SDTFE_ENOMEM, "deferred-pixel counter allocation failed");
  if (entries <= t->defer_cap) return 0;
  dev_free(t->defer_buf, t->stream);
  t->defer_buf = nullptr;
  t->defer_cap = 0;
  if (dev_malloc((void **)&t->defer_buf, sizeof(Deferred) * entries, t->stream) != hipSuccess)
    return fail(SDTFE_ENOMEM, "deferred-pixel queue allocation failed");
  t->defer_cap = entries;
  return 0;
}

// The tessellation's own second-pass buffers, as they stand.
MixedWork tess_work(sdtfe_tess *t) {
  MixedWork w{};
  w.stats = t->stats;
  w.q = t->defer_buf;
  w.q_count = t->defer_count;
  w.q_cap = t->defer_cap;
  w.off = t->dq_off;
  w.tmp = t->dq_tmp;
  w.off_cap = t->dq_off_cap;
  w.tmp_ints = t->dq_off_cap ? scan_tmp_ints((int64_t)t->dq_off_cap) : 0;
  w.val = t->dq_val;
  w.ok = t->dq_ok;
  w.val_cap = t->dq_val_cap;
  return w;
}

// Queues up to this long get a wave per pixel (1024 blocks of 4 waves hold
// them all at once); longer ones a lane per sample.  SDTFE_DEFER_WAVE_ENTRIES
// overrides the limit (both forms give the same map; tests run each).
constexpr int kDeferWaveEntries = 4096;

int defer_wave_entries() {
  const char *e = getenv("SDTFE_DEFER_WAVE_ENTRIES");
  return e && *e ? atoi(e) : kDeferWaveEntries;
}

void launch_deferred_waves(sdtfe_tess *t, const MapGeom &g, const MixedConsts &mc, double *out,
                           int *navg, int accumulate, const DeferQueue &dq, int max_entries,
                           unsigned long long *stats, bool wide) {
  if (wide)
    map_deferred_kernel<true><<<1024, 256, 0, t->stream>>>(t->rec32, t->rec64, t->facerec,
                                                           t->grid, g, mc, out, navg, accumulate,
                                                           dq, max_entries, stats);
  else
    map_deferred_kernel<false><<<1024, 256, 0, t->stream>>>(t->rec32, t->rec64, t->facerec,
                                                            t->grid, g, mc, out, navg,
                                                            accumulate, dq, max_entries, stats);
}

// The second pass with a lane per queued sample over the n queued pixels.
// grow: the per-entry and per-sample buffers are the tessellation's and grow
// as needed; otherwise they are fixed, and a queue whose samples exceed them
// is finished by the wave-per-pixel form (the same map).
int defer_by_sample(sdtfe_tess *t, const MapGeom &g, const MixedConsts &mc, double *out,
                    int *navg, int accumulate, int n, bool wide, MixedWork &w, bool grow) {
  hipStream_t s = t->stream;
  if ((size_t)n + 1 > w.off_cap || scan_tmp_ints((int64_t)n + 1) > w.tmp_ints) {
    if (!grow) return fail(SDTFE_EINVAL, "second pass: queue beyond the scratch offsets");
    dev_free(t->dq_off, s);
    dev_free(t->dq_tmp, s);
    t->dq_off = t->dq_tmp = nullptr;
    t->dq_off_cap = 0;
    if (dev_malloc((void **)&t->dq_off, sizeof(int) * ((size_t)n + 1), s) != hipSuccess ||
        dev_malloc((void **)&t->dq_tmp, sizeof(int) * scan_tmp_ints((int64_t)n + 1), s) !=
            hipSuccess)
      return fail(SDTFE_ENOMEM, "second-pass offsets allocation failed");
    t->dq_off_cap = (size_t)n + 1;
    w.off = t->dq_off;
    w.tmp = t->dq_tmp;
    w.off_cap = t->dq_off_cap;
    w.tmp_ints = scan_tmp_ints((int64_t)n + 1);
  }
  // counts in off[0, n), off[n] = 0, scanned in place into offsets
  // (off[n] = the number of samples)
  defer_count_kernel<<<(n + 255) / 256, 256, 0, s>>>(w.q, n, g.samples, w.off);
  set_int_kernel<<<1, 1, 0, s>>>(w.off + n, 0);
  int local = 0;
  int *ht = w.h_count ? w.h_count : &local;
  int rc = scan_exclusive(w.off, w.off, (int64_t)n + 1, w.tmp, ht, s);
  if (rc) return rc;
  const int total = *ht;
  if (total <= 0) return fail(SDTFE_EHIP, "second pass: queued pixels list no samples");
  if ((size_t)total > w.val_cap) {
    if (!grow) {
      const DeferQueue dq{w.q, w.q_count, (int)w.q_cap};
      launch_deferred_waves(t, g, mc, out, navg, accumulate, dq, INT_MAX, w.stats, wide);
      SDTFE_HIP(hipGetLastError());
      return 0;
    }
    dev_free(t->dq_val, s);
    dev_free(t->dq_ok, s);
    t->dq_val = nullptr;
    t->dq_ok = nullptr;
    t->dq_val_cap = 0;
    if (dev_malloc((void **)&t->dq_val, sizeof(double) * (size_t)total, s) != hipSuccess ||
        dev_malloc((void **)&t->dq_ok, (size_t)total, s) != hipSuccess)
      return fail(SDTFE_ENOMEM, "second-pass values allocation failed");
    t->dq_val_cap = (size_t)total;
    w.val = t->dq_val;
    w.ok = t->dq_ok;
    w.val_cap = t->dq_val_cap;
  }
  const unsigned nb = (unsigned)((total + 255) / 256);
  if (wide)
    defer_walk_kernel<true><<<nb, 256, 0, s>>>(t->rec32, t->rec64, t->facerec, t->grid, g, mc,
                                               w.q, n, w.off, total, w.val, w.ok, w.stats);
  else
    defer_walk_kernel<false><<<nb, 256, 0, s>>>(t->rec32, t->rec64, t->facerec, t->grid, g, mc,
                                                w.q, n, w.off, total, w.val, w.ok, w.stats);
  defer_combine_kernel<<<(n + 255) / 256, 256, 0, s>>>(w.q, n, w.off, w.val, w.ok, out, navg,
                                                       accumulate);
  SDTFE_HIP(hipGetLastError());
  return 0;
}

}  // namespace

__global__ void zero_u64_kernel(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

template <int L>
void launch_first_pass_l(sdtfe_tess *t, const MapGeom &g, const MixedConsts &mc, double *out,
                         int *navg, int accumulate, const DeferQueue &dq,
                         unsigned long long *stats) {
  constexpr int tw = 2 * wave_bw(L), th = 2 * wave_bh(L);
  const int nb = ((g.grid + tw - 1) / tw) * ((g.grid + th - 1) / th);
  const bool wide = wide_records(t->nt);
  if (mc.flt.full_slab && !wide)
    map_mixed_kernel<true, false, L><<<nb, 256, 0, t->stream>>>(
        t->rec32, t->rec64, t->facerec, t->grid, g, mc, out, navg, accumulate, dq, stats);
  else if (!wide)
    map_mixed_kernel<false, false, L><<<nb, 256, 0, t->stream>>>(
        t->rec32, t->rec64, t->facerec, t->grid, g, mc, out, navg, accumulate, dq, stats);
  else if (mc.flt.full_slab)
    map_mixed_kernel<true, true, L><<<nb, 256, 0, t->stream>>>(
        t->rec32, t->rec64, t->facerec, t->grid, g, mc, out, navg, accumulate, dq, stats);
  else
    map_mixed_kernel<false, true, L><<<nb, 256, 0, t->stream>>>(
        t->rec32, t->rec64, t->facerec, t->grid, g, mc, out, navg, accumulate, dq, stats);
}

void launch_first_pass(sdtfe_tess *t, const MapGeom &g, const MixedConsts &mc, double *out,
                       int *navg, int accumulate, const DeferQueue &dq,
                       unsigned long long *stats) {
  switch (map_lanes(g.samples)) {
    case 16: launch_first_pass_l<16>(t, g, mc, out, navg, accumulate, dq, stats); break;
    case 4: launch_first_pass_l<4>(t, g, mc, out, navg, accumulate, dq, stats); break;
    default: launch_first_pass_l<1>(t, g, mc, out, navg, accumulate, dq, stats); break;
  }
}

// Both passes of a mixed map with the buffers of w.  The per-pixel second
// pass is queued behind the first and acts only on a short queue; the host
// reads the queue's length meanwhile and runs the per-sample form for a long
// one.  A queue that overflows w's is grown and the map repeated when grow
// is set (a map that accumulates must have room for every pixel from the
// start, since its first pass cannot be repeated), else it is an error.
int run_map_mixed(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate,
                  MixedWork &w, bool grow) {
  const size_t npix = (size_t)g.grid * (size_t)g.grid;
  const MixedConsts mc = mixed_consts(t, g);
  const bool wide = wide_records(t->nt);
  for (;;) {
    const DeferQueue dq{w.q, w.q_count, (int)std::min(w.q_cap, npix)};
    defer_reset_kernel<<<1, 1, 0, t->stream>>>(w.q_count);
    SDTFE_HIP(hipEventRecord(t->evk0, t->stream));
    launch_first_pass(t, g, mc, out, navg, accumulate, dq, w.stats);
    SDTFE_HIP(hipEventRecord(t->evk1, t->stream));
    const int max_wave = defer_wave_entries();
    launch_deferred_waves(t, g, mc, out, navg, accumulate, dq, max_wave, w.stats, wide);
    SDTFE_HIP(hipGetLastError());
    int local = 0;
    int *hc = w.h_count ? w.h_count : &local;
    SDTFE_HIP(hipMemcpyAsync(hc, w.q_count, sizeof(int), hipMemcpyDeviceToHost, t->stream));
    SDTFE_HIP(hipStreamSynchronize(t->stream));
    const int count = *hc;
    if (count > dq.cap) {
      // overflow: the queued pixels beyond the queue were not written; the
      // repeat writes every pixel again
      if (!grow || accumulate)
        return fail(SDTFE_EHIP, "second pass: deferred-pixel queue overflowed");
      zero_u64_kernel<<<1, 64, 0, t->stream>>>(w.stats, ST_COUNT);
      int rc = ensure_defer(t, npix);
      if (rc) return rc;
      int *const h_count = w.h_count;
      w = tess_work(t);
      w.h_count = h_count;
      continue;
    }
    if (count > max_wave)
      return defer_by_sample(t, g, mc, out, navg, accumulate, count, wide, w, grow);
    return 0;  // done by the per-pixel pass above (or nothing queued)
  }
}

// The queue holds a quarter of the pixels (all of them for a map that
// accumulates); a map that overflows it is run again with room for all.
int launch_map_mixed(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate,
                     int *h_count) {
  const size_t npix = (size_t)g.grid * (size_t)g.grid;
  if (npix > (size_t)INT_MAX) return fail(SDTFE_EINVAL, "map too large for the mixed walk");
  const size_t want = accumulate ? npix : std::max(npix / 4, std::min<size_t>(npix, 65536));
  int rc = ensure_defer(t, want);
  if (rc) return rc;
  MixedWork w = tess_work(t);
  w.h_count = h_count;
  return run_map_mixed(t, g, out, navg, accumulate, w, true);
}

// ----------------------------------------------- maps without allocation
// sdtfe_map_device_noalloc keeps its counters, its queue of pixels for the
// second pass and the second pass's per-entry and per-sample arrays in one
// file-scope array per GPU, sized for maps of up to kNoAllocPixels pixels
// (the queue then holds every pixel, so it cannot overflow) and for up to
// kNoAllocSamples handed-over samples in the per-sample second pass (more
// are finished by the wave-per-pixel form, which needs no further memory).
namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaStats = 0;
constexpr size_t kNaCount = kNaStats + align256(sizeof(unsigned long long) * ST_COUNT);
constexpr size_t kNaQueue = kNaCount + 256;
constexpr size_t kNaOff = kNaQueue + align256(sizeof(Deferred) * kNoAllocPixels);
constexpr size_t kNaTmpInts = 1024;  // >= scan_tmp_ints(kNoAllocPixels + 1)
constexpr size_t kNaTmp = kNaOff + align256(sizeof(int) * (kNoAllocPixels + 1));
constexpr size_t kNaVal = kNaTmp + align256(sizeof(int) * kNaTmpInts);
constexpr size_t kNaOk = kNaVal + align256(sizeof(double) * kNoAllocSamples);
constexpr size_t kNaBytes = kNaOk + align256(kNoAllocSamples);
static_assert(kNaBytes == kNoAllocScratchBytes, "scratch layout");

}  // namespace

// The scratch memory (device side).  Host code never passes this name to a
// kernel: each device's address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int noalloc_work(MixedWork *w) {
  static std::mutex mu;
  static void *addr[64] = {};
  int dev = 0;
  SDTFE_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(SDTFE_EINVAL, "device ordinal beyond 63");
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
      // once per device, with that device current
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(SDTFE_EHIP, std::string("scratch address: ") + hipGetErrorString(e));
      addr[dev] = p;
    }
    base = addr[dev];
  }
  if (scan_tmp_ints((int64_t)kNoAllocPixels + 1) > kNaTmpInts)
    return fail(SDTFE_EINVAL, "scratch scan space too small");
  unsigned char *b = static_cast<unsigned char *>(base);
  *w = MixedWork{};
  w->stats = reinterpret_cast<unsigned long long *>(b + kNaStats);
  w->q_count = reinterpret_cast<int *>(b + kNaCount);
  w->q = reinterpret_cast<Deferred *>(b + kNaQueue);
  w->q_cap = kNoAllocPixels;
  w->off = reinterpret_cast<int *>(b + kNaOff);
  w->off_cap = kNoAllocPixels + 1;
  w->tmp = reinterpret_cast<int *>(b + kNaTmp);
  w->tmp_ints = kNaTmpInts;
  w->val = reinterpret_cast<double *>(b + kNaVal);
  w->ok = b + kNaOk;
  w->val_cap = kNoAllocSamples;
  // SDTFE_NOALLOC_SAMPLES lowers the per-sample limit (tests use it to reach
  // the wave-per-pixel finish; both give the same map)
  const char *e = getenv("SDTFE_NOALLOC_SAMPLES");
  if (e && *e) w->val_cap = std::min<size_t>(w->val_cap, (size_t)std::max(0, atoi(e)));
  return 0;
}

void launch_map_fp64(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate,
                     unsigned long long *stats);

int run_map_noalloc(sdtfe_tess *t, const MapGeom &g, int precision, double *out,
                    MixedWork &w) {
  zero_u64_kernel<<<1, 64, 0, t->stream>>>(w.stats, ST_COUNT);
  SDTFE_HIP(hipEventRecord(t->ev0, t->stream));
  if (precision == SDTFE_PREC_MIXED) {
    const int rc = run_map_mixed(t, g, out, nullptr, 1, w, false);
    if (rc) return rc;
  } else {
    launch_map_fp64(t, g, out, nullptr, 1, w.stats);
  }
  SDTFE_HIP(hipGetLastError());
  SDTFE_HIP(hipEventRecord(t->ev1, t->stream));
  return 0;
}

void launch_trace_mixed(sdtfe_tess *t, const MapGeom &g, int n, const double *x,
                        const double *y, const int *entry, const uint64_t *pixel,
                        const uint32_t *sample, double *value, int *crossed) {
  trace_mixed_kernel<<<(n + 255) / 256, 256, 0, t->stream>>>(
      t->rec32, t->rec64, t->facerec, t->grid, g, mixed_consts(t, g), n, x, y, entry, pixel,
      sample, value, crossed, t->stats);
}

}  // namespace sdtfe
