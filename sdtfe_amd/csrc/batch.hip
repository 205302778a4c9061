// batch.hip -- many maps (cutouts of uneven size) in one persistent launch.
//
// Work items are 256-lane tiles of pixels.  Item i's map is cut into tiles
// of (256 / L_i) pixels, where L_i ("lanes per pixel") splits each pixel's
// samples over L_i lanes: L = 1 gives one pixel per lane (16x16 pixels per
// tile), L = 4 gives 8x8 pixels whose samples are spread over 4 lanes each
// (more parallelism for small maps and short columns).  Per-sample values
// land in LDS and one lane per pixel sums them in sample order, so the FP64
// map is bit-identical to the one-pixel-per-lane launch.
//
// The tile space is item-major and cut into eight contiguous shares, one per
// XCD; a block (dealt round-robin to the XCDs) claims tiles of its own
// share from that share's counter, so neighbouring tiles run on one L2, and
// takes tiles from the other shares when its own is exhausted.  Every wave
// leaves once all counters pass their ends.
//
// The launch (counter reset kernel + persistent kernel) is captured once in
// a hipGraph and replayed; item descriptors are copied in before each
// replay, so the graph stays valid as long as the buffers do.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "sdtfe_host.h"
#include "walk_defer.h"
#include "walk_fp64_face.h"

namespace sdtfe {

constexpr int kXcd = 8;
constexpr int kLdsBytes = 32 * 1024;

struct BatchItem {
  const Rec64 *rec64;
  const Rec32 *rec32;
  const FaceRec *facerec;
  const FRec64 *frec64;
  EntryGrid grid;
  MapGeom geom;
  MixedConsts mc;
  double *out;
  int lpp;          // lanes per pixel
  int tile_w;       // pixels per tile row
  int tile_h;
  int tiles_x;
  int first_tile;   // prefix over items
  int n_tiles;      // tiles of this item in this launch
  int tile_offset;  // first tile of the item's map covered by this launch
};

struct BatchHdr {
  int n_items;
  int total_tiles;
  int active_blocks;    // blocks of the persistent kernel that take tiles
  int share[kXcd + 1];  // tile-space cut points
  DeferQueue dq;        // mixed walk: pixels for the FP64 pass (count: counters[kXcd])
};

// Wave sub-tile shape for L lanes per pixel: 64/L pixels as bw x bh.
SDTFE_HD inline void wave_block(int lpp, int &bw, int &bh) {
  bw = wave_bw(lpp);
  bh = wave_bh(lpp);
}

__global__ void batch_reset_kernel(const BatchHdr *hdr, int *counters,
                                   unsigned long long *item_stats, int n_stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kXcd) counters[i] = hdr->share[i];
  if (i == kXcd) counters[kXcd] = 0;
  for (int k = i; k < n_stats; k += gridDim.x * blockDim.x) item_stats[k] = 0ull;
}

// Per-sample slots in LDS (lanes per pixel > 1): the value, or one of two
// NaN codes for a sample that adds nothing (a miss or a failed walk) and one
// handed to the FP64 walk.
constexpr uint64_t kSlotNone = 0x7ff8dead00000001ull;
constexpr uint64_t kSlotHanded = 0x7ff8dead00000002ull;

// One sample of a pixel: value and tet count, false if handed over.
template <bool kMixed, bool kWide>
__device__ inline bool sample_one(const BatchItem &it, const uint8_t *faces, int px, int py,
                                  uint64_t pix, int s, double &v, int &crossed,
                                  LaneCounters &cnt) {
  const MapGeom &g = it.geom;
  const uint64_t key = sample_key(g.seed, pix, (uint32_t)s);
  double xt, yt;
  sample_point(g, px, py, key, xt, yt);
  cnt.samples++;
  v = 0.0;
  crossed = 0;
  const int entry = kMixed ? locate_entry(it.grid, xt, yt) : locate_entry_rid(it.grid, xt, yt);
  if (entry < 0) {
    cnt.missed++;
    return true;
  }
  if (kMixed) {
    Resume rs;  // the second pass walks the sample again and resumes there
    const bool done = it.mc.flt.full_slab
                          ? walk_mixed<true, kWide>(it.rec32, it.rec64, it.facerec, faces, entry,
                                                    xt, yt, g, it.mc.flt, v, crossed, cnt, rs)
                          : walk_mixed<false, kWide>(it.rec32, it.rec64, it.facerec, faces, entry,
                                                     xt, yt, g, it.mc.flt, v, crossed, cnt, rs);
    if (!done) return false;
  } else {
    v = walk64f(it.rec64, it.frec64, entry, xt, yt, g, key, crossed, cnt);
  }
  cnt.tet_steps += (unsigned)crossed;
  if (crossed == 0) cnt.failed++;
  return true;
}

// The persistent loop; lds: per-sample values for lanes-per-pixel > 1;
// kWide: 64-bit face-record offsets (tessellations beyond 2^25 tets).
template <bool kMixed, bool kWide>
__device__ __forceinline__ void batch_body(const BatchItem *__restrict__ items,
                                           const BatchHdr *__restrict__ hdr, int *counters,
                                           unsigned long long *item_stats, double *lds,
                                           int &s_tile, uint8_t *faces) {
  // blocks beyond active_blocks leave at once, freeing their slots for
  // other streams' kernels (the same for every thread of a block)
  if ((int)blockIdx.x >= hdr->active_blocks) return;
  if (kMixed && threadIdx.x < kFaceCodes) faces[threadIdx.x] = (uint8_t)face_code((int)threadIdx.x);
  const int home = blockIdx.x % kXcd;
  const int n_items = hdr->n_items;
  int xcd = home, tried = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int t = -1;
      while (tried < kXcd) {
        const int c = atomicAdd(counters + xcd, 1);
        if (c < hdr->share[xcd + 1]) {
          t = c;
          break;
        }
        xcd = (xcd + 1) % kXcd;
        ++tried;
      }
      s_tile = t;
    }
    __syncthreads();
    const int tile = s_tile;
    __syncthreads();
    if (tile < 0) break;
    // item holding this tile (items sorted by first_tile)
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].first_tile <= tile) lo = mid; else hi = mid - 1;
    }
    const BatchItem &it = items[lo];
    const MapGeom &g = it.geom;
    const int lt = tile - it.first_tile + it.tile_offset;
    const int L = it.lpp;
    int bw, bh;
    wave_block(L, bw, bh);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = lane / L, j = lane % L;  // pixel within the wave block, sample slice
    const int px = (lt % it.tiles_x) * it.tile_w + (w & 1) * bw + q % bw;
    const int py = (lt / it.tiles_x) * it.tile_h + (w >> 1) * bh + q / bw;
    const bool active = px < g.grid && py < g.grid;
    const uint64_t pix = (uint64_t)py * (uint64_t)g.grid + (uint64_t)px;
    LaneCounters cnt;
    PixelSamples<1> ps(g.samples);
    const int slot = threadIdx.x / L;  // pixel slot in the block
    if (active) {
      for (int s = j; s < g.samples; s += L) {
        double v;
        int crossed;
        const bool done = sample_one<kMixed, kWide>(it, faces, px, py, pix, s, v, crossed, cnt);
        if (L == 1) {
          if (done)
            ps.add(v, crossed);  // a miss or a failed walk adds 0 (src/dtfe.c:626-636)
          else if (!ps.hand_over(s))
            break;
        } else {
          lds[slot * g.samples + s] =
              done ? (crossed > 0 ? v : __builtin_bit_cast(double, kSlotNone))
                   : __builtin_bit_cast(double, kSlotHanded);
        }
      }
    }
    if (L > 1) {
      // lanes of one pixel are adjacent within the wave; the first sums in
      // sample order (a window of kDeferWindow >= 16 L samples fits)
      __syncthreads();
      if (active && j == 0) {
        PixelSamples<4> pw(g.samples);
        for (int s = 0; s < g.samples; ++s) {
          const double v = lds[slot * g.samples + s];
          const uint64_t u = __builtin_bit_cast(uint64_t, v);
          if (u == kSlotHanded)
            pw.hand_over(s);
          else if (u != kSlotNone)
            pw.add(v, 1);
        }
        if (pw.clean(g.samples))
          it.out[pix] += pw.n_ok > 0 ? pw.sum / pw.n_ok : 0.0;
        else
          push_deferred(hdr->dq, pw, pix, lo);
      }
    } else if (active) {
      if (ps.clean(g.samples))
        it.out[pix] += ps.n_ok > 0 ? ps.sum / ps.n_ok : 0.0;
      else
        push_deferred(hdr->dq, ps, pix, lo);
    }
    cnt.flush(item_stats + (size_t)lo * ST_COUNT);
    __syncthreads();
  }
}

// Built for 3 waves per SIMD, as the single-map FP64 kernel is (left to
// itself the allocator takes 178 VGPRs, which allows only 2).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void batch_kernel_fp64(
    const BatchItem *__restrict__ items, const BatchHdr *__restrict__ hdr, int *counters,
    unsigned long long *item_stats) {
  extern __shared__ double lds[];
  __shared__ int s_tile;
  batch_body<false, false>(items, hdr, counters, item_stats, lds, s_tile, nullptr);
}

// Built for 4 waves per SIMD: the persistent loop holds more state than the
// single-map kernel, and at 5 the allocator spills inside the walk loop.
#ifndef SDTFE_BATCH_MIXED_WAVES
#define SDTFE_BATCH_MIXED_WAVES 4
#endif
template <bool kWide>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SDTFE_BATCH_MIXED_WAVES))) void
batch_kernel_mixed(const BatchItem *__restrict__ items, const BatchHdr *__restrict__ hdr,
                   int *counters, unsigned long long *item_stats) {
  extern __shared__ double lds[];
  __shared__ int s_tile;
  __shared__ uint8_t faces[kFaceCodes];
  batch_body<true, kWide>(items, hdr, counters, item_stats, lds, s_tile, faces);
}

template <bool kMixed, bool kWide>
constexpr auto batch_kernel() {
  if constexpr (kMixed)
    return batch_kernel_mixed<kWide>;
  else
    return batch_kernel_fp64;
}

// Second pass of a mixed batch: one wave per queued pixel of any item.
template <bool kWide>
__global__ __launch_bounds__(256) void batch_deferred_kernel(const BatchItem *__restrict__ items,
                                                            const BatchHdr *__restrict__ hdr,
                                                            const int *counters,
                                                            unsigned long long *item_stats) {
  __shared__ uint8_t faces[kFaceCodes];
  if (threadIdx.x < kFaceCodes) faces[threadIdx.x] = (uint8_t)face_code((int)threadIdx.x);
  __syncthreads();
  const DeferQueue dq = hdr->dq;
  const int n = min(counters[kXcd], dq.cap);
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
    const Deferred *d = dq.q + i;
    const BatchItem &it = items[d->item];
    LaneCounters cnt;
    int n_ok;
    const DeferCtx c{it.rec32, it.rec64, it.facerec, faces, it.grid, it.geom, it.mc.flt};
    const double sum = finish_deferred_wave<kWide>(d, c, n_ok, cnt);
    if ((threadIdx.x & 63) == 0) it.out[d->pix] += n_ok > 0 ? sum / n_ok : 0.0;
    cnt.flush(item_stats + (size_t)d->item * ST_COUNT);  // the wave's lanes share the item
  }
}

}  // namespace sdtfe

// ------------------------------------------------------------ host side

struct sdtfe_batch {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  sdtfe::BatchItem *d_items = nullptr;
  sdtfe::BatchHdr *d_hdr = nullptr;
  int *d_counters = nullptr;
  unsigned long long *d_stats = nullptr;
  sdtfe::BatchItem *h_items = nullptr;    // pinned
  sdtfe::BatchHdr *h_hdr = nullptr;       // pinned
  unsigned long long *h_stats = nullptr;  // pinned, the items' counters read back
  int cap = 0;
  // captured launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_mixed = false, graph_wide = false;
  int graph_blocks = 0;
  size_t graph_lds = 0;
  int blocks_fp64 = 0, blocks_mixed = 0, blocks_wide = 0;
  int cus = 0;  // compute units of the device
  sdtfe::Deferred *d_defer = nullptr;  // mixed walk: pixels for the FP64 pass
  size_t defer_cap = 0;
};

namespace sdtfe {
namespace {

void drop_graph(sdtfe_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int ensure_cap(sdtfe_batch *b, int n) {
  if (n <= b->cap) return 0;
  drop_graph(b);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_stats) (void)hipFree(b->d_stats);
  if (b->h_items) (void)hipHostFree(b->h_items);
  if (b->h_stats) (void)hipHostFree(b->h_stats);
  b->d_items = nullptr;
  b->d_stats = nullptr;
  b->h_items = nullptr;
  b->h_stats = nullptr;
  b->cap = 0;
  const int cap = std::max(n, 64);
  const size_t stat_bytes = sizeof(unsigned long long) * ST_COUNT * cap;
  if (hipMalloc((void **)&b->d_items, sizeof(BatchItem) * cap) != hipSuccess ||
      hipMalloc((void **)&b->d_stats, stat_bytes) != hipSuccess ||
      hipHostMalloc((void **)&b->h_items, sizeof(BatchItem) * cap, hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void **)&b->h_stats, stat_bytes, hipHostMallocDefault) != hipSuccess)
    return fail(SDTFE_ENOMEM, "batch buffers");
  b->cap = cap;
  return 0;
}

template <bool kMixed, bool kWide>
int blocks_for(sdtfe_batch *b, size_t lds) {
  int per_cu = 0;
  SDTFE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<kMixed, kWide>(),
                                                         256, lds));
  int dev = 0;
  SDTFE_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SDTFE_HIP(hipGetDeviceProperties(&prop, dev));
  b->cus = prop.multiProcessorCount;
  return std::max(1, per_cu) * prop.multiProcessorCount;
}

// Capture reset + persistent kernel on the batch's stream.
template <bool kMixed, bool kWide>
int capture(sdtfe_batch *b, int blocks, size_t lds) {
  drop_graph(b);
  SDTFE_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  batch_reset_kernel<<<64, 256, 0, b->stream>>>(b->d_hdr, b->d_counters, b->d_stats,
                                                ST_COUNT * b->cap);
  batch_kernel<kMixed, kWide>()<<<blocks, 256, lds, b->stream>>>(b->d_items, b->d_hdr,
                                                                b->d_counters, b->d_stats);
  if (kMixed)
    batch_deferred_kernel<kWide><<<1024, 256, 0, b->stream>>>(b->d_items, b->d_hdr,
                                                              b->d_counters, b->d_stats);
  hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
  if (e != hipSuccess) return hip_check(e, "hipStreamEndCapture");
  SDTFE_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_mixed = kMixed;
  b->graph_wide = kWide;
  b->graph_blocks = blocks;
  b->graph_lds = lds;
  return 0;
}

}  // namespace

}  // namespace sdtfe

namespace sdtfe {

int batch_reserve(sdtfe_batch *b, int n) { return ensure_cap(b, n); }

int batch_tiles(const sdtfe_map_params &p, int lanes_per_pixel) {
  int L = lanes_per_pixel;
  if (L != 1 && L != 2 && L != 4 && L != 8 && L != 16) L = 1;
  while (L > 1 && (size_t)(256 / L) * (size_t)std::max(1, p.n_mc_samp) * 8 > (size_t)kLdsBytes)
    L >>= 1;
  int bw, bh;
  wave_block(L, bw, bh);
  const int tx = (p.grid_dim + 2 * bw - 1) / (2 * bw), ty = (p.grid_dim + 2 * bh - 1) / (2 * bh);
  return tx * ty;
}

int batch_run_ranges(sdtfe_batch *b, int n, sdtfe_tess *const *tess,
                     const sdtfe_map_params *params, double *const *rho_device,
                     const int *lanes_per_pixel, const int *tile_begin, const int *tile_count,
                     sdtfe_stats *item_stats, sdtfe_stats *total, bool leave_room) {
  if (!b || n < 0 || (n && (!tess || !params || !rho_device)))
    return fail(SDTFE_EINVAL, "bad argument");
  if (n == 0) return 0;
  const int prec = params[0].precision;
  for (int i = 0; i < n; ++i) {
    if (!tess[i] || !rho_device[i]) return fail(SDTFE_EINVAL, "null item");
    if (params[i].precision != prec) return fail(SDTFE_EINVAL, "mixed precisions in one batch");
    if (prec == SDTFE_PREC_FP64 && !tess[i]->frec64)
      return fail(SDTFE_EINVAL, "batched FP64 maps take tessellations of fewer than 2^29 tets");
    if (params[i].grid_dim <= 0 || params[i].grid_dim > 65536 || params[i].n_mc_samp < 0 ||
        !(params[i].box_len > 0) || !(params[i].box_len_z > 0))
      return fail(SDTFE_EINVAL, "bad map parameters");
  }
  if (prec != SDTFE_PREC_FP64 && prec != SDTFE_PREC_MIXED)
    return fail(SDTFE_EINVAL, "unknown precision");
  // face records of all items addressed with 64-bit offsets when any item
  // needs them
  bool wide = false;
  if (prec == SDTFE_PREC_MIXED)
    for (int i = 0; i < n; ++i) wide |= wide_records(tess[i]->nt);
  int rc = ensure_cap(b, n);
  if (rc) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  // the tessellations' setup is complete: sdtfe_tess_create returns after
  // its stream has finished the setup (the stream itself may be a server
  // pool's, busy with other requests, and is not waited for here)

  int tiles = 0;
  for (int i = 0; i < n; ++i) {
    const sdtfe_map_params &p = params[i];
    BatchItem &it = b->h_items[i];
    std::memset(&it, 0, sizeof it);
    it.rec64 = tess[i]->rec64;
    it.rec32 = tess[i]->rec32;
    it.facerec = tess[i]->facerec;
    it.frec64 = tess[i]->frec64;
    it.grid = tess[i]->grid;
    it.geom = make_geom_pub(tess[i], p);
    it.mc = mixed_consts(tess[i], it.geom);
    it.out = rho_device[i];
    int L = lanes_per_pixel ? lanes_per_pixel[i] : 1;
    if (L != 1 && L != 2 && L != 4 && L != 8 && L != 16) L = 1;
    // per-sample values of a tile must fit in LDS
    while (L > 1 && (size_t)(256 / L) * (size_t)std::max(1, p.n_mc_samp) * 8 > (size_t)kLdsBytes)
      L >>= 1;
    int bw, bh;
    wave_block(L, bw, bh);
    it.lpp = L;
    it.tile_w = 2 * bw;
    it.tile_h = 2 * bh;
    it.tiles_x = (p.grid_dim + it.tile_w - 1) / it.tile_w;
    const int tiles_y = (p.grid_dim + it.tile_h - 1) / it.tile_h;
    const int all = it.tiles_x * tiles_y;
    const int b0 = tile_begin ? std::min(std::max(tile_begin[i], 0), all) : 0;
    const int nt = tile_count ? std::min(std::max(tile_count[i], 0), all - b0) : all - b0;
    it.tile_offset = b0;
    it.n_tiles = nt;
    it.first_tile = tiles;
    tiles += nt;
  }
  const bool mixed = prec == SDTFE_PREC_MIXED;
  if (mixed) {
    // room for every pixel of the launch: a batch accumulates, so its first
    // pass cannot be repeated
    size_t pixels = 0;
    for (int i = 0; i < n; ++i)
      pixels += (size_t)b->h_items[i].n_tiles * b->h_items[i].tile_w * b->h_items[i].tile_h;
    if (pixels > (size_t)INT_MAX) return fail(SDTFE_EINVAL, "batch too large for the mixed walk");
    if (pixels > b->defer_cap) {
      if (b->d_defer) (void)hipFree(b->d_defer);
      b->d_defer = nullptr;
      b->defer_cap = 0;
      const size_t cap = std::max<size_t>(pixels, 65536);
      if (hipMalloc((void **)&b->d_defer, sizeof(Deferred) * cap) != hipSuccess)
        return fail(SDTFE_ENOMEM, "deferred-pixel queue allocation failed");
      b->defer_cap = cap;
    }
  }
  BatchHdr &h = *b->h_hdr;
  h.dq = DeferQueue{b->d_defer, b->d_counters + kXcd, (int)std::min<size_t>(b->defer_cap, INT_MAX)};
  h.n_items = n;
  h.total_tiles = tiles;
  for (int x = 0; x <= kXcd; ++x) h.share[x] = (int)((int64_t)tiles * x / kXcd);

  int &blocks = !mixed ? b->blocks_fp64 : (wide ? b->blocks_wide : b->blocks_mixed);
  if (!blocks)
    blocks = !mixed ? blocks_for<false, false>(b, kLdsBytes)
                    : (wide ? blocks_for<true, true>(b, kLdsBytes)
                            : blocks_for<true, false>(b, kLdsBytes));
  if (blocks <= 0) return fail(SDTFE_EHIP, "occupancy query failed");
  h.active_blocks = leave_room && blocks > b->cus ? blocks - b->cus : blocks;
  // one LDS size for every launch, so one captured graph serves every batch
  if (!b->exec || b->graph_mixed != mixed || b->graph_wide != wide ||
      b->graph_blocks != blocks) {
    rc = !mixed ? capture<false, false>(b, blocks, kLdsBytes)
                : (wide ? capture<true, true>(b, blocks, kLdsBytes)
                        : capture<true, false>(b, blocks, kLdsBytes));
    if (rc) return rc;
  }
  SDTFE_HIP(hipMemcpyAsync(b->d_items, b->h_items, sizeof(BatchItem) * n, hipMemcpyHostToDevice,
                           b->stream));
  SDTFE_HIP(hipMemcpyAsync(b->d_hdr, b->h_hdr, sizeof(BatchHdr), hipMemcpyHostToDevice,
                           b->stream));
  SDTFE_HIP(hipEventRecord(b->ev0, b->stream));
  SDTFE_HIP(hipGraphLaunch(b->exec, b->stream));
  SDTFE_HIP(hipEventRecord(b->ev1, b->stream));
  // the items' counters follow the launch on its stream into pinned memory,
  // so one wait covers both (the null stream would order the copy against
  // other work of the process, such as a server's set-up threads)
  SDTFE_HIP(hipMemcpyAsync(b->h_stats, b->d_stats, sizeof(unsigned long long) * ST_COUNT * n,
                           hipMemcpyDeviceToHost, b->stream));
  SDTFE_HIP(hipStreamSynchronize(b->stream));
  float ms = 0;
  SDTFE_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  sdtfe_stats sum{};
  for (int i = 0; i < n; ++i) {
    sdtfe_stats s{};
    fill_stats_pub(b->h_stats + (size_t)i * ST_COUNT, &s);
    if (item_stats) item_stats[i] = s;
    sum.samples += s.samples;
    sum.missed += s.missed;
    sum.failed += s.failed;
    sum.no_intersection += s.no_intersection;
    sum.perturb_exhausted += s.perturb_exhausted;
    sum.no_isect_perturb += s.no_isect_perturb;
    sum.perturb_steps += s.perturb_steps;
    sum.tet_steps += s.tet_steps;
    sum.step_cap += s.step_cap;
    sum.fp64_redo += s.fp64_redo;
  }
  sum.t_walk_ms = ms;
  sum.t_total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (total) *total = sum;
  return 0;
}

}  // namespace sdtfe

using namespace sdtfe;

extern "C" {

int sdtfe_batch_create(sdtfe_batch **out) {
  if (!out) return fail(SDTFE_EINVAL, "null output handle");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SDTFE_ENODEV, "no HIP device visible");
  sdtfe_batch *b = new sdtfe_batch();
  int rc = hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
  if (!rc) rc = hip_check(hipEventCreate(&b->ev0), "event");
  if (!rc) rc = hip_check(hipEventCreate(&b->ev1), "event");
  if (!rc && (hipMalloc((void **)&b->d_hdr, sizeof(BatchHdr)) != hipSuccess ||
              hipMalloc((void **)&b->d_counters, sizeof(int) * (kXcd + 1)) != hipSuccess ||
              hipHostMalloc((void **)&b->h_hdr, sizeof(BatchHdr), hipHostMallocDefault) !=
                  hipSuccess))
    rc = fail(SDTFE_ENOMEM, "batch header");
  if (!rc) rc = ensure_cap(b, 64);
  if (rc) {
    sdtfe_batch_destroy(b);
    return rc;
  }
  *out = b;
  return 0;
}

void sdtfe_batch_destroy(sdtfe_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_stats) (void)hipFree(b->d_stats);
  if (b->d_hdr) (void)hipFree(b->d_hdr);
  if (b->d_counters) (void)hipFree(b->d_counters);
  if (b->d_defer) (void)hipFree(b->d_defer);
  if (b->h_items) (void)hipHostFree(b->h_items);
  if (b->h_hdr) (void)hipHostFree(b->h_hdr);
  if (b->h_stats) (void)hipHostFree(b->h_stats);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int sdtfe_batch_run(sdtfe_batch *b, int n, sdtfe_tess *const *tess,
                    const sdtfe_map_params *params, double *const *rho_device,
                    const int *lanes_per_pixel, sdtfe_stats *item_stats, sdtfe_stats *total) {
  return sdtfe::batch_run_ranges(b, n, tess, params, rho_device, lanes_per_pixel, nullptr,
                                 nullptr, item_stats, total);
}

}  // extern "C"
