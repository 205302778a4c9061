// capi.cpp -- extern "C" entry points of libsdtfe_amd.so (declared in
// include/sdtfe_amd.h).  Host C++ over the HIP runtime; no torch types.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sdtfe_host.h"
#include "walk_fp64_core.h"

namespace sdtfe {

static thread_local std::string g_err;
static thread_local int g_dropin_status = 0;

void set_dropin_status(int rc) { g_dropin_status = rc; }

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  return fail(SDTFE_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void launch_map_fp64(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate,
                     unsigned long long *stats);
void launch_trace_fp64(sdtfe_tess *t, const MapGeom &g, int n, const double *x, const double *y,
                       const int *entry, const uint64_t *pixel, const uint32_t *sample,
                       double *value, int *crossed);
int launch_map_mixed(sdtfe_tess *t, const MapGeom &g, double *out, int *navg, int accumulate,
                     int *h_count);
void launch_trace_mixed(sdtfe_tess *t, const MapGeom &g, int n, const double *x,
                        const double *y, const int *entry, const uint64_t *pixel,
                        const uint32_t *sample, double *value, int *crossed);
void launch_locate(sdtfe_tess *t, int n, const double *x, const double *y, int *entry);

extern const bool kHaveMixed;

namespace {

std::atomic<uint64_t> g_seed{0};

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

}  // namespace

// Map geometry exactly as src/dtfe.c:579-589 computes it.
MapGeom make_geom_pub(const sdtfe_tess *t, const sdtfe_map_params &p) {
  MapGeom g{};
  g.dxy = p.box_len / p.grid_dim;
  g.off_x = -(p.box_len / 2.0) + p.center[0];
  g.off_y = -(p.box_len / 2.0) + p.center[1];
  g.off_z = -(p.box_len_z / 2.0) + p.center[2];
  g.lz = p.box_len_z;
  g.scale = p.delta_sample > g.dxy ? p.delta_sample : g.dxy;
  g.soff = p.delta_sample > g.dxy ? (p.delta_sample - g.dxy) / 2.0 : 0.0;
  g.org_x = t->org[0];
  g.org_y = t->org[1];
  g.org_z = t->org[2];
  g.seed = p.seed;
  g.grid = p.grid_dim;
  g.samples = p.n_mc_samp;
  g.max_steps = p.max_steps > 0 ? p.max_steps : SDTFE_MAX_STEPS_DEFAULT;
  // samples fall in [off - soff, off - soff + (G - 1) dxy + scale]
  const double lo_x = g.off_x - g.soff, hi_x = lo_x + (p.grid_dim - 1) * g.dxy + g.scale;
  const double lo_y = g.off_y - g.soff, hi_y = lo_y + (p.grid_dim - 1) * g.dxy + g.scale;
  g.ray_span = std::max(std::max(std::fabs(lo_x - g.org_x), std::fabs(hi_x - g.org_x)),
                        std::max(std::fabs(lo_y - g.org_y), std::fabs(hi_y - g.org_y)));
  return g;
}

void fill_stats_pub(const unsigned long long *c, sdtfe_stats *s) {
  s->samples = (int64_t)c[ST_SAMPLES];
  s->missed = (int64_t)c[ST_MISSED];
  s->failed = (int64_t)c[ST_FAILED];
  s->no_intersection = (int64_t)c[ST_NO_ISECT];
  s->perturb_exhausted = (int64_t)c[ST_PERT_EXHAUSTED];
  s->no_isect_perturb = (int64_t)c[ST_NO_ISECT_PERT];
  s->perturb_steps = (int64_t)c[ST_PERT_STEPS];
  s->tet_steps = (int64_t)c[ST_TET_STEPS];
  s->step_cap = (int64_t)c[ST_STEP_CAP];
  s->fp64_redo = (int64_t)c[ST_FP64_REDO];
}

namespace {

int check_params(const sdtfe_map_params *p) {
  if (!p) return fail(SDTFE_EINVAL, "null map parameters");
  if (p->grid_dim <= 0 || p->grid_dim > 65536) return fail(SDTFE_EINVAL, "grid_dim out of range");
  if (p->n_mc_samp < 0) return fail(SDTFE_EINVAL, "n_mc_samp < 0");
  if (!(p->box_len > 0) || !(p->box_len_z > 0)) return fail(SDTFE_EINVAL, "box sizes must be > 0");
  if (p->precision != SDTFE_PREC_FP64 && p->precision != SDTFE_PREC_MIXED)
    return fail(SDTFE_EINVAL, "unknown precision");
  if (p->precision == SDTFE_PREC_MIXED && !kHaveMixed)
    return fail(SDTFE_EINVAL, "mixed-precision walk not built");
  return 0;
}

}  // namespace

bool wide_records(int64_t n_tetra) {
  if (n_tetra > kFaceRecNarrowTets) return true;
  const char *e = std::getenv("SDTFE_WIDE_RECORDS");
  return e && std::strcmp(e, "1") == 0;
}

namespace {

int check_volume_params(const sdtfe_volume_params *p) {
  if (!p) return fail(SDTFE_EINVAL, "null volume parameters");
  if (p->grid_dim <= 0 || p->grid_dim > 65536) return fail(SDTFE_EINVAL, "grid_dim out of range");
  if (p->supersampling == 0 || p->supersampling > 1024)
    return fail(SDTFE_EINVAL, "supersampling must be 1 .. 1024");  // src/dtfe.c:669
  if ((int64_t)p->grid_dim * (int64_t)p->supersampling > INT_MAX)
    return fail(SDTFE_EINVAL, "grid_dim * supersampling too large");
  if (!(p->box_len > 0)) return fail(SDTFE_EINVAL, "box_len must be > 0");
  return 0;
}

int run_volume(sdtfe_tess *t, const sdtfe_volume_params *p, double *grid, double *t_walk_ms) {
  const VolGeom v = make_vol_geom(*p);
  SDTFE_HIP(hipMemsetAsync(t->stats, 0, sizeof(unsigned long long) * ST_COUNT, t->stream));
  SDTFE_HIP(hipEventRecord(t->ev0, t->stream));
  SDTFE_HIP(hipEventRecord(t->evk0, t->stream));
  launch_volume_fp64(t, v, grid);
  SDTFE_HIP(hipGetLastError());
  SDTFE_HIP(hipEventRecord(t->evk1, t->stream));
  SDTFE_HIP(hipEventRecord(t->ev1, t->stream));
  if (t_walk_ms) {
    SDTFE_HIP(hipEventSynchronize(t->ev1));
    float ms = 0;
    SDTFE_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    *t_walk_ms = ms;
  }
  return 0;
}

int ensure_map_buf(sdtfe_tess *t, size_t npix) {
  if (npix <= t->map_cap) return 0;
  dev_free(t->map_buf, t->stream);
  dev_free(t->navg_buf, t->stream);
  t->map_buf = nullptr;
  t->navg_buf = nullptr;
  t->map_cap = 0;
  if (dev_malloc((void **)&t->map_buf, sizeof(double) * npix, t->stream) != hipSuccess ||
      dev_malloc((void **)&t->navg_buf, sizeof(int) * npix, t->stream) != hipSuccess)
    return fail(SDTFE_ENOMEM, "map buffer allocation failed");
  t->map_cap = npix;
  return 0;
}

// Pinned host memory for a call's small read-backs (a pageable destination
// costs the runtime a staged copy of its own): the first buffer of a pooled
// staging, or null when the pool could not pin memory.
void *pinned_words(const PooledStaging &ps) { return ps.st ? ps.st->buf[0] : nullptr; }

// The counters, read through host (pinned, ST_COUNT words) when given.
int read_stats(sdtfe_tess *t, sdtfe_stats *s, void *host = nullptr) {
  unsigned long long local[ST_COUNT];
  unsigned long long *c = host ? static_cast<unsigned long long *>(host) : local;
  SDTFE_HIP(hipMemcpyAsync(c, t->stats, sizeof local, hipMemcpyDeviceToHost, t->stream));
  SDTFE_HIP(hipStreamSynchronize(t->stream));
  if (s) fill_stats_pub(c, s);
  return 0;
}

template <typename T>
struct DevArray {
  T *p = nullptr;
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
  int upload(const T *h, size_t n, hipStream_t s) {
    if (hipMalloc((void **)&p, sizeof(T) * (n ? n : 1)) != hipSuccess)
      return fail(SDTFE_ENOMEM, "scratch allocation failed");
    if (h && n) SDTFE_HIP(hipMemcpyAsync(p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return 0;
  }
};

std::once_flag g_pool_kept[64];

int create(const double *pd, int n, const int *tetra, int nt, float p_mass, bool given_field,
           Staging *st, hipStream_t stream, sdtfe_tess **out, bool fp64_faces = true);

}  // namespace

hipError_t dev_malloc(void **p, size_t bytes, hipStream_t s) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::call_once(g_pool_kept[dev & 63], [dev] {
    // keep freed blocks in the pool: later tessellations reuse them
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  });
  return hipMallocAsync(p, bytes ? bytes : 1, s);
}

void dev_free(void *p, hipStream_t s) {
  if (p) (void)hipFreeAsync(p, s);
}

int staging_init(Staging &st, size_t bytes) {
  st.bytes = bytes;
  for (int k = 0; k < 2; ++k) {
    if (hipHostMalloc(&st.buf[k], bytes, hipHostMallocDefault) != hipSuccess)
      return fail(SDTFE_ENOMEM, "pinned staging allocation failed");
    if (hipEventCreateWithFlags(&st.ev[k], hipEventDisableTiming) != hipSuccess)
      return fail(SDTFE_EHIP, "staging event");
  }
  return 0;
}

void staging_free(Staging &st) {
  for (int k = 0; k < 2; ++k) {
    if (st.ev[k]) {
      (void)hipEventSynchronize(st.ev[k]);
      (void)hipEventDestroy(st.ev[k]);
    }
    if (st.buf[k]) (void)hipHostFree(st.buf[k]);
    st.ev[k] = nullptr;
    st.buf[k] = nullptr;
  }
  st.bytes = 0;  // uploads then go straight from pageable memory
}

namespace {

// Uploads of at least this many bytes are split into kUploadParts ranges,
// each copied through its own pinned buffers by its own host thread (the
// host copy into pinned memory, not the transfer, limits one thread).
constexpr size_t kParallelUpload = (size_t)64 << 20;
constexpr int kUploadParts = 4;

int upload_serial(void *dst, const void *src, size_t bytes, hipStream_t s, Staging *st);

int upload_parallel(void *dst, const void *src, size_t bytes, hipStream_t s, Staging *st) {
  std::vector<std::unique_ptr<PooledStaging>> more;
  for (int k = 1; k < kUploadParts; ++k) more.push_back(std::make_unique<PooledStaging>());
  int rc[kUploadParts] = {};
  std::string err[kUploadParts];
  int dev = 0;
  (void)hipGetDevice(&dev);
  auto part = [&](int k) {
    if (k > 0) (void)hipSetDevice(dev);  // a new thread starts on device 0
    const size_t lo = bytes * k / kUploadParts, hi = bytes * (k + 1) / kUploadParts;
    Staging *sk = k == 0 ? st : more[k - 1]->st;
    rc[k] = upload_serial(static_cast<char *>(dst) + lo, static_cast<const char *>(src) + lo,
                          hi - lo, s, sk);
    if (rc[k]) err[k] = g_err;  // the worker's message, for the caller's thread
  };
  std::vector<std::thread> th;
  for (int k = 1; k < kUploadParts; ++k) {
    try {
      th.emplace_back(part, k);
    } catch (const std::exception &) {
      part(k);  // no thread to be had: this thread copies the part
    }
  }
  part(0);
  for (std::thread &x : th) x.join();
  for (int k = 0; k < kUploadParts; ++k)
    if (rc[k]) return fail(rc[k], err[k]);
  return 0;
}

}  // namespace

int upload(void *dst, const void *src, size_t bytes, hipStream_t s, Staging *st) {
  if (!st || !st->bytes || !st->buf[0] || !st->buf[1]) {
    SDTFE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return 0;
  }
  if (bytes >= kParallelUpload) return upload_parallel(dst, src, bytes, s, st);
  return upload_serial(dst, src, bytes, s, st);
}

namespace {

int upload_serial(void *dst, const void *src, size_t bytes, hipStream_t s, Staging *st) {
  if (!st->bytes || !st->buf[0] || !st->buf[1]) {
    SDTFE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return 0;
  }
  for (size_t off = 0, k = 0; off < bytes; off += st->bytes, k ^= 1) {
    const size_t n = std::min(st->bytes, bytes - off);
    SDTFE_HIP(hipEventSynchronize(st->ev[k]));  // the buffer's previous copy is done
    std::memcpy(st->buf[k], static_cast<const char *>(src) + off, n);
    SDTFE_HIP(hipMemcpyAsync(static_cast<char *>(dst) + off, st->buf[k], n,
                             hipMemcpyHostToDevice, s));
    SDTFE_HIP(hipEventRecord(st->ev[k], s));
  }
  return 0;
}

}  // namespace

int download(const void *src, size_t bytes, hipStream_t s, Staging *st,
             const std::function<void(const char *, size_t, size_t)> &consume) {
  if (!bytes) return 0;
  if (!st || !st->bytes || !st->buf[0] || !st->buf[1]) {
    std::vector<char> tmp(bytes);
    SDTFE_HIP(hipMemcpyAsync(tmp.data(), src, bytes, hipMemcpyDeviceToHost, s));
    SDTFE_HIP(hipStreamSynchronize(s));
    consume(tmp.data(), 0, bytes);
    return 0;
  }
  const size_t slice = st->bytes, n = (bytes + slice - 1) / slice;
  auto issue = [&](size_t k) -> int {
    const size_t off = k * slice;
    SDTFE_HIP(hipEventSynchronize(st->ev[k & 1]));  // the buffer's previous copy is done
    SDTFE_HIP(hipMemcpyAsync(st->buf[k & 1], static_cast<const char *>(src) + off,
                             std::min(slice, bytes - off), hipMemcpyDeviceToHost, s));
    SDTFE_HIP(hipEventRecord(st->ev[k & 1], s));
    return 0;
  };
  if (int rc = issue(0)) return rc;
  for (size_t k = 0; k < n; ++k) {
    // the next slice goes into the buffer the host finished with last time
    if (k + 1 < n)
      if (int rc = issue(k + 1)) return rc;
    SDTFE_HIP(hipEventSynchronize(st->ev[k & 1]));
    const size_t off = k * slice;
    consume(static_cast<const char *>(st->buf[k & 1]), off, std::min(slice, bytes - off));
  }
  return 0;
}

namespace {

constexpr size_t kPoolSlice = (size_t)8 << 20;
std::mutex g_staging_mu;
std::vector<std::pair<int, Staging *>> g_staging_idle;  // (device, buffers)

}  // namespace

// Buffers and events of the device current when taken (events belong to a
// device).
PooledStaging::PooledStaging() : st(nullptr), dev(0) {
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> g(g_staging_mu);
    for (size_t i = g_staging_idle.size(); i-- > 0;)
      if (g_staging_idle[i].first == dev) {
        st = g_staging_idle[i].second;
        g_staging_idle.erase(g_staging_idle.begin() + (long)i);
        return;
      }
  }
  st = new Staging();
  const std::string keep = g_err;  // a failed pinned allocation is not the call's error
  if (staging_init(*st, kPoolSlice)) staging_free(*st);  // pageable copies then
  g_err = keep;
}

PooledStaging::~PooledStaging() {
  for (int k = 0; k < 2; ++k)
    if (st->ev[k]) (void)hipEventSynchronize(st->ev[k]);  // back to the pool idle
  std::lock_guard<std::mutex> g(g_staging_mu);
  g_staging_idle.emplace_back(dev, st);
}

int tess_create_staged(const double *particle_data, int n, const int *tetra_data, int nt,
                       float p_mass, Staging *st, hipStream_t stream, bool fp64_faces,
                       sdtfe_tess **out) {
  return create(particle_data, n, tetra_data, nt, p_mass, false, st, stream, out, fp64_faces);
}

void link_hull_rows(int *rows, int f) {
  // Neighbours across shared edges (src/triangulate.c:358-392): slot 5 for
  // edge (a, b), 4 for (a, c), 3 for (b, c); the reference compares all
  // pairs, so of several triangles on one edge the highest-numbered other
  // one is kept.  Here the edges are sorted by their vertex pair instead.
  struct Edge {
    int lo, hi, tri, slot;
  };
  std::vector<Edge> e;
  e.reserve(3 * (size_t)f);
  static const int pair[3][3] = {{0, 1, 5}, {0, 2, 4}, {1, 2, 3}};
  for (int k = 0; k < f; ++k)
    for (const auto &q : pair) {
      const int u = rows[8 * (size_t)k + q[0]], v = rows[8 * (size_t)k + q[1]];
      e.push_back({std::min(u, v), std::max(u, v), k, q[2]});
    }
  std::sort(e.begin(), e.end(), [](const Edge &x, const Edge &y) {
    return x.lo != y.lo ? x.lo < y.lo : (x.hi != y.hi ? x.hi < y.hi : x.tri < y.tri);
  });
  for (size_t s0 = 0; s0 < e.size();) {
    size_t s1 = s0;
    while (s1 < e.size() && e[s1].lo == e[s0].lo && e[s1].hi == e[s0].hi) ++s1;
    // sorted by tri within the group: the highest other is the last, or
    // the one before it for the last itself
    for (size_t a = s0; a < s1; ++a) {
      const int best = e[s1 - 1].tri != e[a].tri ? e[s1 - 1].tri
                                                  : (s1 - s0 > 1 ? e[s1 - 2].tri : -1);
      if (best >= 0 && best != e[a].tri) rows[8 * (size_t)e[a].tri + e[a].slot] = best;
    }
    s0 = s1;
  }
}

namespace {

void free_tess(sdtfe_tess *t) {
  if (!t) return;
  // pos4 .. frec64 and tri_xy .. tri_abc live in field_block and tri_block
  void *bufs[] = {t->field_block, t->tri_block, t->cell_off, t->cell_tri,    t->stats,
                  t->map_buf,     t->navg_buf,  t->defer_buf, t->defer_count, t->dq_off,
                  t->dq_tmp,      t->dq_val,    t->dq_ok};
  if (t->stream) {
    // The frees are stream-ordered: work still queued on the stream ends
    // first.  A stream of the caller's pool is not drained, so freeing never
    // waits for other work the caller has queued on it; its owner has
    // finished any work on the tessellation issued on other streams.
    if (t->own_stream) (void)hipStreamSynchronize(t->stream);
    for (void *b : bufs) dev_free(b, t->stream);
    if (t->own_stream) (void)hipStreamSynchronize(t->stream);
  }
  if (t->ev0) (void)hipEventDestroy(t->ev0);
  if (t->ev1) (void)hipEventDestroy(t->ev1);
  if (t->evk0) (void)hipEventDestroy(t->evk0);
  if (t->evk1) (void)hipEventDestroy(t->evk1);
  if (t->stream && t->own_stream) (void)hipStreamDestroy(t->stream);
  delete t;
}

int create(const double *pd, int n, const int *tetra, int nt, float p_mass, bool given_field,
           Staging *st, hipStream_t stream, sdtfe_tess **out, bool fp64_faces) {
  if (!out) return fail(SDTFE_EINVAL, "null output handle");
  *out = nullptr;
  if (n < 0 || nt < 0) return fail(SDTFE_EINVAL, "negative size");
  if ((n > 0 && !pd) || (nt > 0 && !tetra)) return fail(SDTFE_EINVAL, "null input array");
  // the device count once per process: the query costs a round trip to the
  // driver, and a server creates a tessellation per request
  static const int ndev = [] {
    int k = 0;
    return hipGetDeviceCount(&k) == hipSuccess ? k : 0;
  }();
  if (ndev == 0) return fail(SDTFE_ENODEV, "no HIP device visible");
  auto t0 = Clock::now();
  sdtfe_tess *t = new sdtfe_tess();
  t->n = n;
  t->nt = nt;
  t->p_mass = p_mass;
  t->given_field = given_field;
  t->fp64_faces = fp64_faces;
  int rc = 0;
  if (stream) {
    t->stream = stream;  // the caller's; it outlives the tessellation
    t->own_stream = false;
  } else {
    rc = hip_check(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking), "stream");
  }
  if (!rc) rc = hip_check(hipEventCreate(&t->ev0), "event");
  if (!rc) rc = hip_check(hipEventCreate(&t->ev1), "event");
  if (!rc) rc = hip_check(hipEventCreate(&t->evk0), "event");
  if (!rc) rc = hip_check(hipEventCreate(&t->evk1), "event");
  if (!rc && dev_malloc((void **)&t->stats, sizeof(unsigned long long) * ST_COUNT, t->stream) !=
                 hipSuccess)
    rc = fail(SDTFE_ENOMEM, "stats allocation failed");
  std::unique_ptr<PooledStaging> pooled;  // when the caller brings no staging
  if (!rc && !st) {
    pooled = std::make_unique<PooledStaging>();
    st = pooled->st;
  }
  if (!rc) rc = build_field(t, pd, tetra, st);
  if (!rc) rc = build_entry_grid(t, st);
  if (rc) {
    free_tess(t);
    return rc;
  }
  t->t_setup_ms = ms_since(t0);
  *out = t;
  return 0;
}

int run_map(sdtfe_tess *t, const sdtfe_map_params *p, double *out, int *navg, int accumulate,
            double *t_walk_ms, double *t_kernel_ms, void *pinned) {
  const MapGeom g = make_geom_pub(t, *p);
  SDTFE_HIP(hipMemsetAsync(t->stats, 0, sizeof(unsigned long long) * ST_COUNT, t->stream));
  SDTFE_HIP(hipEventRecord(t->ev0, t->stream));
  if (p->precision == SDTFE_PREC_MIXED) {
    const int rc = launch_map_mixed(t, g, out, navg, accumulate, static_cast<int *>(pinned));
    if (rc) return rc;
  } else {
    launch_map_fp64(t, g, out, navg, accumulate, t->stats);
  }
  SDTFE_HIP(hipGetLastError());
  SDTFE_HIP(hipEventRecord(t->ev1, t->stream));
  if (t_walk_ms) {
    SDTFE_HIP(hipEventSynchronize(t->ev1));
    float ms = 0;
    SDTFE_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    *t_walk_ms = ms;
    SDTFE_HIP(hipEventElapsedTime(&ms, t->evk0, t->evk1));
    *t_kernel_ms = ms;
  }
  return 0;
}

}  // namespace
}  // namespace sdtfe

using namespace sdtfe;

namespace {

// Column 3 of the particle rows (x, y, z, rho) into a contiguous array.
__global__ void density_column_kernel(const double *pos4, int n, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = pos4[4 * (size_t)i + 3];
}

}  // namespace

extern "C" {

const char *sdtfe_last_error(void) { return g_err.c_str(); }
int sdtfe_dropin_status(void) { return g_dropin_status; }
void sdtfe_set_seed(uint64_t seed) { g_seed.store(seed); }
uint64_t sdtfe_get_seed(void) { return g_seed.load(); }

int sdtfe_tess_create(const double *particle_data, int n_particles, const int *tetra_data,
                      int n_tetra, float p_mass, sdtfe_tess **out) {
  return create(particle_data, n_particles, tetra_data, n_tetra, p_mass, false, nullptr, nullptr,
                out);
}

int sdtfe_tess_create_field(const double *particle_data, int n_particles, const int *tetra_data,
                            int n_tetra, sdtfe_tess **out) {
  return create(particle_data, n_particles, tetra_data, n_tetra, 1.0f, true, nullptr, nullptr,
                out);
}

void sdtfe_tess_destroy(sdtfe_tess *t) { free_tess(t); }

int sdtfe_tess_get_info(const sdtfe_tess *t, sdtfe_tess_info *info) {
  if (!t || !info) return fail(SDTFE_EINVAL, "null argument");
  info->n_particles = t->n;
  info->n_tetra = t->nt;
  info->n_tri = t->n_tri;
  info->grid_nx = t->grid.nx;
  info->grid_ny = t->grid.ny;
  info->device_bytes = t->device_bytes;
  info->t_setup_ms = t->t_setup_ms;
  return 0;
}

int sdtfe_tess_vertex_density(const sdtfe_tess *t, double *rho) {
  if (!t || (!rho && t->n)) return fail(SDTFE_EINVAL, "null argument");
  if (t->n == 0) return 0;
  // the density column alone goes back (a quarter of the particle rows)
  double *col = nullptr;
  if (dev_malloc((void **)&col, sizeof(double) * (size_t)t->n, t->stream) != hipSuccess)
    return fail(SDTFE_ENOMEM, "density read-back buffer allocation failed");
  density_column_kernel<<<(unsigned)((t->n + 255) / 256), 256, 0, t->stream>>>(t->pos4, t->n,
                                                                              col);
  int rc = hip_check(hipGetLastError(), "density_column_kernel");
  if (!rc) {
    PooledStaging ps;
    rc = download(col, sizeof(double) * (size_t)t->n, t->stream, ps.st,
                  [rho](const char *h, size_t off, size_t n) {
                    std::memcpy(reinterpret_cast<char *>(rho) + off, h, n);
                  });
  }
  dev_free(col, t->stream);
  return rc;
}

int sdtfe_tess_gradients(const sdtfe_tess *t, double *nabla) {
  if (!t || (!nabla && t->nt)) return fail(SDTFE_EINVAL, "null argument");
  if (t->nt == 0) return 0;
  PooledStaging ps;
  return download(t->nabla, sizeof(double) * 3 * (size_t)t->nt, t->stream, ps.st,
                  [nabla](const char *h, size_t off, size_t n) {
                    std::memcpy(reinterpret_cast<char *>(nabla) + off, h, n);
                  });
}

int sdtfe_tess_triangles(const sdtfe_tess *t, int *rows, int capacity) {
  if (!t) return fail(SDTFE_EINVAL, "null tessellation");
  if (capacity < t->n_tri || (!rows && t->n_tri)) return fail(SDTFE_EINVAL, "capacity too small");
  const int f = t->n_tri;
  if (f == 0) return 0;
  std::vector<int> abc(3 * (size_t)f), tet((size_t)f);
  SDTFE_HIP(hipMemcpyAsync(abc.data(), t->tri_abc, sizeof(int) * abc.size(),
                           hipMemcpyDeviceToHost, t->stream));
  SDTFE_HIP(hipMemcpyAsync(tet.data(), t->tri_tet, sizeof(int) * tet.size(),
                           hipMemcpyDeviceToHost, t->stream));
  SDTFE_HIP(hipStreamSynchronize(t->stream));
  for (int k = 0; k < f; ++k) {
    int *r = rows + 8 * (size_t)k;
    r[0] = abc[3 * (size_t)k];
    r[1] = abc[3 * (size_t)k + 1];
    r[2] = abc[3 * (size_t)k + 2];
    r[3] = r[4] = r[5] = -1;
    r[6] = tet[k];
    r[7] = 0;
  }
  link_hull_rows(rows, f);
  return 0;
}

int sdtfe_tess_walk_records(const sdtfe_tess *t, void *records, size_t bytes) {
  if (!t || bytes < sizeof(Rec64) * (size_t)t->nt || (!records && t->nt))
    return fail(SDTFE_EINVAL, "bad argument");
  if (t->nt == 0) return 0;
  SDTFE_HIP(hipMemcpyAsync(records, t->rec64, sizeof(Rec64) * (size_t)t->nt,
                           hipMemcpyDeviceToHost, t->stream));
  SDTFE_HIP(hipStreamSynchronize(t->stream));
  return 0;
}

int sdtfe_tess_locate(sdtfe_tess *t, int n, const double *x, const double *y, int *entry) {
  if (!t || n < 0 || (n && (!x || !y || !entry))) return fail(SDTFE_EINVAL, "bad argument");
  if (n == 0) return 0;
  DevArray<double> dx, dy;
  DevArray<int> de;
  int rc;
  if ((rc = dx.upload(x, n, t->stream)) || (rc = dy.upload(y, n, t->stream)) ||
      (rc = de.upload(nullptr, n, t->stream)))
    return rc;
  launch_locate(t, n, dx.p, dy.p, de.p);
  SDTFE_HIP(hipGetLastError());
  SDTFE_HIP(hipMemcpyAsync(entry, de.p, sizeof(int) * n, hipMemcpyDeviceToHost, t->stream));
  SDTFE_HIP(hipStreamSynchronize(t->stream));
  return 0;
}

int sdtfe_tess_trace(sdtfe_tess *t, int precision, int n, const double *x, const double *y,
                     const int *entry, const uint64_t *pixel, const uint32_t *sample,
                     double off_z, double lz, double delta_xy, uint64_t seed, double *value,
                     int *crossed, sdtfe_stats *stats) {
  if (!t || n < 0 || (n && (!x || !y || !pixel || !sample || !value || !crossed)))
    return fail(SDTFE_EINVAL, "bad argument");
  if (precision != SDTFE_PREC_FP64 && precision != SDTFE_PREC_MIXED)
    return fail(SDTFE_EINVAL, "unknown precision");
  auto t0 = Clock::now();
  MapGeom g{};
  g.dxy = delta_xy;
  g.off_z = off_z;
  g.lz = lz;
  g.org_x = t->org[0];
  g.org_y = t->org[1];
  g.org_z = t->org[2];
  g.seed = seed;
  g.max_steps = SDTFE_MAX_STEPS_DEFAULT;
  g.ray_span = 0;
  for (int i = 0; i < n; ++i)
    g.ray_span = std::max(g.ray_span, std::max(std::fabs(x[i] - g.org_x), std::fabs(y[i] - g.org_y)));
  SDTFE_HIP(hipMemsetAsync(t->stats, 0, sizeof(unsigned long long) * ST_COUNT, t->stream));
  if (n > 0) {
    DevArray<double> dx, dy, dv;
    DevArray<int> de, dc;
    DevArray<uint64_t> dp;
    DevArray<uint32_t> ds;
    int rc;
    if ((rc = dx.upload(x, n, t->stream)) || (rc = dy.upload(y, n, t->stream)) ||
        (rc = dp.upload(pixel, n, t->stream)) || (rc = ds.upload(sample, n, t->stream)) ||
        (rc = dv.upload(nullptr, n, t->stream)) || (rc = dc.upload(nullptr, n, t->stream)))
      return rc;
    if (entry && (rc = de.upload(entry, n, t->stream))) return rc;
    if (precision == SDTFE_PREC_MIXED)
      launch_trace_mixed(t, g, n, dx.p, dy.p, entry ? de.p : nullptr, dp.p, ds.p, dv.p, dc.p);
    else
      launch_trace_fp64(t, g, n, dx.p, dy.p, entry ? de.p : nullptr, dp.p, ds.p, dv.p, dc.p);
    SDTFE_HIP(hipGetLastError());
    SDTFE_HIP(hipMemcpyAsync(value, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost, t->stream));
    SDTFE_HIP(hipMemcpyAsync(crossed, dc.p, sizeof(int) * n, hipMemcpyDeviceToHost, t->stream));
    SDTFE_HIP(hipStreamSynchronize(t->stream));
  }
  int rc = read_stats(t, stats);
  if (stats) stats->t_total_ms = ms_since(t0);
  return rc;
}

int sdtfe_map(sdtfe_tess *t, const sdtfe_map_params *p, double *rho, int *n_avg,
              sdtfe_stats *stats) {
  if (!t) return fail(SDTFE_EINVAL, "null tessellation");
  int rc = check_params(p);
  if (rc) return rc;
  if (!rho) return fail(SDTFE_EINVAL, "null rho");
  auto t0 = Clock::now();
  const size_t npix = (size_t)p->grid_dim * (size_t)p->grid_dim;
  if ((rc = ensure_map_buf(t, npix))) return rc;
  double t_walk = 0, t_kernel = 0;
  {
    PooledStaging ps;
    void *words = pinned_words(ps);
    if ((rc = run_map(t, p, t->map_buf, t->navg_buf, 0, &t_walk, &t_kernel, words))) return rc;
    rc = download(t->map_buf, sizeof(double) * npix, t->stream, ps.st,
                  [rho](const char *h, size_t off, size_t n) {
                    const double *v = reinterpret_cast<const double *>(h);
                    double *r = rho + off / sizeof(double);
                    for (size_t i = 0; i < n / sizeof(double); ++i) r[i] += v[i];  // src/dtfe.c:642-645
                  });
    if (!rc && n_avg)
      rc = download(t->navg_buf, sizeof(int) * npix, t->stream, ps.st,
                    [n_avg](const char *h, size_t off, size_t n) {
                      std::memcpy(reinterpret_cast<char *>(n_avg) + off, h, n);
                    });
    if (!rc) rc = read_stats(t, stats, words);  // the downloads have finished with words
    if (rc) return rc;
  }
  if (stats) {
    stats->t_walk_ms = t_walk;
    stats->t_kernel_ms = t_kernel;
    stats->t_total_ms = ms_since(t0);
  }
  return 0;
}

int sdtfe_map_device(sdtfe_tess *t, const sdtfe_map_params *p, double *rho_device,
                     sdtfe_stats *stats) {
  if (!t) return fail(SDTFE_EINVAL, "null tessellation");
  int rc = check_params(p);
  if (rc) return rc;
  if (!rho_device) return fail(SDTFE_EINVAL, "null rho");
  auto t0 = Clock::now();
  double t_walk = 0, t_kernel = 0;
  PooledStaging ps;
  void *words = pinned_words(ps);
  if ((rc = run_map(t, p, rho_device, nullptr, 1, stats ? &t_walk : nullptr, &t_kernel, words)))
    return rc;
  if (stats) {
    if ((rc = read_stats(t, stats, words))) return rc;
    stats->t_walk_ms = t_walk;
    stats->t_kernel_ms = t_kernel;
    stats->t_total_ms = ms_since(t0);
  }
  return 0;
}

int sdtfe_map_device_noalloc(sdtfe_tess *t, const sdtfe_map_params *p, double *rho_device,
                             sdtfe_stats *stats) {
  if (!t) return fail(SDTFE_EINVAL, "null tessellation");
  int rc = check_params(p);
  if (rc) return rc;
  if (!rho_device) return fail(SDTFE_EINVAL, "null rho");
  const size_t npix = (size_t)p->grid_dim * (size_t)p->grid_dim;
  if (p->precision == SDTFE_PREC_MIXED && npix > kNoAllocPixels)
    return fail(SDTFE_EINVAL, "sdtfe_map_device_noalloc: a mixed map holds at most " +
                                  std::to_string(kNoAllocPixels) + " pixels");
  auto t0 = Clock::now();
  MixedWork w;
  if ((rc = noalloc_work(&w))) return rc;
  const MapGeom g = make_geom_pub(t, *p);
  if ((rc = run_map_noalloc(t, g, p->precision, rho_device, w))) return rc;
  if (stats) {
    unsigned long long c[ST_COUNT];
    SDTFE_HIP(hipMemcpyAsync(c, w.stats, sizeof c, hipMemcpyDeviceToHost, t->stream));
    SDTFE_HIP(hipStreamSynchronize(t->stream));
    fill_stats_pub(c, stats);
    float ms = 0;
    SDTFE_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    stats->t_walk_ms = ms;
    SDTFE_HIP(hipEventElapsedTime(&ms, t->evk0, t->evk1));
    stats->t_kernel_ms = ms;
    stats->t_total_ms = ms_since(t0);
  }
  return 0;
}

int sdtfe_volume(sdtfe_tess *t, const sdtfe_volume_params *p, double *grid,
                 sdtfe_stats *stats) {
  if (!t) return fail(SDTFE_EINVAL, "null tessellation");
  int rc = check_volume_params(p);
  if (rc) return rc;
  if (!grid) return fail(SDTFE_EINVAL, "null grid");
  auto t0 = Clock::now();
  const size_t cells = (size_t)p->grid_dim * (size_t)p->grid_dim * (size_t)p->grid_dim;
  double *d = nullptr;
  if (dev_malloc((void **)&d, sizeof(double) * cells, t->stream) != hipSuccess)
    return fail(SDTFE_ENOMEM, "volumetric grid allocation failed");
  // the walk adds into the caller's values in the reference's order, so
  // they go up first
  double t_walk = 0;
  {
    PooledStaging ps;
    rc = upload(d, grid, sizeof(double) * cells, t->stream, ps.st);
    if (!rc) rc = run_volume(t, p, d, &t_walk);
    if (!rc)
      rc = download(d, sizeof(double) * cells, t->stream, ps.st,
                    [grid](const char *h, size_t off, size_t n) {
                      std::memcpy(reinterpret_cast<char *>(grid) + off, h, n);
                    });
  }
  dev_free(d, t->stream);
  if (!rc) rc = hip_check(hipStreamSynchronize(t->stream), "sync");
  if (rc) return rc;
  if ((rc = read_stats(t, stats))) return rc;
  if (stats) {
    stats->t_walk_ms = t_walk;
    stats->t_kernel_ms = t_walk;
    stats->t_total_ms = ms_since(t0);
  }
  return 0;
}

int sdtfe_volume_device(sdtfe_tess *t, const sdtfe_volume_params *p, double *grid_device,
                        sdtfe_stats *stats) {
  if (!t) return fail(SDTFE_EINVAL, "null tessellation");
  int rc = check_volume_params(p);
  if (rc) return rc;
  if (!grid_device) return fail(SDTFE_EINVAL, "null grid");
  auto t0 = Clock::now();
  double t_walk = 0;
  if ((rc = run_volume(t, p, grid_device, stats ? &t_walk : nullptr))) return rc;
  if (stats) {
    if ((rc = read_stats(t, stats))) return rc;
    stats->t_walk_ms = t_walk;
    stats->t_kernel_ms = t_walk;
    stats->t_total_ms = ms_since(t0);
  }
  return 0;
}

int sdtfe_tetra_from_simplices(const double *particle_data, int n_particles,
                               const int *simplices, int n_tetra, int *tetra_data) {
  return tetra_from_simplices(particle_data, n_particles, simplices, n_tetra, tetra_data);
}

int sdtfe_tess_sync(sdtfe_tess *t) {
  if (!t) return fail(SDTFE_EINVAL, "null tessellation");
  SDTFE_HIP(hipStreamSynchronize(t->stream));
  return 0;
}

void *sdtfe_device_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    fail(SDTFE_ENOMEM, "hipMalloc failed");
    return nullptr;
  }
  return p;
}
int sdtfe_device_free(void *p) {
  SDTFE_HIP(hipFree(p));
  return 0;
}
int sdtfe_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  SDTFE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}
int sdtfe_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  SDTFE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  // from pageable memory the call may return before the device copy lands
  SDTFE_HIP(hipStreamSynchronize(nullptr));
  return 0;
}
int sdtfe_memset_d(void *p, int value, size_t bytes) {
  // waits for the fill: tessellation streams do not order against the
  // null stream, so a later launch on one must not overtake it
  SDTFE_HIP(hipMemsetAsync(p, value, bytes, nullptr));
  SDTFE_HIP(hipStreamSynchronize(nullptr));
  return 0;
}
int sdtfe_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// ------------------------------------------------------------ drop-in

static int dropin_precision() {
  const char *e = std::getenv("SDTFE_PRECISION");
  return (e && std::strcmp(e, "mixed") == 0) ? SDTFE_PREC_MIXED : SDTFE_PREC_FP64;
}

void compute_density(double *particle_data, int n_particles, int *tetra_data, int n_tetra,
                     int grid_dim, double box_len, double box_len_z, double *rho, float p_mass,
                     double center_x, double center_y, double center_z, const int n_mc_samp,
                     const double delta_sample) {
  sdtfe_tess *t = nullptr;
  const int precision = dropin_precision();
  // a mixed map never walks the FP64 face records: they are not built
  int rc = tess_create_staged(particle_data, n_particles, tetra_data, n_tetra, p_mass, nullptr,
                              nullptr, precision == SDTFE_PREC_FP64, &t);
  if (rc == 0) {
    sdtfe_map_params p{};
    p.grid_dim = grid_dim;
    p.n_mc_samp = n_mc_samp;
    p.box_len = box_len;
    p.box_len_z = box_len_z;
    p.center[0] = center_x;
    p.center[1] = center_y;
    p.center[2] = center_z;
    p.delta_sample = delta_sample;
    p.seed = g_seed.fetch_add(1);
    p.precision = precision;
    rc = sdtfe_map(t, &p, rho, nullptr, nullptr);
    if (rc == 0 && n_particles > 0) {
      std::vector<double> r((size_t)n_particles);
      rc = sdtfe_tess_vertex_density(t, r.data());
      if (rc == 0)
        for (int i = 0; i < n_particles; ++i) particle_data[4 * (size_t)i + 3] = r[i];
    }
  }
  set_dropin_status(rc);
  if (rc) std::fprintf(stderr, "compute_density (sdtfe_amd): %s\n", g_err.c_str());
  sdtfe_tess_destroy(t);
}

void pre_compute_vol(double **particle_data, int n_particles, int *tetra_data, int n_tetra,
                     float p_mass) {
  if (!particle_data) return;
  sdtfe_tess *t = nullptr;
  int rc = sdtfe_tess_create(*particle_data, n_particles, tetra_data, n_tetra, p_mass, &t);
  if (rc == 0 && n_particles > 0) {
    std::vector<double> r((size_t)n_particles);
    rc = sdtfe_tess_vertex_density(t, r.data());
    if (rc == 0)
      for (int i = 0; i < n_particles; ++i) (*particle_data)[4 * (size_t)i + 3] = r[i];
  }
  set_dropin_status(rc);
  if (rc) std::fprintf(stderr, "pre_compute_vol (sdtfe_amd): %s\n", g_err.c_str());
  sdtfe_tess_destroy(t);
}

static int dropin_volume(sdtfe_tess *t, int grid_dim, double box_len, double *grid,
                         double center_x, double center_y, double center_z,
                         unsigned supersampling) {
  sdtfe_volume_params p{};
  p.grid_dim = grid_dim;
  p.supersampling = supersampling;
  p.box_len = box_len;
  p.center[0] = center_x;
  p.center[1] = center_y;
  p.center[2] = center_z;
  p.seed = g_seed.fetch_add(1);
  return sdtfe_volume(t, &p, grid, nullptr);
}

void interpolate_to_3d_grid(double *particle_data, int n_particles, int *tetra_data,
                            int n_tetra, int grid_dim, double box_len, double *grid,
                            double center_x, double center_y, double center_z,
                            unsigned supersampling) {
  sdtfe_tess *t = nullptr;
  int rc = sdtfe_tess_create_field(particle_data, n_particles, tetra_data, n_tetra, &t);
  if (rc == 0)
    rc = dropin_volume(t, grid_dim, box_len, grid, center_x, center_y, center_z,
                       supersampling);
  set_dropin_status(rc);
  if (rc) std::fprintf(stderr, "interpolate_to_3d_grid (sdtfe_amd): %s\n", g_err.c_str());
  sdtfe_tess_destroy(t);
}

void compute_3d_density(double *particle_data, int n_particles, int *tetra_data, int n_tetra,
                        int grid_dim, double box_len, double *rho, float p_mass,
                        double center_x, double center_y, double center_z,
                        unsigned supersampling) {
  sdtfe_tess *t = nullptr;
  int rc = sdtfe_tess_create(particle_data, n_particles, tetra_data, n_tetra, p_mass, &t);
  if (rc == 0 && n_particles > 0) {
    std::vector<double> r((size_t)n_particles);
    rc = sdtfe_tess_vertex_density(t, r.data());
    if (rc == 0)
      for (int i = 0; i < n_particles; ++i) particle_data[4 * (size_t)i + 3] = r[i];
  }
  if (rc == 0)
    rc = dropin_volume(t, grid_dim, box_len, rho, center_x, center_y, center_z, supersampling);
  set_dropin_status(rc);
  if (rc) std::fprintf(stderr, "compute_3d_density (sdtfe_amd): %s\n", g_err.c_str());
  sdtfe_tess_destroy(t);
}

double *gradients(int *tetra_data, int n_tetra, double *particle_data) {
  double *nabla = (double *)std::malloc(sizeof(double) * 3 * (size_t)(n_tetra > 0 ? n_tetra : 1));
  if (!nabla) {
    set_dropin_status(fail(SDTFE_ENOMEM, "gradients: host allocation failed"));
    return nullptr;
  }
  // vertex count: the largest index the tets name
  int n = 0;
  for (int64_t i = 0; i < 4 * (int64_t)n_tetra; ++i) {
    const int v = tetra_data[8 * (i / 4) + i % 4];
    if (v + 1 > n) n = v + 1;
  }
  sdtfe_tess *t = nullptr;
  int rc = sdtfe_tess_create_field(particle_data, n, tetra_data, n_tetra, &t);
  if (rc == 0) rc = sdtfe_tess_gradients(t, nabla);
  set_dropin_status(rc);
  if (rc) {
    std::fprintf(stderr, "gradients (sdtfe_amd): %s\n", g_err.c_str());
    std::free(nabla);
    nabla = nullptr;
  }
  sdtfe_tess_destroy(t);
  return nabla;
}

// src/triangulate.h:28 (src/triangulate.c:263-316): sets bit 4 t + f of
// bitmap (bitmap.h layout: bit f of word t) for every hull face f of tet t
// (neighbour -1) whose (t1 x t2)_z < 0, and returns their count; the test
// runs on the GPU.  Bits already set stay set, as set_bit leaves them.
int find_forward_hull(int *tetra_data, int n_tetra, double *particle_data, int n_particles,
                      uint32_t *bitmap) {
  int rc = 0, count = 0;
  if (n_tetra < 0 || n_particles < 0 || (n_tetra > 0 && (!tetra_data || !bitmap)) ||
      (n_particles > 0 && !particle_data))
    rc = fail(SDTFE_EINVAL, "find_forward_hull: bad argument");
  std::vector<unsigned char> mask(n_tetra > 0 ? (size_t)n_tetra : 0);
  if (!rc) rc = hull_mask_host(particle_data, n_particles, tetra_data, n_tetra, mask.data());
  if (!rc)
    for (int t = 0; t < n_tetra; ++t) {
      bitmap[t] |= mask[t];
      count += __builtin_popcount(mask[t]);
    }
  set_dropin_status(rc);
  if (rc) std::fprintf(stderr, "find_forward_hull (sdtfe_amd): %s\n", g_err.c_str());
  return count;
}

// src/triangulate.h:29 (src/triangulate.c:318-395): a malloc'd n_tri x 8
// array (free() it) of the faces bitmap marks, in tet then face order:
// (a, b, c, n_CB, n_AC, n_BA, tet, 0) with the face's vertices in the
// reference's order per face and the triangle across each edge (-1 where
// none).  The rows are emitted on the GPU and linked on the host.  A bitmap
// marking more than n_tri faces is an error (the reference writes past its
// array); rows beyond the marked faces are zero.
int *copy_fh_triangles(int *tetra_data, int n_tetra, double *particle_data, int n_particles,
                       int n_tri, uint32_t *bitmap) {
  int rc = 0;
  int *rows = (int *)std::calloc(8 * (size_t)(n_tri > 0 ? n_tri : 1), sizeof(int));
  if (!rows) rc = fail(SDTFE_ENOMEM, "copy_fh_triangles: host allocation failed");
  if (!rc && (n_tetra < 0 || n_particles < 0 || n_tri < 0 ||
              (n_tetra > 0 && (!tetra_data || !bitmap)) || (n_particles > 0 && !particle_data)))
    rc = fail(SDTFE_EINVAL, "copy_fh_triangles: bad argument");
  int marked = 0;
  std::vector<unsigned char> mask(!rc && n_tetra > 0 ? (size_t)n_tetra : 0);
  if (!rc)
    for (int t = 0; t < n_tetra; ++t) {
      mask[t] = (unsigned char)(bitmap[t] & 15u);
      marked += __builtin_popcount(mask[t]);
    }
  if (!rc && marked > n_tri)
    rc = fail(SDTFE_EINVAL, "copy_fh_triangles: the bitmap marks more than n_tri faces");
  if (!rc && marked > 0) {
    std::vector<int> abc(3 * (size_t)marked), tet((size_t)marked);
    rc = hull_rows_host(particle_data, n_particles, tetra_data, n_tetra, mask.data(), abc.data(),
                        tet.data(), marked);
    if (!rc) {
      for (int k = 0; k < marked; ++k) {
        int *r = rows + 8 * (size_t)k;
        r[0] = abc[3 * (size_t)k];
        r[1] = abc[3 * (size_t)k + 1];
        r[2] = abc[3 * (size_t)k + 2];
        r[3] = r[4] = r[5] = -1;
        r[6] = tet[k];
        r[7] = 0;
      }
      link_hull_rows(rows, marked);
    }
  }
  set_dropin_status(rc);
  if (rc) std::fprintf(stderr, "copy_fh_triangles (sdtfe_amd): %s\n", g_err.c_str());
  return rows;
}

}  // extern "C"
