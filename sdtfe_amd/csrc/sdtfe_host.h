// sdtfe_host.h -- host-side state and launch functions of libsdtfe_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>

#include "../../include/sdtfe_amd.h"
#include "sdtfe_internal.h"

namespace sdtfe {

// Sets the calling thread's last error; returns code.
int fail(int code, const std::string &msg);
// Maps a HIP status to SDTFE_EHIP with context; 0 if ok.
int hip_check(hipError_t e, const char *what);

#define SDTFE_HIP(call)                                     \
  do {                                                      \
    int _rc = ::sdtfe::hip_check((call), #call);            \
    if (_rc) return _rc;                                    \
  } while (0)

// Device-wide exclusive scan of n ints (n < 2^28); in and out may alias.
// tmp needs scan_tmp_ints(n) ints.  Returns the total through *total_host
// when non-null (synchronises the stream).
size_t scan_tmp_ints(int64_t n);
int scan_exclusive(const int *in, int *out, int64_t n, int *tmp, int *total_host,
                   hipStream_t s);
// The total of a scan of n > 0 items in device memory (inside tmp), for a
// caller that reads it back with other values.
const int *scan_total_device(const int *tmp, int64_t n);

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

}  // namespace sdtfe

// Device-resident tessellation (opaque to callers).
struct sdtfe_tess {
  int n = 0, nt = 0, n_tri = 0;
  float p_mass = 1.f;
  bool given_field = false;  // vertex values taken from particle_data[:,3]
  bool fp64_faces = true;    // build the FP64 walk's face records
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: a stream of the caller's pool (the server's)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evk0 = nullptr, evk1 = nullptr;  // around the walk kernel proper
  // particles and tets as uploaded, plus the field; pos4 .. frec64 are
  // arrays of field_block, one allocation
  void *field_block = nullptr;
  double *pos4 = nullptr;     // N x 4 (x, y, z, rho)
  int *tets = nullptr;        // T x 8
  double *vol = nullptr;      // T
  double *nabla = nullptr;    // T x 3
  sdtfe::Rec64 *rec64 = nullptr;
  sdtfe::Rec32 *rec32 = nullptr;
  sdtfe::FaceRec *facerec = nullptr;  // 4 per tet
  sdtfe::FRec64 *frec64 = nullptr;    // 4 per tet; null beyond 2^29 tets or when not built
  double org[3] = {0, 0, 0};  // centre of the particles' bounding box
  double span = 0;            // its half-extent (max over axes)
  double z_min = 0, z_max = 0;  // the particles' (and the hull's) z range
  // entry structure; tri_xy .. tri_abc are arrays of tri_block
  void *tri_block = nullptr;
  double *tri_xy = nullptr;
  int *tri_tet = nullptr;
  int *tri_rid = nullptr;     // F: 4 x tet + face, with face records
  int *tri_abc = nullptr;     // F x 3 vertex ids, for inspection
  int *cell_off = nullptr;
  int *cell_tri = nullptr;
  sdtfe::EntryGrid grid{};
  // per-call scratch
  unsigned long long *stats = nullptr;  // ST_COUNT counters
  double *map_buf = nullptr;            // G*G pixel values
  int *navg_buf = nullptr;
  size_t map_cap = 0;
  sdtfe::Deferred *defer_buf = nullptr;  // pixels queued for the FP64 pass
  size_t defer_cap = 0;
  int *defer_count = nullptr;
  // second pass, one lane per queued sample: per-entry counts and offsets
  // (+ scan scratch), per-sample values and flags
  int *dq_off = nullptr, *dq_tmp = nullptr;
  size_t dq_off_cap = 0;
  double *dq_val = nullptr;
  unsigned char *dq_ok = nullptr;
  size_t dq_val_cap = 0;
  int64_t device_bytes = 0;
  double t_setup_ms = 0;
};

namespace sdtfe {
// capi.cpp
MapGeom make_geom_pub(const sdtfe_tess *t, const sdtfe_map_params &p);
void fill_stats_pub(const unsigned long long *c, sdtfe_stats *s);
// batch.hip: tiles of one map in a batch, and a batched launch over tile
// ranges [tile_begin[i], tile_begin[i] + tile_count[i]) of each item (null
// ranges: whole maps).  leave_room: the persistent kernel leaves one block
// slot per CU unused, so kernels of other streams (a server's set-ups) run
// beside it instead of after it.
int batch_tiles(const sdtfe_map_params &p, int lanes_per_pixel);
int batch_run_ranges(sdtfe_batch *b, int n, sdtfe_tess *const *tess,
                     const sdtfe_map_params *params, double *const *rho_device,
                     const int *lanes_per_pixel, const int *tile_begin, const int *tile_count,
                     sdtfe_stats *item_stats, sdtfe_stats *total, bool leave_room = false);
// Sizes a batch's item buffers for n items up front, so later launches of up
// to n items neither reallocate them nor capture the graph again.
int batch_reserve(sdtfe_batch *b, int n);
// Stream-ordered device memory from the device's default pool, which keeps
// freed blocks for reuse (capi.cpp): unlike hipMalloc / hipFree these do not
// synchronise the device, so tessellations set up or freed on one stream do
// not wait for kernels on others.
hipError_t dev_malloc(void **p, size_t bytes, hipStream_t s);
void dev_free(void *p, hipStream_t s);
// Pinned double buffers for uploads from pageable memory by one host
// thread: chunks are copied into one buffer while the other is in flight
// (several threads then upload concurrently).
struct Staging {
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  size_t bytes = 0;
};
int staging_init(Staging &st, size_t bytes);
void staging_free(Staging &st);
int upload(void *dst, const void *src, size_t bytes, hipStream_t s, Staging *st);
// Device-to-host read-back through st's pinned buffers: the copy of one
// slice is in flight while consume(host, offset, n) takes the one before
// (offset and n in bytes).  Without pinned buffers, one pageable copy.
int download(const void *src, size_t bytes, hipStream_t s, Staging *st,
             const std::function<void(const char *, size_t, size_t)> &consume);
// Pinned staging for calls that bring none (the drop-ins, sdtfe_tess_create,
// maps and grids read back to the host): taken from a process-wide pool (per
// device) for the call and returned idle; the pool's buffers are kept until
// the process ends, one pair per call (four per large upload) running at
// once.  Uploads of 64 MB or more are split over four host threads.
struct PooledStaging {
  Staging *st;
  int dev;  // the device its events belong to
  PooledStaging();
  ~PooledStaging();
  PooledStaging(const PooledStaging &) = delete;
  PooledStaging &operator=(const PooledStaging &) = delete;
};
// A tessellation as sdtfe_tess_create makes it, uploading through st, on
// stream (the caller's, kept when the tessellation is destroyed) or, when
// null, on a stream of its own; without the FP64 walk's face records when
// fp64_faces is false (FP64 maps then walk the per-tet records).
int tess_create_staged(const double *particle_data, int n, const int *tetra_data, int nt,
                       float p_mass, Staging *st, hipStream_t stream, bool fp64_faces,
                       sdtfe_tess **out);
// volume.hip: the volumetric grid's geometry (src/dtfe.c:386-409,
// 686-695) and its launch (the tessellation's stream)
VolGeom make_vol_geom(const sdtfe_volume_params &p);
void launch_volume_fp64(sdtfe_tess *t, const VolGeom &v, double *grid);
// adapter.hip: the reference's tetra_data from Delaunay simplices
int tetra_from_simplices(const double *particle_data, int n, const int *simplices, int nt,
                         int *tetra_data);
// Whether a tessellation's face records need 64-bit offsets (more than
// 2^25 tets); SDTFE_WIDE_RECORDS=1 takes the 64-bit kernels for every size
// (they give the same maps; tests run both).
bool wide_records(int64_t n_tetra);
// field.hip
int build_field(sdtfe_tess *t, const double *particle_data, const int *tetra_data,
                Staging *st);
int build_entry_grid(sdtfe_tess *t, Staging *st);
// Forward-hull faces of host arrays, on the GPU (no tessellation handle):
// mask[t] bit f marks face f of tet t (src/triangulate.c:274-311); the
// faces a mask marks as vertex triples and tets in tet, then face order
// (src/triangulate.c:325-355), n_faces being the mask's count.
int hull_mask_host(const double *pd, int n, const int *tetra, int nt, unsigned char *mask);
int hull_rows_host(const double *pd, int n, const int *tetra, int nt, const unsigned char *mask,
                   int *abc, int *tet, int n_faces);
// Neighbour slots 3-5 of forward-hull triangle rows (a, b, c, n_CB, n_AC,
// n_BA, tet, 0) across shared edges, as copy_fh_triangles links them
// (src/triangulate.c:358-392).
void link_hull_rows(int *rows, int n_rows);
// The status of the calling thread's last drop-in call (0: success).
void set_dropin_status(int rc);

// walk_mixed.hip: buffers of a mixed map's second pass and the map's
// counters -- a tessellation's own (grown on demand) or the fixed scratch
// memory of the entry point that allocates nothing.
struct MixedWork {
  unsigned long long *stats;  // ST_COUNT counters
  Deferred *q;                // queued pixels
  int *q_count;
  size_t q_cap;
  int *off, *tmp;             // per-entry offsets (entries + 1) and scan scratch
  size_t off_cap, tmp_ints;
  double *val;                // per handed-over sample: value and whether it counts
  unsigned char *ok;
  size_t val_cap;
  int *h_count = nullptr;     // pinned host word for the queue's length (null: pageable)
};
// Limits of sdtfe_map_device_noalloc: mixed maps of up to kNoAllocPixels
// pixels; up to kNoAllocSamples handed-over samples take the per-sample
// second pass (more take the wave-per-pixel form).  The scratch memory is
// kNoAllocScratchBytes per GPU.
constexpr size_t kNoAllocPixels = (size_t)1 << 20;
constexpr size_t kNoAllocSamples = (size_t)1 << 21;
constexpr size_t kNoAllocScratchBytes = 256 + 256 + 64 * kNoAllocPixels +
                                        (((4 * (kNoAllocPixels + 1)) + 255) & ~(size_t)255) +
                                        4096 + 8 * kNoAllocSamples + kNoAllocSamples;
// The scratch memory of the current device (its address found once per
// device) as a MixedWork.
int noalloc_work(MixedWork *w);
// A map into device memory (accumulated into) on the tessellation's stream
// with w's buffers and no allocation: counters cleared by a kernel, then the
// FP64 map or the mixed map's two passes (grow = false).
int run_map_noalloc(sdtfe_tess *t, const MapGeom &g, int precision, double *out,
                    MixedWork &w);
}  // namespace sdtfe
