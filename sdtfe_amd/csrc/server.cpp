// server.cpp -- a map-serving loop with continuous batching (host C++ over
// the batched launch of batch.hip).
//
//   submit()      copies a request (particles, tets, map parameters) into
//                 host memory owned by the server and queues it.
//   setup threads admit queued requests in arrival order while their HBM
//                 footprint fits the budget and fewer than max_active are
//                 resident, upload them and build their tessellations
//                 (sdtfe_tess_create), several requests at a time (uploads
//                 overlap each other's kernels).  Each thread sets its
//                 requests up on two streams of a pool of twice the setup
//                 threads (at most 16) in turn, so the streams a server holds
//                 stay bounded however many requests are resident, and one
//                 thread's set-up never waits for another's.
//   walk thread   repeatedly launches one persistent batch over the next
//                 tiles of every resident request (at most tile_budget tiles
//                 per launch, shared round-robin), so a request admitted
//                 while others are on the GPU joins the next launch; while
//                 requests are being set up, a launch leaves one block slot
//                 per CU to their kernels, which then do not wait for it; a
//                 request whose tiles are all done goes to the completion
//                 thread, and the next launch does not wait for it.
//   completion    copies a finished map out on a stream of its own (through
//   thread        pinned memory), adds it into the caller's map, frees the
//                 request's device memory and wakes its waiters.  Nothing
//                 there waits for the streams other requests set up on.
//
// Every pixel is computed by the same code with the same stream whatever
// else shares the launch, so a map is identical alone and in a mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "sdtfe_host.h"
#include "walk_common.h"

namespace sdtfe {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  std::vector<double> pd;
  std::vector<int> tets;
  float p_mass = 1.f;
  sdtfe_map_params params{};
  double *out = nullptr;  // caller's host map, accumulated into
  int64_t footprint = 0;
  sdtfe_tess *tess = nullptr;
  hipStream_t stream = nullptr;  // of the server's pool: set-up, map buffer
  double *d_map = nullptr;
  int n_tiles = 0, next_tile = 0, lpp = 1;
  Clock::time_point t_submit, t_admit, t_start, t_done;
  bool started = false, done = false;
  int status = 0;
  std::string error;  // why status is not 0 (the failing thread's message)
  sdtfe_stats stats{};
};

// A request whose walk is over, for the completion thread: the launch's
// status and, when it failed, the walk thread's message.
struct Finished {
  std::shared_ptr<Request> r;
  int status = 0;
  std::string error;
};

// The completion thread's pinned buffer for copying maps out.
struct Pinned {
  double *p = nullptr;
  size_t n = 0;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

int64_t estimate_footprint(int n, int nt, int grid, bool fp64) {
  // resident: pos4 32 B/vertex; per tet: row 32, volume 8, gradient 24,
  // records 144 (FP64) + 96 (mixed) + 4 x 32 (faces) and, for FP64 maps,
  // 4 x 144 (FP64 faces); map 8 B/pixel; setup scratch ~ 16 B/tet + 12 B/vertex
  return (int64_t)n * 44 + (int64_t)nt * (fp64 ? 1024 : 448) + (int64_t)grid * grid * 8 +
         (1 << 20);
}

}  // namespace
}  // namespace sdtfe

struct sdtfe_server {
  sdtfe_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_setup, cv_walk, cv_done, cv_complete;
  std::deque<std::shared_ptr<sdtfe::Request>> queued;
  std::vector<std::shared_ptr<sdtfe::Request>> active;  // resident, tiles pending
  std::deque<sdtfe::Finished> completing;                // walked, not yet copied out
  std::map<int64_t, std::shared_ptr<sdtfe::Request>> all;
  int64_t next_id = 1;
  int64_t used_bytes = 0;
  int n_setup = 0;       // being set up (counted in used_bytes)
  int n_completing = 0;  // handed to the completion thread, not yet done
  bool stop = false;
  sdtfe_batch *batch = nullptr;
  std::vector<std::thread> setup_thr;
  std::thread walk_thr, complete_thr;
  hipStream_t copy_stream = nullptr;  // the completion thread's
  sdtfe::Clock::time_point t0;
  int64_t launches = 0;
  double walk_ms = 0;
  std::vector<hipStream_t> streams;  // the pool requests are set up on
};

namespace sdtfe {
namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

void add_stats(sdtfe_stats &a, const sdtfe_stats &s) {
  a.samples += s.samples;
  a.missed += s.missed;
  a.failed += s.failed;
  a.no_intersection += s.no_intersection;
  a.perturb_exhausted += s.perturb_exhausted;
  a.no_isect_perturb += s.no_isect_perturb;
  a.perturb_steps += s.perturb_steps;
  a.tet_steps += s.tet_steps;
  a.step_cap += s.step_cap;
  a.fp64_redo += s.fp64_redo;
}

// Adds a walked request's map into the caller's host map: slices of at most
// 2^20 pixels go through the pinned buffer on the copy stream, which carries
// nothing else.  The walk that wrote the map has finished.
int copy_out(sdtfe_server *sv, Request &r, Pinned &pb) {
  const size_t npix = (size_t)r.params.grid_dim * r.params.grid_dim;
  const size_t slice = std::min(npix, (size_t)1 << 20);
  if (pb.n < slice) {
    if (pb.p) (void)hipHostFree(pb.p);
    pb.p = nullptr;
    pb.n = 0;
    if (hipHostMalloc((void **)&pb.p, sizeof(double) * slice, hipHostMallocDefault) !=
        hipSuccess) {
      pb.p = nullptr;
      return fail(SDTFE_ENOMEM, "pinned copy-out buffer allocation failed");
    }
    pb.n = slice;
  }
  for (size_t off = 0; off < npix; off += slice) {
    const size_t n = std::min(slice, npix - off);
    SDTFE_HIP(hipMemcpyAsync(pb.p, r.d_map + off, sizeof(double) * n, hipMemcpyDeviceToHost,
                             sv->copy_stream));
    SDTFE_HIP(hipStreamSynchronize(sv->copy_stream));
    double *out = r.out + off;
    for (size_t i = 0; i < n; ++i) out[i] += pb.p[i];
  }
  return 0;
}

// Frees a request's device memory and wakes its waiters; status and error
// are what the waiter sees.  The caller holds no lock, and the request's
// device work is complete (its frees are ordered on its stream).
void finish(sdtfe_server *sv, const std::shared_ptr<Request> &r, int status, std::string error) {
  if (r->d_map) dev_free(r->d_map, r->stream);
  r->d_map = nullptr;
  sdtfe_tess_destroy(r->tess);
  r->tess = nullptr;
  std::vector<double>().swap(r->pd);
  std::vector<int>().swap(r->tets);
  std::lock_guard<std::mutex> g(sv->mu);
  r->status = status;
  r->error = std::move(error);
  r->done = true;
  r->t_done = Clock::now();
  sv->used_bytes -= r->footprint;
  sv->cv_done.notify_all();
  sv->cv_setup.notify_all();
}

void setup_loop(sdtfe_server *sv, int thread_no) {
  Staging st;  // this thread's pinned upload buffers
  if (staging_init(st, (size_t)8 << 20)) staging_free(st);  // unpinned uploads then
  // this thread's two streams of the pool, in turn (shared with another
  // thread only beyond 8 set-up threads)
  const size_t n_streams = sv->streams.size();
  size_t turn = 0;
  for (;;) {
    std::shared_ptr<Request> r;
    hipStream_t stream = nullptr;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_setup.wait(lk, [&] {
        if (sv->stop) return true;
        if (sv->queued.empty()) return false;
        const auto &q = sv->queued.front();
        const int resident = (int)sv->active.size() + sv->n_setup + sv->n_completing;
        const bool fits = sv->used_bytes + q->footprint <= sv->cfg.hbm_budget_bytes ||
                          (resident == 0);  // a lone request is always admitted
        return resident < sv->cfg.max_active && fits;
      });
      if (sv->stop) {
        lk.unlock();
        staging_free(st);
        return;
      }
      r = sv->queued.front();
      sv->queued.pop_front();
      sv->used_bytes += r->footprint;
      sv->n_setup++;
      r->t_admit = Clock::now();
      stream = sv->streams[(2 * (size_t)thread_no + (turn++ & 1)) % n_streams];
    }
    // the map buffer first, cleared on the same stream: the tessellation's
    // set-up ends by synchronising that stream, which covers the clearing
    r->stream = stream;
    const size_t npix = (size_t)r->params.grid_dim * r->params.grid_dim;
    int rc = 0;
    if (dev_malloc((void **)&r->d_map, sizeof(double) * npix, stream) != hipSuccess) {
      r->d_map = nullptr;
      rc = fail(SDTFE_ENOMEM, "map buffer");
    } else if (hipMemsetAsync(r->d_map, 0, sizeof(double) * npix, stream) != hipSuccess) {
      rc = fail(SDTFE_EHIP, "clearing the map buffer failed");
    }
    // a mixed-precision server never walks the FP64 face records
    if (rc == 0)
      rc = tess_create_staged(r->pd.data(), (int)(r->pd.size() / 4), r->tets.data(),
                              (int)(r->tets.size() / 8), r->p_mass, &st, stream,
                              sv->cfg.precision == SDTFE_PREC_FP64, &r->tess);
    // the host copies are uploaded: release them now
    std::vector<double>().swap(r->pd);
    std::vector<int>().swap(r->tets);
    if (rc == 0) {
      r->lpp = sv->cfg.lanes_per_pixel > 0 ? sv->cfg.lanes_per_pixel
                                           : map_lanes(r->params.n_mc_samp);
      r->n_tiles = batch_tiles(r->params, r->lpp);
    }
    if (rc) {
      finish(sv, r, rc, sdtfe_last_error());
      std::lock_guard<std::mutex> g(sv->mu);
      sv->n_setup--;
      sv->cv_done.notify_all();
      sv->cv_setup.notify_all();
      continue;
    }
    // from being set up to resident in one step: the drain in
    // sdtfe_server_destroy never sees the request in neither state
    std::lock_guard<std::mutex> g(sv->mu);
    sv->n_setup--;
    sv->active.push_back(r);
    sv->cv_walk.notify_all();
  }
}

void walk_loop(sdtfe_server *sv) {
  size_t rr = 0;  // round-robin start
  for (;;) {
    std::vector<std::shared_ptr<Request>> items;
    bool setups = false;  // set-up work may come while this launch runs
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_walk.wait(lk, [&] { return sv->stop || !sv->active.empty(); });
      if (sv->stop) return;
      items = sv->active;
      setups = sv->n_setup > 0 || !sv->queued.empty();
    }
    // share the tile budget: every request gets an equal slice, leftovers
    // go round-robin
    const int n = (int)items.size();
    std::vector<int> begin(n), count(n, 0);
    int budget = std::max(sv->cfg.tile_budget, n);
    int remaining = 0;
    for (int i = 0; i < n; ++i) remaining += items[i]->n_tiles - items[i]->next_tile;
    budget = std::min(budget, remaining);
    const int fair = std::max(1, budget / n);
    int given = 0;
    for (int k = 0; k < n; ++k) {
      const int i = (int)((rr + k) % n);
      const int left = items[i]->n_tiles - items[i]->next_tile;
      count[i] = std::min(left, fair);
      given += count[i];
    }
    for (int k = 0; k < n && given < budget; ++k) {
      const int i = (int)((rr + k) % n);
      const int extra = std::min(items[i]->n_tiles - items[i]->next_tile - count[i],
                                 budget - given);
      count[i] += extra;
      given += extra;
    }
    rr = (rr + 1) % std::max(n, 1);
    std::vector<sdtfe_tess *> ts;
    std::vector<sdtfe_map_params> ps;
    std::vector<double *> outs;
    std::vector<int> lp, b0, cnt;
    std::vector<std::shared_ptr<Request>> in_launch;
    const auto now = Clock::now();
    for (int i = 0; i < n; ++i) {
      if (count[i] == 0) continue;
      Request &r = *items[i];
      if (!r.started) {
        r.started = true;
        r.t_start = now;
      }
      ts.push_back(r.tess);
      ps.push_back(r.params);
      outs.push_back(r.d_map);
      lp.push_back(r.lpp);
      b0.push_back(r.next_tile);
      cnt.push_back(count[i]);
      in_launch.push_back(items[i]);
    }
    std::vector<sdtfe_stats> st(ts.size());
    sdtfe_stats tot{};
    const int rc = ts.empty() ? 0
                              : batch_run_ranges(sv->batch, (int)ts.size(), ts.data(), ps.data(),
                                                 outs.data(), lp.data(), b0.data(), cnt.data(),
                                                 st.data(), &tot, setups);
    std::vector<std::shared_ptr<Request>> finished, empty;  // empty: had no tiles
    {
      std::lock_guard<std::mutex> g(sv->mu);
      sv->launches++;
      sv->walk_ms += tot.t_walk_ms;
      for (size_t k = 0; k < in_launch.size(); ++k) {
        Request &r = *in_launch[k];
        add_stats(r.stats, st[k]);
        r.next_tile += cnt[k];
        if (rc || r.next_tile >= r.n_tiles) finished.push_back(in_launch[k]);
      }
      for (auto &f : finished)
        sv->active.erase(std::remove(sv->active.begin(), sv->active.end(), f), sv->active.end());
      // requests with no tiles at all
      for (auto it = sv->active.begin(); it != sv->active.end();) {
        if ((*it)->n_tiles == 0) {
          empty.push_back(*it);
          it = sv->active.erase(it);
        } else {
          ++it;
        }
      }
      // to the completion thread; a request in no launch does not take the
      // launch's status
      const std::string error = rc ? std::string(sdtfe_last_error()) : std::string();
      for (auto &f : finished) sv->completing.push_back(Finished{f, rc, error});
      for (auto &f : empty) sv->completing.push_back(Finished{f, 0, std::string()});
      sv->n_completing += (int)(finished.size() + empty.size());
    }
    if (!finished.empty() || !empty.empty()) sv->cv_complete.notify_one();
  }
}

void complete_loop(sdtfe_server *sv) {
  Pinned pb;
  for (;;) {
    Finished f;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_complete.wait(lk, [&] { return sv->stop || !sv->completing.empty(); });
      if (sv->completing.empty()) return;  // stopped, nothing left
      f = std::move(sv->completing.front());
      sv->completing.pop_front();
    }
    if (f.status == 0 && f.r->d_map && copy_out(sv, *f.r, pb)) {
      f.status = SDTFE_EHIP;
      f.error = std::string("copying the map back failed: ") + sdtfe_last_error();
    }
    finish(sv, f.r, f.status, std::move(f.error));
    std::lock_guard<std::mutex> g(sv->mu);
    sv->n_completing--;
    sv->cv_done.notify_all();
    sv->cv_setup.notify_all();
  }
}

}  // namespace
}  // namespace sdtfe

using namespace sdtfe;

extern "C" {

int sdtfe_server_create(const sdtfe_server_config *cfg, sdtfe_server **out) {
  if (!cfg || !out) return fail(SDTFE_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->hbm_budget_bytes <= 0 || cfg->max_active <= 0 || cfg->tile_budget <= 0)
    return fail(SDTFE_EINVAL, "bad server configuration");
  if (cfg->precision != SDTFE_PREC_FP64 && cfg->precision != SDTFE_PREC_MIXED)
    return fail(SDTFE_EINVAL, "unknown precision");
  sdtfe_server *sv = new sdtfe_server();
  sv->cfg = *cfg;
  // item buffers for max_active requests now: the walk thread then never
  // reallocates them or captures the batch graph again while set-up threads
  // are issuing work
  int rc = sdtfe_batch_create(&sv->batch);
  if (!rc) rc = batch_reserve(sv->batch, cfg->max_active);
  if (rc) {
    if (sv->batch) sdtfe_batch_destroy(sv->batch);
    delete sv;
    return rc;
  }
  sv->t0 = Clock::now();
  const int n_setup = cfg->setup_threads > 0 ? std::min(cfg->setup_threads, 64) : 4;
  const int n_streams = std::min(2 * n_setup, 16);
  for (int i = 0; i <= n_streams; ++i) {  // the pool, then the copy stream
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t x : sv->streams) (void)hipStreamDestroy(x);
      sdtfe_batch_destroy(sv->batch);
      delete sv;
      return fail(SDTFE_EHIP, "server stream pool");
    }
    if (i < n_streams)
      sv->streams.push_back(st);
    else
      sv->copy_stream = st;
  }
  for (int i = 0; i < n_setup; ++i) sv->setup_thr.emplace_back(setup_loop, sv, i);
  sv->walk_thr = std::thread(walk_loop, sv);
  sv->complete_thr = std::thread(complete_loop, sv);
  *out = sv;
  return 0;
}

int sdtfe_server_submit(sdtfe_server *sv, const double *particle_data, int n_particles,
                        const int *tetra_data, int n_tetra, float p_mass,
                        const sdtfe_map_params *p, double *rho, int64_t *request_id) {
  if (!sv || !p || !rho || n_particles < 0 || n_tetra < 0 ||
      (n_particles && !particle_data) || (n_tetra && !tetra_data))
    return fail(SDTFE_EINVAL, "bad argument");
  if (p->precision != sv->cfg.precision)
    return fail(SDTFE_EINVAL, "request precision differs from the server's");
  if (p->grid_dim <= 0 || p->grid_dim > 65536 || p->n_mc_samp < 0 || !(p->box_len > 0) ||
      !(p->box_len_z > 0))
    return fail(SDTFE_EINVAL, "bad map parameters");
  auto r = std::make_shared<Request>();
  r->pd.assign(particle_data, particle_data + 4 * (size_t)n_particles);
  r->tets.assign(tetra_data, tetra_data + 8 * (size_t)n_tetra);
  r->p_mass = p_mass;
  r->params = *p;
  r->out = rho;
  r->footprint = estimate_footprint(n_particles, n_tetra, p->grid_dim,
                                    p->precision == SDTFE_PREC_FP64);
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(sv->mu);
  r->id = sv->next_id++;
  sv->all[r->id] = r;
  sv->queued.push_back(r);
  if (request_id) *request_id = r->id;
  sv->cv_setup.notify_all();
  return 0;
}

int sdtfe_server_wait(sdtfe_server *sv, int64_t id, sdtfe_request_info *info) {
  if (!sv) return fail(SDTFE_EINVAL, "null server");
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return fail(SDTFE_EINVAL, "unknown request id");
    r = it->second;
    sv->cv_done.wait(lk, [&] { return r->done; });
    sv->all.erase(id);
  }
  if (info) {
    info->status = r->status;
    info->queue_ms = ms_between(r->t_submit, r->t_admit);
    info->setup_ms = r->started ? ms_between(r->t_admit, r->t_start) : 0.0;
    info->latency_ms = ms_between(r->t_submit, r->t_done);
    info->submit_ms = ms_between(sv->t0, r->t_submit);
    info->done_ms = ms_between(sv->t0, r->t_done);
    info->footprint_bytes = r->footprint;
    info->stats = r->stats;
  }
  if (r->status)
    return fail(r->status, "request " + std::to_string(id) + " failed: " + r->error);
  return 0;
}

int sdtfe_server_get_info(sdtfe_server *sv, sdtfe_server_info *info) {
  if (!sv || !info) return fail(SDTFE_EINVAL, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  info->queued = (int)sv->queued.size();
  info->active = (int)sv->active.size() + sv->n_setup + sv->n_completing;
  info->used_bytes = sv->used_bytes;
  info->launches = sv->launches;
  info->walk_ms = sv->walk_ms;
  return 0;
}

void sdtfe_server_destroy(sdtfe_server *sv) {
  if (!sv) return;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    // drain: wait until nothing is queued, being set up, resident or
    // being copied out
    sv->cv_done.wait(lk, [&] {
      return sv->queued.empty() && sv->active.empty() && sv->n_setup == 0 &&
             sv->n_completing == 0;
    });
    sv->stop = true;
    sv->cv_setup.notify_all();
    sv->cv_walk.notify_all();
    sv->cv_complete.notify_all();
  }
  for (std::thread &th : sv->setup_thr)
    if (th.joinable()) th.join();
  if (sv->walk_thr.joinable()) sv->walk_thr.join();
  if (sv->complete_thr.joinable()) sv->complete_thr.join();
  sdtfe_batch_destroy(sv->batch);
  for (hipStream_t st : sv->streams) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  if (sv->copy_stream) {
    (void)hipStreamSynchronize(sv->copy_stream);
    (void)hipStreamDestroy(sv->copy_stream);
  }
  delete sv;
}

}  // extern "C"
