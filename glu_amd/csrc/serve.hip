// serve.hip -- a serving loop for streams of gauge-fixing requests.
//
// GLU fixes one configuration per call (Wrappers/GF_wrap.c:179-212) and
// checks beforehand that its temporaries fit in memory
// (Utils/GLU_memcheck.c:72-95).  The serving loop keeps that admission rule
// for a whole device: requests wait in a FIFO queue until the device memory
// their context needs (glu_gfix_footprint) fits under a budget, then run
// concurrently, each on its own context and stream.  One worker thread
// drives every admitted Landau run batch by batch (landau_begin /
// landau_advance), so a long run never holds up a short one
// (continuous batching).  Contexts are pooled by (dims, gftype,
// precision) and reused, and idle ones are evicted when a new shape needs
// their memory.  A request's result does not depend on what it was served
// with: every context owns its buffers and every reduction is
// deterministic.
//
// With glu_gfix_server_set_batching, FP64 Landau CG requests of small
// lattices (every extent 2..16) skip the contexts: every queued request of
// one shape and settings is fixed in one shared launch, one workgroup per
// lattice (facg_batch.hip), on a host thread of its own, and requests of
// that key arriving meanwhile gather for the next launch; a lattice whose
// FACG fails goes back to the queue head and takes Landau()'s restarts on
// a context.
//
// Landau requests follow Landau()'s restart policy (glu_gfix_landau).
// Coulomb requests (the batched-slice driver is synchronous) run on a host
// thread of their own, so the worker keeps advancing the Landau runs in
// flight meanwhile.  A request may carry GLU's own input-file keys
// (GFTYPE, ACCURACY, MAX_ITERS, GF_TUNE, SEED, RANDOM_TRANSFORM;
// IO/input_reader.c:188-193, 444-481, 510-517).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using Clock = std::chrono::steady_clock;
using gfix::launch_pack;
using gfix::launch_unpack;

size_t gfix_footprint(const size_t dims[4], int gftype, int precision);

namespace {

enum ReqState { RQ_QUEUED = 0, RQ_RUNNING, RQ_DONE, RQ_FAILED };

struct Request {
  size_t dims[4];
  int gftype, precision, method;
  double gf, acc;
  size_t max_iters;
  double *links;  // device buffer, GLU layout double[V][4][18], fixed in place
  size_t bytes;
  Clock::time_point t_submit, t_done;
  int state = RQ_QUEUED;  // changed under the server lock only
  int copying = 0;        // worker-only: the output copy is in flight
  int waited = 0;         // a caller is waiting for this ticket
  double switch_theta = 0.0;
  int use_graph = 1;
  uint64_t restart_seed = 0;
  // GLU's RANDOM_TRANSFORM = YES: the input is first transformed with the
  // streams seeded SEED = seed, and the restarts continue those streams
  int random_transform = 0;
  uint32_t seed = 0;
  // a Coulomb run on its own thread: 1 while it runs (worker-visible)
  std::shared_ptr<std::atomic<int>> async;
  int ticket = 0;
  // FACG failed in a shared small-lattice launch: the restarts of Landau()
  // follow on a context, from the input the launch left untouched
  int from_batch = 0;
  int attempt = 0;   // Landau() restarts taken so far
  size_t total = 0;  // iterations counted across restarts
  int rc = 0;
  std::string err;
  double theta = 0.0;
  size_t iters = 0;
  glu_gfix_ctx *ctx = nullptr;
  hipEvent_t ev = nullptr;
};

struct Pooled {
  glu_gfix_ctx *ctx;
  size_t dims[4];
  int gftype, precision;
  size_t bytes;
  bool busy;
  unsigned long long last_use;
};

bool same_shape(const size_t a[4], const size_t b[4]) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && a[3] == b[3];
}

// extents the shared small-lattice launch takes (facg_batch.hip)
constexpr size_t kBatchMaxExtent = 16;
constexpr size_t kBatchMaxLattices = 65535;
// its scratch per site of each lattice: 40 complex numbers
constexpr size_t kBatchScratchPerSite = 40 * 16;

// a shared launch in flight: its requests and the flag its host thread
// clears when the results are in
struct Batch {
  std::vector<Request *> reqs;
  std::shared_ptr<std::atomic<int>> running;
  size_t bytes = 0;
};

}  // namespace

struct glu_gfix_server {
  int device = 0;
  size_t budget = 0;
  int max_active = 0;
  double switch_theta = 0.0;  // 0: the context default
  int use_graph = 1;
  uint64_t restart_seed = 1;
  int batch_small = 0;  // small lattices in shared launches
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<int, Request> reqs;
  std::deque<int> queue;
  std::vector<Pooled> pool;
  size_t used = 0, peak = 0;
  unsigned long long tick = 0;
  int next_ticket = 1;
  bool stop = false;
  std::vector<double> lat_ms;
  std::thread worker;
};

namespace {

bool batchable(const glu_gfix_server *s, const Request &r) {
  if (!s->batch_small || r.gftype != GLU_GFIX_LANDAU || r.method != GLU_GFIX_CG || r.precision != GLU_GFIX_FP64 ||
      r.random_transform || r.from_batch)
    return false;
  for (int mu = 0; mu < 4; mu++)
    if (r.dims[mu] < 2 || r.dims[mu] > kBatchMaxExtent) return false;
  return true;
}

// context for a request: an idle pooled one of the same shape, else a new
// one once idle contexts of other shapes have made room.  Called with the
// lock held; returns nullptr when the request has to wait.
glu_gfix_ctx *acquire(glu_gfix_server *s, const Request &r, int *err) {
  for (Pooled &p : s->pool)
    if (!p.busy && same_shape(p.dims, r.dims) && p.gftype == r.gftype && p.precision == r.precision) {
      p.busy = true;
      p.last_use = ++s->tick;
      return p.ctx;
    }
  while (s->used + r.bytes > s->budget) {  // evict the least recently used idle context
    int victim = -1;
    for (int i = 0; i < (int)s->pool.size(); i++)
      if (!s->pool[i].busy && (victim < 0 || s->pool[i].last_use < s->pool[victim].last_use)) victim = i;
    if (victim < 0) return nullptr;
    glu_gfix_destroy(s->pool[victim].ctx);
    s->used -= s->pool[victim].bytes;
    s->pool.erase(s->pool.begin() + victim);
  }
  glu_gfix_ctx *c = nullptr;
  *err = glu_gfix_create(&c, r.dims, r.gftype, r.precision);
  if (*err) return nullptr;
  s->pool.push_back(Pooled{c, {r.dims[0], r.dims[1], r.dims[2], r.dims[3]}, r.gftype, r.precision, r.bytes, true,
                           ++s->tick});
  s->used += r.bytes;
  s->peak = std::max(s->peak, s->used);
  return c;
}

void release(glu_gfix_server *s, glu_gfix_ctx *c, bool broken) {
  for (int i = 0; i < (int)s->pool.size(); i++)
    if (s->pool[i].ctx == c) {
      if (broken) {
        glu_gfix_destroy(c);
        s->used -= s->pool[i].bytes;
        s->pool.erase(s->pool.begin() + i);
      } else {
        s->pool[i].busy = false;
      }
      return;
    }
}

struct Ended {
  Request *r;
  int rc;
  std::string err;
};

// finish a request (lock held)
void complete(glu_gfix_server *s, Request &r, int rc, const std::string &err) {
  r.t_done = Clock::now();
  if (rc) {
    r.state = RQ_FAILED;
    r.rc = rc;
    r.err = err;
  } else {
    r.state = RQ_DONE;
    s->lat_ms.push_back(std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count());
  }
  if (r.ev) {
    (void)hipEventDestroy(r.ev);
    r.ev = nullptr;
  }
  release(s, r.ctx, rc != 0);
  r.ctx = nullptr;
  s->cv_done.notify_all();
}

int after_attempt(Request &r, size_t loc, double theta);
int after_run(Request &r);
int copy_out(Request &r);

// start a request on its context: load the links, begin the run.  Returns
// 0 and sets *finished when nothing is left to poll.
int start(glu_gfix_server *s, Request &r, bool *finished) {
  glu_gfix_ctx *c = r.ctx;
  *finished = false;
  if (hipEventCreateWithFlags(&r.ev, hipEventDisableTiming) != hipSuccess)
    return gfix_fail(-2, "serving: event creation failed");
  c->use_graph = r.use_graph;
  if (r.switch_theta > 0.0) c->switch_theta = r.switch_theta;
  c->mwc_live = 0;  // a pooled context's streams never carry over
  c->rng_seed = r.seed ? r.seed : 1;
  launch_pack<double>(r.links, c->U64, c->g.V, 4, 0, c->g.V, c->stream);
  if (r.random_transform) {  // GLU's read_and_fix with RANDOM_TRANSFORM = YES
    if (int rc = mwc_seed(c, c->rng_seed, c->rng_threads)) return rc;
    if (int rc = mwc_random_transform(c)) return rc;
  }
  if (r.gftype == GLU_GFIX_COULOMB) {
    // Coulomb_FA on a thread of its own; it ends with the output copy
    // enqueued, and the worker takes over from there
    r.async = std::make_shared<std::atomic<int>>(1);
    Request *rp = &r;
    const int dev = s->device;
    std::thread([rp, c, dev, flag = r.async] {
      (void)hipSetDevice(dev);
      int rc = coulomb_fix(c, rp->method, rp->gf, rp->acc, rp->max_iters, &rp->theta, &rp->iters);
      if (!rc) rc = copy_out(*rp);
      rp->rc = rc;
      if (rc) rp->err = glu_gfix_last_error();
      flag->store(0, std::memory_order_release);
    }).detach();
    return 0;
  }
  if (r.from_batch) {  // attempt 0 failed in a shared launch: the restarts
    const int more = after_attempt(r, 123456789, r.theta);
    if (more < 0) return more;
    *finished = more == 0;
    return 0;
  }
  if (int rc = landau_begin(c, r.method, r.gf, r.acc, r.max_iters)) return rc;
  if (c->run.finished) {
    const int more = after_run(r);
    if (more < 0) return more;
    *finished = more == 0;
  }
  return 0;
}

// a Landau attempt ended with FACG()'s return loc: apply the restart
// policy of Landau() (Gfix/Landau.c:301-345, as glu_gfix_landau).  0: the
// request is finished, 1: a restarted run is in flight, < 0: error
int after_attempt(Request &r, size_t loc, double theta) {
  glu_gfix_ctx *c = r.ctx;
  r.theta = theta;
  if (r.attempt == 0) {
    r.total = loc;
    if (loc != 123456789) {
      r.iters = loc;
      return 0;
    }
  } else if (landau_restart_policy(r.total, loc, r.max_iters, &r.total)) {
    r.iters = r.total;
    return 0;
  }
  if (++r.attempt >= 7)
    return gfix_fail(-5, "no convergence after %d random restarts (GF_TUNE %g)", r.attempt - 1, r.gf);
  // restart from the request's input, which stays untouched until copy_out,
  // under random_transform drawn from GLU's streams, seeded with the
  // request's seed at its first restart as glu_gfix_landau does (a pooled
  // context's streams never carry over from another request)
  launch_pack<double>(r.links, c->U64, c->g.V, 4, 0, c->g.V, c->stream);
  if (r.attempt == 1 && !r.random_transform)
    if (int rc = mwc_seed(c, (uint32_t)(r.restart_seed ? r.restart_seed : 1), c->rng_threads)) return rc;
  if (int rc = mwc_random_transform(c)) return rc;
  if (int rc = landau_begin(c, r.method, r.gf, r.acc, r.max_iters)) return rc;
  if (c->run.finished) return after_run(r);
  return 1;
}

int after_run(Request &r) { return after_attempt(r, r.ctx->run.iters, r.ctx->run.theta); }

// enqueue the copy of the fixed links into the request's buffer
int copy_out(Request &r) {
  glu_gfix_ctx *c = r.ctx;
  launch_unpack<double>(c->U64, r.links, c->g.V, 4, 0, c->g.V, c->stream);
  if (hipEventRecord(r.ev, c->stream) != hipSuccess) return gfix_fail(-2, "serving: event record failed");
  r.copying = 1;
  return 0;
}

void worker_main(glu_gfix_server *s) {
  (void)hipSetDevice(s->device);
  // requests this thread drives; std::map nodes stay put while other
  // threads insert, so the pointers stay valid until a request completes
  std::vector<Request *> active;
  std::vector<Batch> batches;  // shared small-lattice launches in flight
  std::unique_lock<std::mutex> lk(s->mu);
  // a request entering service takes the server's settings (lock held)
  auto settle = [&](Request &r) {
    r.state = RQ_RUNNING;
    r.switch_theta = s->switch_theta;
    r.use_graph = s->use_graph;
    if (!r.seed) {
      r.restart_seed = s->restart_seed;
      r.seed = (uint32_t)s->restart_seed;
    } else {
      r.restart_seed = r.seed;
    }
  };
  for (;;) {
    // small lattices: the first queued request that can share a launch,
    // with every queued request of its shape and settings, as many as the
    // memory budget holds.  A launch takes one slot; while one runs for a
    // shape and settings, the requests of that key gather for the next
    std::vector<Batch> launched;
    auto same_key = [](const Request &a, const Request &b) {
      return same_shape(a.dims, b.dims) && a.gf == b.gf && a.acc == b.acc && a.max_iters == b.max_iters;
    };
    auto key_running = [&](const Request &h) {
      for (const Batch &b : batches)
        if (same_key(*b.reqs[0], h)) return true;
      for (const Batch &b : launched)
        if (same_key(*b.reqs[0], h)) return true;
      return false;
    };
    for (size_t qi = 0; qi < s->queue.size() && (int)(active.size() + batches.size() + launched.size()) <
                                                      s->max_active;) {
      const Request &h = s->reqs[s->queue[qi]];
      if (!batchable(s, h) || key_running(h)) {
        ++qi;
        continue;
      }
      const size_t per = h.dims[0] * h.dims[1] * h.dims[2] * h.dims[3] * kBatchScratchPerSite;
      const size_t room = std::min(s->budget > s->used ? (s->budget - s->used) / per : 0, kBatchMaxLattices);
      if (room == 0) break;  // waits for memory
      const Request key = h;
      Batch b;
      for (auto it = s->queue.begin(); it != s->queue.end() && b.reqs.size() < room;) {
        Request &r = s->reqs[*it];
        if (batchable(s, r) && same_key(r, key)) {
          settle(r);
          b.reqs.push_back(&r);
          it = s->queue.erase(it);
        } else {
          ++it;
        }
      }
      b.bytes = b.reqs.size() * per;
      s->used += b.bytes;
      s->peak = std::max(s->peak, s->used);
      b.running = std::make_shared<std::atomic<int>>(1);
      launched.push_back(std::move(b));
    }
    // admission of the others: FIFO among themselves, as many as the budget
    // and the slot count allow
    std::vector<Request *> admitted;
    for (size_t qi = 0; qi < s->queue.size() &&
                        (int)(active.size() + batches.size() + launched.size() + admitted.size()) < s->max_active;) {
      Request &r = s->reqs[s->queue[qi]];
      if (batchable(s, r)) {  // gathers for a shared launch
        ++qi;
        continue;
      }
      int err = 0;
      glu_gfix_ctx *c = acquire(s, r, &err);
      if (err) {  // creation failed: fail the request, keep serving
        s->queue.erase(s->queue.begin() + qi);
        r.t_done = Clock::now();
        r.state = RQ_FAILED;
        r.rc = err;
        r.err = glu_gfix_last_error();
        s->cv_done.notify_all();
        continue;
      }
      if (!c) break;  // waits for memory
      r.ctx = c;
      settle(r);
      admitted.push_back(&r);
      s->queue.erase(s->queue.begin() + qi);
    }
    if (active.empty() && admitted.empty() && batches.empty() && launched.empty()) {
      if (s->stop && s->queue.empty()) break;
      s->cv_work.wait(lk);
      continue;
    }
    lk.unlock();
    // device work happens without the lock; only this thread touches a
    // running request
    std::vector<Ended> ended;
    auto fail = [&](Request *r, int rc) { ended.push_back({r, rc, glu_gfix_last_error()}); };
    for (Batch &b : launched) {
      // FACG of every lattice of the launch; a failed one keeps its input
      // links for the restarts on a context
      std::vector<double *> lat;
      for (Request *r : b.reqs) lat.push_back(r->links);
      const Request &h = *b.reqs[0];
      std::thread([reqs = b.reqs, lat, dims = std::vector<size_t>(h.dims, h.dims + 4), gf = h.gf, acc = h.acc,
                   mi = h.max_iters, flag = b.running, dev = s->device] {
        (void)hipSetDevice(dev);
        const size_t n = reqs.size();
        std::vector<size_t> it(n);
        std::vector<double> th(n);
        const int rc = facg_batch_run(dims.data(), n, lat.data(), gf, acc, mi, it.data(), th.data(), nullptr, 0, 1);
        const std::string err = rc ? glu_gfix_last_error() : "";
        for (size_t k = 0; k < n; k++) {
          Request *r = reqs[k];
          r->rc = rc;
          r->err = err;
          if (!rc) {
            r->theta = th[k];
            r->iters = it[k];
            r->from_batch = it[k] == 123456789;
          }
        }
        flag->store(0, std::memory_order_release);
      }).detach();
      batches.push_back(std::move(b));
    }
    // finished shared launches: done, failed, or back to the queue head for
    // Landau()'s restarts
    std::vector<Request *> requeue;
    size_t freed = 0;
    for (auto it = batches.begin(); it != batches.end();) {
      if (it->running->load(std::memory_order_acquire)) {
        ++it;
        continue;
      }
      for (Request *r : it->reqs) {
        if (r->rc)
          ended.push_back({r, r->rc, r->err});
        else if (r->from_batch)
          requeue.push_back(r);
        else
          ended.push_back({r, 0, std::string()});
      }
      freed += it->bytes;
      it = batches.erase(it);
    }
    bool progressed = !admitted.empty() || !launched.empty() || freed;
    for (Request *r : admitted) {
      bool fin = false;
      int rc = start(s, *r, &fin);
      if (!rc && fin) rc = copy_out(*r);
      if (rc)
        fail(r, rc);
      else
        active.push_back(r);
    }
    for (Request *r : active) {
      if (r->async) {  // a Coulomb run on its own thread
        if (r->async->load(std::memory_order_acquire)) continue;
        const int rc = r->rc;
        r->async.reset();
        if (rc) {
          ended.push_back({r, rc, r->err});
          continue;
        }
        progressed = true;
      }
      if (!r->copying) {
        const int st = landau_advance(r->ctx, 0);
        if (st < 0) {
          fail(r, st);
          continue;
        }
        if (st == 1) {
          const int more = after_run(*r);
          if (more < 0) {
            fail(r, more);
            continue;
          }
          if (more == 0) {
            if (int rc = copy_out(*r)) {
              fail(r, rc);
              continue;
            }
          }
        }
        if (st != 0) progressed = true;
      }
      if (r->copying) {
        const hipError_t e = hipEventQuery(r->ev);
        if (e == hipSuccess) {
          ended.push_back({r, 0, std::string()});
        } else if (e != hipErrorNotReady) {
          gfix_fail(-2, "serving: output copy failed: %s", hipGetErrorString(e));
          fail(r, -2);
        }
      }
    }
    lk.lock();
    s->used -= freed;
    for (auto r = requeue.rbegin(); r != requeue.rend(); ++r) {  // in their order, ahead of the rest
      (*r)->state = RQ_QUEUED;
      s->queue.push_front((*r)->ticket);
    }
    for (Ended &e : ended) {
      complete(s, *e.r, e.rc, e.err);
      auto f = std::find(active.begin(), active.end(), e.r);
      if (f != active.end()) active.erase(f);  // requests of a shared launch were never in it
    }
    if (ended.empty() && !progressed) {
      // every run in flight is busy on the device: poll again shortly
      s->cv_work.wait_for(lk, std::chrono::microseconds(50));
    }
  }
}

}  // namespace

extern "C" int glu_gfix_server_create(glu_gfix_server **out, size_t hbm_budget, int max_active) {
  if (!out) return gfix_fail(-1, "glu_gfix_server_create: null argument");
  *out = nullptr;
  if (max_active < 1) return gfix_fail(-1, "max_active must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return gfix_fail(-3, "no HIP device: the gauge-fixing library has no CPU path");
  glu_gfix_server *s = new glu_gfix_server();
  (void)hipGetDevice(&s->device);
  if (hbm_budget == 0) {  // default: 90% of the memory free now
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return gfix_fail(-2, "hipMemGetInfo failed");
    }
    hbm_budget = fr / 10 * 9;
  }
  s->budget = hbm_budget;
  s->max_active = max_active;
  s->worker = std::thread(worker_main, s);
  *out = s;
  return 0;
}

extern "C" int glu_gfix_server_configure(glu_gfix_server *s, double switch_theta, int use_graph,
                                         uint64_t restart_seed) {
  if (!s) return gfix_fail(-1, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->switch_theta = switch_theta;
  s->use_graph = use_graph ? 1 : 0;
  s->restart_seed = restart_seed;
  return 0;
}

extern "C" int glu_gfix_server_set_batching(glu_gfix_server *s, int enable) {
  if (!s) return gfix_fail(-1, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->batch_small = enable ? 1 : 0;
  s->cv_work.notify_one();
  return 0;
}

static int submit(glu_gfix_server *s, const size_t dims[4], int gftype, int precision, int method, double gf_tune,
                  double accuracy, size_t max_iters, double *links_dev, int *ticket, uint32_t seed, int rtrans) {
  if (!s || !dims || !links_dev || !ticket) return gfix_fail(-1, "glu_gfix_server_submit: null argument");
  if (gftype != GLU_GFIX_LANDAU && gftype != GLU_GFIX_COULOMB) return gfix_fail(-1, "unknown gftype %d", gftype);
  if (precision != GLU_GFIX_FP64 && precision != GLU_GFIX_MIXED)
    return gfix_fail(-1, "unknown precision %d", precision);
  if (method != GLU_GFIX_CG && method != GLU_GFIX_SD) return gfix_fail(-1, "unknown method %d", method);
  if (max_iters == 0) return gfix_fail(-1, "max_iters must be positive");
  Request r{};
  memcpy(r.dims, dims, sizeof(r.dims));
  r.gftype = gftype;
  r.precision = precision;
  r.method = method;
  r.gf = gf_tune;
  r.acc = accuracy;
  r.max_iters = max_iters;
  r.links = links_dev;
  r.seed = seed;
  r.random_transform = rtrans;
  r.bytes = gfix_footprint(dims, gftype, precision);
  std::lock_guard<std::mutex> lk(s->mu);
  if (r.bytes > s->budget)
    return gfix_fail(-4, "request needs %zu bytes of device memory, above the budget of %zu", r.bytes, s->budget);
  if (s->stop) return gfix_fail(-1, "server is shutting down");
  r.t_submit = Clock::now();
  const int t = s->next_ticket++;
  r.ticket = t;
  s->reqs.emplace(t, r);
  s->queue.push_back(t);
  *ticket = t;
  s->cv_work.notify_one();
  return 0;
}

extern "C" int glu_gfix_server_submit(glu_gfix_server *s, const size_t dims[4], int gftype, int precision,
                                      int method, double gf_tune, double accuracy, size_t max_iters,
                                      double *links_dev, int *ticket) {
  return submit(s, dims, gftype, precision, method, gf_tune, accuracy, max_iters, links_dev, ticket, 0, 0);
}

// GLU's gauge-fixing keys from the text of an input file (KEY = VALUE
// lines; other keys are ignored): read_gf_struct and the SEED /
// RANDOM_TRANSFORM tags of IO/input_reader.c:188-193, 444-481, 510-517.
// ACCURACY is the exponent: the accuracy is 10^-ACCURACY.
extern "C" int glu_gfix_server_submit_glu(glu_gfix_server *s, const size_t dims[4], int precision,
                                          const char *options, double *links_dev, int *ticket) {
  if (!s || !dims || !options || !links_dev || !ticket)
    return gfix_fail(-1, "glu_gfix_server_submit_glu: null argument");
  std::map<std::string, std::string> kv;
  const char *p = options;
  while (*p) {
    const char *e = strchr(p, '\n');
    std::string line(p, e ? (size_t)(e - p) : strlen(p));
    p = e ? e + 1 : p + line.size();
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string t) {
      const size_t a = t.find_first_not_of(" \t\r"), b = t.find_last_not_of(" \t\r");
      return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
    };
    kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  auto need = [&](const char *k, std::string *v) {
    auto it = kv.find(k);
    if (it == kv.end() || it->second.empty()) return gfix_fail(-1, "[IO] tag %s not found", k);
    *v = it->second;
    return 0;
  };
  std::string gft, acc, mit, tune;
  if (int rc = need("GFTYPE", &gft)) return rc;
  if (int rc = need("ACCURACY", &acc)) return rc;
  if (int rc = need("MAX_ITERS", &mit)) return rc;
  if (int rc = need("GF_TUNE", &tune)) return rc;
  int gftype;
  if (gft == "LANDAU")
    gftype = GLU_GFIX_LANDAU;
  else if (gft == "COULOMB")
    gftype = GLU_GFIX_COULOMB;
  else
    return gfix_fail(-1, "GFTYPE %s is not provided by this library (LANDAU, COULOMB)", gft.c_str());
  const long long maxit = atoll(mit.c_str());
  if (maxit <= 0) return gfix_fail(-1, "MAX_ITERS %s must be positive", mit.c_str());
  unsigned seed = 0;
  auto sit = kv.find("SEED");
  if (sit != kv.end() && sscanf(sit->second.c_str(), "%u", &seed) != 1)
    return gfix_fail(-1, "SEED %s is not an unsigned integer", sit->second.c_str());
  auto rit = kv.find("RANDOM_TRANSFORM");
  const int rtrans = rit != kv.end() && rit->second == "YES";
  if (seed == 0 && (rtrans || sit != kv.end())) {  // SEED = 0: GLU reads the entropy pool
    FILE *f = fopen("/dev/urandom", "r");
    if (!f || fread(&seed, sizeof(seed), 1, f) != 1) {
      if (f) fclose(f);
      return gfix_fail(-1, "[RNG] /dev/urandom not read");
    }
    fclose(f);
    if (!seed) seed = 1;
  }
  return submit(s, dims, gftype, precision, GLU_GFIX_CG, atof(tune.c_str()), pow(10.0, -atof(acc.c_str())),
                (size_t)maxit, links_dev, ticket, seed, rtrans);
}

extern "C" int glu_gfix_server_wait(glu_gfix_server *s, int ticket, double *theta, size_t *iters,
                                    double *latency_ms) {
  if (!s) return gfix_fail(-1, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return gfix_fail(-1, "unknown ticket %d", ticket);
  if (it->second.waited) return gfix_fail(-1, "ticket %d is already being waited for", ticket);
  it->second.waited = 1;
  s->cv_done.wait(lk, [&] { return it->second.state == RQ_DONE || it->second.state == RQ_FAILED; });
  const Request r = it->second;
  s->reqs.erase(it);
  if (r.state == RQ_FAILED) return gfix_fail(r.rc, "request %d failed: %s", ticket, r.err.c_str());
  if (theta) *theta = r.theta;
  if (iters) *iters = r.iters;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  return 0;
}

extern "C" int glu_gfix_server_stats(glu_gfix_server *s, size_t *completed, double *p50_ms, double *p99_ms,
                                     size_t *peak_bytes) {
  if (!s) return gfix_fail(-1, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> v = s->lat_ms;
  std::sort(v.begin(), v.end());
  // nearest-rank percentiles
  auto pct = [&](double p) {
    if (v.empty()) return 0.0;
    size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
    return v[std::min(v.size(), std::max<size_t>(k, 1)) - 1];
  };
  if (completed) *completed = v.size();
  if (p50_ms) *p50_ms = pct(50.0);
  if (p99_ms) *p99_ms = pct(99.0);
  if (peak_bytes) *peak_bytes = s->peak;
  return 0;
}

extern "C" int glu_gfix_server_destroy(glu_gfix_server *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  s->worker.join();
  for (Pooled &p : s->pool) glu_gfix_destroy(p.ctx);
  for (auto &kv : s->reqs)
    if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
  delete s;
  return 0;
}
