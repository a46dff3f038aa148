// phaser_serve.cc -- continuous-batching request loop over the phaser_amd C
// ABI (include/phaser_serve.h).  Host C++ only: every device operation goes
// through phz_* calls, one context (stream + resident tables) per worker.
//
// Scheduling, all under one mutex:
//  * requests queue per shape (bandwidth, n_voxels);
//  * each context has a worker thread that takes up to max_batch queued
//    requests of its shape whenever it is free, so arrivals during a batch
//    form the next one and contexts of different shapes overlap on the GPU;
//  * shapes are admitted oldest-request first; a shape with no context gets
//    one if its phz_estimate_memory bytes fit the budget, else contexts of
//    shapes with younger (or no) queued work are retired after their current
//    batch until it fits; extra contexts for a busy shape are only added
//    from memory that is free and not claimed by a waiting shape.
#include "phaser_serve.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;
thread_local std::string g_error;

int fail(int rc, const std::string& msg) {
  g_error = msg;
  return rc;
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  uint64_t ticket;
  phs_request req;
  Clock::time_point submitted;
};

struct Shape;

struct Lane {
  Shape* shape = nullptr;
  std::thread th;
  size_t bytes = 0;
  bool ready = false;   // context built
  bool busy = false;    // a batch is on the device
  bool retire = false;  // exit after the current batch
  Clock::time_point last_used{};
};

struct Shape {
  int B = 0, N = 0;
  size_t bytes = 0;  // per context
  bool no_extra = false;  // building an extra context failed once
  std::deque<Request> queue;
  std::vector<Lane*> lanes;
};

}  // namespace

struct phs_server {
  phs_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work;  // workers: queue changed or retire
  std::condition_variable cv_done;  // callers: responses or lanes changed
  std::map<std::pair<int, int>, std::unique_ptr<Shape>> shapes;
  std::map<uint64_t, phs_response> done;
  std::map<uint64_t, bool> outstanding;  // submitted, response not yet returned
  std::vector<std::thread> reap;         // exited workers, joined outside the lock
  uint64_t next_ticket = 1;
  size_t bytes_reserved = 0;
  int lanes_live = 0;
  bool stopping = false;
  phs_stats st{};
};

namespace {

void worker(phs_server* s, Lane* L);

phz_config shape_config(const phs_config& sc, int B, int N) {
  phz_config c;
  phz_default_config(&c);
  c.bandwidth = B;
  c.n_voxels = N;
  c.discretize_lower = sc.discretize_lower;
  c.discretize_upper = sc.discretize_upper;
  c.max_points = sc.max_points;
  c.max_batch = sc.max_batch;
  c.device = sc.device;
  c.precision = sc.precision;
  c.bingham_peak_neighbors = sc.bingham_peak_neighbors;
  c.gaussian_peak_neighbors = sc.gaussian_peak_neighbors;
  c.method = sc.method;
  c.skip_rot_axes = sc.skip_rot_axes;
  return c;
}

Clock::time_point age(const Shape& sh) {
  return sh.queue.empty() ? Clock::time_point::max() : sh.queue.front().submitted;
}

int active_lanes(const Shape& sh) {
  int k = 0;
  for (const Lane* l : sh.lanes) k += !l->retire;
  return k;
}

void add_lane(phs_server* s, Shape* sh) {
  Lane* L = new Lane();
  L->shape = sh;
  L->bytes = sh->bytes;
  L->last_used = Clock::now();
  sh->lanes.push_back(L);
  s->bytes_reserved += L->bytes;
  s->st.bytes_peak = std::max<uint64_t>(s->st.bytes_peak, s->bytes_reserved);
  ++s->lanes_live;
  L->th = std::thread(worker, s, L);
}

// Admission and eviction; called with the lock held after every event.
void schedule_locked(phs_server* s) {
  std::vector<Shape*> waiting;
  for (auto& kv : s->shapes)
    if (!kv.second->queue.empty()) waiting.push_back(kv.second.get());
  std::sort(waiting.begin(), waiting.end(),
            [](const Shape* a, const Shape* b) { return age(*a) < age(*b); });
  size_t pending_release = 0;
  for (auto& kv : s->shapes)
    for (const Lane* l : kv.second->lanes)
      if (l->retire) pending_release += l->bytes;
  bool claimed = false;  // an older shape is waiting for memory
  for (Shape* sh : waiting) {
    const int nr = active_lanes(*sh);
    if (nr == 0) {
      if (s->bytes_reserved + sh->bytes <= s->cfg.hbm_budget) {
        if (!claimed) add_lane(s, sh);
        continue;
      }
      if (claimed) continue;
      claimed = true;
      // retire contexts of shapes with younger or no queued work: idle ones
      // first, then least recently used
      std::vector<Lane*> victims;
      for (auto& kv : s->shapes) {
        Shape* o = kv.second.get();
        if (o == sh || age(*o) <= age(*sh)) continue;
        for (Lane* l : o->lanes)
          if (l->ready && !l->retire) victims.push_back(l);
      }
      std::sort(victims.begin(), victims.end(), [](const Lane* a, const Lane* b) {
        if (a->busy != b->busy) return !a->busy;
        return a->last_used < b->last_used;
      });
      for (Lane* v : victims) {
        if (s->bytes_reserved - pending_release + sh->bytes <= s->cfg.hbm_budget) break;
        v->retire = true;
        pending_release += v->bytes;
      }
      s->cv_work.notify_all();
    } else if (!claimed && !sh->no_extra && nr < s->cfg.contexts_per_shape) {
      bool all_busy = true;
      for (const Lane* l : sh->lanes)
        if (!l->retire && (!l->ready || !l->busy)) all_busy = false;
      if (all_busy && s->bytes_reserved + sh->bytes <= s->cfg.hbm_budget) add_lane(s, sh);
    }
  }
}

void finish_locked(phs_server* s, const Request& r, int status, const std::string& err,
                   const phz_result* res, int pairs, bool shared, Clock::time_point t0,
                   Clock::time_point t1) {
  phs_response out;
  std::memset(&out, 0, sizeof(out));
  out.ticket = r.ticket;
  out.tag = r.req.tag;
  out.status = status;
  out.batch_pairs = pairs;
  out.shared_target = shared ? 1 : 0;
  if (res) out.result = *res;
  out.queue_ms = ms_between(r.submitted, t0);
  out.service_ms = ms_between(t0, t1);
  out.latency_ms = ms_between(r.submitted, t1);
  std::snprintf(out.error, sizeof(out.error), "%s", err.c_str());
  s->done[r.ticket] = out;
  if (status == 0)
    ++s->st.completed;
  else
    ++s->st.failed;
}

// Takes the next batch of sh's queue.  If the oldest request's target key
// is shared by at least two queued requests, the batch is that group (one
// target, spectra computed once).  Otherwise it is the oldest max_batch
// requests in arrival order, leaving out requests whose key group has two
// or more members: those wait for a batch of their own group.
bool take_batch(phs_server* s, Shape* sh, std::vector<Request>* batch) {
  batch->clear();
  const size_t maxb = (size_t)s->cfg.max_batch;
  std::map<uint64_t, int> group;
  for (const Request& r : sh->queue)
    if (r.req.target_key != 0) ++group[r.req.target_key];
  const uint64_t key = sh->queue.front().req.target_key;
  const bool shared = key != 0 && group[key] >= 2;
  std::deque<Request> rest;
  for (Request& r : sh->queue) {
    const uint64_t k = r.req.target_key;
    const bool take = batch->size() < maxb &&
                      (shared ? k == key : (k == 0 || group[k] < 2));
    if (take)
      batch->push_back(r);
    else
      rest.push_back(r);
  }
  sh->queue.swap(rest);
  return shared;
}

void worker(phs_server* s, Lane* L) {
  Shape* sh = L->shape;
  const phz_config cfg = shape_config(s->cfg, sh->B, sh->N);
  phz_ctx* ctx = nullptr;
  int crc = phz_create(&cfg, &ctx);
  if (!crc && s->cfg.autotune) crc = phz_autotune(ctx, 3, nullptr);
  if (!crc && s->cfg.graphs) crc = phz_set_graphs(ctx, 1);
  if (crc && ctx) {
    phz_destroy(ctx);
    ctx = nullptr;
  }
  const std::string cerr = crc ? std::string(phz_last_error()) : std::string();
  std::unique_lock<std::mutex> lk(s->mu);
  if (crc) {
    // a shape whose context cannot be built fails its queued requests,
    // unless another context of the shape is serving them
    bool served = false;
    for (const Lane* l : sh->lanes) served |= l != L && l->ready && !l->retire;
    sh->no_extra = true;
    if (!served) {
      const Clock::time_point now = Clock::now();
      for (const Request& r : sh->queue)
        finish_locked(s, r, crc, "context creation failed: " + cerr, nullptr, 0, false, now, now);
      sh->queue.clear();
    }
    L->retire = true;
  } else {
    L->ready = true;
    ++s->st.contexts_created;
  }
  std::vector<Request> batch;
  std::vector<const float*> tg, sr;
  std::vector<int64_t> pt, ps;
  std::vector<phz_result> res;
  while (ctx) {
    s->cv_work.wait(lk, [&] { return L->retire || !sh->queue.empty(); });
    if (L->retire) break;
    if (s->cfg.batch_window_ms > 0 && (int)sh->queue.size() < s->cfg.max_batch && !s->stopping) {
      const auto until =
          Clock::now() + std::chrono::microseconds((int64_t)(s->cfg.batch_window_ms * 1000.0));
      s->cv_work.wait_until(lk, until, [&] {
        return L->retire || s->stopping || (int)sh->queue.size() >= s->cfg.max_batch;
      });
      if (L->retire) break;
      if (sh->queue.empty()) continue;
    }
    const bool shared = take_batch(s, sh, &batch);
    L->busy = true;
    schedule_locked(s);  // the queue shrank; a waiting shape may now claim memory
    lk.unlock();

    const int n = (int)batch.size();
    tg.clear();
    sr.clear();
    pt.clear();
    ps.clear();
    res.assign(n, phz_result{});
    for (int i = 0; i < n; ++i) {
      if (!shared || i == 0) {
        tg.push_back(batch[i].req.target);
        pt.push_back(batch[i].req.Pt);
      }
      sr.push_back(batch[i].req.source);
      ps.push_back(batch[i].req.Ps);
    }
    const Clock::time_point t0 = Clock::now();
    const int rc = phz_register_pairs(ctx, n, shared ? 1 : 0, tg.data(), pt.data(), sr.data(),
                                      ps.data(), res.data());
    const Clock::time_point t1 = Clock::now();
    const std::string err = rc ? std::string(phz_last_error()) : std::string();

    lk.lock();
    for (int i = 0; i < n; ++i) {
      if (rc == -4) {
        // the registration ran; only the pairs that failed one of the
        // reference's peak-window checks fail (their results say which)
        const bool bad = res[i].status != 0;
        finish_locked(s, batch[i], bad ? -4 : 0,
                      bad ? "a peak-window check failed (phz_result.status " +
                                std::to_string(res[i].status) + ")"
                          : std::string(),
                      &res[i], n, shared, t0, t1);
      } else {
        finish_locked(s, batch[i], rc, err, rc ? nullptr : &res[i], n, shared, t0, t1);
      }
    }
    ++s->st.batches;
    s->st.pairs_in_batches += (uint64_t)n;
    L->busy = false;
    L->last_used = t1;
    if (rc == -2) L->retire = true;  // a runtime error may leave the context unusable
    s->cv_done.notify_all();
    schedule_locked(s);
  }
  lk.unlock();
  if (ctx) phz_destroy(ctx);
  lk.lock();
  s->bytes_reserved -= L->bytes;
  if (L->ready) ++s->st.contexts_released;
  sh->lanes.erase(std::find(sh->lanes.begin(), sh->lanes.end(), L));
  s->reap.push_back(std::move(L->th));
  delete L;
  --s->lanes_live;
  schedule_locked(s);
  s->cv_work.notify_all();
  s->cv_done.notify_all();
}

void join_reaped(phs_server* s) {
  std::vector<std::thread> r;
  {
    std::lock_guard<std::mutex> g(s->mu);
    r.swap(s->reap);
  }
  for (auto& t : r) t.join();
}

}  // namespace

extern "C" {

void phs_default_config(phs_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->max_batch = 16;
  cfg->contexts_per_shape = 2;
  cfg->discretize_lower = -50;
  cfg->discretize_upper = 50;
  cfg->max_points = 262144;
  cfg->hbm_budget = 64ull << 30;
  cfg->batch_window_ms = 0.0;
  cfg->graphs = 1;
}

const char* phs_last_error(void) { return g_error.c_str(); }

int phs_create(const phs_config* cfg, phs_server** out) {
  if (!cfg || !out) return fail(-1, "null argument");
  *out = nullptr;
  if (cfg->max_batch < 1 || cfg->max_batch > 1024) return fail(-1, "max_batch out of range");
  if (cfg->contexts_per_shape < 1 || cfg->contexts_per_shape > 16)
    return fail(-1, "contexts_per_shape out of range");
  if (cfg->max_points < 1) return fail(-1, "max_points out of range");
  if (cfg->discretize_lower >= cfg->discretize_upper)
    return fail(-1, "discretize bounds must be increasing");
  if (!(cfg->batch_window_ms >= 0.0)) return fail(-1, "batch_window_ms must be >= 0");
  phs_server* s = new phs_server();
  s->cfg = *cfg;
  *out = s;
  return 0;
}

int phs_submit(phs_server* s, const phs_request* req, uint64_t* ticket) {
  if (!s || !req || !ticket) return fail(-1, "null argument");
  if (req->Pt < 0 || req->Ps < 0 || req->Pt > s->cfg.max_points || req->Ps > s->cfg.max_points)
    return fail(-3, "point count outside [0, max_points]");
  if ((req->Pt > 0 && !req->target) || (req->Ps > 0 && !req->source))
    return fail(-1, "null point buffer");
  const phz_config c = shape_config(s->cfg, req->bandwidth, req->n_voxels);
  size_t bytes = 0;
  if (int rc = phz_estimate_memory(&c, &bytes)) return fail(rc, phz_last_error());
  if (bytes > s->cfg.hbm_budget)
    return fail(-3, "a context for bandwidth " + std::to_string(req->bandwidth) + ", " +
                        std::to_string(req->n_voxels) + " voxels needs " + std::to_string(bytes) +
                        " bytes, more than hbm_budget");
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stopping) return fail(-1, "server is shutting down");
    auto& slot = s->shapes[{req->bandwidth, req->n_voxels}];
    if (!slot) {
      slot.reset(new Shape());
      slot->B = req->bandwidth;
      slot->N = req->n_voxels;
      slot->bytes = bytes;
    }
    Request r;
    r.ticket = s->next_ticket++;
    r.req = *req;
    r.submitted = Clock::now();
    slot->queue.push_back(r);
    s->outstanding[r.ticket] = true;
    ++s->st.submitted;
    *ticket = r.ticket;
    schedule_locked(s);
  }
  s->cv_work.notify_all();
  join_reaped(s);
  return 0;
}

int phs_wait(phs_server* s, uint64_t ticket, phs_response* out) {
  if (!s || !out) return fail(-1, "null argument");
  {
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->outstanding.count(ticket)) return fail(-1, "unknown or already returned ticket");
    s->cv_done.wait(lk, [&] { return s->done.count(ticket) > 0; });
    *out = s->done[ticket];
    s->done.erase(ticket);
    s->outstanding.erase(ticket);
  }
  join_reaped(s);
  return 0;
}

int phs_poll(phs_server* s, phs_response* out, int max, int* count, int wait_ms) {
  if (!s || !count || (max > 0 && !out)) return fail(-1, "null argument");
  *count = 0;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->done.empty() || s->done.size() == s->outstanding.size(); };
    if (wait_ms < 0)
      s->cv_done.wait(lk, ready);
    else if (wait_ms > 0)
      s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    while (*count < max && !s->done.empty()) {
      auto it = s->done.begin();
      out[(*count)++] = it->second;
      s->outstanding.erase(it->first);
      s->done.erase(it);
    }
  }
  join_reaped(s);
  return 0;
}

int phs_drain(phs_server* s) {
  if (!s) return fail(-1, "null argument");
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->st.completed + s->st.failed == s->st.submitted; });
  }
  join_reaped(s);
  return 0;
}

int phs_get_stats(phs_server* s, phs_stats* out) {
  if (!s || !out) return fail(-1, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->st;
  out->bytes_in_use = s->bytes_reserved;
  out->contexts_live = s->lanes_live;
  int q = 0;
  for (auto& kv : s->shapes) q += (int)kv.second->queue.size();
  out->queued = q;
  return 0;
}

void phs_destroy(phs_server* s) {
  if (!s) return;
  phs_drain(s);
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->stopping = true;
    for (auto& kv : s->shapes)
      for (Lane* l : kv.second->lanes) l->retire = true;
    s->cv_work.notify_all();
    s->cv_done.wait(lk, [&] { return s->lanes_live == 0; });
  }
  join_reaped(s);
  delete s;
}

}  // extern "C"
