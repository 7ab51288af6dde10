// Request loop over many stacks of mixed sizes on one GPU (the serving part
// of include/simr.h).  The reference reconstructs one stack per process with
// process-global state (cudaSireconImpl.h:70-82, _hh.cu:20-46); here every
// in-flight stack owns a context (stream + arena), so stacks run side by side
// and their kernels overlap on the device.
//
// Host-only code on top of the stage ABI: worker threads (one per in-flight
// slot) take requests in arrival order, skipping over those that cannot be
// admitted yet, run the per-time-point stage sequence on a context cached for
// the request's (setup, shape), and hand the context back for reuse.
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

#include "../../include/simr.h"

namespace {

double now_us() {
  using namespace std::chrono;
  return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

struct Setup {
  simr_params p;
  std::vector<float> otf;
};

struct Key {
  int setup, nx, ny, nz;
  bool operator==(const Key& o) const { return setup == o.setup && nx == o.nx && ny == o.ny && nz == o.nz; }
};

struct Slot {
  simr_ctx* ctx = nullptr;
  Key key{};
  // the key's setup, taken under the server lock when the slot is made: the
  // request path reads it without the lock while setups may still be added
  std::shared_ptr<const Setup> setup;
  size_t bytes = 0;
  bool busy = false;
  uint64_t last_use = 0;
};

struct Request {
  int64_t id = 0;
  simr_request rq{};
  simr_request_info info{};
  bool done = false;
};

struct Series {
  bool has_t0 = false;  // t = 0 submitted (later members may only be submitted after it)
  bool ready = false, failed = false;
  std::vector<unsigned char> state;
};

void set_error(simr_request_info& info, int status, const std::string& msg) {
  info.status = status;
  std::snprintf(info.error, sizeof(info.error), "%s", msg.c_str());
}

}  // namespace

struct simr_server {
  int device = 0;
  size_t budget = 0;
  std::mutex m;
  std::condition_variable cv;
  std::vector<std::shared_ptr<const Setup>> setups;  // read and grown under m only
  std::deque<Request*> queue;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  std::vector<std::unique_ptr<Slot>> slots;
  std::map<int, Series> series;
  size_t used = 0;
  int busy = 0;
  simr_server_stats st{};
  int64_t next_id = 1;
  uint64_t tick = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  static simr_params params_of(const Setup& su, const Key& k) {
    simr_params p = su.p;
    p.nx = k.nx;
    p.ny = k.ny;
    p.nz = k.nz;
    if (p.ntimes < 2) p.ntimes = 2;  // lets t > 0 of a series reuse the time-0 k0; no effect at t = 0
    return p;
  }

  void complete(Request* r) {  // under m
    r->done = true;
    r->info.t_end_us = now_us();
    if (r->info.status == SIMR_OK)
      ++st.completed;
    else
      ++st.failed;
  }

  void release(Slot* s, bool keep) {  // under m
    --busy;
    if (keep && s->ctx) {
      s->busy = false;
      s->last_use = ++tick;
      return;
    }
    if (s->ctx) simr_destroy(s->ctx);
    used -= s->bytes;
    st.bytes_in_use = used;
    slots.erase(std::find_if(slots.begin(), slots.end(), [&](const std::unique_ptr<Slot>& q) { return q.get() == s; }));
  }

  // Next admissible request and its slot (under m).  Requests that can never
  // run complete here with an error.
  Request* pick(Slot*& slot) {
    for (auto it = queue.begin(); it != queue.end();) {
      Request* r = *it;
      const simr_request& q = r->rq;
      if (q.series >= 0 && q.t > 0) {
        const Series& sr = series[q.series];
        if (sr.failed) {
          set_error(r->info, SIMR_ERR_STATE, "time 0 of this series failed");
          complete(r);
          it = queue.erase(it);
          cv.notify_all();
          continue;
        }
        if (!sr.ready) {
          ++it;
          continue;
        }
      }
      const Key key{q.setup, q.nx, q.ny, q.nz};
      bool key_busy = false;  // a context of this key is running another request
      for (auto& s : slots) {
        if (!s->busy && s->key == key) {
          s->busy = true;
          ++busy;
          slot = s.get();
          r->info.context_reused = 1;
          queue.erase(it);
          return r;
        }
        key_busy = key_busy || s->key == key;
      }
      const simr_params p = params_of(*setups[key.setup], key);
      const size_t need = simr_footprint(&p);
      if (need == 0 || need > budget) {
        if (need == 0)
          set_error(r->info, SIMR_ERR_ARG, std::string("invalid stack shape for this setup: ") + simr_last_error());
        else
          set_error(r->info, SIMR_ERR_NOMEM,
                    "stack needs " + std::to_string(need) + " bytes of HBM, more than the budget of " +
                        std::to_string(budget));
        if (q.series >= 0 && q.t == 0) series[q.series].failed = true;
        complete(r);
        it = queue.erase(it);
        cv.notify_all();
        continue;
      }
      // A second context of a key that has a running one is made only when it
      // fits as it is: evicting other keys' contexts for it costs more (their
      // re-creation) than waiting for the running one to come free.
      if (key_busy && used + need > budget) {
        ++it;
        continue;
      }
      // evict idle contexts, least recently used first, until the new one fits
      while (used + need > budget) {
        Slot* lru = nullptr;
        for (auto& s : slots)
          if (!s->busy && (!lru || s->last_use < lru->last_use)) lru = s.get();
        if (!lru) break;
        ++busy;  // release() counts it back
        release(lru, false);
        ++st.contexts_evicted;
      }
      if (used + need <= budget) {
        auto s = std::make_unique<Slot>();
        s->key = key;
        s->setup = setups[key.setup];
        s->bytes = need;
        s->busy = true;
        used += need;
        st.bytes_in_use = used;
        st.bytes_peak = std::max(st.bytes_peak, used);
        ++busy;
        slot = s.get();
        slots.push_back(std::move(s));
        queue.erase(it);
        return r;
      }
      ++it;  // not admissible yet; a later, smaller request may be
    }
    return nullptr;
  }

  // Runs one request on its slot (without m).  Returns false when the context
  // must not be reused.
  bool run(Request* r, Slot* s) {
    simr_request_info& info = r->info;
    const simr_request& q = r->rq;
    const int t = q.series >= 0 ? q.t : 0;
    auto fail = [&](int rc) {
      set_error(info, rc, simr_last_error());
      return false;
    };
    int rc;
    if (!s->ctx) {
      const simr_params p = params_of(*s->setup, s->key);
      rc = simr_create(&p, s->setup->otf.data(), device, &s->ctx);
      if (rc) return fail(rc);
      std::lock_guard<std::mutex> lk(m);
      ++st.contexts_created;
    }
    if (q.raw_u16)
      rc = q.raw_on_device ? simr_set_raw_u16_dev(s->ctx, (const uint16_t*)q.raw)
                           : simr_set_raw_u16_host(s->ctx, (const uint16_t*)q.raw);
    else
      rc = q.raw_on_device ? simr_set_raw_f32_dev(s->ctx, (const float*)q.raw)
                           : simr_set_raw_f32_host(s->ctx, (const float*)q.raw);
    if (rc) return fail(rc);
    if (q.series >= 0 && t > 0) {
      std::vector<unsigned char> state;
      {
        std::lock_guard<std::mutex> lk(m);
        state = series[q.series].state;
      }
      rc = simr_set_time0_state(s->ctx, state.data(), state.size());
      if (rc) return fail(rc);
    }
    if ((rc = simr_load_and_rescale(s->ctx, t))) return fail(rc);
    if ((rc = simr_set_time_index(s->ctx, t))) return fail(rc);
    if ((rc = simr_find_modulation(s->ctx))) return fail(rc);
    if ((rc = simr_filter_assemble(s->ctx))) return fail(rc);
    if (q.series >= 0 && t == 0) {
      std::vector<unsigned char> state(simr_time0_state_size(s->ctx));
      if ((rc = simr_get_time0_state(s->ctx, state.data(), state.size()))) return fail(rc);
      std::lock_guard<std::mutex> lk(m);
      Series& sr = series[q.series];
      sr.state = std::move(state);
      sr.ready = true;
    }
    info.out_elems = simr_result_elems(s->ctx);
    if (q.out) {
      rc = q.out_on_device ? simr_get_result_dev(s->ctx, q.out) : simr_get_result_host(s->ctx, q.out);
      if (rc) return fail(rc);
    } else if ((rc = simr_sync(s->ctx))) {
      return fail(rc);
    }
    const int ndirs = s->setup->p.ndirs;
    for (int d = 0; d < ndirs && d < SIMR_MAX_DIRS; ++d)
      if ((rc = simr_get_fit(s->ctx, d, &info.fit[d]))) return fail(rc);
    info.status = SIMR_OK;
    return true;
  }

  void worker() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      Request* r = nullptr;
      Slot* s = nullptr;
      cv.wait(lk, [&] {
        if (stop) return true;
        r = pick(s);
        return r != nullptr;
      });
      if (!r) return;  // stop
      r->info.t_start_us = now_us();
      lk.unlock();
      const bool ok = run(r, s);
      lk.lock();
      if (!ok && r->rq.series >= 0 && r->rq.t == 0) series[r->rq.series].failed = true;
      release(s, ok);
      complete(r);
      cv.notify_all();
    }
  }
};

extern "C" {

int simr_server_create(int device, size_t hbm_budget, int max_inflight, simr_server** out) {
  if (!out || max_inflight < 1 || hbm_budget == 0) {
    simr_set_last_error("simr_server_create: need an output pointer, max_inflight >= 1 and a non-zero budget");
    return SIMR_ERR_ARG;
  }
  auto* s = new simr_server;
  s->device = device;
  s->budget = hbm_budget;
  for (int i = 0; i < max_inflight; ++i) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return SIMR_OK;
}

int simr_server_add_setup(simr_server* srv, const simr_params* p, const float* otf, int* setup_id) {
  if (!srv || !p || !otf || !setup_id) {
    simr_set_last_error("simr_server_add_setup: null argument");
    return SIMR_ERR_ARG;
  }
  const size_t n = simr_otf_total(p);
  if (n == 0) {
    simr_set_last_error("simr_server_add_setup: empty OTF for these parameters");
    return SIMR_ERR_ARG;
  }
  auto su = std::make_shared<Setup>();
  su->p = *p;
  su->otf.assign(otf, otf + 2 * n);
  std::lock_guard<std::mutex> lk(srv->m);
  srv->setups.push_back(std::move(su));
  *setup_id = (int)srv->setups.size() - 1;
  return SIMR_OK;
}

int simr_server_submit(simr_server* srv, const simr_request* rq, int64_t* ticket) {
  if (!srv || !rq || !ticket || !rq->raw) {
    simr_set_last_error("simr_server_submit: null argument");
    return SIMR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(srv->m);
  if (rq->setup < 0 || rq->setup >= (int)srv->setups.size()) {
    simr_set_last_error("simr_server_submit: unknown setup id");
    return SIMR_ERR_ARG;
  }
  if (rq->series >= 0) {
    // t = 0 first, exactly once: a later member then always has a t = 0
    // that is queued, running or finished to wait for
    Series& sr = srv->series[rq->series];
    if (rq->t < 0 || (rq->t == 0 && sr.has_t0) || (rq->t > 0 && !sr.has_t0)) {
      simr_set_last_error(rq->t < 0    ? "simr_server_submit: negative time index"
                          : rq->t == 0 ? "simr_server_submit: time 0 of this series was already submitted"
                                       : "simr_server_submit: submit time 0 of a series before its later time points");
      return SIMR_ERR_ARG;
    }
    if (rq->t == 0) sr.has_t0 = true;
  }
  auto r = std::make_unique<Request>();
  r->id = srv->next_id++;
  r->rq = *rq;
  r->info.status = SIMR_ERR_STATE;
  r->info.t_submit_us = now_us();
  srv->queue.push_back(r.get());
  *ticket = r->id;
  srv->reqs[r->id] = std::move(r);
  ++srv->st.submitted;
  srv->cv.notify_all();
  return SIMR_OK;
}

int simr_server_wait(simr_server* srv, int64_t ticket, simr_request_info* info) {
  if (!srv) {
    simr_set_last_error("simr_server_wait: null server");
    return SIMR_ERR_ARG;
  }
  std::unique_lock<std::mutex> lk(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) {
    simr_set_last_error("simr_server_wait: unknown or already collected ticket");
    return SIMR_ERR_ARG;
  }
  Request* r = it->second.get();
  srv->cv.wait(lk, [&] { return r->done; });
  const int status = r->info.status;
  if (info) *info = r->info;
  if (status != SIMR_OK) simr_set_last_error(r->info.error);
  srv->reqs.erase(it);
  return status;
}

int simr_server_stats_get(simr_server* srv, simr_server_stats* st) {
  if (!srv || !st) {
    simr_set_last_error("simr_server_stats_get: null argument");
    return SIMR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(srv->m);
  *st = srv->st;
  return SIMR_OK;
}

int simr_server_destroy(simr_server* srv) {
  if (!srv) return SIMR_OK;
  {
    std::unique_lock<std::mutex> lk(srv->m);
    srv->cv.wait(lk, [&] { return srv->queue.empty() && srv->busy == 0; });
    srv->stop = true;
    srv->cv.notify_all();
  }
  for (auto& t : srv->workers) t.join();
  for (auto& s : srv->slots)
    if (s->ctx) simr_destroy(s->ctx);
  delete srv;
  return SIMR_OK;
}

}  // extern "C"
