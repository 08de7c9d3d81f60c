// Serving: a request loop over resident ensembles (include/piafs_amd.h,
// piafs_server_*).
//
// Requests (solves of a deck directory to its n_iter steps) are parsed on the
// submitting thread and queued.  Each class of cases (1D, 2D) has a lane: an
// engine with its own stream and worker thread.  A lane admits queued
// requests of its class while the device memory they would hold fits the
// server's budget, keeps their state resident, and advances all its members
// together (one launch per kernel class and RK stage covering every member,
// replayed from a captured graph).  Members join and leave between rounds:
// a round advances every member by the same number of steps, never past the
// step count of the member closest to completion, so each result is exactly
// that of the case run alone.  The lanes run concurrently on the device.
// The reference has no counterpart (it runs one simulation, or a serial loop
// over an ensemble: src/Simulation/EnsembleSimulationsDefine.cpp:31-119).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/piafs_amd.h"
#include "case.hpp"

using namespace piafs;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int id = 0;
  std::unique_ptr<Case> c;
  size_t bytes = 0;
  long cells = 0;       // interior points
  int lane = 0;
  int target = 0;       // steps to take (n_iter - restart_iter)
  int done = 0;         // steps taken
  int state = 0;        // 0 queued, 1 running, 2 done, < 0 failed
  double t_submit = 0, t_start = 0, t_done = 0;
  std::vector<double> result;
  piafs_info info{};
  std::string err;
};

}  // namespace

struct piafs_server {
  int device = 0;
  size_t budget = 0, in_use = 0;
  double round_ms = 4.0;  // target device time of one round of a lane
  // interior points advanced together in one lane beyond which a lane admits
  // no further request (a single 2D case of about 700^2 already fills the
  // device; more members would only share it)
  long inflight_cells[2] = {1L << 22, 1L << 19};
  std::mutex mu;       // request table, queue and budget
  std::mutex dev_mu;   // device memory management, plan builds, graph capture
  std::condition_variable cv;
  std::deque<Request *> pending;
  std::map<int, std::unique_ptr<Request>> reqs;
  int next_id = 0;
  bool stop = false;
  std::string err;
  struct Lane {
    Engine eng;
    std::vector<Request *> members;
    double ms_per_step = 0.0;  // moving estimate of one step of the current member set
    std::thread th;
  } lanes[2];

  void run_lane(int L);
};

// Lane loop.  Device memory management, plan building, graph capture and
// result downloads of the two lanes never overlap (dev_mu): an allocation,
// free or synchronous copy issued by one thread while the other is capturing
// its step graph can invalidate that capture.  Graph replays and waits run
// concurrently.
void piafs_server::run_lane(int L) {
  Lane &ln = lanes[L];
  hipSetDevice(device);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  auto finish = [&](Request *r, int state, const std::string &msg) {  // called with mu held
    r->state = state;
    if (state < 0) r->err = msg;
    r->t_done = now_s();
    in_use -= r->bytes;
    cv.notify_all();
  };
  // rebuild the plan for ln.members (dev_mu held); on failure every member fails
  auto rebuild = [&]() {
    ln.eng.cases.clear();
    for (Request *r : ln.members) ln.eng.cases.push_back(r->c.get());
    int rc = ln.eng.build();
    if (!rc) rc = ln.eng.prepare();
    ln.ms_per_step = 0.0;
    if (!rc) return true;
    std::vector<Request *> failed = ln.members;
    ln.members.clear();
    ln.eng.cases.clear();
    ln.eng.build();
    {
      std::lock_guard<std::mutex> lk(mu);
      err = ln.eng.err;
      for (Request *r : failed) finish(r, -1, "ensemble setup: " + ln.eng.err);
    }
    for (Request *r : failed) r->c.reset();
    return false;
  };
  for (;;) {
    // admission: queued requests of this lane, in arrival order, while they fit
    std::vector<Request *> admit;
    {
      std::unique_lock<std::mutex> lk(mu);
      if (ln.members.empty())
        cv.wait(lk, [&] {
          if (stop) return true;
          for (Request *r : pending)
            if (r->lane == L) return true;
          return false;
        });
      if (stop) break;
      // arrival order within the lane; a request that does not fit stops the
      // admission (no overtaking), except into an empty lane
      long cells = 0;
      for (Request *r : ln.members) cells += r->cells;
      for (auto it = pending.begin(); it != pending.end();) {
        Request *r = *it;
        if (r->lane != L) {
          ++it;
          continue;
        }
        const bool empty = ln.members.empty() && admit.empty();
        if (in_use + r->bytes > budget || (!empty && cells + r->cells > inflight_cells[L])) break;
        in_use += r->bytes;
        cells += r->cells;
        admit.push_back(r);
        it = pending.erase(it);
      }
    }
    if (!admit.empty()) {
      std::lock_guard<std::mutex> dk(dev_mu);
      for (Request *r : admit) {
        const int rc = r->c->setup(device, ln.eng.stream);
        if (rc) {
          const std::string msg = r->c->err;
          r->c.reset();
          std::lock_guard<std::mutex> lk(mu);
          finish(r, -1, msg);
          continue;
        }
        {
          std::lock_guard<std::mutex> lk(mu);
          r->state = 1;
          r->t_start = now_s();
        }
        ln.members.push_back(r);
      }
      rebuild();
    }
    if (ln.members.empty()) continue;
    // one round: the same number of steps for every member
    int k = 1 << 20;
    for (Request *r : ln.members) k = std::min(k, r->target - r->done);
    if (ln.ms_per_step > 0.0) k = std::min(k, std::max(1, (int)(round_ms / ln.ms_per_step)));
    else k = std::min(k, 1);
    hipEventRecord(e0, ln.eng.stream);
    int rc = ln.eng.step(k);  // graph replays of the captured step
    hipEventRecord(e1, ln.eng.stream);
    if (!rc) rc = ln.eng.sync();
    if (!rc) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      const double per = ms / k;
      ln.ms_per_step = ln.ms_per_step > 0.0 ? 0.5 * (ln.ms_per_step + per) : per;
    }
    // completions (and failures)
    std::vector<Request *> keep, gone;
    for (Request *r : ln.members) {
      if (!rc) r->done += k;
      if (!rc && r->done < r->target) keep.push_back(r);
      else gone.push_back(r);
    }
    if (gone.empty()) continue;
    std::lock_guard<std::mutex> dk(dev_mu);
    for (Request *r : gone) {
      int drc = rc;
      std::string msg = ln.eng.err;
      if (!rc) {
        const Config &cf = r->c->cfg;
        r->result.resize((size_t)cf.npts_int * cf.nvars);
        drc = r->c->download(r->result.data(), ln.eng.stream);
        if (drc) msg = r->c->err;
        fill_info(r->c.get(), &r->info);
      }
      std::lock_guard<std::mutex> lk(mu);
      finish(r, drc ? -1 : 2, msg);
    }
    // the engine forgets the finished members before their memory is released
    ln.members = keep;
    rebuild();
    for (Request *r : gone) r->c.reset();
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
}

extern "C" {

int piafs_server_create(int device, size_t hbm_budget, piafs_server **out) {
  if (!out) return 1;
  *out = nullptr;
  piafs_server *s = new piafs_server();
  s->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete s;
    return -1;
  }
  if (hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    hbm_budget = fr / 10 * 8;
  }
  s->budget = hbm_budget;
  for (int L = 0; L < 2; L++) {
    auto &ln = s->lanes[L];
    ln.eng.owns_cases = false;
    ln.eng.use_graph = true;
    if (ln.eng.init(device)) {
      s->err = ln.eng.err;
      *out = s;
      return -1;
    }
  }
  for (int L = 0; L < 2; L++) s->lanes[L].th = std::thread([s, L] { s->run_lane(L); });
  *out = s;
  return 0;
}

int piafs_server_submit(piafs_server *s, const char *dir, int *id) {
  if (!s) return 1;  // no server
  auto r = std::make_unique<Request>();
  r->t_submit = now_s();
  r->c = std::make_unique<Case>();
  std::string err;
  if (r->c->cfg.parse(dir ? dir : "", err)) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = err;
    return 1;
  }
  const Config &cf = r->c->cfg;
  r->bytes = Case::plan_bytes(cf);
  r->lane = cf.ndims == 1 ? 0 : 1;
  r->cells = (long)cf.npts_int;
  r->target = cf.n_iter - cf.restart_iter;
  if (r->target < 0) r->target = 0;
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->bytes > s->budget) {
    s->err = "request needs " + std::to_string(r->bytes) + " bytes of device memory, above the server budget of " +
             std::to_string(s->budget);
    return 2;
  }
  r->id = s->next_id++;
  if (id) *id = r->id;
  Request *p = r.get();
  s->reqs[r->id] = std::move(r);
  if (p->target == 0) {
    fill_info(p->c.get(), &p->info);
    p->state = 2;
    p->t_done = p->t_submit;
    return 0;  // nothing to integrate; the result is read from the parsed state
  }
  s->pending.push_back(p);
  s->cv.notify_all();
  return 0;
}

int piafs_server_poll(piafs_server *s, int id, int *state, double *latency) {
  if (!s) return 1;  // no server
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return 1;
  Request *r = it->second.get();
  if (state) *state = r->state;
  if (r->state < 0) s->err = r->err;
  if (latency) *latency = (r->state == 2 || r->state < 0) ? r->t_done - r->t_submit : -1.0;
  return 0;
}

int piafs_server_result(piafs_server *s, int id, double *u_host, piafs_info *info) {
  if (!s) return 1;  // no server
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return 1;
  Request *r = it->second.get();
  if (r->state < 0) {
    s->err = r->err;
    return -1;
  }
  if (r->state != 2) return 2;
  if (r->target == 0) {
    // no steps: the initial state, already in WriteBinary order
    const Config &cf = r->c->cfg;
    if (u_host) std::copy(cf.u0.begin(), cf.u0.end(), u_host);
    if (info) *info = r->info;
    return 0;
  }
  if (u_host) std::copy(r->result.begin(), r->result.end(), u_host);
  if (info) *info = r->info;
  return 0;
}

int piafs_server_release(piafs_server *s, int id) {
  if (!s) return 1;  // no server
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return 1;
  if (it->second->state == 0 || it->second->state == 1) return 2;
  s->reqs.erase(it);
  return 0;
}

int piafs_server_wait_all(piafs_server *s) {
  if (!s) return 1;  // no server
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] {
    for (auto &kv : s->reqs)
      if (kv.second->state == 0 || kv.second->state == 1) return false;
    return true;
  });
  return 0;
}

int piafs_server_stats(piafs_server *s, size_t *bytes_in_use, int *running, int *queued) {
  if (!s) return 1;  // no server
  std::lock_guard<std::mutex> lk(s->mu);
  if (bytes_in_use) *bytes_in_use = s->in_use;
  int run = 0, q = 0;
  for (auto &kv : s->reqs) {
    run += kv.second->state == 1;
    q += kv.second->state == 0;
  }
  if (running) *running = run;
  if (queued) *queued = q;
  return 0;
}

const char *piafs_server_last_error(const piafs_server *s) { return s ? s->err.c_str() : ""; }

int piafs_server_destroy(piafs_server *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  for (auto &ln : s->lanes)
    if (ln.th.joinable()) ln.th.join();
  for (auto &ln : s->lanes) {
    ln.eng.sync();
    ln.eng.cases.clear();
    ln.members.clear();
  }
  delete s;
  return 0;
}

}  // extern "C"
