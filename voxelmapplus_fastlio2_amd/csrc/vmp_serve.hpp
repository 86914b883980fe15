// Continuous-batching server (include/vmp_serve.h).  Host C++ over the batch
// step of vmp_capi.hip: included at the end of that file so it can drive
// run_step_on / enqueue_motion on its own executor contexts.
#pragma once
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <vector>

#include "vmp_serve.h"

struct ServSess {
  vmp_session* s = nullptr;  // resident session, or nullptr
  vmp_state x{};             // host copy: initial state or the last posterior
  std::vector<double> P;     // 23x23 column-major
  bool has_state = false;    // vmp_server_session_init was called
  bool built = false;        // the resident map holds a MAP_INIT
  bool seen = false;         // at least one scan was processed
  bool busy = false;         // in a batch in flight
  double last_use = 0.0;
  uint64_t mark = 0;         // batch-formation epoch that visited it
};

struct vmp_server {
  vmp_server_cfg cfg{};
  vmp_map_cfg scfg{};
  vmp_ctx* home = nullptr;         // allocation / admission context
  std::vector<vmp_ctx*> ex;        // executors, one stream each
  std::vector<ServSess> sess;
  int64_t sess_bytes = 0;
  int64_t resident_bytes = 0;
  int32_t resident = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::string err;
  // run state
  vmp_request* reqs = nullptr;
  int32_t nreq = 0, next = 0, inflight = 0, batch_id = 0;
  std::deque<int32_t> queue;
  std::chrono::steady_clock::time_point t0;
  double q[23]{};
  int first_err = VMP_OK;
  uint64_t epoch = 0;
  // counters of the last run
  int64_t st_batches = 0, st_requests = 0, st_evictions = 0, st_admissions = 0, st_maxbatch = 0;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// Make session `id` resident (sv->mu held).  Evicts least recently used idle
// sessions while the budget is exceeded.  Returns false when no memory can be
// freed now (every resident session is busy).
static bool serve_admit(vmp_server* sv, int32_t id, int* rc) {
  ServSess& t = sv->sess[size_t(id)];
  if (t.s) return true;
  while (sv->resident_bytes + sv->sess_bytes > sv->cfg.hbm_budget) {
    int32_t victim = -1;
    for (size_t i = 0; i < sv->sess.size(); ++i) {
      const ServSess& c = sv->sess[i];
      if (c.s && !c.busy && (victim < 0 || c.last_use < sv->sess[size_t(victim)].last_use)) victim = int32_t(i);
    }
    if (victim < 0) return false;
    ServSess& v = sv->sess[size_t(victim)];
    int e = vmp_state_get(v.s, &v.x, v.P.data());
    if (e == VMP_OK) e = vmp_session_destroy(v.s);
    v.s = nullptr;
    v.built = false;
    sv->resident_bytes -= sv->sess_bytes;
    sv->resident -= 1;
    sv->st_evictions += 1;
    if (e != VMP_OK) {
      *rc = e;
      return false;
    }
  }
  int e = vmp_session_create(sv->home, &sv->scfg, &t.s);
  if (e == VMP_OK) e = vmp_state_set(t.s, &t.x, t.P.data());
  if (e != VMP_OK) {
    if (t.s) vmp_session_destroy(t.s);
    t.s = nullptr;
    *rc = e;
    return false;
  }
  t.built = false;
  sv->resident_bytes += sv->sess_bytes;
  sv->resident += 1;
  sv->st_admissions += 1;
  return true;
}

static void serve_worker(vmp_server* sv, int e) {
  hipSetDevice(sv->cfg.device);
  vmp_ctx* c = sv->ex[size_t(e)];
  std::vector<int32_t> pick, breq, rreq;
  std::vector<vmp_session*> bs, rs;
  std::vector<double> mR, mp;
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    if (sv->first_err != VMP_OK && sv->inflight == 0) break;
    const double now = sv->now();
    while (sv->next < sv->nreq && sv->reqs[sv->next].arrival <= now) sv->queue.push_back(sv->next++);
    // ---- form a batch: eligible requests in arrival order, one per idle session
    pick.clear();
    int64_t pts = 0;
    const uint64_t ep = ++sv->epoch;
    bool memory_wait = false;
    for (auto it = sv->queue.begin(); it != sv->queue.end();) {
      vmp_request& r = sv->reqs[*it];
      ServSess& ss = sv->sess[size_t(r.session)];
      const bool first_visit = ss.mark != ep;
      ss.mark = ep;
      if (!first_visit || ss.busy || int(pick.size()) >= sv->cfg.max_batch_sessions ||
          (!pick.empty() && pts + r.n > sv->cfg.max_batch_points)) {
        ++it;
        continue;
      }
      int arc = VMP_OK;
      if (!serve_admit(sv, r.session, &arc)) {
        if (arc != VMP_OK) {
          r.status = arc;
          if (sv->first_err == VMP_OK) sv->first_err = arc;
          it = sv->queue.erase(it);
        } else {
          memory_wait = true;
          ++it;
        }
        continue;
      }
      ss.busy = true;
      pick.push_back(*it);
      pts += r.n;
      it = sv->queue.erase(it);
    }
    if (pick.empty()) {
      if (sv->next >= sv->nreq && sv->queue.empty() && sv->inflight == 0) {
        sv->cv.notify_all();
        break;
      }
      if (sv->first_err != VMP_OK && sv->inflight == 0) break;
      if (sv->next < sv->nreq && !memory_wait && sv->queue.empty()) {
        const auto wake = sv->t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                       std::chrono::duration<double>(sv->reqs[sv->next].arrival));
        sv->cv.wait_until(lk, wake);
      } else if (sv->next < sv->nreq) {
        const auto wake = sv->t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                       std::chrono::duration<double>(sv->reqs[sv->next].arrival));
        sv->cv.wait_until(lk, wake);
      } else {
        sv->cv.wait(lk);
      }
      continue;
    }
    const int bid = sv->batch_id++;
    sv->inflight += 1;
    // ---- split into MAP_INIT scans and registered scans
    breq.clear();
    rreq.clear();
    bs.clear();
    rs.clear();
    for (int32_t ri : pick) {
      vmp_request& r = sv->reqs[ri];
      ServSess& ss = sv->sess[size_t(r.session)];
      if (ss.built) {
        rreq.push_back(ri);
        rs.push_back(ss.s);
        r.kind = 2;
      } else {
        breq.push_back(ri);
        bs.push_back(ss.s);
        r.kind = ss.seen ? 3 : 1;
      }
      bind_points_device(ss.s, r.x, r.y, r.z, r.n);  // checked by vmp_server_run
    }
    const double tstart = sv->now();
    lk.unlock();
    int rc = VMP_OK;
    auto motion = [&](const std::vector<int32_t>& ids, std::vector<vmp_session*>& ss) {
      mR.resize(9 * ids.size());
      mp.resize(3 * ids.size());
      for (size_t i = 0; i < ids.size(); ++i) {
        std::memcpy(&mR[9 * i], sv->reqs[ids[i]].dR, 9 * sizeof(double));
        std::memcpy(&mp[3 * i], sv->reqs[ids[i]].dp, 3 * sizeof(double));
      }
      std::lock_guard<std::mutex> g(c->mu);
      return enqueue_motion(c, ss.data(), int(ss.size()), mR.data(), mp.data(), sv->q, 0);
    };
    if (!bs.empty()) {
      // re-admitted sessions first advance their kept posterior by the motion
      std::vector<int32_t> mids;
      std::vector<vmp_session*> ms;
      for (size_t i = 0; i < breq.size(); ++i)
        if (sv->reqs[breq[i]].kind == 3) {
          mids.push_back(breq[i]);
          ms.push_back(bs[i]);
        }
      if (!ms.empty()) rc = motion(mids, ms);
      if (rc == VMP_OK) rc = run_step_on(c, bs.data(), int(bs.size()), kStepMap | kStepBuild, nullptr, nullptr);
      for (size_t i = 0; i < breq.size(); ++i) {
        const BatchOut& o = c->ws.h_out[i];
        vmp_request& r = sv->reqs[breq[i]];
        std::memcpy(r.rot, o.x.rot, sizeof(r.rot));
        std::memcpy(r.pos, o.x.pos, sizeof(r.pos));
        r.iterations = 0;
      }
    }
    if (!rs.empty() && rc == VMP_OK) {
      rc = motion(rreq, rs);
      if (rc == VMP_OK) rc = run_step_on(c, rs.data(), int(rs.size()), kStepIeskf | kStepMap, nullptr, nullptr);
      for (size_t i = 0; i < rreq.size(); ++i) {
        const BatchOut& o = c->ws.h_out[i];
        vmp_request& r = sv->reqs[rreq[i]];
        std::memcpy(r.rot, o.x.rot, sizeof(r.rot));
        std::memcpy(r.pos, o.x.pos, sizeof(r.pos));
        r.iterations = o.iters;
      }
    }
    lk.lock();
    const double tdone = sv->now();
    for (int32_t ri : pick) {
      vmp_request& r = sv->reqs[ri];
      ServSess& ss = sv->sess[size_t(r.session)];
      r.start = tstart;
      r.done = tdone;
      r.status = rc;
      r.batch = bid;
      r.batch_sessions = int32_t(pick.size());
      ss.busy = false;
      ss.seen = true;
      ss.built = rc == VMP_OK;
      ss.last_use = tdone;
    }
    if (rc != VMP_OK && sv->first_err == VMP_OK) {
      sv->first_err = rc;
      sv->err = vmp_last_error(pick.empty() ? nullptr : sv->sess[size_t(sv->reqs[pick[0]].session)].s);
    }
    sv->st_batches += 1;
    sv->st_requests += int64_t(pick.size());
    sv->st_maxbatch = std::max<int64_t>(sv->st_maxbatch, int64_t(pick.size()));
    sv->inflight -= 1;
    sv->cv.notify_all();
  }
}

extern "C" int vmp_server_create(const vmp_server_cfg* cfg, const vmp_map_cfg* scfg, int32_t nsessions,
                                 vmp_server** out) {
  if (!cfg || !scfg || !out || nsessions <= 0 || cfg->executors < 1 || cfg->executors > 8 ||
      cfg->max_batch_sessions < 1 || cfg->max_batch_points < 1 || check_cfg(scfg) != VMP_OK)
    return VMP_EINVAL;
  vmp_server* sv = new (std::nothrow) vmp_server();
  if (!sv) return VMP_ENOMEM;
  sv->cfg = *cfg;
  sv->scfg = *scfg;
  sv->sess.resize(size_t(nsessions));
  for (ServSess& s : sv->sess) s.P.assign(529, 0.0);
  sv->sess_bytes = vmp_session_bytes(scfg);
  int rc = vmp_ctx_create(cfg->device, &sv->home);
  for (int e = 0; rc == VMP_OK && e < cfg->executors; ++e) {
    vmp_ctx* c = nullptr;
    rc = vmp_ctx_create(cfg->device, &c);
    if (rc == VMP_OK) sv->ex.push_back(c);
  }
  if (rc != VMP_OK) {
    vmp_server_destroy(sv);
    return rc;
  }
  if (sv->sess_bytes > cfg->hbm_budget) {
    vmp_server_destroy(sv);
    return VMP_ENOMEM;
  }
  *out = sv;
  return VMP_OK;
}

extern "C" int vmp_server_destroy(vmp_server* sv) {
  if (!sv) return VMP_EINVAL;
  for (ServSess& s : sv->sess)
    if (s.s) vmp_session_destroy(s.s);
  for (vmp_ctx* c : sv->ex) vmp_ctx_destroy(c);
  if (sv->home) vmp_ctx_destroy(sv->home);
  delete sv;
  return VMP_OK;
}

extern "C" int vmp_server_session_init(vmp_server* sv, int32_t id, const vmp_state* x0, const double* P0) {
  if (!sv || !x0 || !P0 || id < 0 || id >= int32_t(sv->sess.size())) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  ServSess& s = sv->sess[size_t(id)];
  if (s.busy) return VMP_ESTATE;
  if (s.s) {
    vmp_session_destroy(s.s);
    s.s = nullptr;
    sv->resident_bytes -= sv->sess_bytes;
    sv->resident -= 1;
  }
  s.x = *x0;
  std::memcpy(s.P.data(), P0, 529 * sizeof(double));
  s.has_state = true;
  s.built = false;
  s.seen = false;
  // resident at once while the budget has room (no eviction here)
  int rc = VMP_OK;
  if (sv->resident_bytes + sv->sess_bytes <= sv->cfg.hbm_budget && !serve_admit(sv, id, &rc)) return rc;
  return rc;
}

extern "C" int vmp_server_run(vmp_server* sv, vmp_request* reqs, int32_t n, const double* qdiag) {
  if (!sv || (n > 0 && !reqs) || !qdiag || n < 0) return VMP_EINVAL;
  for (int32_t i = 0; i < n; ++i) {
    const vmp_request& r = reqs[i];
    if (r.session < 0 || r.session >= int32_t(sv->sess.size()) || !sv->sess[size_t(r.session)].has_state ||
        r.n < 0 || r.n > sv->scfg.max_scan_points || (i > 0 && r.arrival < reqs[i - 1].arrival))
      return VMP_EINVAL;
  }
  {
    // every request's points are checked once, before the run's clock starts
    std::vector<PtsRef> refs(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) refs[size_t(i)] = PtsRef{reqs[i].x, reqs[i].y, reqs[i].z, nullptr, reqs[i].n, 0};
    vmp_ctx* c = sv->home;
    std::lock_guard<std::mutex> lk(c->mu);
    hipSetDevice(c->device);
    const int bad = check_device_points(c, refs.data(), n);
    if (bad == -2) return VMP_EDEVICE;
    if (bad >= 0) {
      reqs[bad].status = VMP_EINVAL;
      std::lock_guard<std::mutex> g(sv->mu);
      sv->err = "request " + std::to_string(bad) + ": scan holds a non-finite coordinate";
      return VMP_EINVAL;
    }
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->reqs = reqs;
    sv->nreq = n;
    sv->next = 0;
    sv->inflight = 0;
    sv->batch_id = 0;
    sv->queue.clear();
    std::memcpy(sv->q, qdiag, sizeof(sv->q));
    sv->first_err = VMP_OK;
    sv->err.clear();
    sv->st_batches = sv->st_requests = sv->st_evictions = sv->st_admissions = sv->st_maxbatch = 0;
    for (int32_t i = 0; i < n; ++i) {
      reqs[i].status = VMP_ESTATE;  // not processed
      reqs[i].kind = 0;
      reqs[i].batch = -1;
    }
    sv->t0 = std::chrono::steady_clock::now();
  }
  std::vector<std::thread> th;
  for (int e = 0; e < int(sv->ex.size()); ++e) th.emplace_back(serve_worker, sv, e);
  for (std::thread& t : th) t.join();
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->reqs = nullptr;
  return sv->first_err;
}

extern "C" vmp_session* vmp_server_session(vmp_server* sv, int32_t id) {
  if (!sv || id < 0 || id >= int32_t(sv->sess.size())) return nullptr;
  std::lock_guard<std::mutex> lk(sv->mu);
  return sv->sess[size_t(id)].s;
}

extern "C" int vmp_server_stats(vmp_server* sv, int64_t out[8]) {
  if (!sv || !out) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  out[0] = sv->st_batches;
  out[1] = sv->st_requests;
  out[2] = sv->st_evictions;
  out[3] = sv->st_admissions;
  out[4] = sv->resident;
  out[5] = sv->resident_bytes;
  out[6] = sv->st_maxbatch;
  out[7] = int64_t(sv->ex.size());
  return VMP_OK;
}

extern "C" const char* vmp_server_last_error(const vmp_server* sv) { return sv ? sv->err.c_str() : "null server"; }
