// Served form: a request server over the batched engine (no counterpart in the
// reference, which solves one problem per process run).
//
// Requests (a problem as COO F and P, its gamma, the solve's mode / maxiter /
// eps and precision) are queued by size class. Each size class is a lane: one
// Engine (batch) on its own stream, driven by its own host thread, so that small
// problems advance at their own pace instead of at the pace of the large ones.
// A lane admits the head of its queue when a slot is free and the request's
// device footprint fits the server's device-byte budget: the footprint is
// computed from F's pattern before the context exists (what its creation, the
// Gram matrix, mixed precision and the Gram split allocate, plus the transient
// dense matrices of the Gram formation), then corrected to the context's actual
// bytes once created. Admitted problems join the lane's next step (continuous
// batching); a finished problem's Pi is copied to the host and its device memory
// freed at once. Every problem computes the same bits as when solved alone.
//
// Included into pqdts_amd.hip (one translation unit).
#include <condition_variable>
#include <deque>
#include <thread>

namespace {

struct SrvRequest {
  int64_t id = 0, M = 0, N = 0, D = 0;
  std::vector<int32_t> Fr, Fc, Pr, Pc;
  std::vector<double> Fv, Pv, X0;
  double gamma = 0.0, eps = 0.0;
  int mode = 2, maxiter = 0, precision = 0, lane = 0;
  int64_t reserve = 0;  // device bytes held against the budget while admitted
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
  bool done = false;
  int status = 0;
  std::string err;
  std::vector<double> X;  // row-major result
  double conv = 0.0;
  int stats[7] = {0, 0, 0, 0, 0, 0, 0};  // finished, Newton, PCG, trials, polish status, iterations, rounds
  pqdts_ctx::Polish pol;                 // the server's polish settings when submitted
};

}  // namespace

struct pqdts_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::deque<SrvRequest *>> queue;  // per lane, FIFO
  std::map<int64_t, std::unique_ptr<SrvRequest>> reqs;
  std::deque<int64_t> finished;  // finished, not yet returned by pqdts_server_wait
  int64_t next_id = 1, budget = 0, in_use = 0, peak = 0, submitted = 0, completed = 0, running = 0;
  std::vector<int64_t> bounds;  // M*N upper bounds of lanes 0 .. nlanes-2
  std::vector<std::unique_ptr<Engine>> lanes;
  std::vector<std::thread> workers;
  bool stop = false;
  int device = 0;
  pqdts_ctx::Polish pol;  // applied to requests submitted from now on
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

void server_finish(pqdts_server *S, SrvRequest *r, int status, const std::string &err) {
  std::lock_guard<std::mutex> lk(S->mu);
  S->in_use -= r->reserve;
  r->reserve = 0;
  r->status = status;
  r->err = err;
  r->done = true;
  r->t_done = S->now();
  S->running--;
  S->completed++;
  S->finished.push_back(r->id);
  S->cv_done.notify_all();
  S->cv_work.notify_all();  // the freed bytes may let another lane admit
}

void server_lane(pqdts_server *S, int g) {
  (void)hipSetDevice(S->device);
  Engine &e = *S->lanes[g];
  std::map<int, std::pair<SrvRequest *, pqdts_ctx *>> live;
  auto drop_all = [&](int status, const std::string &msg) {
    for (auto &kv : live) {
      e.detach(kv.first);
      destroy_ctx(kv.second.second);
      server_finish(S, kv.second.first, status, msg);
    }
    live.clear();
  };
  for (;;) {
    std::vector<SrvRequest *> take;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->stop) break;
        auto &q = S->queue[g];
        while (!q.empty() && (int)(live.size() + take.size()) < e.cap &&
               S->in_use + q.front()->reserve <= S->budget) {
          SrvRequest *r = q.front();
          q.pop_front();
          S->in_use += r->reserve;
          S->peak = std::max(S->peak, S->in_use);
          S->running++;
          r->t_admit = S->now();
          take.push_back(r);
        }
        if (!take.empty() || !live.empty()) break;
        S->cv_work.wait_for(lk, std::chrono::milliseconds(50));
      }
      if (S->stop) {
        lk.unlock();
        for (SrvRequest *r : take) server_finish(S, r, -1, "server destroyed");
        drop_all(-1, "server destroyed");
        return;
      }
    }
    for (SrvRequest *r : take) {
      pqdts_ctx *c = nullptr;
      const int rc = guarded([&] {
        c = create_ctx(r->M, r->N, r->D, (int64_t)r->Fv.size(), r->Fr.data(), r->Fc.data(), r->Fv.data(),
                       (int64_t)r->Pv.size(), r->Pr.data(), r->Pc.data(), r->Pv.data(), r->gamma);
        c->pol = r->pol;
        if (r->X0.empty()) r->X0.assign((size_t)r->M * r->N, 1.0 / (double)r->N);
        copy_slot(c, PQDTS_X, nullptr, r->X0.data(), PQDTS_ROW_MAJOR, true);
        const int slot = engine_admit(e, c, r->mode, r->maxiter, r->eps, 1, 0, r->precision, nullptr, nullptr);
        live[slot] = {r, c};
      });
      if (rc != 0) {
        if (c) destroy_ctx(c);
        server_finish(S, r, rc, g_err);
        continue;
      }
      // the context's actual bytes replace the estimate (which also covered the
      // transient Gram formation); the host copies of the inputs are no longer needed
      const int64_t actual = ctx_device_bytes(c);
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->in_use += actual - r->reserve;
        r->reserve = actual;
      }
      for (auto *v : {&r->Fr, &r->Fc, &r->Pr, &r->Pc}) std::vector<int32_t>().swap(*v);
      for (auto *v : {&r->Fv, &r->Pv, &r->X0}) std::vector<double>().swap(*v);
    }
    if (live.empty()) continue;
    const int rc = guarded([&] {
      if (e.any_active()) e.tick();
      HIPC(hipStreamSynchronize(e.s));
    });
    if (rc != 0) {
      drop_all(rc, g_err);
      continue;
    }
    std::vector<int> fin;
    for (auto &kv : live)
      if (e.st[kv.first] && e.st[kv.first]->phase == 2) fin.push_back(kv.first);
    for (int p : fin) {
      SrvRequest *r = live[p].first;
      pqdts_ctx *c = live[p].second;
      const PState &q = *e.st[p];
      r->conv = q.conv;
      r->stats[0] = 1;
      r->stats[1] = q.newton;
      r->stats[2] = q.pcg;
      r->stats[3] = q.trials;
      r->stats[4] = c->pol_out.status;
      r->stats[5] = c->pol_out.iters;
      r->stats[6] = c->pol_out.rounds;
      std::string msg;
      const int rc2 = guarded([&] {
        r->X.resize((size_t)r->M * r->N);
        copy_slot(c, PQDTS_X, r->X.data(), nullptr, PQDTS_ROW_MAJOR, false);
      });
      if (rc2 != 0) msg = g_err;
      e.detach(p);
      destroy_ctx(c);
      live.erase(p);
      server_finish(S, r, rc2, msg);
    }
    if (live.empty() || e.deferred.size() > Engine::kDeferMax) e.release_deferred();
  }
}

int lane_of(const pqdts_server *S, int64_t MN) {
  int g = 0;
  for (int64_t b : S->bounds)
    if (MN >= b) g++;
  return g;
}

}  // namespace

extern "C" {

int pqdts_footprint(int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *F_row, const int32_t *F_col,
                    int precision, int64_t *persistent, int64_t *transient) {
  return guarded([&] {
    require(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (mixed)");
    require(persistent != nullptr && transient != nullptr, "null output");
    const Plan pl = make_plan(M, N, D, nF, F_row, F_col);
    plan_footprint(pl, env_hvp_pref(), 0, precision != 0, persistent, transient);
  });
}

int pqdts_server_create(int nlanes, const int64_t *lane_bounds, const int32_t *lane_cus, int capacity,
                        int64_t budget_bytes, pqdts_server **out) {
  return guarded([&] {
    require(out != nullptr, "null output pointer");
    *out = nullptr;
    require(nlanes >= 1 && nlanes <= 16, "1 to 16 lanes");
    require(nlanes == 1 || lane_bounds != nullptr, "lane bounds needed for more than one lane");
    require(capacity >= 1 && capacity <= 4096, "capacity per lane must be 1 to 4096");
    require(budget_bytes >= 0, "negative budget");
    for (int g = 0; g + 1 < nlanes; g++)
      require(lane_bounds[g] > 0 && (g == 0 || lane_bounds[g] > lane_bounds[g - 1]),
              "lane bounds must be positive and increasing");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Fail(-2, "no GPU visible to pqdts_amd");
    std::unique_ptr<pqdts_server> S(new pqdts_server());
    HIPC(hipGetDevice(&S->device));
    if (budget_bytes == 0) {
      size_t fr = 0, tot = 0;
      HIPC(hipMemGetInfo(&fr, &tot));
      budget_bytes = (int64_t)(0.9 * (double)fr);
    }
    S->budget = budget_bytes;
    for (int g = 0; g + 1 < nlanes; g++) S->bounds.push_back(lane_bounds[g]);
    S->queue.resize(nlanes);
    const int ncu = pqdts_device_cu_count();
    for (int g = 0; g < nlanes; g++) {
      S->lanes.emplace_back(new Engine());
      S->lanes.back()->init(capacity, nullptr, true);
      // a lane's Newton iterations replay captured graphs (PQDTS_GRAPHS=2) unless
      // the environment chooses: the server times no kernels, and the graphs give
      // the eager bits
      if (!getenv("PQDTS_GRAPHS")) S->lanes.back()->graphs_on = 2;
      // restricted to CUs lo..hi (a consecutive range), before any worker starts
      if (lane_cus && lane_cus[2 * g] >= 0) {
        const int lo = lane_cus[2 * g], hi = lane_cus[2 * g + 1];
        require(lo <= hi && hi < ncu, "lane CU range outside the device");
        std::vector<uint32_t> words((size_t)(ncu + 31) / 32, 0u);
        for (int cu = lo; cu <= hi; cu++) words[cu / 32] |= 1u << (cu % 32);
        S->lanes.back()->set_cu_mask(words.data(), (int)words.size());
      }
    }
    for (int g = 0; g < nlanes; g++) S->workers.emplace_back(server_lane, S.get(), g);
    *out = S.release();
  });
}

int pqdts_server_destroy(pqdts_server *s) {
  return guarded([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
      s->cv_work.notify_all();
    }
    for (auto &t : s->workers) t.join();
    delete s;
  });
}

int pqdts_server_submit(pqdts_server *s, int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *F_row,
                        const int32_t *F_col, const double *F_val, int64_t nP, const int32_t *P_row,
                        const int32_t *P_col, const double *P_val, double gamma, int mode, int maxiter,
                        double eps, int precision, const double *X0, int64_t *id) {
  return guarded([&] {
    require(s != nullptr, "null server");
    require(mode == 1 || mode == 2, "mode must be 1 or 2");
    require(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (mixed)");
    check_shape(M, N, D, nF, nP);
    require(nF == 0 || (F_row && F_col && F_val), "null F arrays");
    require(nP == 0 || (P_row && P_col && P_val), "null P arrays");
    for (int64_t e = 0; e < nP; e++)
      require(P_row[e] >= 0 && P_row[e] < D && P_col[e] >= 0 && P_col[e] < N,
              "P index out of range at entry " + std::to_string(e));
    std::unique_ptr<SrvRequest> r(new SrvRequest());
    const Plan pl = make_plan(M, N, D, nF, F_row, F_col);
    int64_t pers = 0, trans = 0;
    plan_footprint(pl, env_hvp_pref(), 0, precision != 0, &pers, &trans);
    r->reserve = pers + trans;
    require(r->reserve <= s->budget, "the request needs " + std::to_string(r->reserve) +
                                         " device bytes, more than the server's budget of " +
                                         std::to_string(s->budget));
    r->M = M;
    r->N = N;
    r->D = D;
    r->Fr.assign(F_row, F_row + nF);
    r->Fc.assign(F_col, F_col + nF);
    r->Fv.assign(F_val, F_val + nF);
    r->Pr.assign(P_row, P_row + nP);
    r->Pc.assign(P_col, P_col + nP);
    r->Pv.assign(P_val, P_val + nP);
    if (X0) r->X0.assign(X0, X0 + M * N);
    r->gamma = gamma;
    r->mode = mode;
    r->maxiter = maxiter;
    r->eps = eps;
    r->precision = precision;
    std::lock_guard<std::mutex> lk(s->mu);
    require(!s->stop, "server stopped");
    r->id = s->next_id++;
    r->pol = s->pol;
    r->lane = lane_of(s, M * N);
    r->t_submit = s->now();
    if (id) *id = r->id;
    s->queue[r->lane].push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->submitted++;
    s->cv_work.notify_all();
  });
}

int pqdts_server_wait(pqdts_server *s, int64_t *ids, int cap, int timeout_ms) {
  int n = 0;
  const int rc = guarded([&] {
    require(s != nullptr, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->finished.empty() && timeout_ms != 0) {
      auto ready = [&] { return !s->finished.empty() || s->completed == s->submitted; };
      if (timeout_ms < 0)
        s->cv_done.wait(lk, ready);
      else
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    }
    while (!s->finished.empty() && n < cap) {
      if (ids) ids[n] = s->finished.front();
      s->finished.pop_front();
      n++;
    }
  });
  return rc == 0 ? n : -1;
}

int pqdts_server_result(pqdts_server *s, int64_t id, double *X, double *conv, int *stats, double *times) {
  int status = 0;
  const int rc = guarded([&] {
    require(s != nullptr, "null server");
    std::unique_ptr<SrvRequest> r;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(id);
      require(it != s->reqs.end(), "unknown request id " + std::to_string(id));
      require(it->second->done, "request " + std::to_string(id) + " has not finished");
      r = std::move(it->second);
      s->reqs.erase(it);
    }
    if (X && !r->X.empty()) std::memcpy(X, r->X.data(), sizeof(double) * r->X.size());
    if (conv) *conv = r->conv;
    if (stats)
      for (int k = 0; k < 7; k++) stats[k] = r->stats[k];
    if (times) {
      times[0] = r->t_admit - r->t_submit;
      times[1] = r->t_done - r->t_admit;
      times[2] = r->t_done - r->t_submit;
    }
    status = r->status;
    if (status != 0) throw Fail(status, "request " + std::to_string(id) + " failed: " + r->err);
  });
  return rc;
}

int pqdts_server_set_polish(pqdts_server *s, int max_outer, double cg_rtol, double tol, double budget, int fp32) {
  return guarded([&] {
    require(s != nullptr, "null server");
    require(max_outer >= 0 && cg_rtol > 0.0 && cg_rtol < 1.0 && tol > 0.0 && (fp32 == 0 || fp32 == 1),
            "bad polish settings");
    std::lock_guard<std::mutex> lk(s->mu);
    s->pol.max_outer = max_outer;
    s->pol.cg_rtol = cg_rtol;
    s->pol.tol = tol;
    s->pol.budget = budget;
    s->pol.fp32 = fp32;
  });
}

int pqdts_server_set_polish_method(pqdts_server *s, int method) {
  return guarded([&] {
    require(s != nullptr, "null server");
    require(method == 0 || method == 1, "bad polish method (0 Newton-CG, 1 direct face solves)");
    std::lock_guard<std::mutex> lk(s->mu);
    s->pol.method = method;
  });
}

int pqdts_server_info(pqdts_server *s, int64_t *info) {
  return guarded([&] {
    require(s != nullptr && info != nullptr, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t queued = 0;
    for (auto &q : s->queue) queued += (int64_t)q.size();
    info[0] = queued;
    info[1] = s->running;
    info[2] = (int64_t)s->finished.size();
    info[3] = s->in_use;
    info[4] = s->peak;
    info[5] = s->budget;
    info[6] = s->submitted;
    info[7] = s->completed;
  });
}

}  // extern "C"
