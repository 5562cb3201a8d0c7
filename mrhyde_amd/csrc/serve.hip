// serve.hip — request loop for many independent solves of mixed mesh size:
// continuous batching with admission by device-memory footprint, resident
// contexts reused across requests of the same mesh, and per-request latency.
//
// The reference has no serving layer; its closest pattern is the sequential
// sample loop of the UQ analysis (AnalysisManager, analysisManager_solve.hpp),
// one forwardModel / SolverManager::nonlinearSolver per sample.  Here each
// request is one such steady solve (mhx_nonlinear_solve), and requests run
// concurrently: one host worker and one HIP stream per active request, so
// their kernels overlap on the device.  A request is admitted as soon as a
// worker is free and its device footprint (mhx_footprint) fits the budget
// next to the contexts already held; a request whose mesh matches an idle
// cached context reuses it (mesh, CRS graph, Jacobian and Krylov buffers stay
// in HBM; only the coefficients are uploaded, mhx_ctx_update).  The queue is
// scanned first-fit, so a small request may pass a large one that does not
// fit yet.  Results are independent of the mix: every reduction runs in a
// fixed order on the request's own stream.

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mhx_internal.hpp"

using namespace mhx;

namespace {

using clk = std::chrono::steady_clock;

// mesh identity of a request: contexts with equal keys can be reused
struct Key {
  mhx_problem_desc d{};
  bool operator<(const Key& o) const { return std::memcmp(&d, &o.d, sizeof(d)) < 0; }
  bool operator==(const Key& o) const { return std::memcmp(&d, &o.d, sizeof(d)) == 0; }
};

Key key_of(const mhx_problem_desc& in) {
  Key k;
  std::memset(&k.d, 0, sizeof(k.d));
  mhx_problem_desc& d = k.d;
  d.dim = in.dim; d.order = in.order; d.quadrature = in.quadrature;
  d.nx = in.nx; d.ny = in.ny; d.nz = in.dim == 3 ? in.nz : 1;
  d.xmin = in.xmin; d.xmax = in.xmax; d.ymin = in.ymin; d.ymax = in.ymax; d.zmin = in.zmin; d.zmax = in.zmax;
  d.perturb = in.perturb; d.perturb_seed = in.perturb_seed; d.physics = in.physics;
  d.dbc_sides = in.dbc_sides; d.ns_lid_sides = in.ns_lid_sides; d.ns_pin_pressure = in.ns_pin_pressure;
  for (int a = 0; a < 3; ++a) d.ns_comp_sides[a] = in.ns_comp_sides[a];
  return k;
}

struct Request {
  int64_t id = 0;
  mhx_problem_desc desc{};
  std::vector<double> k0, src, dbc;  // owned copies of the description's arrays
  std::vector<std::string> fn_name, fn_expr;  // and of its coefficient functions
  std::vector<const char*> fn_name_p, fn_expr_p;
  Key key;
  int64_t n_dof = 0, bytes = 0;
  clk::time_point t_submit, t_start, t_end;
  bool done = false;
  int status = MHX_OK;
  std::string err;
  std::vector<double> x;
  mhx_solve_stats st{};
};

struct Slot {
  Key key;
  mhx_ctx* ctx = nullptr;
  int64_t bytes = 0;
  clk::time_point last_used;
};

}  // namespace

struct mhx_server {
  mhx_server_desc d{};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> queue;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  std::list<Slot> idle;  // least recently used first
  int64_t used = 0;      // bytes held by live contexts (active + idle)
  int64_t next_id = 1, n_done = 0, n_reused = 0, n_created = 0, n_active = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  std::string err;

  // Under mu: pick the first queued request that can run now.  Returns the
  // request and its slot (ctx == nullptr: create one); contexts evicted to
  // make room are appended to `drop` and destroyed by the caller unlocked.
  Request* admit(Slot& slot, std::vector<mhx_ctx*>& drop) {
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      Request* r = *it;
      for (auto s = idle.begin(); s != idle.end(); ++s)
        if (s->key == r->key) {
          slot = *s;
          idle.erase(s);
          queue.erase(it);
          ++n_reused;
          return r;
        }
      int64_t freeable = 0;
      for (const Slot& s : idle) freeable += s.bytes;
      if (used + r->bytes <= d.hbm_budget || used - freeable + r->bytes <= d.hbm_budget) {
        while (used + r->bytes > d.hbm_budget && !idle.empty()) {
          used -= idle.front().bytes;
          drop.push_back(idle.front().ctx);
          idle.pop_front();
        }
        slot = Slot{r->key, nullptr, r->bytes, clk::now()};
        used += r->bytes;
        queue.erase(it);
        return r;
      }
    }
    return nullptr;
  }

  void run(Request* r, Slot& slot) {
    int rc = MHX_OK;
    if (!slot.ctx) {
      rc = mhx_ctx_create(&r->desc, &slot.ctx);
      if (rc == MHX_OK) {
        std::lock_guard<std::mutex> lk(mu);
        ++n_created;
      }
    } else {
      rc = mhx_ctx_update(slot.ctx, &r->desc);
    }
    if (rc == MHX_OK) {
      r->x.assign(r->n_dof, 0.0);
      rc = mhx_set_dirichlet(slot.ctx, r->x.data(), 0);
      if (rc == MHX_OK) {
        mhx_newton_opts o = d.newton;
        o.flags = 0;
        o.linear.flags = 0;
        rc = mhx_nonlinear_solve(slot.ctx, r->x.data(), &o, &r->st);
      }
    }
    r->status = rc;
    if (rc < 0 && rc != MHX_ERR_NOT_CONVERGED) r->err = mhx_last_error(slot.ctx);
  }

  void worker() {
    for (;;) {
      Slot slot;
      std::vector<mhx_ctx*> drop;
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || (r = admit(slot, drop)) != nullptr || !drop.empty(); });
        if (!r && stop) return;
        if (r) {
          ++n_active;
          r->t_start = clk::now();
        }
      }
      for (mhx_ctx* c : drop) mhx_ctx_destroy(c);
      drop.clear();
      if (!r) continue;
      run(r, slot);
      {
        std::lock_guard<std::mutex> lk(mu);
        --n_active;
        r->t_end = clk::now();
        r->done = true;
        ++n_done;
        if (slot.ctx && r->status != MHX_ERR_DEVICE) {
          slot.last_used = r->t_end;
          idle.push_back(slot);
          // keep at most max_cached idle contexts, trimmed once the queue has
          // drained: destroying a context frees device memory, which
          // synchronises the device and would stall every running request
          // (under memory pressure admit() evicts as needed)
          while (queue.empty() && (int)idle.size() > d.max_cached) {
            used -= idle.front().bytes;
            drop.push_back(idle.front().ctx);
            idle.pop_front();
          }
        } else {
          used -= slot.bytes;
          if (slot.ctx) drop.push_back(slot.ctx);
        }
        cv.notify_all();
      }
      for (mhx_ctx* c : drop) mhx_ctx_destroy(c);
    }
  }
};

extern "C" {

int mhx_server_create(const mhx_server_desc* desc, mhx_server** out) {
  if (!desc || !out) return MHX_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MHX_ERR_DEVICE;
  auto* s = new mhx_server();
  s->d = *desc;
  if (s->d.max_active <= 0) s->d.max_active = 8;
  if (s->d.max_cached < 0) s->d.max_cached = 0;
  if (s->d.hbm_budget <= 0) {
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    s->d.hbm_budget = (int64_t)(fr / 2);
  }
  mhx_newton_opts& o = s->d.newton;
  if (o.max_iters <= 0) o.max_iters = 2;
  if (o.nl_tol <= 0) o.nl_tol = 1e-8;
  if (o.linear.solver == 0) o.linear.solver = MHX_KRYLOV_CG;
  if (o.linear.tol <= 0) o.linear.tol = 1e-10;
  if (o.linear.max_iters <= 0) o.linear.max_iters = 20000;
  for (int i = 0; i < s->d.max_active; ++i) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return MHX_OK;
}

int mhx_server_submit(mhx_server* s, const mhx_problem_desc* desc, int64_t* ticket) {
  if (!s || !desc || !ticket) return MHX_ERR_ARG;
  auto r = std::make_unique<Request>();
  int64_t nnz = 0;
  int rc = mhx_footprint(desc, &s->d.newton.linear, &r->n_dof, &nnz, &r->bytes);
  if (rc != MHX_OK) return rc;
  if (r->bytes > s->d.hbm_budget) return MHX_ERR_NOMEM;  // could never be admitted
  r->desc = *desc;
  long ne = (long)desc->nx * desc->ny * (desc->dim == 3 ? desc->nz : 1);
  if (desc->k0) {
    r->k0.assign(desc->k0, desc->k0 + ne);
    r->desc.k0 = r->k0.data();
  }
  if (desc->src_elem) {
    r->src.assign(desc->src_elem, desc->src_elem + ne);
    r->desc.src_elem = r->src.data();
  }
  if (desc->dbc_values) {
    r->dbc.assign(desc->dbc_values, desc->dbc_values + r->n_dof);
    r->desc.dbc_values = r->dbc.data();
  }
  if (desc->n_functions > 0) {
    if (!desc->function_names || !desc->function_exprs) return MHX_ERR_ARG;
    for (int i = 0; i < desc->n_functions; ++i) {
      if (!desc->function_names[i] || !desc->function_exprs[i]) return MHX_ERR_ARG;
      r->fn_name.emplace_back(desc->function_names[i]);
      r->fn_expr.emplace_back(desc->function_exprs[i]);
    }
    for (int i = 0; i < desc->n_functions; ++i) {
      r->fn_name_p.push_back(r->fn_name[i].c_str());
      r->fn_expr_p.push_back(r->fn_expr[i].c_str());
    }
    r->desc.function_names = r->fn_name_p.data();
    r->desc.function_exprs = r->fn_expr_p.data();
  }
  r->key = key_of(*desc);
  r->t_submit = clk::now();
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  *ticket = r->id;
  s->queue.push_back(r.get());
  s->reqs[r->id] = std::move(r);
  s->cv.notify_all();
  return MHX_OK;
}

int mhx_server_poll(mhx_server* s, int64_t ticket, int32_t wait, double* x, mhx_solve_stats* stats,
                    double* latency_ms) {
  if (!s) return MHX_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return MHX_ERR_ARG;
  Request* r = it->second.get();
  if (!r->done) {
    if (!wait) return 0;
    s->cv.wait(lk, [&] { return r->done; });
  }
  if (x && !r->x.empty()) std::memcpy(x, r->x.data(), r->x.size() * sizeof(double));
  if (stats) *stats = r->st;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_submit).count();
  const int status = r->status;
  if (status < 0) s->err = r->err;
  s->reqs.erase(it);
  return status < 0 && status != MHX_ERR_NOT_CONVERGED ? status : 1;
}

int mhx_server_info(mhx_server* s, int64_t* out8) {
  if (!s || !out8) return MHX_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  out8[0] = s->n_done;
  out8[1] = s->n_active;
  out8[2] = (int64_t)s->queue.size();
  out8[3] = s->used;
  out8[4] = s->n_reused;
  out8[5] = s->n_created;
  out8[6] = (int64_t)s->idle.size();
  out8[7] = s->d.hbm_budget;
  return MHX_OK;
}

// Copied under the server lock (workers assign s->err from their threads)
// into a buffer of the calling thread; valid until that thread's next call.
const char* mhx_server_last_error(mhx_server* s) {
  thread_local std::string buf;
  if (!s) return "";
  std::lock_guard<std::mutex> lk(s->mu);
  buf = s->err;
  return buf.c_str();
}

void mhx_server_destroy(mhx_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->queue.clear();  // queued requests are dropped; running ones finish
  }
  s->cv.notify_all();
  for (auto& t : s->workers) t.join();
  for (Slot& sl : s->idle) mhx_ctx_destroy(sl.ctx);
  delete s;
}

}  // extern "C"
