// Batched solves and the continuous-batching request loop.
//
// No reference counterpart on the device: the reference's only request loop is
// the viewer's SolveRequest::run (crates/studio/src/solve.rs:41-75), one
// solve at a time on the host. Each request here is one Hodge-Laplace source
// solve (elliptic.rs:132-182) on a jittered Kuhn box.
//
// Batch: every request's mesh, numbering and KKT plans are built by a pool of
// worker threads, each with its own context (HIP stream); then one persistent
// launch computes the element blocks of all requests, of every dimension,
// from a (mesh, 64-cell range) work queue; then the workers scatter values and
// upload sources, and the MINRES runs of all requests advance in lockstep, one
// launch per phase covering every request that is still running (each with
// the blocks and partial sums of a solve on its own, so results equal single
// solves bit for bit).
//
// Server: `max_inflight` slots, each a worker thread with its own stream. A
// queued request is admitted when a slot is free and its estimated resident
// footprint fits the remaining HBM budget (requests that fit are admitted past
// one that does not); its mesh, CSR and Krylov state stay on the device until it
// completes. A request's result depends only on its own inputs: every
// reduction has a fixed order and no buffer is shared between requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "feec_internal.hpp"

namespace feec {
struct Problem;
void problem_build(Problem& P);
void problem_assemble(Problem& P);
void problem_assemble_from_elements(Problem& P);
void problem_element_outputs(Problem& P, double** outs);
Problem* problem_new(Geom* g, int bc, int k, const feec_solve_opts& o);
void problem_delete(Problem* P);
feec_report problem_solve_host(Problem* P, const double* src);
void problem_prepare_rhs(Problem* P, const double* src);
void problem_solve_batch(std::vector<Problem*>& probs, Ctx* lead, std::vector<feec_report>& reps);
void problem_set_report(Problem* P, const feec_report& r);
void problem_u(Problem* P, double* u);
int64_t problem_ndofs(Problem* P);
}  // namespace feec

using namespace feec;

namespace {

using clk = std::chrono::steady_clock;

double seconds_since(clk::time_point t0) {
  return std::chrono::duration<double>(clk::now() - t0).count();
}

int64_t kuhn_cells(int d, int n) {
  int64_t c = 1, f = 1;
  for (int a = 0; a < d; ++a) {
    c *= n;
    f *= (a + 1);
  }
  return c * f;
}

// Resident bytes of one request, estimated from its element block size.
int64_t estimate_bytes(const feec_request& r) {
  int nfm = r.grade >= 1 ? binom(r.dim + 1, r.grade) : 0;
  int nf = binom(r.dim + 1, r.grade + 1);
  int64_t per = (int64_t)(nfm + nf) * (nfm + nf);
  return kuhn_cells(r.dim, r.n) * per * 48 + (64ll << 20);
}

void check_request(const feec_request& r) {
  FEEC_REQUIRE(r.dim >= 2 && r.dim <= 4, FEEC_E_ARG, "request dimension must be 2..4");
  FEEC_REQUIRE(r.grade >= 0 && r.grade <= r.dim, FEEC_E_ARG, "request grade out of range");
  FEEC_REQUIRE(r.n >= 1 && r.n <= 4096, FEEC_E_ARG, "request grid size out of range");
  FEEC_REQUIRE(r.bc == FEEC_FULL || r.bc == FEEC_RELATIVE, FEEC_E_ARG, "bad boundary condition");
}

struct Instance {
  feec_request req{};
  feec_mesh* mesh = nullptr;
  feec_geom* geom = nullptr;
  Problem* prob = nullptr;
  int64_t ncells = 0, nk = 0;
  ~Instance() {
    if (prob) problem_delete(prob);
    feec_geom_destroy(geom);
    feec_mesh_destroy(mesh);
  }
};

void check_status(feec_status s) {
  if (s != FEEC_OK) fail(s, feec_last_error());
}

void instance_build(Instance& I, feec_ctx* ctx, const feec_request& r, const feec_solve_opts& base,
                    bool assemble) {
  check_request(r);
  I.req = r;
  int64_t nc = 0, nv = 0;
  check_status(feec_kuhn_counts(r.dim, r.n, &nc, &nv));
  std::vector<int64_t> cells(nc * (r.dim + 1));
  std::vector<double> coords(nv * r.dim);
  check_status(feec_kuhn_grid(r.dim, r.n, cells.data(), coords.data()));
  if (r.jitter != 0.0)
    check_status(feec_jitter_interior(r.dim, r.n, coords.data(), r.jitter, r.mesh_seed,
                                      (1u << r.dim) - 1u));
  check_status(feec_mesh_create(ctx, r.dim, nv, nc, cells.data(), &I.mesh));
  check_status(feec_geom_from_coords(I.mesh, coords.data(), FEEC_EUCLIDEAN, &I.geom));
  feec_solve_opts o = base;
  o.precision = r.precision;
  I.prob = problem_new(I.geom, r.bc, r.grade, o);
  I.ncells = nc;
  I.nk = I.mesh->n[r.grade];
  if (assemble) problem_assemble(*I.prob);
}

void instance_result(Instance& I, const feec_report& rep, feec_result& res,
                     std::vector<double>* u_out);

void instance_solve(Instance& I, feec_result& res, std::vector<double>* u_out) {
  std::vector<double> src(I.nk);
  check_status(feec_normal_vector(I.nk, I.req.source_seed, src.data()));
  const auto t0 = clk::now();
  feec_report rep = problem_solve_host(I.prob, src.data());
  const double solve = seconds_since(t0);
  instance_result(I, rep, res, u_out);
  res.solve_seconds = solve;
}

void instance_result(Instance& I, const feec_report& rep, feec_result& res,
                     std::vector<double>* u_out) {
  problem_set_report(I.prob, rep);
  std::vector<double> u(I.nk);
  problem_u(I.prob, u.data());
  double nrm = 0.0, chk = 0.0;
  for (int64_t i = 0; i < I.nk; ++i) {
    nrm += u[i] * u[i];
    chk += u[i] * (double)((i % 7) - 3);
  }
  res.ndofs = problem_ndofs(I.prob);
  res.ncells = I.ncells;
  res.iters = rep.iters;
  res.refinements = rep.refinements;
  res.residual = rep.residual;
  res.status = rep.converged ? FEEC_OK : FEEC_E_NOT_CONVERGED;
  res.u_norm = std::sqrt(nrm);
  res.u_checksum = chk;
  if (u_out) *u_out = std::move(u);
}

}  // namespace

// ===========================================================================
// server
// ===========================================================================
struct feec_server_s {
  feec_ctx* ctx = nullptr;
  feec_server_opts opts{};
  std::mutex mu;
  std::condition_variable cv;
  struct Entry {
    feec_request req{};
    int state = 0;  // 0 queued, 1 running, 2 done
    double submit = 0.0;
    feec_result result{};
    std::vector<double> u;
    int64_t bytes = 0;
  };
  std::map<int64_t, Entry> entries;
  std::deque<int64_t> queue;
  int64_t next_ticket = 1;
  int64_t hbm_in_use = 0;
  int running = 0;
  bool stop = false;
  clk::time_point t0;
  std::vector<std::thread> workers;
  std::vector<feec_ctx*> worker_ctx;

  void worker(int w) {
    feec_ctx* wctx = worker_ctx[w];
    (void)hipSetDevice(wctx->device);
    for (;;) {
      int64_t ticket = -1;
      feec_request req{};
      int64_t bytes = 0;
      {
        std::unique_lock<std::mutex> lock(mu);
        cv.wait(lock, [&] {
          if (stop) return true;
          for (int64_t t : queue)
            if (entries[t].bytes + hbm_in_use <= opts.hbm_budget_bytes) return true;
          return false;
        });
        if (stop) return;
        for (auto it = queue.begin(); it != queue.end(); ++it) {
          Entry& e = entries[*it];
          if (e.bytes + hbm_in_use <= opts.hbm_budget_bytes) {
            ticket = *it;
            queue.erase(it);
            e.state = 1;
            e.result.admit_time = seconds_since(t0);
            hbm_in_use += e.bytes;
            ++running;
            req = e.req;
            bytes = e.bytes;
            break;
          }
        }
      }
      feec_result res{};
      std::vector<double> u;
      feec_status st = guard([&] {
        Instance I;
        const auto tb = clk::now();
        instance_build(I, wctx, req, opts.solve, true);
        const double build = seconds_since(tb);
        instance_solve(I, res, &u);
        res.build_seconds = build;
      });
      if (st != FEEC_OK) res.status = st;
      {
        std::lock_guard<std::mutex> lock(mu);
        Entry& e = entries[ticket];
        const double admit = e.result.admit_time;
        e.result = res;
        e.result.ticket = ticket;
        e.result.admit_time = admit;
        e.result.finish_time = seconds_since(t0);
        e.result.latency = e.result.finish_time - e.submit;
        e.u = std::move(u);
        e.state = 2;
        hbm_in_use -= bytes;
        --running;
      }
      cv.notify_all();
    }
  }
};

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

feec_status feec_batch_run(feec_ctx* ctx, int nreq, const feec_request* reqs,
                           const feec_solve_opts* opts, feec_result* results) {
  return guard([&] {
    FEEC_REQUIRE(ctx && (nreq == 0 || (reqs && results)), FEEC_E_ARG, "null argument");
    if (nreq == 0) return;
    feec_solve_opts base;
    if (opts)
      base = *opts;
    else
      feec_solve_opts_default(&base);
    // the lockstep runs advance the polynomial preconditioner of every system
    // in one launch per phase; the multigrid cycle is per system
    if (base.preconditioner == FEEC_PC_AUTO) base.preconditioner = FEEC_PC_CHEBYSHEV;
    for (int i = 0; i < nreq; ++i) check_request(reqs[i]);
    const int W = std::min(nreq, 8);
    std::vector<feec_ctx*> wctx(W, nullptr);
    for (int w = 0; w < W; ++w) check_status(feec_ctx_create(ctx->device, &wctx[w]));
    std::vector<std::unique_ptr<Instance>> inst(nreq);
    std::vector<feec_status> status(nreq, FEEC_OK);
    std::vector<std::string> errors(nreq);
    auto for_workers = [&](auto&& body) {
      std::vector<std::thread> th;
      for (int w = 0; w < W; ++w)
        th.emplace_back([&, w] {
          (void)hipSetDevice(ctx->device);
          for (int i = w; i < nreq; i += W) {
            if (status[i] != FEEC_OK) continue;
            status[i] = guard([&] { body(i, wctx[w]); });
            if (status[i] != FEEC_OK) errors[i] = feec_last_error();
          }
        });
      for (auto& t : th) t.join();
    };
    auto t0 = clk::now();
    // 1. meshes, numbering and plans
    std::vector<double> t_build(nreq, 0.0);
    for_workers([&](int i, feec_ctx* c) {
      const auto tb = clk::now();
      inst[i] = std::make_unique<Instance>();
      instance_build(*inst[i], c, reqs[i], base, false);
      t_build[i] = seconds_since(tb);
    });
    // 2. one persistent launch for every element block of every request
    std::vector<BatchElementRequest> br;
    for (int i = 0; i < nreq; ++i) {
      if (status[i] != FEEC_OK) continue;
      BatchElementRequest r{};
      r.g = inst[i]->geom;
      r.grade = reqs[i].grade;
      problem_element_outputs(*inst[i]->prob, r.outs);
      br.push_back(r);
    }
    for (auto* c : wctx) c->sync();
    kkt_element_blocks_batched(ctx, br);
    ctx->sync();
    const double t_elements = seconds_since(t0);
    // 3. values, preconditioner and sources per request on the worker
    //    streams; then every solve with its MINRES runs in lockstep (one
    //    launch per phase for all requests); then the results
    std::vector<double> t_done(nreq, 0.0);
    for_workers([&](int i, feec_ctx*) {
      problem_assemble_from_elements(*inst[i]->prob);
      std::vector<double> src(inst[i]->nk);
      check_status(feec_normal_vector(inst[i]->nk, inst[i]->req.source_seed, src.data()));
      problem_prepare_rhs(inst[i]->prob, src.data());
    });
    for (auto* c : wctx) c->sync();
    const double t_assembled = seconds_since(t0);
    std::vector<Problem*> probs;
    std::vector<int> owner;
    for (int i = 0; i < nreq; ++i)
      if (status[i] == FEEC_OK) {
        probs.push_back(inst[i]->prob);
        owner.push_back(i);
      }
    std::vector<feec_report> reps;
    if (!probs.empty()) {
      feec_status bs = guard([&] { problem_solve_batch(probs, ctx, reps); });
      if (bs != FEEC_OK) {
        const std::string msg = feec_last_error();
        for (int i : owner) {
          status[i] = bs;
          errors[i] = msg;
        }
      }
    }
    const double t_solved = seconds_since(t0);
    std::vector<feec_report> rep_of(nreq);
    for (size_t a = 0; a < owner.size() && a < reps.size(); ++a) rep_of[owner[a]] = reps[a];
    for_workers([&](int i, feec_ctx*) {
      results[i] = feec_result{};
      instance_result(*inst[i], rep_of[i], results[i], nullptr);
      results[i].solve_seconds = t_solved - t_assembled;
      t_done[i] = seconds_since(t0);
    });
    for (int i = 0; i < nreq; ++i) {
      if (status[i] != FEEC_OK) {
        results[i] = feec_result{};
        results[i].status = status[i];
      }
      results[i].ticket = i;
      // batch phases: meshes, plans and all element blocks done by admit_time,
      // this request's values and solve by finish_time
      results[i].admit_time = t_elements;
      results[i].finish_time = t_done[i];
      results[i].latency = t_done[i];
      results[i].build_seconds = t_build[i];
    }
    // tear down on the workers too (hundreds of stream-ordered frees per request)
    for_workers([&](int i, feec_ctx*) { inst[i].reset(); });
    inst.clear();
    std::vector<std::thread> th;
    for (auto* c : wctx) th.emplace_back([c] { feec_ctx_destroy(c); });
    for (auto& t : th) t.join();
  });
}

feec_status feec_server_create(feec_ctx* ctx, const feec_server_opts* opts, feec_server** out) {
  return guard([&] {
    FEEC_REQUIRE(ctx && opts && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(opts->max_inflight >= 1 && opts->max_inflight <= 16, FEEC_E_ARG,
                 "max_inflight must be 1..16");
    auto* s = new feec_server_s();
    s->ctx = ctx;
    s->opts = *opts;
    if (s->opts.hbm_budget_bytes <= 0) s->opts.hbm_budget_bytes = 64ll << 30;
    s->t0 = clk::now();
    for (int w = 0; w < opts->max_inflight; ++w) {
      feec_ctx* c = nullptr;
      feec_status st = feec_ctx_create(ctx->device, &c);
      if (st != FEEC_OK) {
        for (auto* cc : s->worker_ctx) feec_ctx_destroy(cc);
        delete s;
        fail(st, feec_last_error());
      }
      s->worker_ctx.push_back(c);
    }
    for (int w = 0; w < opts->max_inflight; ++w) s->workers.emplace_back([s, w] { s->worker(w); });
    *out = s;
  });
}

feec_status feec_server_submit(feec_server* s, const feec_request* req, int64_t* ticket) {
  return guard([&] {
    FEEC_REQUIRE(s && req && ticket, FEEC_E_ARG, "null argument");
    check_request(*req);
    int64_t bytes = estimate_bytes(*req);
    FEEC_REQUIRE(bytes <= s->opts.hbm_budget_bytes, FEEC_E_ADMISSION,
                 "request footprint exceeds the server's HBM budget");
    {
      std::lock_guard<std::mutex> lock(s->mu);
      int64_t t = s->next_ticket++;
      auto& e = s->entries[t];
      e.req = *req;
      e.bytes = bytes;
      e.submit = seconds_since(s->t0);
      s->queue.push_back(t);
      *ticket = t;
    }
    s->cv.notify_all();
  });
}

feec_status feec_server_poll(feec_server* s, int64_t ticket, feec_result* result) {
  feec_status st = FEEC_OK;
  feec_status g = guard([&] {
    FEEC_REQUIRE(s && result, FEEC_E_ARG, "null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->entries.find(ticket);
    FEEC_REQUIRE(it != s->entries.end(), FEEC_E_ARG, "unknown ticket");
    if (it->second.state != 2) {
      st = FEEC_E_AGAIN;
      return;
    }
    *result = it->second.result;
  });
  return g != FEEC_OK ? g : st;
}

feec_status feec_server_solution(feec_server* s, int64_t ticket, double* u, int64_t* n) {
  return guard([&] {
    FEEC_REQUIRE(s, FEEC_E_ARG, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->entries.find(ticket);
    FEEC_REQUIRE(it != s->entries.end(), FEEC_E_ARG, "unknown ticket");
    FEEC_REQUIRE(it->second.state == 2, FEEC_E_AGAIN, "request not finished");
    if (n) *n = (int64_t)it->second.u.size();
    if (u) std::copy(it->second.u.begin(), it->second.u.end(), u);
  });
}

feec_status feec_server_drain(feec_server* s) {
  return guard([&] {
    FEEC_REQUIRE(s, FEEC_E_ARG, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv.wait(lock, [&] { return s->queue.empty() && s->running == 0; });
  });
}

feec_status feec_server_release(feec_server* s, int64_t ticket) {
  return guard([&] {
    FEEC_REQUIRE(s, FEEC_E_ARG, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->entries.find(ticket);
    FEEC_REQUIRE(it != s->entries.end(), FEEC_E_ARG, "unknown ticket");
    FEEC_REQUIRE(it->second.state == 2, FEEC_E_AGAIN, "request not finished");
    s->entries.erase(it);
  });
}

feec_status feec_server_held(feec_server* s, int64_t* entries, int64_t* solution_doubles) {
  return guard([&] {
    FEEC_REQUIRE(s, FEEC_E_ARG, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    int64_t held = 0;
    for (auto& kv : s->entries) held += (int64_t)kv.second.u.size();
    if (entries) *entries = (int64_t)s->entries.size();
    if (solution_doubles) *solution_doubles = held;
  });
}

void feec_server_destroy(feec_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto* c : s->worker_ctx) feec_ctx_destroy(c);
  delete s;
}

}  // extern "C"
