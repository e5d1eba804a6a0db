// serve.cpp -- continuous batching of build requests on one device.
//
// Requests are queued by svo_server_submit.  n_lanes worker threads each own
// a pipeline lane (a HIP stream with its own workspace).  A free worker forms
// a batch from the head of the queue: it waits up to max_wait_us after the
// oldest pending request for more to arrive, then admits requests in arrival
// order while upper bounds on their device memory fit the budget together
// with everything already resident or in flight.  Batches built by different
// workers run concurrently on the device.  Finished trees stay resident until
// the caller frees them; the budget check sees that through the context's
// byte counter.
//
// What a request reserves: the grid bound (every level full; O(1), an upper
// bound) when that is at most 1/32 of the budget; else the count-only
// estimate for a closed surface (O(1), not a bound) when that is below 1/8 of
// the budget; else the bound from its triangles (estimate_bytes_mesh, a host
// pass of a few ns per triangle plus the marking of its level-2 cells),
// which intake threads compute while earlier requests build -- the request
// is admitted once it is known.  Computing that bound for every request would
// cost milliseconds of host time per large-resolution request, more than its
// build, so small requests are admitted on the estimate and a pass that runs
// out of memory is retried.  Without a budget nothing is reserved.
//
// Results are per request: a batch can run as several passes, and when one of
// them fails the requests of the others still complete with their trees.  A
// request whose pass ran out of memory (an estimate below the real need, or
// device memory taken by other users of the context) goes back to the head of
// the queue and is retried while it is less than a second old; after that it
// completes with SVO_E_OOM.
#include <chrono>
#include <deque>
#include <thread>

#include "svo_host.h"

using namespace svo;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  uint64_t id;
  std::vector<float> pts;
  std::vector<int32_t> tris;
  svo_params params;
  uint64_t estimate;  // upper bound on its device bytes, once `ready`
  bool ready;
  double t_submit;
};
using RequestPtr = std::shared_ptr<Request>;

}  // namespace

struct svo_server {
  std::shared_ptr<Context> ctx;
  svo_server_cfg cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<RequestPtr> pending;
  std::deque<RequestPtr> intake;  // requests whose bound is still to compute
  std::deque<svo_completion> done;
  std::vector<std::thread> workers, intakers;
  std::condition_variable cv_intake;
  bool stop = false;
  uint64_t next_id = 1;
  uint64_t in_flight_bytes = 0;  // estimates of batches being built
  uint32_t in_flight = 0;
  uint64_t batches_run = 0;      // passes' batches finished
  uint64_t retried = 0;          // requests re-queued after running out of memory

  void worker_loop();
  void intake_loop();
};

void svo_server::intake_loop() {
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    cv_intake.wait(lk, [&] { return stop || !intake.empty(); });
    if (stop) return;
    RequestPtr r = intake.front();
    intake.pop_front();
    lk.unlock();
    svo_mesh_desc d{r->pts.data(), r->pts.size() / 3, r->tris.data(), r->tris.size() / 3};
    const uint64_t e = estimate_bytes_mesh(&d, &r->params);
    lk.lock();
    r->estimate = e;
    r->ready = true;
    cv_work.notify_all();
  }
}

void svo_server::worker_loop() {
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    cv_work.wait(lk, [&] { return stop || !pending.empty(); });
    if (stop && pending.empty()) return;
    // batching window measured from the oldest pending request
    const double deadline = pending.front()->t_submit + cfg.max_wait_us * 1e-6;
    while (!stop && pending.size() < cfg.max_batch) {
      const double left = deadline - now_s();
      if (left <= 0) break;
      cv_work.wait_for(lk, std::chrono::duration<double>(left));
      if (pending.empty()) break;
    }
    if (pending.empty()) continue;
    // admission against the budget, in arrival order, of requests whose
    // bound is known
    std::vector<RequestPtr> batch;
    uint64_t est = 0;
    // resident trees; lane workspaces are reused by the batch and are part
    // of its bound
    const uint64_t used = ctx->in_use.load(), ws = ctx->ws_bytes.load();
    const uint64_t resident = used > ws ? used - ws : 0;
    while (!pending.empty() && batch.size() < cfg.max_batch && pending.front()->ready) {
      const uint64_t e = pending.front()->estimate;
      if (cfg.mem_budget && resident + in_flight_bytes + est + e > cfg.mem_budget &&
          !batch.empty())
        break;
      if (cfg.mem_budget && resident + in_flight_bytes + est + e > cfg.mem_budget &&
          batch.empty() && in_flight > 0) {
        break;  // wait for the batches in flight
      }
      est += e;
      batch.push_back(std::move(pending.front()));
      pending.pop_front();
    }
    if (batch.empty()) {
      // the head's bound is being computed, or memory is short
      cv_work.wait_for(lk, std::chrono::microseconds(200));
      continue;
    }
    in_flight_bytes += est;
    ++in_flight;
    lk.unlock();
    const double t_start = now_s();
    std::vector<MeshInput> in(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) {
      in[i].desc.xyz = batch[i]->pts.data();
      in[i].desc.n_points = batch[i]->pts.size() / 3;
      in[i].desc.tri = batch[i]->tris.data();
      in[i].desc.n_tris = batch[i]->tris.size() / 3;
      in[i].params = batch[i]->params;
      in[i].on_device = false;
    }
    std::vector<svo_tree*> trees;
    std::vector<int> st;
    // per-request results: a batch split into several passes can fail in
    // one pass while the others build their trees
    run_batch(*ctx, in, trees, st);
    const double t_done = now_s();
    lk.lock();
    in_flight_bytes -= est;
    --in_flight;
    // requests whose pass ran out of memory go back to the head of the
    // queue, in arrival order, while they are less than a second old (other
    // batches or freed trees may release memory meanwhile); the others
    // complete now
    std::vector<RequestPtr> retry;
    for (size_t i = 0; i < batch.size(); ++i) {
      if (st[i] == SVO_E_OOM && !stop && now_s() - batch[i]->t_submit < 1.0) {
        if (trees[i]) svo_tree_free(trees[i]);
        retry.push_back(std::move(batch[i]));
        continue;
      }
      svo_completion c{};
      c.request_id = batch[i]->id;
      c.status = st[i];
      c.batch_size = (uint32_t)batch.size();
      c.tree = st[i] == SVO_OK ? trees[i] : nullptr;
      if (!c.tree && trees[i]) svo_tree_free(trees[i]);
      c.t_submit_s = batch[i]->t_submit;
      c.t_start_s = t_start;
      c.t_done_s = t_done;
      done.push_back(c);
    }
    ++batches_run;
    if (!retry.empty()) {
      retried += retry.size();
      for (auto it = retry.rbegin(); it != retry.rend(); ++it) pending.push_front(std::move(*it));
      if (retry.size() == batch.size()) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(200));
        lk.lock();
      }
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

int svo_server_create(svo_ctx* ctx, const svo_server_cfg* cfg, svo_server** out) {
  if (!ctx || !cfg || !out) return SVO_E_INVALID;
  svo_server* s = new svo_server();
  s->ctx = ctx->impl;
  s->ctx->servers.fetch_add(1);
  s->cfg = *cfg;
  if (!s->cfg.max_batch) s->cfg.max_batch = 64;
  if (!s->cfg.max_queue) s->cfg.max_queue = 1 << 16;
  if (!s->cfg.n_lanes) s->cfg.n_lanes = 2;
  s->cfg.n_lanes = std::min<uint32_t>(s->cfg.n_lanes, 4);
  for (uint32_t i = 0; i < s->cfg.n_lanes; ++i) s->workers.emplace_back([s] { s->worker_loop(); });
  if (s->cfg.mem_budget)
    for (int i = 0; i < 4; ++i) s->intakers.emplace_back([s] { s->intake_loop(); });
  *out = s;
  return SVO_OK;
}

int svo_server_submit(svo_server* srv, const svo_mesh_desc* mesh, const svo_params* params,
                      uint64_t* request_id) {
  if (!srv || !mesh || !params) return SVO_E_INVALID;
  const int v = validate_params(mesh, params, nullptr);
  if (v != SVO_OK) return v;
  int kb = 32;
  validate_params(mesh, params, &kb);
  auto r = std::make_shared<Request>();
  r->pts.assign(mesh->xyz, mesh->xyz + 3 * mesh->n_points);
  r->tris.assign(mesh->tri, mesh->tri + 3 * mesh->n_tris);
  r->params = *params;
  // the grid bound does not hold when the context forces the sort path
  bool forced_sort = false;
  {
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    forced_sort = srv->ctx->tuning.force_path == 2;
  }
  const uint64_t budget = srv->cfg.mem_budget;
  const uint64_t grid = grid_bound_bytes(params, kb, mesh->n_points, mesh->n_tris);
  const uint64_t quick = estimate_bytes(mesh->n_points, mesh->n_tris, params);
  r->ready = true;
  if (!budget) {
    r->estimate = quick;
  } else if (!forced_sort && grid <= budget / 32) {
    r->estimate = grid;
  } else if (quick < budget / 8) {
    r->estimate = quick;
  } else {
    r->estimate = grid;  // replaced by the mesh bound before admission
    r->ready = false;
  }
  r->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->pending.size() >= srv->cfg.max_queue) return SVO_E_BUSY;
    r->id = srv->next_id++;
    if (request_id) *request_id = r->id;
    srv->pending.push_back(r);
    if (!r->ready) srv->intake.push_back(r);
  }
  if (!r->ready) srv->cv_intake.notify_one();
  srv->cv_work.notify_one();
  return SVO_OK;
}

int svo_server_poll(svo_server* srv, svo_completion* out, uint32_t cap, uint32_t* n,
                    int64_t timeout_us) {
  if (!srv || !n || (cap && !out)) return SVO_E_INVALID;
  std::unique_lock<std::mutex> lk(srv->mu);
  if (srv->done.empty() && timeout_us > 0)
    srv->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us),
                          [&] { return !srv->done.empty(); });
  uint32_t k = 0;
  while (k < cap && !srv->done.empty()) {
    out[k++] = srv->done.front();
    srv->done.pop_front();
  }
  *n = k;
  return SVO_OK;
}

int svo_server_stats(svo_server* srv, uint64_t* bytes_reserved, uint32_t* pending,
                     uint32_t* in_flight) {
  if (!srv) return SVO_E_INVALID;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (bytes_reserved) *bytes_reserved = srv->in_flight_bytes + srv->ctx->in_use.load();
  if (pending) *pending = (uint32_t)srv->pending.size();
  if (in_flight) *in_flight = srv->in_flight;
  return SVO_OK;
}

int svo_server_destroy(svo_server* srv) {
  if (!srv) return SVO_E_INVALID;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_intake.notify_all();
  for (auto& t : srv->intakers) t.join();
  {
    // requests whose bound was never computed are admitted on the grid bound
    std::lock_guard<std::mutex> lk(srv->mu);
    for (auto& r : srv->pending) r->ready = true;
  }
  srv->cv_work.notify_all();
  for (auto& t : srv->workers) t.join();
  for (auto& c : srv->done)
    if (c.tree) svo_tree_free(c.tree);
  srv->ctx->servers.fetch_sub(1);
  delete srv;
  return SVO_OK;
}

}  // extern "C"
