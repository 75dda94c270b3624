// Serving engine: continuous batching of alignment requests on one GPU.
//
// Requests (odometry scan-to-map, loop-closure verification, ...) arrive at
// any time from any thread.  An engine thread runs "steps": each step is
// kStepIters Gauss-Newton iterations (one while a request waits) for every
// in-flight alignment, each iteration one tile launch plus one finish launch
// per aligner kind over the concatenated tile list, and one host
// synchronization.  Between steps the engine retires finished alignments
// (their results become visible to waiters) and admits queued requests while
// slots and the HBM budget allow.  A request admitted while others are mid-way joins at the
// next step, so arrivals overlap in-flight work instead of waiting for a
// batch boundary.
//
// Each alignment's arithmetic depends only on its own tiles: its partial rows
// are reduced in its own tile order in FP64, so the result is bit-identical
// whether it runs alone or in any mix.
//
// Maps (fixed trees) are uploaded by the caller and stay resident; the caller
// declares them with kdh_server_add_resident so they count against the
// budget.  Moving trees are copied at submit and uploaded at admission, and
// freed when the request retires.  The reference serves nothing; it aligns one
// pair per call on the tracker / loop-closure threads
// (tracker/tracker_proc_frame_impl_.h:78, slam/slam_proc_loops_impl_.h:95-115).
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kdh.h"
#include "kdh_capi_internal.h"
#include "kdh_kernels.h"

using namespace kdh;

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kEmpty = 9;  // status of an unused slot (finish and tile kernels skip it)
// Gauss-Newton iterations per engine step while no request waits for
// admission (a waiting request gets a one-iteration step, so it joins at
// the next iteration boundary); alignments that finish inside a step skip
// the step's remaining launches.
constexpr int kStepIters = 4;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Req {
  int64_t id = 0;
  std::vector<kdh_node3f> nodes;
  std::vector<uint64_t> leaves;
  kdh_request rq{};
  kdh_prior prior{};
  bool has_prior = false;
  kdh_tree* moving = nullptr;
  int slot = -1;
  int state = 0;  // 0 queued, 1 active, 2 done
  bool cancel = false;
  bool admitted = false;
  int64_t bytes = 0;
  Clock::time_point t_submit, t_admit, t_done;
  kdh_request_result res{};
};

}  // namespace

struct kdh_server {
  kdh_ctx* ctx = nullptr;  // private context: its own stream
  int64_t budget = 0;
  int max_inflight = 0;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_idle;
  bool stop = false;
  int waiters = 0;  // threads inside kdh_server_wait / kdh_server_poll
  std::string err;
  std::deque<std::shared_ptr<Req>> queue;
  std::map<int64_t, std::shared_ptr<Req>> reqs;  // queued, active and finished-not-collected
  std::vector<std::shared_ptr<Req>> slots;
  std::map<const kdh_tree*, int64_t> resident;
  int64_t bytes_resident = 0, bytes_inflight = 0;
  int64_t next_id = 1, steps = 0, completed = 0;
  // device buffers owned by the engine thread
  int* d_status = nullptr;
  int* h_status = nullptr;
  int* d_tile_first = nullptr;
  std::vector<int> h_tile_first;
  float* d_partials = nullptr;
  size_t partials_cap = 0;
  TileDesc* d_tiles = nullptr;
  size_t tile_cap = 0;
  // the tile list and per-slot first tiles on the device, to skip the upload
  // while the active set is unchanged
  std::vector<TileDesc> dev_tp;
  std::vector<int> dev_tile_first;
  const TileDesc* dev_tp_ptr = nullptr;
};

static int64_t request_bytes(const kdh_request& r) {
  const int64_t tiles = (r.n_moving_leaves + kTile - 1) / kTile;
  // the moving tree as kdh_tree_upload lays it out: node and leaf records, node stamps
  return (int64_t)(align256((size_t)std::max<int64_t>(r.n_moving_nodes, 1) * 32) +
                   align256((size_t)std::max<int64_t>(r.n_moving_leaves, 1) * 32) +
                   align256((size_t)std::max<int64_t>(r.n_moving_nodes, 1) * 4)) +
         tiles * 4 * kCtSlots + (int64_t)sizeof(PairState);
}

template <typename T>
static hipError_t grow_dev(T** p, size_t* cap, size_t need) {
  if (need <= *cap && *p) return hipSuccess;
  size_t nc = std::max(need, 2 * *cap);
  if (*p) {
    hipError_t e = hipFree(*p);
    *p = nullptr;
    if (e != hipSuccess) return e;
  }
  hipError_t e = hipMalloc((void**)p, nc * sizeof(T));
  *cap = (e == hipSuccess) ? nc : 0;
  return e;
}

static void finish_req(kdh_server* S, const std::shared_ptr<Req>& r, int status) {
  // called with S->mu held
  r->t_done = Clock::now();
  r->state = 2;
  r->res.status = status;
  r->res.latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
  r->res.queue_us = r->admitted ? std::chrono::duration<double, std::micro>(r->t_admit - r->t_submit).count()
                                 : r->res.latency_us;
  ++S->completed;
}

static void engine(kdh_server* S) {
  kdh_ctx* ctx = S->ctx;
  hipSetDevice(ctx->device);
  const int M = S->max_inflight;
  PairState* h_pairs = ctx->h_pairs;  // pinned, sized M by kdh_server_create
  for (;;) {
    std::vector<std::shared_ptr<Req>> admitted, cancelled;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto has_active = [&] {
        for (auto& s : S->slots)
          if (s) return true;
        return false;
      };
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || has_active(); });
      if (S->stop) break;
      // admission: FIFO while a slot is free and the budget allows
      while (!S->queue.empty()) {
        auto r = S->queue.front();
        if (r->cancel) {
          S->queue.pop_front();
          finish_req(S, r, KDH_ERR_CANCELLED);
          continue;
        }
        int slot = -1;
        for (int i = 0; i < M; ++i)
          if (!S->slots[i]) {
            slot = i;
            break;
          }
        if (slot < 0) break;
        if (S->bytes_resident + S->bytes_inflight + r->bytes > S->budget) {
          // Nothing in flight will free memory: resident maps added after
          // this request was queued leave it no room, now or later.
          if (S->bytes_inflight == 0) {
            S->queue.pop_front();
            S->err = "request no longer fits beside the resident maps";
            finish_req(S, r, KDH_ERR_BUDGET);
            S->cv_done.notify_all();
            continue;
          }
          break;
        }
        S->queue.pop_front();
        r->slot = slot;
        r->state = 1;
        r->t_admit = Clock::now();
        r->admitted = true;
        S->slots[slot] = r;
        S->bytes_inflight += r->bytes;
        admitted.push_back(r);
      }
      for (int i = 0; i < M; ++i)
        if (S->slots[i] && S->slots[i]->cancel) cancelled.push_back(S->slots[i]);
    }
    // retire cancelled in-flight requests: mark their slots empty on the device
    for (auto& r : cancelled) {
      int st = kEmpty;
      hipMemcpyAsync(&ctx->d_pairs[r->slot].status, &st, sizeof(int), hipMemcpyHostToDevice, ctx->stream);
      hipStreamSynchronize(ctx->stream);
      std::lock_guard<std::mutex> lk(S->mu);
      if (r->moving) kdh_tree_free(ctx, r->moving);
      r->moving = nullptr;
      S->slots[r->slot] = nullptr;
      S->bytes_inflight -= r->bytes;
      finish_req(S, r, KDH_ERR_CANCELLED);
      S->cv_done.notify_all();
    }
    // upload and set up the admitted requests
    for (auto& r : admitted) {
      int rc = kdh_tree_upload(ctx, r->nodes.data(), (int64_t)r->nodes.size(), r->leaves.data(),
                               (int64_t)r->leaves.size(), &r->moving);
      PairState& ps = h_pairs[r->slot];
      if (rc == KDH_OK) rc = kdh_fill_pair(ctx, ps, r->rq.fixed, r->moving, r->rq.ct, r->rq.params, 0, 0);
      if (rc == KDH_OK) {
        memcpy(ps.X, r->rq.X0, sizeof(ps.X));
        memcpy(ps.Xf, r->rq.Xf, sizeof(ps.Xf));
        for (int i = 0; i < 6; ++i) ps.v[i] = r->rq.ct ? r->rq.v0[i] : 0.0;
        kdh_setup_prior(ps, r->has_prior ? &r->prior : nullptr);
        kdh_cache_from_state(ps);
        if (ps.n_tiles == 0) {  // no moving leaves: the first round fails min_good
          ps.iter = 1;
          ps.status = ps.sp.min_good_measurements > 0 ? 2 : 1;
          memset(&ps.stats_log[0], 0, sizeof(DevStats));
        }
        // from pinned memory, ordered before the step's launches on the same
        // stream: no wait here (a failed copy fails the step)
        S->h_status[r->slot] = ps.status;
        hipError_t e = hipMemcpyAsync(&ctx->d_pairs[r->slot], &ps, sizeof(PairState), hipMemcpyHostToDevice,
                                      ctx->stream);
        if (e == hipSuccess)
          e = hipMemcpyAsync(&S->d_status[r->slot], &S->h_status[r->slot], sizeof(int), hipMemcpyHostToDevice,
                             ctx->stream);
        if (e != hipSuccess) rc = KDH_ERR_HIP;
      }
      if (rc != KDH_OK) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->err = ctx->err;
        if (r->moving) kdh_tree_free(ctx, r->moving);
        r->moving = nullptr;
        S->slots[r->slot] = nullptr;
        S->bytes_inflight -= r->bytes;
        finish_req(S, r, rc);
        S->cv_done.notify_all();
      }
    }
    // one engine step: kStepIters Gauss-Newton iterations (one while a
    // request waits) for every in-flight alignment
    std::vector<int> active;
    int iters = 1;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (int i = 0; i < M; ++i)
        if (S->slots[i]) active.push_back(i);
      if (S->queue.empty()) iters = kStepIters;
    }
    if (active.empty()) continue;
    int n_t[2] = {0, 0};
    std::vector<TileDesc> lists[2];
    for (int slot : active) {
      const PairState& ps = h_pairs[slot];
      if (ps.status != 0) continue;  // completed at admission
      const int k = ps.ct;
      S->h_tile_first[slot] = n_t[k];
      for (int t = 0; t < ps.n_tiles; ++t) lists[k].push_back(make_tile_desc(ps, slot, t, ctx->tile_leaves));
      n_t[k] += ps.n_tiles;
    }
    bool ok = true;
    const size_t tiles_all = (size_t)n_t[0] + n_t[1];
    ok = ok && grow_dev(&S->d_tiles, &S->tile_cap, std::max<size_t>(tiles_all, 1)) == hipSuccess;
    const size_t part_all = (size_t)n_t[0] * slots(0) + (size_t)n_t[1] * slots(1);
    ok = ok && grow_dev(&S->d_partials, &S->partials_cap, std::max<size_t>(part_all, 1)) == hipSuccess;
    std::vector<TileDesc> tp;
    tp.reserve(tiles_all);
    tp.insert(tp.end(), lists[0].begin(), lists[0].end());
    tp.insert(tp.end(), lists[1].begin(), lists[1].end());
    if (ok && tiles_all) {
      if (!kdh_same_tiles(tp, S->dev_tp) || S->h_tile_first != S->dev_tile_first || S->dev_tp_ptr != S->d_tiles) {
        ok = hipMemcpyAsync(S->d_tiles, tp.data(), tiles_all * sizeof(TileDesc), hipMemcpyHostToDevice,
                            ctx->stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(S->d_tile_first, S->h_tile_first.data(), M * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream) == hipSuccess;
        // what the device now holds
        S->dev_tp = tp;
        S->dev_tile_first = S->h_tile_first;
        S->dev_tp_ptr = ok ? S->d_tiles : nullptr;
      }
      for (int it = 0; it < iters && ok; ++it) {
      size_t poff = 0, toff = 0;
      for (int k = 0; k < 2 && ok; ++k) {
        if (!n_t[k]) continue;
        LaunchArgs a{};
        a.pairs = ctx->d_pairs;
        a.tiles = S->d_tiles + toff;
        a.n_tiles_total = n_t[k];
        a.partials = S->d_partials + poff;
        a.grid = kdh_tile_grid(ctx, n_t[k]);
        a.tile_leaves = ctx->tile_leaves;  // the private context's, as kdh_fill_pair counted the tiles
        a.ct = k;
        a.lds_nodes = (k && ctx->lds_nodes > 1024) ? 1024 : ctx->lds_nodes;
        a.xcd_remap = ctx->xcd_remap;
        a.pair_first = 0;
        a.n_pairs = M;
        a.phases = 3;
        a.pair_tile_first = S->d_tile_first;
        a.status_out = S->d_status;
        ok = launch_tiles(a, ctx->stream) == hipSuccess;
        toff += n_t[k];
        poff += (size_t)n_t[k] * slots(k);
      }
      }
    }
    if (ok)
      ok = hipMemcpyAsync(S->h_status, S->d_status, M * sizeof(int), hipMemcpyDeviceToHost, ctx->stream) ==
           hipSuccess;
    ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
    // retire finished alignments
    std::vector<int> done;
    for (int slot : active)
      if (!ok || S->h_status[slot] != 0 || h_pairs[slot].status != 0) done.push_back(slot);
    // one pull of every slot's results into the pinned mirror, one wait
    hipError_t e = hipSuccess;
    if (!done.empty()) {
      e = kdh_pull_results(ctx, M) == KDH_OK ? hipSuccess : hipErrorUnknown;
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    for (int slot : done) {
      PairState& ps = h_pairs[slot];
      std::lock_guard<std::mutex> lk(S->mu);
      auto r = S->slots[slot];
      if (!r) continue;
      kdh_pair_result& out = r->res.r;
      memcpy(out.X, ps.X, sizeof(out.X));
      memcpy(out.v, ps.v, sizeof(out.v));
      out.iterations = ps.iter;
      if (ps.iter > 0)
        memcpy(&out.last, &ps.stats_log[ps.iter - 1], sizeof(kdh_stats));
      else
        memset(&out.last, 0, sizeof(kdh_stats));
      out.entry = out.last;
      out.result = kdh_result_code(ps, r->rq.params.max_iterations);
      if (r->moving) kdh_tree_free(ctx, r->moving);
      r->moving = nullptr;
      S->slots[slot] = nullptr;
      S->bytes_inflight -= r->bytes;
      finish_req(S, r, (ok && e == hipSuccess) ? KDH_OK : KDH_ERR_HIP);
      if (!ok) S->err = "step failed";
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      ++S->steps;
    }
    if (!done.empty()) S->cv_done.notify_all();
  }
}

int kdh_server_create(int device, int64_t budget_bytes, int32_t max_inflight, kdh_server** out) {
  if (!out || budget_bytes <= 0 || max_inflight < 1 || max_inflight > 4096) return KDH_ERR_INVALID;
  *out = nullptr;
  kdh_server* S = new kdh_server();
  int rc = kdh_ctx_create(device, &S->ctx);
  if (rc != KDH_OK) {
    delete S;
    return rc;
  }
  S->budget = budget_bytes;
  S->max_inflight = max_inflight;
  S->slots.resize(max_inflight);
  S->h_tile_first.assign(max_inflight, 0);
  rc = kdh_ensure_pairs(S->ctx, (size_t)max_inflight);
  bool ok = rc == KDH_OK;
  ok = ok && hipMalloc((void**)&S->d_status, sizeof(int) * max_inflight) == hipSuccess;
  ok = ok && hipMalloc((void**)&S->d_tile_first, sizeof(int) * max_inflight) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&S->h_status, sizeof(int) * max_inflight, hipHostMallocDefault) == hipSuccess;
  if (ok) {
    // every slot starts empty: status kEmpty so both kernels skip it
    for (int i = 0; i < max_inflight; ++i) {
      memset(&S->ctx->h_pairs[i], 0, sizeof(PairState));
      S->ctx->h_pairs[i].status = kEmpty;
      S->h_status[i] = kEmpty;
    }
    ok = hipMemcpy(S->ctx->d_pairs, S->ctx->h_pairs, sizeof(PairState) * max_inflight, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(S->d_status, S->h_status, sizeof(int) * max_inflight, hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) {
    if (S->d_status) hipFree(S->d_status);
    if (S->d_tile_first) hipFree(S->d_tile_first);
    if (S->h_status) hipHostFree(S->h_status);
    kdh_ctx_destroy(S->ctx);
    delete S;
    return KDH_ERR_NOMEM;
  }
  S->th = std::thread(engine, S);
  *out = S;
  return KDH_OK;
}

int kdh_server_destroy(kdh_server* S) {
  if (!S) return KDH_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  for (auto& kv : S->reqs)
    if (kv.second->moving) kdh_tree_free(S->ctx, kv.second->moving);
  hipSetDevice(S->ctx->device);
  if (S->d_status) hipFree(S->d_status);
  if (S->d_tile_first) hipFree(S->d_tile_first);
  if (S->d_partials) hipFree(S->d_partials);
  if (S->d_tiles) hipFree(S->d_tiles);
  if (S->h_status) hipHostFree(S->h_status);
  kdh_ctx_destroy(S->ctx);
  {
    // wake every kdh_server_wait (it returns KDH_ERR_CANCELLED) and free the
    // server only once the last waiter has left its lock
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.notify_all();
    S->cv_idle.wait(lk, [&] { return S->waiters == 0; });
  }
  delete S;
  return KDH_OK;
}

const char* kdh_server_last_error(kdh_server* S) { return S ? S->err.c_str() : "null server"; }

int kdh_server_add_resident(kdh_server* S, const kdh_tree* map) {
  if (!S || !map) return KDH_ERR_INVALID;
  if (map->device != S->ctx->device) return KDH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->resident.count(map)) return KDH_OK;
  if (S->bytes_resident + (int64_t)map->bytes > S->budget) {
    S->err = "resident maps exceed the budget";
    return KDH_ERR_BUDGET;
  }
  S->resident[map] = (int64_t)map->bytes;
  S->bytes_resident += (int64_t)map->bytes;
  return KDH_OK;
}

int kdh_server_remove_resident(kdh_server* S, const kdh_tree* map) {
  if (!S || !map) return KDH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->resident.find(map);
  if (it == S->resident.end()) return KDH_ERR_NOT_FOUND;
  S->bytes_resident -= it->second;
  S->resident.erase(it);
  S->cv_work.notify_all();
  return KDH_OK;
}

int kdh_server_submit(kdh_server* S, const kdh_request* rq, int64_t* id) {
  if (!S || !rq || !id) return KDH_ERR_INVALID;
  *id = 0;
  if (!rq->fixed || rq->n_moving_nodes < 0 || rq->n_moving_leaves < 0 ||
      (rq->n_moving_nodes > 0 && !rq->moving_nodes) || (rq->n_moving_leaves > 0 && !rq->moving_leaves) ||
      rq->params.max_iterations < 1 || rq->params.max_iterations > kMaxIter) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = "invalid request";
    return KDH_ERR_INVALID;
  }
  auto r = std::make_shared<Req>();
  r->rq = *rq;
  r->nodes.assign(rq->moving_nodes, rq->moving_nodes + rq->n_moving_nodes);
  r->leaves.assign(rq->moving_leaves, rq->moving_leaves + rq->n_moving_leaves);
  if (rq->prior) {
    r->prior = *rq->prior;
    r->has_prior = true;
  }
  r->rq.moving_nodes = nullptr;
  r->rq.moving_leaves = nullptr;
  r->rq.prior = nullptr;
  r->bytes = request_bytes(*rq);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (r->bytes + S->bytes_resident > S->budget) {
      S->err = "request larger than the memory budget";
      return KDH_ERR_BUDGET;
    }
    r->id = S->next_id++;
    S->reqs[r->id] = r;
    S->queue.push_back(r);
    *id = r->id;
  }
  S->cv_work.notify_all();
  return KDH_OK;
}

// Counts a caller inside kdh_server_wait / kdh_server_poll (constructed and
// destroyed with S->mu held), so that kdh_server_destroy frees the server
// only after every such caller has left.
struct WaiterGuard {
  kdh_server* S;
  explicit WaiterGuard(kdh_server* s) : S(s) { ++S->waiters; }
  ~WaiterGuard() {
    if (--S->waiters == 0) S->cv_idle.notify_all();
  }
};

int kdh_server_poll(kdh_server* S, int64_t id, int* done, kdh_request_result* out) {
  if (!S || !done) return KDH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  WaiterGuard guard(S);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return KDH_ERR_NOT_FOUND;
  *done = it->second->state == 2;
  if (*done) {
    if (out) *out = it->second->res;
    S->reqs.erase(it);
  }
  return KDH_OK;
}

int kdh_server_wait(kdh_server* S, int64_t id, double timeout_s, kdh_request_result* out) {
  if (!S) return KDH_ERR_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  WaiterGuard guard(S);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return KDH_ERR_NOT_FOUND;
  auto r = it->second;
  auto pred = [&] { return r->state == 2 || S->stop; };
  if (timeout_s > 0) {
    if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return KDH_ERR_NOT_FOUND;
  } else {
    S->cv_done.wait(lk, pred);
  }
  if (r->state != 2) return KDH_ERR_CANCELLED;
  if (out) *out = r->res;
  S->reqs.erase(id);
  return KDH_OK;
}

int kdh_server_cancel(kdh_server* S, int64_t id) {
  if (!S) return KDH_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return KDH_ERR_NOT_FOUND;
    it->second->cancel = true;
  }
  S->cv_work.notify_all();
  return KDH_OK;
}

int kdh_server_info_get(kdh_server* S, kdh_server_info* info) {
  if (!S || !info) return KDH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  info->bytes_budget = S->budget;
  info->bytes_resident = S->bytes_resident;
  info->bytes_inflight = S->bytes_inflight;
  int inflight = 0;
  for (auto& s : S->slots) inflight += s ? 1 : 0;
  info->inflight = inflight;
  info->queued = (int32_t)S->queue.size();
  info->steps = S->steps;
  info->completed = S->completed;
  return KDH_OK;
}
