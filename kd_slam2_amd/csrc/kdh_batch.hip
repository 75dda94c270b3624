// Many alignments in one launch (kdh_batch_align).
//
// All pairs of a call share one tile list per aligner kind (rigid / CT); each
// Gauss-Newton iteration is one persistent launch per kind over that list,
// and the iteration is captured once as a HIP graph and replayed until every
// pair has finished.  Per-pair results do not depend on the other pairs: the
// tile decomposition, the per-tile reduction and the FP64 tile-order sum of a
// pair are the same alone and in any mix.
//
// The reference runs alignments one after another on the host
// (loop-closure candidates, slam/slam_proc_loops_impl_.h:95-115; bundle
// factors, map/multiview_icp_factor_impl_.h:10-62); this is the batched form
// of that loop.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/kdh.h"
#include "kdh_capi_internal.h"
#include "kdh_kernels.h"
#include "kdh_solve.h"

using namespace kdh;

#define KDH_HIP_TRY(ctx, expr)                                 \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) return kdh_hip_fail(ctx, _e, #expr); \
  } while (0)

namespace {

// The tiles of one aligner kind within one pair range: group ct + 2 * half
// holds the pairs of kind ct in [p_lo, p_hi); a batch run on two streams
// splits its pairs into two halves (half 0, half 1) at a forest boundary.
constexpr int kGroups = 4;
struct Group {
  int ct = 0;
  int n_tiles = 0;
  size_t partial_off = 0;  // floats
  size_t tile_off = 0;     // ints
  int p_lo = 0, p_hi = 0;  // pair range (finish launch)
};

}  // namespace

// Pair states (into hs[0 .. n_pairs)), forests and the per-kind tile lists
// of a batch, on the host.  mode 0: align (solve + update each iteration);
// mode 1: build the form once.  dual: the CT pairs build the dual form, with
// the fixed velocities vf (n_pairs x 6).  halves = 2: the pairs are split
// into two ranges of about equal tile count (group ct + 2 * half), for the
// two-stream replay.  tp: the tile list (one record per tile: its pair, the
// pair's trees and the tile's first leaf; groups one after another);
// n_partials: floats of the partial rows.
static int batch_plan(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, int mode, int stats_mode, Group* g,
                      std::vector<int>& leader, int& max_iter, int dual, const double* vf, int halves,
                      PairState* hs, std::vector<TileDesc>& tp, size_t& n_partials) {
  int rc = KDH_OK;
  for (int k = 0; k < kGroups; ++k) {
    g[k] = Group();
    g[k].ct = k & 1;
  }
  max_iter = 0;
  // Forests: pair i with forest = k leads pairs i+1 .. i+k.
  leader.assign(n_pairs, 0);
  for (int i = 0; i < n_pairs;) {
    const int k = pairs[i].forest;
    if (k < 0 || k >= n_pairs - i)
      return kdh_fail(ctx, KDH_ERR_INVALID, "pair " + std::to_string(i) + ": forest runs past the last pair");
    for (int j = 1; j <= k; ++j) {
      const kdh_pair& m = pairs[i + j];
      if (m.moving != pairs[i].moving || m.ct != pairs[i].ct || m.forest != 0)
        return kdh_fail(ctx, KDH_ERR_INVALID,
                        "pair " + std::to_string(i + j) + ": forest entry must share the moving tree and ct and "
                        "lead no forest");
    }
    for (int j = 0; j <= k; ++j) leader[i + j] = i;
    i += k + 1;
  }
  for (int i = 0; i < n_pairs; ++i) {
    const kdh_pair& lp = pairs[leader[i]];  // pose, velocities, params, prior
    const kdh_pair& pr = pairs[i];           // fixed tree and pose
    PairState& ps = hs[i];
    rc = kdh_fill_pair(ctx, ps, pr.fixed, lp.moving, lp.ct, lp.params, stats_mode, mode);
    if (rc) return kdh_fail(ctx, rc, "pair " + std::to_string(i) + ": " + ctx->err);
    memcpy(ps.X, lp.X0, sizeof(ps.X));
    memcpy(ps.Xf, pr.Xf, sizeof(ps.Xf));
    for (int k = 0; k < 6; ++k) ps.v[k] = lp.ct ? lp.v0[k] : 0.0;
    kdh_setup_prior(ps, leader[i] == i ? lp.prior : nullptr);
    kdh_cache_from_state(ps);
    if (vf)
      for (int k = 0; k < 6; ++k) ps.cache.vf[k] = (float)vf[6 * i + k];
    if (leader[i] != i) {
      ps.group_n = 0;
      PairState& L = hs[leader[i]];
      L.group_n += 1;
      L.group_tiles += ps.n_tiles;
    }
    max_iter = std::max(max_iter, ps.sp.max_iterations);
  }
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& L = hs[leader[i]];
    if (L.group_tiles == 0) {
      // no moving leaves: the first round fails the min_good gate
      PairState& ps = hs[i];
      ps.iter = 1;
      ps.status = L.sp.min_good_measurements > 0 ? 2 : 1;
    }
  }
  // The split: the first forest boundary where the tiles before it reach
  // half of all tiles (no split with one half, or when nothing would follow).
  int split = n_pairs;
  if (halves == 2) {
    long long total = 0, acc = 0;
    for (int i = 0; i < n_pairs; ++i) total += hs[i].n_tiles;
    for (int i = 0; i < n_pairs; ++i) {
      if (leader[i] == i && i > 0 && 2 * acc >= total) {
        split = i;
        break;
      }
      acc += hs[i].n_tiles;
    }
  }
  for (int k = 0; k < kGroups; ++k) {
    g[k].p_lo = k < 2 ? 0 : split;
    g[k].p_hi = k < 2 ? split : n_pairs;
  }
  for (int i = 0; i < n_pairs; ++i) {
    PairState& ps = hs[i];
    Group& G = g[ps.ct + (i >= split ? 2 : 0)];
    ps.tile_first = G.n_tiles;
    G.n_tiles += ps.n_tiles;
  }
  size_t n_tiles_all = 0;
  n_partials = 0;
  for (int k = 0; k < kGroups; ++k) {
    g[k].partial_off = n_partials;
    g[k].tile_off = n_tiles_all;
    n_partials += (size_t)g[k].n_tiles * slots(g[k].ct, g[k].ct ? dual : 0);
    n_tiles_all += (size_t)g[k].n_tiles;
  }
  tp.assign(std::max<size_t>(n_tiles_all, 1), TileDesc{});
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& ps = hs[i];
    const Group& G = g[ps.ct + (i >= split ? 2 : 0)];
    for (int t = 0; t < ps.n_tiles; ++t) tp[G.tile_off + ps.tile_first + t] = make_tile_desc(ps, i, t, ctx->tile_leaves);
  }
  return KDH_OK;
}

// batch_plan into the context's pinned states, then the tile list and the
// states' input sections uploaded into the context's buffers.  Called with
// ctx->mu held and the device set.
static int batch_setup(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, int mode, int stats_mode, Group* g,
                       std::vector<int>& leader, int& max_iter, int dual = 0, const double* vf = nullptr,
                       int halves = 1) {
  int rc = kdh_ensure_pairs(ctx, (size_t)n_pairs);
  if (rc) return rc;
  std::vector<TileDesc> tp;
  size_t n_partials = 0;
  rc = batch_plan(ctx, n_pairs, pairs, mode, stats_mode, g, leader, max_iter, dual, vf, halves, ctx->h_pairs, tp,
                  n_partials);
  if (rc) return rc;
  rc = kdh_ensure_partials(ctx, std::max<size_t>(n_partials, 1));
  if (rc) return rc;
  rc = kdh_ensure_tiles(ctx, tp.size());
  if (rc) return rc;
  rc = kdh_upload_tiles(ctx, tp);
  if (rc) return rc;
  // Only the inputs and the evolving state go up (offsetof(stats_log) of each
  // ~6 KB record): the statistics log, H, b and the dual blocks are outputs
  // the device writes before anything reads them.
  KDH_HIP_TRY(ctx, hipMemcpy2DAsync(ctx->d_pairs, sizeof(PairState), ctx->h_pairs, sizeof(PairState),
                                    kStateBytes, (size_t)n_pairs, hipMemcpyHostToDevice, ctx->stream));
  return KDH_OK;
}

// The device buffers of a batch: pair states, tile list, partial rows.
struct BatchBufs {
  PairState* pairs;
  const TileDesc* tiles;
  float* partials;
};

// Launches of the groups of half h (kinds 0 and 1) on stream st; phases as
// kdh_launch (1 tile, 2 finish, 3 both).
static int launch_half(kdh_ctx* ctx, const BatchBufs& B, const Group* g, int h, int phases, hipStream_t st) {
  for (int k = 2 * h; k < 2 * h + 2; ++k) {
    if (g[k].n_tiles == 0) continue;
    int rc = kdh_launch(ctx, B.pairs, B.tiles + g[k].tile_off, g[k].n_tiles, g[k].ct, nullptr, nullptr,
                        B.partials + g[k].partial_off, g[k].p_hi - g[k].p_lo, phases, 0, g[k].p_lo, st);
    if (rc != KDH_OK) return rc;
  }
  return KDH_OK;
}

static int launch_phase(kdh_ctx* ctx, const BatchBufs& B, const Group* g, int phases) {
  int rc = launch_half(ctx, B, g, 0, phases, ctx->stream);
  return rc != KDH_OK ? rc : launch_half(ctx, B, g, 1, phases, ctx->stream);
}

// Profiling: plain launches on ctx->stream with events around each kernel
// phase of every iteration (the tile phase repeated kProfileReps times).
static int enqueue_profiled(kdh_ctx* ctx, const BatchBufs& B, const Group* g, int iters) {
  KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  for (int it = 0; it < iters; ++it) {
    for (int r = 0; r < kdh_ctx::kProfileReps; ++r) {
      int rc = launch_phase(ctx, B, g, 1);
      if (rc != KDH_OK) return rc;
    }
    KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev[2 * it + 1], ctx->stream));
    int rc = launch_phase(ctx, B, g, 2);
    if (rc != KDH_OK) return rc;
    KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev[2 * it + 2], ctx->stream));
  }
  return KDH_OK;
}

// Every iteration (tile + finish launch per non-empty group), enqueued on
// ctx->stream while it is being captured.  The pair states live in device
// memory, so the same launches replay each Gauss-Newton step, and pairs that
// have finished skip their tiles.  With two streams each half of the pairs
// is its own chain of launches; the second half starts after the first
// half's first tile launch, so from then on one half's finish stage runs
// beside the other half's tile launch.
static int enqueue_iterations(kdh_ctx* ctx, const BatchBufs& B, const Group* g, int iters, bool two) {
  if (!two) {
    for (int it = 0; it < iters; ++it) {
      int rc = launch_phase(ctx, B, g, 3);
      if (rc != KDH_OK) return rc;
    }
    return KDH_OK;
  }
  KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
  KDH_HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
  int rc = launch_half(ctx, B, g, 0, 1, ctx->stream);
  if (rc != KDH_OK) return rc;
  KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev_stag, ctx->stream));
  rc = launch_half(ctx, B, g, 0, 2, ctx->stream);
  if (rc != KDH_OK) return rc;
  KDH_HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev_stag, 0));
  for (int it = 1; it < iters && rc == KDH_OK; ++it) rc = launch_half(ctx, B, g, 0, 3, ctx->stream);
  for (int it = 0; it < iters && rc == KDH_OK; ++it) rc = launch_half(ctx, B, g, 1, 3, ctx->stream2);
  if (rc != KDH_OK) return rc;
  KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev_join, ctx->stream2));
  KDH_HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  return KDH_OK;
}

// The second stream and its fork / stagger / join events, made on first use.
static int ensure_stream2(kdh_ctx* ctx) {
  if (ctx->stream2) return KDH_OK;
  KDH_HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
  KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_stag, hipEventDisableTiming));
  KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
  return KDH_OK;
}

// Ends the capture on ctx->stream begun by the caller and instantiates it;
// on an earlier failure (rc) the partial graph is dropped.
static int end_capture(kdh_ctx* ctx, int rc, hipGraphExec_t* exec) {
  hipGraph_t graph = nullptr;
  hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
  if (rc != KDH_OK) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (ce != hipSuccess) return kdh_hip_fail(ctx, ce, "hipStreamEndCapture");
  hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ie != hipSuccess) return kdh_hip_fail(ctx, ie, "hipGraphInstantiate");
  return KDH_OK;
}

int kdh_batch_align(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, kdh_pair_result* results) {
  if (!ctx) return KDH_ERR_INVALID;
  if (n_pairs < 0 || (n_pairs > 0 && (!pairs || !results))) return kdh_fail(ctx, KDH_ERR_INVALID, "bad arguments");
  if (n_pairs == 0) return KDH_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  Group g[kGroups];
  std::vector<int> leader;
  int max_iter = 0;
  // profiling times each phase on one stream
  const int halves = ctx->profile ? 1 : ctx->batch_streams;
  int rc = batch_setup(ctx, n_pairs, pairs, 0, 0, g, leader, max_iter, 0, nullptr, halves);
  if (rc) return rc;
  const bool two = halves == 2 && (g[2].n_tiles + g[3].n_tiles) > 0 && (g[0].n_tiles + g[1].n_tiles) > 0;
  if (two) {
    rc = ensure_stream2(ctx);
    if (rc) return rc;
  }
  const int iters = std::min(max_iter, kMaxIter);
  const BatchBufs B{ctx->d_pairs, ctx->d_tiles, ctx->d_partials};
  hipError_t le = hipSuccess;
  if (ctx->profile) {
    rc = enqueue_profiled(ctx, B, g, iters);
    if (rc) return rc;
  } else {
    // all iterations in one graph, launched once
    const long long key[16] = {n_pairs, g[0].n_tiles, g[1].n_tiles, iters, ctx->lds_nodes, ctx->grid_override,
                               ctx->xcd_remap, (long long)(intptr_t)ctx->d_pairs,
                               (long long)(intptr_t)ctx->d_partials, (long long)(intptr_t)ctx->d_tiles, 2,
                               ctx->leaves_per_lane | (ctx->tile_leaves << 8), two ? 2 : 1, g[2].n_tiles,
                               g[3].n_tiles,
                               (long long)(intptr_t)ctx->stream2};
    hipGraphExec_t exec = kdh_graph_lookup(ctx->graph_batch, key);
    if (!exec) {
      KDH_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      rc = end_capture(ctx, enqueue_iterations(ctx, B, g, iters, two), &exec);
      if (rc) return rc;
      kdh_graph_store(ctx->graph_batch, key, exec);
    }
    le = hipEventRecord(ctx->ev[0], ctx->stream);
    if (le == hipSuccess) le = hipGraphLaunch(exec, ctx->stream);
    if (le == hipSuccess) le = hipEventRecord(ctx->ev[1], ctx->stream);
    if (le == hipSuccess) le = hipStreamSynchronize(ctx->stream);
  }
  if (le == hipSuccess) {
    rc = kdh_pull_results(ctx, n_pairs);
    if (rc) return rc;
    le = hipStreamSynchronize(ctx->stream);
  }
  if (le != hipSuccess) return kdh_hip_fail(ctx, le, "batch replay");
  if (ctx->profile) {
    kdh_collect_times(ctx, iters, true);
  } else {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->last_kernel_us = iters ? 1000.0 * ms / iters : 0.0;
    ctx->last_tile_us = ctx->last_finish_us = 0.0;
    ctx->last_launches = iters;
  }
  // Per-entry stats of forests: each entry's share of the last round, summed
  // on the host from the partial rows of that round (rows of finished pairs
  // are not rewritten).
  std::vector<kdh_stats> entry(n_pairs);
  std::vector<float> rows;
  // A forest without moving leaves never ran a round on the device: its one
  // failed round has all-zero statistics (the state's log entry was not
  // written and holds whatever an earlier batch left there).
  auto last_of = [](const PairState& L, kdh_stats& st) {
    memset(&st, 0, sizeof(kdh_stats));
    if (L.iter > 0 && L.group_tiles > 0) memcpy(&st, &L.stats_log[L.iter - 1], sizeof(kdh_stats));
  };
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& L = ctx->h_pairs[leader[i]];
    last_of(L, entry[i]);
    entry[i].pert_pose_norm = entry[i].pert_vel_norm = 0.0;
  }
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& L = ctx->h_pairs[i];
    if (leader[i] != i || L.group_n < 2 || L.iter == 0 || L.group_tiles == 0) continue;
    const int N = slots(L.ct), CS = L.ct ? 90 : 27;
    rows.resize((size_t)L.group_tiles * N);
    const float* src = ctx->d_partials + g[L.ct + (i >= g[2].p_lo ? 2 : 0)].partial_off + (size_t)L.tile_first * N;
    KDH_HIP_TRY(ctx, hipMemcpy(rows.data(), src, rows.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int e = 0; e < L.group_n; ++e) {
      const PairState& ps = ctx->h_pairs[i + e];
      double acc[7] = {0, 0, 0, 0, 0, 0, 0};
      for (int t = 0; t < ps.n_tiles; ++t) {
        const float* row = rows.data() + (size_t)(ps.tile_first - L.tile_first + t) * N;
        for (int k = 0; k < 7; ++k) acc[k] += (double)row[CS + k];
      }
      kdh_stats& st = entry[i + e];
      st.chi2 = acc[0];
      st.chi2_kernelized = acc[1];
      st.num_inliers = (int)acc[2 + kInlier];
      st.num_outliers = (int)acc[2 + kOutlier];
      st.num_bad = (int)acc[2 + kBad];
      st.num_err_normal = (int)acc[2 + kErrNormal];
      st.num_err_fn = (int)acc[2 + kErrFn];
    }
  }
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& ps = ctx->h_pairs[leader[i]];
    kdh_pair_result& r = results[i];
    memcpy(r.X, ps.X, sizeof(r.X));
    memcpy(r.v, ps.v, sizeof(r.v));
    r.iterations = ps.iter;
    last_of(ps, r.last);
    r.entry = entry[i];
    if (ps.group_n < 2) r.entry = r.last;
    r.result = kdh_result_code(ps, pairs[leader[i]].params.max_iterations);
  }
  return KDH_OK;
}

int kdh_batch_build(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, int stats_mode, double* H, double* b,
                    kdh_stats* stats) {
  if (!ctx) return KDH_ERR_INVALID;
  if (n_pairs < 0 || (n_pairs > 0 && (!pairs || !H || !b || !stats)))
    return kdh_fail(ctx, KDH_ERR_INVALID, "bad arguments");
  if (n_pairs == 0) return KDH_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  Group g[kGroups];
  std::vector<int> leader;
  int max_iter = 0;
  int rc = batch_setup(ctx, n_pairs, pairs, 1, stats_mode, g, leader, max_iter);
  if (rc) return rc;
  for (int k = 0; k < 2; ++k) {
    if (g[k].n_tiles == 0) continue;
    rc = kdh_launch(ctx, ctx->d_pairs, ctx->d_tiles + g[k].tile_off, g[k].n_tiles, g[k].ct, nullptr, nullptr,
                    ctx->d_partials + g[k].partial_off, n_pairs, 3);
    if (rc) return rc;
  }
  rc = kdh_pull_forms(ctx, n_pairs, 0);
  if (rc) return rc;
  KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& L = ctx->h_pairs[leader[i]];
    const int D = L.ct ? 12 : 6;
    double* Hi = H + (size_t)i * 144;
    double* bi = b + (size_t)i * 12;
    memset(Hi, 0, sizeof(double) * 144);
    memset(bi, 0, sizeof(double) * 12);
    memset(&stats[i], 0, sizeof(kdh_stats));
    if (L.group_tiles == 0) continue;  // no moving leaves: an empty form
    memcpy(Hi, L.H, sizeof(double) * D * D);
    memcpy(bi, L.b, sizeof(double) * D);
    memcpy(&stats[i], &L.stats_log[0], sizeof(kdh_stats));
  }
  return KDH_OK;
}

int kdh_batch_build_dual(kdh_ctx* ctx, int n_factors, const kdh_dual_pair* factors, int stats_mode,
                         kdh_dual_form* out) {
  if (!ctx) return KDH_ERR_INVALID;
  if (n_factors < 0 || (n_factors > 0 && (!factors || !out))) return kdh_fail(ctx, KDH_ERR_INVALID, "bad arguments");
  if (n_factors == 0) return KDH_OK;
  // The dual form is a CT build: each factor becomes a CT pair with its
  // moving state as (X0, v0) and the fixed pose as Xf; the fixed velocities
  // enter the cache beside it.
  std::vector<kdh_pair> pairs(n_factors);
  std::vector<double> vf((size_t)n_factors * 6);
  for (int i = 0; i < n_factors; ++i) {
    const kdh_dual_pair& f = factors[i];
    kdh_pair& p = pairs[i];
    memset(&p, 0, sizeof(p));
    p.fixed = f.fixed;
    p.moving = f.moving;
    p.ct = 1;
    p.forest = 0;
    memcpy(p.X0, f.X_moving, sizeof(p.X0));
    memcpy(p.v0, f.v_moving, sizeof(p.v0));
    memcpy(p.Xf, f.X_fixed, sizeof(p.Xf));
    p.params = f.params;
    p.prior = nullptr;
    memcpy(&vf[(size_t)6 * i], f.v_fixed, sizeof(double) * 6);
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  Group g[kGroups];
  std::vector<int> leader;
  int max_iter = 0;
  int rc = batch_setup(ctx, n_factors, pairs.data(), 1, stats_mode, g, leader, max_iter, 1, vf.data());
  if (rc) return rc;
  if (g[1].n_tiles > 0) {
    rc = kdh_launch(ctx, ctx->d_pairs, ctx->d_tiles + g[1].tile_off, g[1].n_tiles, 1, nullptr, nullptr,
                    ctx->d_partials + g[1].partial_off, n_factors, 3, 1);
    if (rc) return rc;
  }
  rc = kdh_pull_forms(ctx, n_factors, 1);
  if (rc) return rc;
  KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n_factors; ++i) {
    const PairState& ps = ctx->h_pairs[i];
    kdh_dual_form& o = out[i];
    memset(&o, 0, sizeof(o));
    if (ps.group_tiles == 0) continue;  // no moving leaves: an empty form
    memcpy(o.H, ps.H, sizeof(o.H));
    memcpy(o.b, ps.b, sizeof(o.b));
    memcpy(o.H_dual_vel_B, ps.Hd, sizeof(o.H_dual_vel_B));
    memcpy(o.b_dual_vel_B, ps.Hd + 36, sizeof(o.b_dual_vel_B));
    memcpy(o.H_dual_cross_pose_vel_B, ps.Hd + 42, sizeof(o.H_dual_cross_pose_vel_B));
    memcpy(o.H_dual_cross_vel_AB, ps.Hd + 78, sizeof(o.H_dual_cross_vel_AB));
    memcpy(&o.stats, &ps.stats_log[0], sizeof(kdh_stats));
  }
  return KDH_OK;
}

// ---- prepared batches (kdh_batch_prepare / kdh_batch_run) ------------------
//
// The batch of kdh_batch_align validated, planned and uploaded once, with one
// graph that holds the whole run: the state initialisation from the caller's
// initial poses (read by the device from pinned host memory), every
// iteration, and the results written back into pinned host memory.  A run is
// then one graph launch and one synchronisation.

namespace {

// Per pair: its forest leader, its state at the start of a run (iter,
// status: a forest without moving leaves starts finished, as batch_plan sets
// it), and the first float of its group's partial rows.
struct BatchMeta {
  int leader;
  int init_iter;
  int init_status;
  int max_iterations;  // the leader's params.max_iterations (result code)
  long long partial_base;
};

// What a run returns per pair, written by the device into pinned memory.
struct BatchOut {
  double X[12];
  double v[6];
  DevStats last;
  DevStats entry;
  int iter;
  int status;
  int pad[2];
};
static_assert(sizeof(BatchOut) % 16 == 0, "BatchOut rows stay 16-byte aligned");
constexpr int kBatchIn = 18;  // doubles per pair of the run inputs: X0 (3x4), v0

}  // namespace

struct kdh_batch {
  kdh_ctx* ctx = nullptr;
  int n = 0;
  int iters = 0;
  int tile_leaves = 0;  // the context's tile size the batch was planned with
  Group g[kGroups];
  bool two = false;
  PairState* d_pairs = nullptr;
  float* d_partials = nullptr;
  TileDesc* d_tiles = nullptr;
  BatchMeta* d_meta = nullptr;
  std::vector<BatchMeta> meta;
  // Two run slots, so that one run can be enqueued while the previous one
  // executes (kdh_batch_submit): pinned and device inputs (kBatchIn doubles
  // per pair), pinned, mapped results (one BatchOut per pair) and a graph
  // per slot (they differ in the input and result addresses).  A slot's
  // inputs go up on the batch's copy stream, so the copy overlaps the run
  // before it; the run waits for it through the slot's event.
  double* d_in = nullptr;
  double* h_in = nullptr;
  hipStream_t copy = nullptr;
  hipEvent_t in_ready[2] = {};
  BatchOut* h_out = nullptr;
  BatchOut* h_out_dev = nullptr;
  hipGraphExec_t exec[2] = {};
  long long key[2][8] = {};  // the launch configuration each graph was built with
  hipEvent_t ev[2][2] = {};  // per slot: before and after its run
  int in_slot = 0;           // the slot holding the latest inputs
  int first = 0, inflight = 0;  // submitted runs not yet collected: count, and the oldest one's slot
};

// Workgroup of 64 pairs: their leaders' inputs staged through LDS (16-byte
// loads, consecutive lanes on consecutive bytes), then lane l sets pair
// 64 b + l: its pose and velocities from its leader's inputs, its counters
// reset, the FP32 cache of that state (cache_from_state, the expressions of
// the host set-up).  Everything else in the state is input that a run does
// not change.  The lane's own meta record and the state fields the cache
// needs are loaded together with the inputs, so the common case is one
// memory round trip.  Pairs and their leaders may sit in different
// workgroups, so a lane reads its leader's row directly when it lies outside
// the workgroup's range.
constexpr int kInitPairs = 64;
__global__ __launch_bounds__(kInitPairs) void k_batch_init(PairState* __restrict__ d,
                                                           const BatchMeta* __restrict__ meta,
                                                           const double* __restrict__ in, int n) {
  __shared__ __align__(16) double2 s_in[kInitPairs * kBatchIn / 2];
  const int base = blockIdx.x * kInitPairs;
  const int rows = min(kInitPairs, n - base);
  const int i = base + threadIdx.x;
  const int ic = min(i, n - 1);  // lanes past the end load pair n - 1 and store nothing
  const double2* src = reinterpret_cast<const double2*>(in + (size_t)kBatchIn * base);
  // every load in flight before the first LDS store: the inputs, then this
  // lane's meta record and the state fields the cache needs
  constexpr int kPer = kBatchIn / 2;  // 16-byte words per lane (kInitPairs rows of kBatchIn doubles)
  const int last = rows * kPer - 1;  // rows >= 1: the grid covers n pairs exactly
  double2 w[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) w[j] = src[min((int)threadIdx.x + j * kInitPairs, last)];
  const BatchMeta m = meta[ic];
  PairState& ps = d[ic];
  const int ct = ps.ct;
  double Xf[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) Xf[k] = ps.Xf[k];
  TermCache C = ps.cache;  // vf (dual form) is kept
  const V3f mroot = ps.moving_root_mean;
  const int recenter = ps.sp.recenter;
  // the meta record arrives with the inputs: the compiler would otherwise sink
  // its load below the barrier, where its value is first used
  __asm__ volatile("" ::"v"(m.leader));
  // Unconditional stores (the buffer holds kInitPairs rows; a row past `rows`
  // receives a clamped copy that no lane reads): a store under a condition
  // lets the compiler sink its load into the branch, one round trip each.
#pragma unroll
  for (int j = 0; j < kPer; ++j) s_in[threadIdx.x + j * kInitPairs] = w[j];
  __syncthreads();
  if (i >= n) return;
  const int r = m.leader - base;
  double x[kBatchIn];
  if (r >= 0 && r < rows) {
    const double* sx = reinterpret_cast<const double*>(s_in) + kBatchIn * r;
#pragma unroll
    for (int k = 0; k < kBatchIn; ++k) x[k] = sx[k];
  } else {
    const double* gx = in + (size_t)kBatchIn * m.leader;
#pragma unroll
    for (int k = 0; k < kBatchIn; ++k) x[k] = gx[k];
  }
  if (!ct) {
#pragma unroll
    for (int k = 12; k < kBatchIn; ++k) x[k] = 0.0;
  }
  double c64[3];
  cache_from_pose(C, c64, x, Xf, x + 12, mroot, recenter);
#pragma unroll
  for (int k = 0; k < 12; ++k) ps.X[k] = x[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) ps.v[k] = x[12 + k];
  ps.cache = C;
  for (int k = 0; k < 3; ++k) ps.c64[k] = c64[k];
  ps.iter = m.init_iter;
  ps.status = m.init_status;
  ps.prev_score = 0.0f;
}

// One workgroup per pair: the pair's result record (its leader's pose,
// velocities, iteration count and status, the last round's statistics, the
// entry's own share for a forest entry) staged in LDS, then written to pinned
// host memory in 16-byte stores.  The common pair is its own leader and ran
// every one of the batch's `iters` rounds, so its leader's fields and the
// statistics of round iters - 1 are loaded together with its meta record (one
// memory round trip); a forest entry, or a pair that stopped earlier, loads
// what differs after that.
__global__ __launch_bounds__(64) void k_batch_collect(const PairState* __restrict__ d,
                                                      const BatchMeta* __restrict__ meta,
                                                      const float* __restrict__ partials, BatchOut* out, int n,
                                                      int iters) {
  __shared__ __align__(16) BatchOut s;
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n) return;
  const BatchMeta m = meta[i];
  // the pair as its own leader after all iters rounds: every load in flight at once
  const int guess = min(max(iters, 1), kMaxIter) - 1;
  const PairState* L = d + i;
  double xv = t < 12 ? L->X[t] : (t < 18 ? L->v[t - 12] : 0.0);
  int it = L->iter;
  int status = L->status;
  int group_tiles = L->group_tiles;
  int group_n = L->group_n;
  DevStats last = L->stats_log[guess];
  if (m.leader != i) {  // a forest entry: its leader's fields
    L = d + m.leader;
    xv = t < 12 ? L->X[t] : (t < 18 ? L->v[t - 12] : 0.0);
    it = L->iter;
    status = L->status;
    group_tiles = L->group_tiles;
    group_n = L->group_n;
    last = L->stats_log[guess];
  }
  if (it - 1 != guess && it > 0) last = L->stats_log[it - 1];  // stopped before the last round
  if (t < 12) s.X[t] = xv;
  else if (t < 18) s.v[t - 12] = xv;
  if (t == 0) {
    s.iter = it;
    s.status = status;
    s.pad[0] = s.pad[1] = 0;
    if (!(it > 0 && group_tiles > 0)) last = DevStats{};  // no round ran on the device
    s.last = last;
    DevStats e = last;
    if (group_n >= 2) {
      e.pert_pose_norm = e.pert_vel_norm = 0.0;
      if (it > 0 && group_tiles > 0) {
        const PairState& ps = d[i];
        const int N = L->ct ? kCtSlots : kIcpSlots, CS = L->ct ? 90 : 27;
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int tt = 0; tt < ps.n_tiles; ++tt) {
          const float* row = partials + m.partial_base + (size_t)(ps.tile_first + tt) * N;
          for (int k = 0; k < 7; ++k) acc[k] += (double)row[CS + k];
        }
        e.chi2 = acc[0];
        e.chi2_kernelized = acc[1];
        e.num_inliers = (int)acc[2 + kInlier];
        e.num_outliers = (int)acc[2 + kOutlier];
        e.num_bad = (int)acc[2 + kBad];
        e.num_err_normal = (int)acc[2 + kErrNormal];
        e.num_err_fn = (int)acc[2 + kErrFn];
      }
    }
    s.entry = e;
  }
  __syncthreads();
  constexpr int kWords = (int)(sizeof(BatchOut) / 16);
  static_assert(kWords <= 64, "one 16-byte store per lane");
  // the host reads the records after synchronising with the stream; the
  // kernel's completion makes the pinned-memory writes visible to it
  if (t < kWords) reinterpret_cast<float4*>(out + i)[t] = reinterpret_cast<const float4*>(&s)[t];
}

static void batch_free(kdh_batch* b) {
  for (int k = 0; k < 2; ++k) {
    if (b->exec[k]) hipGraphExecDestroy(b->exec[k]);
    for (int j = 0; j < 2; ++j)
      if (b->ev[k][j]) hipEventDestroy(b->ev[k][j]);
    if (b->in_ready[k]) hipEventDestroy(b->in_ready[k]);
  }
  if (b->copy) hipStreamDestroy(b->copy);
  if (b->d_pairs) hipFree(b->d_pairs);
  if (b->d_partials) hipFree(b->d_partials);
  if (b->d_tiles) hipFree(b->d_tiles);
  if (b->d_meta) hipFree(b->d_meta);
  if (b->d_in) hipFree(b->d_in);
  if (b->h_in) hipHostFree(b->h_in);
  if (b->h_out) hipHostFree(b->h_out);
  delete b;
}

int kdh_batch_prepare(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, kdh_batch** out) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!out || n_pairs <= 0 || !pairs) return kdh_fail(ctx, KDH_ERR_INVALID, "bad arguments");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  kdh_batch* b = new kdh_batch();
  b->ctx = ctx;
  b->n = n_pairs;
  b->tile_leaves = ctx->tile_leaves;
  std::vector<PairState> hs((size_t)n_pairs);
  std::vector<int> leader;
  std::vector<TileDesc> tp;
  size_t n_partials = 0;
  int max_iter = 0;
  int rc = batch_plan(ctx, n_pairs, pairs, 0, 0, b->g, leader, max_iter, 0, nullptr, ctx->batch_streams, hs.data(),
                      tp, n_partials);
  if (rc) {
    delete b;
    return rc;
  }
  b->iters = std::min(max_iter, kMaxIter);
  b->two = ctx->batch_streams == 2 && (b->g[2].n_tiles + b->g[3].n_tiles) > 0 &&
           (b->g[0].n_tiles + b->g[1].n_tiles) > 0;
  b->meta.resize(n_pairs);
  for (int i = 0; i < n_pairs; ++i) {
    BatchMeta& m = b->meta[i];
    const PairState& ps = hs[i];
    m.leader = leader[i];
    m.init_iter = ps.iter;
    m.init_status = ps.status;
    m.max_iterations = pairs[leader[i]].params.max_iterations;
    m.partial_base = (long long)b->g[ps.ct + (i >= b->g[2].p_lo ? 2 : 0)].partial_off;  // p_lo of half 1: the split
  }
  hipError_t e = hipMalloc((void**)&b->d_pairs, sizeof(PairState) * (size_t)n_pairs);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_partials, sizeof(float) * std::max<size_t>(n_partials, 1));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_tiles, sizeof(TileDesc) * tp.size());
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_meta, sizeof(BatchMeta) * (size_t)n_pairs);
  // two slots each; coherent: the device writes the results in place
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&b->h_in, 2 * sizeof(double) * kBatchIn * (size_t)n_pairs, hipHostMallocCoherent);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_in, 2 * sizeof(double) * kBatchIn * (size_t)n_pairs);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&b->h_out, 2 * sizeof(BatchOut) * (size_t)n_pairs, hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&b->h_out_dev, b->h_out, 0);
  for (int k = 0; k < 2; ++k) {
    for (int j = 0; j < 2; ++j)
      if (e == hipSuccess) e = hipEventCreate(&b->ev[k][j]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->in_ready[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->copy, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_pairs, hs.data(), sizeof(PairState) * (size_t)n_pairs, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_tiles, tp.data(), sizeof(TileDesc) * tp.size(), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_meta, b->meta.data(), sizeof(BatchMeta) * (size_t)n_pairs, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    batch_free(b);
    return kdh_hip_fail(ctx, e, "kdh_batch_prepare");
  }
  for (int i = 0; i < n_pairs; ++i) {
    memcpy(b->h_in + (size_t)kBatchIn * i, pairs[i].X0, sizeof(double) * 12);
    memcpy(b->h_in + (size_t)kBatchIn * i + 12, pairs[i].v0, sizeof(double) * 6);
  }
  b->in_slot = 0;
  *out = b;
  return KDH_OK;
}

// Enqueues one run from slot `slot` on the context's stream: the latest
// inputs, with X0 / v0 applied, into the slot's pinned inputs, up to the
// device, then the slot's graph (captured on first use and again after a
// launch configuration change) between the slot's two events.  Called with
// ctx->mu held and the device set.
static int batch_enqueue(kdh_batch* b, int slot, const double* X0, const double* v0, bool overlap) {
  kdh_ctx* ctx = b->ctx;
  const int n = b->n;
  const size_t in_doubles = (size_t)kBatchIn * n;
  double* in = b->h_in + in_doubles * slot;
  if (slot != b->in_slot) memcpy(in, b->h_in + in_doubles * b->in_slot, sizeof(double) * in_doubles);
  b->in_slot = slot;
  if (X0)
    for (int i = 0; i < n; ++i) memcpy(in + (size_t)kBatchIn * i, X0 + (size_t)12 * i, sizeof(double) * 12);
  if (v0)
    for (int i = 0; i < n; ++i) memcpy(in + (size_t)kBatchIn * i + 12, v0 + (size_t)6 * i, sizeof(double) * 6);
  // the inputs go up by DMA ahead of the graph (a kernel reading them from
  // pinned memory across PCIe took longer than the copy); with `overlap` on
  // the copy stream, so that the copy runs beside the run in flight before
  // this one (with nothing in flight the wait between streams only adds
  // latency, and the copy goes on the run's stream)
  double* d_in = b->d_in + in_doubles * slot;
  if (overlap) {
    KDH_HIP_TRY(ctx, hipMemcpyAsync(d_in, in, sizeof(double) * in_doubles, hipMemcpyHostToDevice, b->copy));
    KDH_HIP_TRY(ctx, hipEventRecord(b->in_ready[slot], b->copy));
    KDH_HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, b->in_ready[slot], 0));
  } else {
    KDH_HIP_TRY(ctx, hipMemcpyAsync(d_in, in, sizeof(double) * in_doubles, hipMemcpyHostToDevice, ctx->stream));
  }
  BatchOut* out_dev = b->h_out_dev + (size_t)n * slot;
  const BatchBufs B{b->d_pairs, b->d_tiles, b->d_partials};
  const dim3 init_grid((n + kInitPairs - 1) / kInitPairs);
  const bool two = b->two && ctx->batch_streams == 2;
  const long long key[8] = {ctx->lds_nodes, ctx->grid_override, ctx->xcd_remap, ctx->leaves_per_lane,
                            two ? 2 : 1, (long long)(intptr_t)ctx->stream, ctx->tile_leaves, 0};
  hipGraphExec_t& exec = b->exec[slot];
  if (exec && memcmp(key, b->key[slot], sizeof(key)) != 0) {
    hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  if (!exec) {
    int rc = KDH_OK;
    if (two) {
      rc = ensure_stream2(ctx);
      if (rc) return rc;
    }
    KDH_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_batch_init, init_grid, dim3(kInitPairs), 0, ctx->stream, b->d_pairs, b->d_meta, d_in, n);
    rc = hipGetLastError() == hipSuccess ? KDH_OK : kdh_fail(ctx, KDH_ERR_HIP, "k_batch_init launch");
    if (rc == KDH_OK) rc = enqueue_iterations(ctx, B, b->g, b->iters, two);
    if (rc == KDH_OK) {
      hipLaunchKernelGGL(k_batch_collect, dim3(n), dim3(64), 0, ctx->stream, b->d_pairs, b->d_meta, b->d_partials,
                         out_dev, n, b->iters);
      if (hipGetLastError() != hipSuccess) rc = kdh_fail(ctx, KDH_ERR_HIP, "k_batch_collect launch");
    }
    rc = end_capture(ctx, rc, &exec);
    if (rc) return rc;
    memcpy(b->key[slot], key, sizeof(key));
  }
  hipError_t le = hipEventRecord(b->ev[slot][0], ctx->stream);
  if (le == hipSuccess) le = hipGraphLaunch(exec, ctx->stream);
  if (le == hipSuccess) le = hipEventRecord(b->ev[slot][1], ctx->stream);
  if (le != hipSuccess) return kdh_hip_fail(ctx, le, "kdh_batch_run");
  return KDH_OK;
}

// The results of the run of slot `slot` (complete) into `results`.
static void copy_results(const kdh_batch* b, int slot, kdh_pair_result* results) {
  static_assert(sizeof(kdh_stats) == sizeof(DevStats), "kdh_stats mirrors DevStats");
  const BatchOut* out = b->h_out + (size_t)b->n * slot;
  for (int i = 0; i < b->n; ++i) {
    const BatchOut& o = out[i];
    kdh_pair_result& r = results[i];
    memcpy(r.X, o.X, sizeof(r.X));
    memcpy(r.v, o.v, sizeof(r.v));
    memcpy(&r.last, &o.last, sizeof(r.last));
    memcpy(&r.entry, &o.entry, sizeof(r.entry));
    r.iterations = o.iter;
    r.result = kdh_result_code_of(o.status, o.iter, b->meta[i].max_iterations);
  }
}

// Waits for the run of slot `slot`, records its device time per iteration
// in the context, and copies its results out.  Called with ctx->mu held and
// the device set.
static int batch_results(kdh_batch* b, int slot, kdh_pair_result* results) {
  kdh_ctx* ctx = b->ctx;
  hipError_t le = hipEventSynchronize(b->ev[slot][1]);
  if (le != hipSuccess) return kdh_hip_fail(ctx, le, "kdh_batch_run");
  float ms = 0.0f;
  hipEventElapsedTime(&ms, b->ev[slot][0], b->ev[slot][1]);
  ctx->last_kernel_us = b->iters ? 1000.0 * ms / b->iters : 0.0;
  ctx->last_tile_us = ctx->last_finish_us = 0.0;
  ctx->last_launches = b->iters;
  copy_results(b, slot, results);
  return KDH_OK;
}

// Checks shared by the run entry points.  Called with ctx->mu held.
static int batch_check(kdh_batch* b, const char* who) {
  kdh_ctx* ctx = b->ctx;
  if (ctx->tile_leaves != b->tile_leaves)
    return kdh_fail(ctx, KDH_ERR_INVALID,
                    std::string(who) + ": the context's tile size changed since kdh_batch_prepare; prepare the batch again");
  return KDH_OK;
}

int kdh_batch_run(kdh_batch* b, const double* X0, const double* v0, kdh_pair_result* results) {
  if (!b) return KDH_ERR_INVALID;
  kdh_ctx* ctx = b->ctx;
  if (!results) return kdh_fail(ctx, KDH_ERR_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = batch_check(b, "kdh_batch_run");
  if (rc) return rc;
  if (b->inflight)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_batch_run: collect the submitted runs first (kdh_batch_collect)");
  if (!ctx->profile) {
    const int slot = b->first;
    rc = batch_enqueue(b, slot, X0, v0, false);
    return rc ? rc : batch_results(b, slot, results);
  }
  // profiling: plain launches with events around each kernel phase, on one
  // stream, from slot 0
  const int n = b->n;
  const size_t in_doubles = (size_t)kBatchIn * n;
  double* in = b->h_in;
  if (b->in_slot != 0) memcpy(in, b->h_in + in_doubles * b->in_slot, sizeof(double) * in_doubles);
  b->in_slot = 0;
  if (X0)
    for (int i = 0; i < n; ++i) memcpy(in + (size_t)kBatchIn * i, X0 + (size_t)12 * i, sizeof(double) * 12);
  if (v0)
    for (int i = 0; i < n; ++i) memcpy(in + (size_t)kBatchIn * i + 12, v0 + (size_t)6 * i, sizeof(double) * 6);
  KDH_HIP_TRY(ctx, hipMemcpyAsync(b->d_in, in, sizeof(double) * in_doubles, hipMemcpyHostToDevice, ctx->stream));
  const BatchBufs B{b->d_pairs, b->d_tiles, b->d_partials};
  hipLaunchKernelGGL(k_batch_init, dim3((n + kInitPairs - 1) / kInitPairs), dim3(kInitPairs), 0, ctx->stream,
                     b->d_pairs, b->d_meta, b->d_in, n);
  KDH_HIP_TRY(ctx, hipGetLastError());
  rc = enqueue_profiled(ctx, B, b->g, b->iters);
  if (rc) return rc;
  hipLaunchKernelGGL(k_batch_collect, dim3(n), dim3(64), 0, ctx->stream, b->d_pairs, b->d_meta, b->d_partials,
                     b->h_out_dev, n, b->iters);
  KDH_HIP_TRY(ctx, hipGetLastError());
  KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  kdh_collect_times(ctx, b->iters, true);
  copy_results(b, 0, results);
  return KDH_OK;
}

int kdh_batch_submit(kdh_batch* b, const double* X0, const double* v0) {
  if (!b) return KDH_ERR_INVALID;
  kdh_ctx* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = batch_check(b, "kdh_batch_submit");
  if (rc) return rc;
  if (ctx->profile) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_batch_submit: not in profiling mode (kdh_batch_run)");
  if (b->inflight == 2)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_batch_submit: two runs in flight; collect one first (kdh_batch_collect)");
  const int slot = (b->first + b->inflight) & 1;
  rc = batch_enqueue(b, slot, X0, v0, b->inflight > 0);
  if (rc) return rc;
  b->inflight += 1;
  return KDH_OK;
}

int kdh_batch_collect(kdh_batch* b, kdh_pair_result* results) {
  if (!b) return KDH_ERR_INVALID;
  kdh_ctx* ctx = b->ctx;
  if (!results) return kdh_fail(ctx, KDH_ERR_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (b->inflight == 0) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_batch_collect: no run submitted");
  const int slot = b->first;
  b->first ^= 1;
  b->inflight -= 1;
  return batch_results(b, slot, results);
}

int kdh_batch_release(kdh_batch* b) {
  if (!b) return KDH_ERR_INVALID;
  kdh_ctx* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  batch_free(b);
  return KDH_OK;
}
